"""The benchmark workloads at their stated sizes against the oracle, through the C-ABI.

  1D Sod, 1000 cells, exact solver, moving mesh, t_end 0.2: the whole run (8176 steps) within 1e-9,
    bins and step count exact.
  2D Sedov, 256^2 periodic, HLLC: per-face records and one step within 1e-12 at steps 0 and 7 and
    after a multi-bin step; the whole run to t_end 0.05 (4590 steps) within 1e-9; conservation.
  3D Kelvin-Helmholtz, 128^3 (2.1M cells, 16.7M faces), AIRS: per-face records and one step within
    1e-12 after the initial step, conservation over 12 device steps.
  Ensemble of 512 1D/2D Riemann problems (6.3M cells) in one FP64 batch: every 32nd member against the
    oracle at its own step count after 300 launches, within 1e-9, bins exact.
  Served stream: 2D Sedov (t_end 0.05) and isentropic-vortex jobs with sides log-uniform in [32, 512],
    admitted by HBM budget: each result equal alone and in the mix, and L1(rho) < 1e-5 of the FP32-flux
    result against the FP64 oracle.
"""
import sys

import numpy as np
import pytest

from hydro_rs_amd import ensemble, fvs, ics, serve
from tests.helpers import compare_states, face_flux_rel, oracle_sim, synced

pytestmark = pytest.mark.gpu

TOL_STEP = 1e-12
TOL_RUN = 1e-9


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _device():
    if fvs.device_count() < 1:
        pytest.skip("no HIP device")


def _full_run(prob, steps=1 << 40):
    osim = oracle_sim(prob)
    n_ref = osim.advance(steps)
    dev = ics.build_device_sim(prob)
    n_dev = dev.run(steps)
    errs = compare_states(dev.get_state(), osim.get_fields(), prob.n)
    return osim, dev, n_ref, n_dev, errs


def _face_and_step(prob, k):
    """Per-face records and one fused step of the device against the oracle from the oracle's state after
    k engine steps (same mesh, same state)."""
    osim, dev, inf = synced(prob, k)
    ti_old, ti_cur, tb = inf["ti_old"], inf["ti_next"], inf["time_base"]
    osim.set_engine_info(ti_old, ti_cur, 0)
    f = osim.get_fields()
    from oracle import ora

    active = (f["timebin"] <= ora.get_max_active_bin(ti_cur)).astype(np.uint8)
    ref = osim.compute_fluxes(active)
    info = dev.timestep_info(ti_old, ti_cur, tb)
    dev.compute_fluxes(info)
    face_err = face_flux_rel(dev.get_fluxes(), ref)
    gated = int(np.count_nonzero(np.any(ref[:, :5] != 0.0, axis=1)))
    # the step from the same state (the trait-phase call above changed nothing the step reads back)
    osim2, dev2, _ = synced(prob, k)
    osim2.set_engine_info(ti_old, ti_cur, 0)
    t_ref = osim2.runner_step()
    t_dev = dev2.step(dev2.timestep_info(ti_old, ti_cur, tb))
    errs = compare_states(dev2.get_state(), osim2.get_fields(), prob.n)
    return int(active.sum()), gated, face_err, t_ref, t_dev, errs


# ----------------------------------------------------------------------------------- 1D Sod, 1000 cells
def test_c1_sod1d_1000_full_run(ora):
    prob = ics.sod_1d(n=1000, t_end=0.2)
    say(f"[C1] sod1d 1000 cells, moving mesh, exact solver, t_end 0.2")
    osim, dev, n_ref, n_dev, errs = _full_run(prob)
    say(f"  steps ref={n_ref} dev={n_dev} errs={ {k: f'{v:.1e}' for k, v in errs.items()} }")
    assert n_dev == n_ref > 8000
    assert errs.pop("timebin_mismatch") == 0
    assert max(errs.values()) <= TOL_RUN, errs
    assert dev.engine_info()["ti_current"] == osim.info()["ti_current"]
    assert dev.engine_info()["done"]


# ----------------------------------------------------------------------------------- 2D Sedov, 256^2
@pytest.fixture(scope="module")
def sedov256():
    return ics.sedov_2d(side=256, t_end=0.05)


@pytest.mark.parametrize("k", [0, 7, 40])
def test_c2_sedov256_face_and_step(ora, sedov256, k):
    n_act, gated, face_err, t_ref, t_dev, errs = _face_and_step(sedov256, k)
    say(f"[C2] sedov 256^2 k={k}: active {n_act}, computed faces {gated}, worst face {face_err:.1e}, "
        f"step errs { {q: f'{v:.1e}' for q, v in errs.items()} }")
    assert face_err <= TOL_STEP
    assert t_dev == t_ref
    assert errs.pop("timebin_mismatch") == 0
    assert max(errs.values()) <= TOL_STEP, errs


def test_c2_sedov256_full_run_and_conservation(ora, sedov256):
    say("[C2] sedov 256^2 to t_end 0.05")
    osim, dev, n_ref, n_dev, errs = _full_run(sedov256)
    say(f"  steps ref={n_ref} dev={n_dev} errs={ {k: f'{v:.1e}' for k, v in errs.items()} }")
    assert n_dev == n_ref == 4590
    assert errs.pop("timebin_mismatch") == 0
    assert errs["prim"] <= TOL_RUN and errs["cons"] <= TOL_RUN, errs
    # conservation of the device run: the initial state's sums against the final ones
    fresh = ics.build_device_sim(sedov256)
    c0, c1 = fresh.conservation(), dev.conservation()
    say(f"  conservation: mass {c1[0] - c0[0]:.2e}, energy {c1[4] - c0[4]:.2e}, momentum {np.abs(c1[1:4] - c0[1:4]).max():.2e}")
    assert abs(c1[0] - c0[0]) <= 1e-13 * abs(c0[0])
    assert abs(c1[4] - c0[4]) <= 1e-12 * abs(c0[4])
    assert np.all(np.abs(c1[1:4] - c0[1:4]) <= 1e-13)


# ----------------------------------------------------------------------------------- 3D KH, 128^3
@pytest.fixture(scope="module")
def kh128():
    return ics.kelvin_helmholtz_3d(side=128, t_end=1.0)


@pytest.mark.parametrize("k", [1])
def test_c3_kh128_face_and_step(ora, kh128, k):
    n_act, gated, face_err, t_ref, t_dev, errs = _face_and_step(kh128, k)
    say(f"[C3] kh 128^3 k={k}: active {n_act}, computed faces {gated}, worst face {face_err:.1e}, "
        f"step errs { {q: f'{v:.1e}' for q, v in errs.items()} }")
    assert gated > 1_000_000
    assert face_err <= TOL_STEP
    assert t_dev == t_ref
    assert errs.pop("timebin_mismatch") == 0
    assert max(errs.values()) <= TOL_STEP, errs


def test_c3_kh128_conservation(kh128):
    dev = ics.build_device_sim(kh128)
    c0 = dev.conservation()
    dev.run(12)
    c1 = dev.conservation()
    say(f"[C3] kh 128^3, 12 steps: mass {c1[0] - c0[0]:.2e}, energy {c1[4] - c0[4]:.2e}, "
        f"momentum {np.abs(c1[1:4] - c0[1:4]).max():.2e}")
    assert abs(c1[0] - c0[0]) <= 1e-13 * abs(c0[0])
    assert abs(c1[4] - c0[4]) <= 1e-12 * abs(c0[4])
    assert np.all(np.abs(c1[1:4] - c0[1:4]) <= 1e-12 * abs(c0[0]))


# ----------------------------------------------------------------------------------- ensemble of 512
def test_c4_ensemble512_members_vs_oracle(ora):
    probs = ensemble.ensemble_problems(count=512, seed=4, precision=fvs.FP64)
    say(f"[C4] 512 members, {sum(p.n for p in probs)} cells, FP64 batch, 300 launches")
    b = fvs.Batch.for_problems(probs)
    b.run(300)
    checked = 0
    for s in range(0, 512, 32):
        p = probs[s]
        ei = b.engine_info(s)
        osim = oracle_sim(p)
        try:
            n_ref = osim.advance(ei["step_count"])
            ora_err = 0
        except RuntimeError as e:
            n_ref, ora_err = osim.info()["step_count"], int(str(e).split()[-1])
        assert ei["error_code"] == ora_err, (s, ei["error_code"], ora_err)
        assert n_ref == ei["step_count"], s
        if ora_err:
            continue
        assert ei["ti_current"] == osim.info()["ti_current"], s
        errs = compare_states(b.get_state(s), osim.get_fields(), p.n)
        assert errs.pop("timebin_mismatch") == 0, s
        assert max(errs.values()) <= TOL_RUN, (s, errs)
        checked += 1
        say(f"  member {s}: {p.cfg.dim}D {p.n} cells, {n_ref} steps, worst {max(errs.values()):.1e}")
    assert checked >= 14


# ----------------------------------------------------------------------------------- served stream
def test_c5_served_stream_full_sizes(ora):
    jobs = serve.served_jobs(count=6, rate=400.0, seed=22, side_min=32, side_max=512, sedov_t_end=0.05,
                             vortex_t_end=0.5)
    sides = [int(round(np.sqrt(j.prob.n))) for j in jobs]
    say(f"[C5] {len(jobs)} jobs, sides {sides}, kinds {[j.kind for j in jobs]}")
    assert max(sides) >= 256 and min(sides) <= 64
    srv = serve.Server(hbm_budget_bytes=6 << 30, launches_per_poll=16)
    done = srv.serve(jobs, fetch=("prim", "cons", "timebin"))
    summ = serve.latency_summary(done, srv.wall)
    say(f"  {summ}")
    assert summ["failed_jobs"] == 0 and len(done) == len(jobs)
    overlapped = sum(1 for a in done for b in done if a is not b and a.admitted < b.finished and b.admitted < a.finished)
    assert overlapped > 0
    l1s, kinds = {}, set()
    for j in sorted(done, key=lambda j: j.prob.n):
        d = ics.build_device_sim(j.prob)
        d.run()
        alone = d.get_state(["prim", "cons", "timebin"])
        assert d.engine_info()["step_count"] == j.steps
        d.close()
        for k in ("prim", "cons", "timebin"):
            assert np.array_equal(j.result[k], alone[k]), (j.jid, k)
        # L1(rho) of the FP32-flux result at t_end against the FP64 oracle (the three smallest jobs and the
        # smallest of each kind)
        if len(l1s) < 3 or j.kind not in kinds:
            kinds.add(j.kind)
            osim = oracle_sim(ics.with_precision(j.prob, fvs.FP64))
            osim.advance(1 << 40)
            ref = osim.get_fields()["prim"][: j.prob.n]
            l1s[j.jid] = float(np.mean(np.abs(j.result["prim"][: j.prob.n] - ref)))
            say(f"  job {j.jid} ({j.kind}, side {int(round(np.sqrt(j.prob.n)))}): L1(rho) vs oracle {l1s[j.jid]:.2e}")
    assert len(l1s) >= 3
    assert max(l1s.values()) < 1e-5, l1s
