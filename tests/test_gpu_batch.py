"""Batched ensembles on the device: every simulation of a mixed batch produces exactly its standalone
result (same kernels, per-simulation engines, deterministic per-cell accumulation), slots can be
reloaded while others keep running, and the batch matches the oracle."""
import sys

import numpy as np
import pytest

from hydro_rs_amd import fvs, ics
from tests.helpers import compare_states, oracle_sim

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _device():
    if fvs.device_count() < 1:
        pytest.skip("no HIP device")


def _mix():
    return [
        ics.sod_1d(n=150, t_end=0.05),  # moving 1D mesh, exact solver
        ics.riemann_tube_1d(300, 1.0, 1.0, 0.3, 0.2, 0.1, -0.2, 5.0 / 3.0, t_end=0.05),
        ics.sedov_2d(side=24, t_end=0.01),
        ics.isentropic_vortex_2d(side=20, t_end=0.2),
        ics.kelvin_helmholtz_3d(side=8, t_end=0.05),
    ]


def _alone(prob, steps):
    d = ics.build_device_sim(prob)
    d.run(steps)
    return d.get_state(), d.engine_info()


@pytest.mark.parametrize("static", [1, 0])
def test_batch_equals_standalone(static):
    probs = _mix()
    say(f"[batch] {len(probs)} simulations: {[p.name for p in probs]}, static geometry {static}")
    b = fvs.Batch.for_problems(probs, headroom=1.25)
    b.set_geometry_mode(static)
    launches = b.run(1 << 40)
    say(f"  batch launches until all done: {launches}")
    for s, p in enumerate(probs):
        ei = b.engine_info(s)
        assert ei["done"], (p.name, ei)
        alone, ai = _alone(p, 1 << 40)
        got = b.get_state(s)
        assert ei["step_count"] == ai["step_count"], p.name
        assert ei["ti_current"] == ai["ti_current"], p.name
        for k in alone:
            assert np.array_equal(got[k], alone[k]), (p.name, k)
        say(f"  {p.name}: {ei['step_count']} steps, identical to standalone")


@pytest.mark.parametrize("precision", [fvs.FP64, fvs.FP32_FLUX])
def test_run_chunking_does_not_change_results(precision):
    """Finished simulations leave the launch grids when hfv_batch_run polls, so the grids (and the
    captured step graph) depend on how a caller splits its runs; the results must not."""
    probs = [ics.with_precision(p, precision) for p in _mix()]
    a = fvs.Batch.for_problems(probs)
    na = a.run(1 << 40)
    b = fvs.Batch.for_problems(probs)
    nb, calls = 0, 0
    while not all(b.engine_info(s)["done"] for s in range(len(probs))):
        nb += b.run(3)
        calls += 1
    say(f"[chunking] one call: {na} launches; {calls} calls of 3: {nb} launches")
    for s, p in enumerate(probs):
        ea, eb = a.engine_info(s), b.engine_info(s)
        assert ea["done"] and eb["done"] and ea["step_count"] == eb["step_count"], p.name
        ga, gb = a.get_state(s), b.get_state(s)
        for k in ga:
            assert np.array_equal(ga[k], gb[k]), (p.name, k)


def test_batch_2d_moving_member_equals_alone():
    """A 2D moving-mesh simulation in a batch (beside fixed 1D/2D ones and a moving 1D one) is rebuilt on
    the device in its slot and ends bit for bit where it ends alone."""
    v = ics.isentropic_vortex_2d(side=16, t_end=0.2)
    moving = ics.Problem(v.name, _with_motion(v.cfg, fvs.FLUID), v.loc, v.mass, v.vel, v.u_int, v.mesh)
    probs = [ics.sod_1d(n=150, t_end=0.05), moving, ics.sedov_2d(side=20, t_end=0.01), moving]
    b = fvs.Batch.for_problems(probs, headroom=1.5)
    b.run(1 << 40)
    for s in (1, 3):
        got = b.get_state(s)
        want, info = _alone(moving, 1 << 40)
        assert b.engine_info(s)["step_count"] == info["step_count"] and info["done"]
        for k in want:
            assert np.array_equal(got[k], want[k]), (s, k)


def _with_motion(cfg, motion):
    from dataclasses import replace

    return replace(cfg, motion=motion)


def test_batch_slot_reload_and_oracle(ora):
    probs = _mix()
    b = fvs.Batch.for_problems(probs[:3], headroom=2.0)
    b.run(5)
    # replace slot 1 while the others are part way through
    newp = ics.riemann_tube_1d(200, 0.5, 2.0, 0.0, 1.0, 0.2, 0.0, 1.4, t_end=0.03, seed=4)
    b.load(1, newp)
    b.run(1 << 40)
    got = b.get_state(1)
    ref = oracle_sim(newp)
    ref.advance(1 << 40)
    errs = compare_states(got, ref.get_fields(), newp.n)
    say(f"  reloaded slot vs oracle: { {k: f'{v:.1e}' for k, v in errs.items()} }")
    assert errs.pop("timebin_mismatch") == 0
    assert max(errs.values()) <= 1e-9
    assert b.engine_info(1)["step_count"] == ref.info()["step_count"]
    # the other slots still equal their standalone runs
    for s in (0, 2):
        alone, _ = _alone(probs[s], 1 << 40)
        got = b.get_state(s)
        for k in alone:
            assert np.array_equal(got[k], alone[k]), (s, k)


def test_batch_fp32_flux_matches_fp64_l1():
    probs64 = [ics.sedov_2d(side=32, t_end=0.01, seed=k) for k in range(4)]
    probs32 = [ics.sedov_2d(side=32, t_end=0.01, seed=k, precision=fvs.FP32_FLUX) for k in range(4)]
    b64 = fvs.Batch.for_problems(probs64)
    b32 = fvs.Batch.for_problems(probs32)
    b64.run(1 << 40)
    b32.run(1 << 40)
    for s, p in enumerate(probs64):
        r64 = b64.get_state(s, ["prim"])["prim"][: p.n]
        r32 = b32.get_state(s, ["prim"])["prim"][: p.n]
        l1 = float(np.mean(np.abs(r32 - r64)))
        say(f"  slot {s}: L1(rho) fp32 vs fp64 = {l1:.2e}")
        assert l1 < 1e-5


def test_failed_member_stops_alone():
    """A hard error in one simulation (here: more dump sync points between two steps than the engine
    drifts through, HFV_ERR_UNSUPPORTED) stops that slot only; the other slots run to the end."""
    good = [ics.riemann_tube_1d(300, 1.0, 1.0, 0.3, 0.2, 0.1, -0.2, 5.0 / 3.0, t_end=0.02),
            ics.sedov_2d(side=16, t_end=0.005)]
    bad = ics.riemann_tube_1d(300, 1.0, 1.0, 0.0, 0.125, 0.1, 0.0, 1.4, t_end=0.02, dt_snap=1e-6)
    probs = [good[0], bad, good[1]]
    b = fvs.Batch.for_problems(probs)
    b.run(1 << 40)
    ei = b.engine_info(1)
    say(f"  failed slot: error_code={ei['error_code']} after {ei['step_count']} steps")
    assert ei["done"] and ei["error_code"] == -7  # HFV_ERR_UNSUPPORTED
    for s in (0, 2):
        assert b.engine_info(s)["error_code"] == 0
        alone, ai = _alone(probs[s], 1 << 40)
        assert b.engine_info(s)["step_count"] == ai["step_count"]
        got = b.get_state(s)
        for k in alone:
            assert np.array_equal(got[k], alone[k]), (s, k)


@pytest.fixture(scope="module")
def ensemble64():
    from hydro_rs_amd import ensemble

    return ensemble.ensemble_problems(count=512, seed=4, precision=fvs.FP64)


@pytest.mark.parametrize("member", [20, 112, 470])
def test_hard_error_matches_oracle_step(ora, ensemble64, member):
    """The members of the 512-simulation ensemble whose flows drive a cell to negative pressure (1D
    members 20, 112 and 470: 7810, 25193 and 1283 cells): the reference's c2p assert
    (physical_quantities.rs:220-221) fires; the device path raises the same error in the same step as
    the oracle, alone and as a batch slot, and the batch's other slot finishes."""
    probs = ensemble64
    p = probs[member]
    say(f"[hard error] member {member}: {p.cfg.dim}D, {p.n} cells")
    ref = oracle_sim(p)
    with pytest.raises(RuntimeError) as oe:
        ref.advance(1 << 40)
    ora_code = int(str(oe.value).split()[-1])
    ri = ref.info()
    say(f"  oracle: error after {ri['step_count']} completed steps, t={ri['t_current']}")
    d = ics.build_device_sim(p)
    with pytest.raises(fvs.HfvError) as ex:
        d.run(1 << 40)
    assert ex.value.code == ora_code
    di = d.engine_info()
    assert di["step_count"] == ri["step_count"]
    b = fvs.Batch.for_problems([p, probs[0]])
    b.run(1 << 40)
    bi = b.engine_info(0)
    assert bi["error_code"] == ex.value.code and bi["step_count"] == ri["step_count"]
    assert b.engine_info(1)["done"] and b.engine_info(1)["error_code"] == 0


@pytest.mark.parametrize("precision", [fvs.FP64, fvs.FP32_FLUX])
@pytest.mark.parametrize("staged,static_grad,gstaged", [(s, g, n) for s in (0, 1) for g in (0, 1) for n in (0, 1)])
def test_tuning_choices_are_bit_identical(staged, static_grad, gstaged, precision):
    """Each flux tiling / gradient geometry choice, forced on every slot, ends a run on the same bits
    as the defaults (1D reflective, 2D periodic, 3D periodic)."""
    probs = [ics.sedov_2d(side=40, t_end=0.004, seed=1, precision=precision),
             ics.riemann_tube_1d(600, 1.0, 1.0, 0.0, 0.125, 0.1, 0.0, 1.4, t_end=0.02, precision=precision),
             ics.kelvin_helmholtz_3d(side=16, t_end=0.03, precision=precision)]
    a = fvs.Batch.for_problems(probs)
    for s in range(len(probs)):
        a.set_tuning(s, staged, static_grad, gstaged)
    # the 2D and 3D members have staging lists, so staged / gstaged = 1 run the LDS-staged kernels
    for s in (0, 2):
        assert a.tuning(s) == {"staged": staged, "static_grad": static_grad, "gstaged": gstaged}, probs[s].name
    a.run(1 << 40)
    b = fvs.Batch.for_problems(probs)
    b.run(1 << 40)
    for s, p in enumerate(probs):
        ea, eb = a.engine_info(s), b.engine_info(s)
        say(f"  {p.name}: tuning {a.tuning(s)} steps {ea['step_count']} vs {eb['step_count']} "
            f"err {ea['error_code']} / {eb['error_code']}")
        assert ea["error_code"] == 0 and eb["error_code"] == 0, p.name
        assert ea["step_count"] == eb["step_count"], p.name
        ga, gb = a.get_state(s), b.get_state(s)
        for k in ga:
            assert np.array_equal(ga[k], gb[k]), (p.name, k)


@pytest.mark.parametrize("precision", [fvs.FP64, fvs.FP32_FLUX])
@pytest.mark.parametrize("fpt", [2, 4])
def test_flux_launch_shapes_are_bit_identical(fpt, precision):
    """2 or 4 face chunks per flux block (AIRS with its exact fallback, staged and gathered tilings, a
    single-simulation context and a batch) end a run on the same bits as one chunk per block."""
    probs = [ics.sedov_2d(side=40, t_end=0.004, seed=1, precision=precision),
             ics.riemann_tube_1d(300, 1.0, 10.0, 0.0, 0.125, 0.1, 0.0, 1.4, solver=fvs.AIRS, t_end=0.02,
                                 precision=precision),
             ics.kelvin_helmholtz_3d(side=12, t_end=0.03, precision=precision)]
    a = fvs.Batch.for_problems(probs)
    a.set_launch_shape(fpt)
    assert a.launch_shape()["flux_faces_per_thread"] == fpt
    a.run(1 << 40)
    b = fvs.Batch.for_problems(probs)
    b.run(1 << 40)
    for s, p in enumerate(probs):
        ea, eb = a.engine_info(s), b.engine_info(s)
        assert ea["error_code"] == 0 and eb["error_code"] == 0 and ea["step_count"] == eb["step_count"], p.name
        ga, gb = a.get_state(s), b.get_state(s)
        for k in ga:
            assert np.array_equal(ga[k], gb[k]), (p.name, k)
    d1 = ics.build_device_sim(probs[2])
    d2 = ics.build_device_sim(probs[2])
    d2.set_launch_shape(fpt)
    d1.run(1 << 40)
    d2.run(1 << 40)
    s1, s2 = d1.get_state(), d2.get_state()
    for k in s1:
        assert np.array_equal(s1[k], s2[k]), k
    with pytest.raises(fvs.HfvError):
        d2.set_launch_shape(3)


def test_autotune_is_bit_identical():
    """hfv_autotune picks a tiling/geometry per size class by timing real launches; every candidate
    gives the same bits, so a tuned run ends exactly where an untuned one does."""
    # one size class each, in slot order; each simulation outlives the tuning of its class (277, 318,
    # 1025 and 516 steps against at most 8 candidates x 9 launches per class)
    probs = [ics.sedov_2d(side=40, t_end=0.05, seed=1), ics.sedov_2d(side=12, t_end=0.5, seed=2),
             ics.riemann_tube_1d(600, 1.0, 1.0, 0.0, 0.125, 0.1, 0.0, 1.4, t_end=0.08),
             ics.kelvin_helmholtz_3d(side=10, t_end=1.5)]
    a = fvs.Batch.for_problems(probs)
    report = a.autotune(launches=2, rounds=1)
    say(f"[autotune] {report}")
    assert len(report) == 4, report
    dims = sorted({r["dim"] for r in report})
    assert dims == [1, 2, 3]
    for r in report:
        assert r["staged"] in (0, 1) and r["static_grad"] in (0, 1) and r["gstaged"] in (0, 1)
        tried = [x for x in r["ns_per_update"] if x >= 0]
        assert len(tried) >= 2, r
        # the choice follows from the measured costs: the fastest candidate, unless the class's default
        # (staged tiling in 2D/3D where the class can stage; precomputed geometry and staged neighbours
        # in 3D) is within 1% of it
        ns = r["ns_per_update"]
        can_stage, can_gstage = ns[2] >= 0, ns[4] >= 0
        default = (2 if can_stage and r["dim"] != 1 else 0) | (1 if r["dim"] == 3 else 0) | \
            (4 if r["dim"] == 3 and can_gstage else 0)
        best = min((c for c in range(8) if ns[c] >= 0), key=lambda c: ns[c])
        chosen = r["static_grad"] + 2 * r["staged"] + 4 * r["gstaged"]
        expect = default if 0 <= ns[default] <= ns[best] * 1.01 else best
        assert chosen == expect, (r, default, best)
    # the flux launch shape, timed after the classes: 1 face per thread unless another beats it by 1%
    shape = a.launch_shape()
    say(f"[autotune] launch shape {shape}")
    ns = shape["ns_per_update"]
    assert all(x > 0 for x in ns), shape
    q = min(range(3), key=lambda k: ns[k])
    want = 1 if ns[0] <= ns[q] * 1.01 else 1 << q
    assert shape["flux_faces_per_thread"] == want, shape
    a.run(1 << 40)
    b = fvs.Batch.for_problems(probs)
    b.run(1 << 40)
    for s, p in enumerate(probs):
        ea, eb = a.engine_info(s), b.engine_info(s)
        say(f"  {p.name}: steps {ea['step_count']} vs {eb['step_count']} err {ea['error_code']} / {eb['error_code']}")
        assert ea["done"] and eb["done"], p.name
        assert ea["error_code"] == 0 and eb["error_code"] == 0, p.name
        assert ea["step_count"] == eb["step_count"], p.name
        ga, gb = a.get_state(s), b.get_state(s)
        for k in ga:
            assert np.array_equal(ga[k], gb[k]), (p.name, k)
