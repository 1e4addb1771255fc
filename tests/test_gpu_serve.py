"""Serving and ensembles on the device: each served job's result is the same alone and in the mix
(bit-identical: the same kernels and per-simulation engines, deterministic accumulation), and the
FP32-flux result stays within L1(rho) < 1e-5 of the FP64 path."""
import sys

import numpy as np
import pytest

from hydro_rs_amd import ensemble, fvs, ics, serve

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _device():
    if fvs.device_count() < 1:
        pytest.skip("no HIP device")


def _alone(prob):
    d = ics.build_device_sim(prob)
    d.run()
    st = d.get_state(["prim", "cons", "timebin"])
    info = d.engine_info()
    d.close()
    return st, info


def test_served_results_equal_alone_and_l1():
    jobs = serve.served_jobs(count=10, rate=400.0, seed=9, side_min=12, side_max=40, sedov_t_end=0.004,
                             vortex_t_end=0.05)
    srv = serve.Server(hbm_budget_bytes=1 << 30, launches_per_poll=4, classes=(1024, 4096), max_slots=6)
    done = srv.serve(jobs, fetch=("prim", "cons", "timebin"))
    summ = serve.latency_summary(done, srv.wall)
    say(f"[serve] {summ}")
    assert len(done) == len(jobs)
    assert summ["failed_jobs"] == 0 and all(j.error_code == 0 for j in done)
    overlapped = sum(1 for a in done for b in done if a is not b and a.admitted < b.finished and b.admitted < a.finished)
    assert overlapped > 0  # jobs were resident together
    for j in done:
        alone, info = _alone(j.prob)
        assert j.steps == info["step_count"], j.jid
        for k in ("prim", "cons", "timebin"):
            assert np.array_equal(j.result[k], alone[k]), (j.jid, k)
        # FP32-flux result vs the FP64 path on the same job
        p64 = j.prob
        p64.cfg.precision = fvs.FP64
        ref, _ = _alone(p64)
        p64.cfg.precision = fvs.FP32_FLUX
        l1 = float(np.mean(np.abs(j.result["prim"][: p64.n] - ref["prim"][: p64.n])))
        assert l1 < 1e-5, (j.jid, j.kind, l1)
    say(f"  {len(done)} jobs identical alone and in the mix")


def test_served_moving_mesh_jobs_equal_alone():
    """Moving-mesh jobs (2D steered Sedov, 2D vortex with fluid motion, 3D shear layer with fluid motion)
    served beside fixed-mesh ones: each job's mesh is rebuilt on the device in its slot, and every result
    equals the job run alone bit for bit."""
    import dataclasses

    def moving(p, motion):
        return dataclasses.replace(p, cfg=dataclasses.replace(p.cfg, motion=motion))

    probs = [("sedov_steer", moving(ics.sedov_2d(side=20, t_end=0.004), fvs.STEER)),
             ("vortex", ics.isentropic_vortex_2d(side=16, t_end=0.05)),
             ("vortex_fluid", moving(ics.isentropic_vortex_2d(side=18, t_end=0.05), fvs.FLUID)),
             ("kh3d_fluid", moving(ics.kelvin_helmholtz_3d(side=8, t_end=0.02), fvs.FLUID)),
             ("sedov", ics.sedov_2d(side=24, t_end=0.004))]
    jobs = [serve.Job(k, 0.001 * k, p, kind) for k, (kind, p) in enumerate(probs)]
    srv = serve.Server(hbm_budget_bytes=1 << 30, launches_per_poll=4, classes=(1024,), max_slots=3,
                       face_per_cell=9.0, conn_per_cell=18.0, headroom=1.5)
    done = srv.serve(jobs, fetch=("prim", "cons", "loc", "timebin"))
    assert len(done) == len(jobs) and all(j.error_code == 0 for j in done)
    for j in done:
        say(f"[serve moving] job {j.jid} {j.kind}: {j.steps} steps")
        alone, info = _alone(j.prob)
        d = ics.build_device_sim(j.prob)
        d.run()
        loc = d.get_state(["loc"])["loc"]
        d.close()
        assert j.steps == info["step_count"], j.jid
        for k in ("prim", "cons", "timebin"):
            assert np.array_equal(j.result[k], alone[k]), (j.jid, k)
        assert np.array_equal(j.result["loc"], loc), j.jid


def test_small_ensemble_all_members_finish():
    probs = ensemble.ensemble_problems(count=16, seed=3, n_min=64, n_max=2048, t_end=0.01)
    b, st = ensemble.run_ensemble(probs)
    say(f"[ensemble] {st['launches']} launches, {st['cell_updates']} cell updates, all done={st['all_done']}")
    assert st["all_done"]
    for s in (0, 5, 11):
        alone, info = _alone(probs[s])
        got = b.get_state(s, ["prim", "cons", "timebin"])
        for k in got:
            assert np.array_equal(got[k], alone[k]), (s, k)
