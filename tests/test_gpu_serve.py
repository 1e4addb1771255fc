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
