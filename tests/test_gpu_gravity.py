"""Gravity solvers through the C-ABI against the oracle (src/gravity.rs; the runner's use of them,
optimal_order.rs:47-50, godunov.rs:41-44).

GravitySolver::compute_timesteps and ::compute_accelerations of the device equal the oracle's within
1e-14 (norm-relative; both evaluate the same correctly rounded operations in the same order); the
Keplerian potential's timestep in the fused step gives the oracle's whole run within 1e-9 with timebins and
step count exact, and the phase-by-phase step with the caller taking the minimum equals the fused step bit
for bit.
"""
import dataclasses
import sys

import numpy as np
import pytest

from hydro_rs_amd import fvs, ics
from tests.helpers import compare_states, oracle_sim, rel_err

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _device():
    if fvs.device_count() < 1:
        pytest.skip("no HIP device")


def _disc(gravity=fvs.GRAV_KEPLERIAN, **kw):
    prob = ics.keplerian_disc_2d(**kw)
    return dataclasses.replace(prob, cfg=dataclasses.replace(prob.cfg, gravity=gravity,
                                                             grav_acceleration=(0.25, -1.5, 0.0)))


@pytest.mark.parametrize("gravity", [fvs.GRAV_SELF, fvs.GRAV_CONSTANT, fvs.GRAV_KEPLERIAN])
def test_gravity_solver_entries(ora, gravity):
    """compute_timesteps (+inf for self gravity and the constant potential while a_grav = 0) and
    compute_accelerations at the initial state and after 9 steps (masses and positions of a run)."""
    say(f"[gravity entries] kind={gravity} disc side=24")
    prob = _disc(gravity)
    osim = oracle_sim(prob)
    dev = ics.build_device_sim(prob)
    n = prob.n
    for steps in (0, 9):
        if steps:
            assert osim.advance(steps) == steps
            assert dev.run(steps) == steps
        dt_d, dt_o = dev.gravity_timesteps(), osim.gravity_timesteps()
        if gravity == fvs.GRAV_KEPLERIAN:
            assert np.all(np.isfinite(dt_o))
            assert rel_err(dt_d, dt_o) <= 1e-14, rel_err(dt_d, dt_o)
        else:
            assert np.all(np.isinf(dt_d)) and np.all(np.isinf(dt_o))
        a_d = dev.gravity_accelerations()
        a_o = osim.gravity_accelerations().reshape(3, n)
        errs = [rel_err(a_d[k], a_o[k]) for k in range(2)]
        say(f"  steps={steps} accel errs {errs}")
        assert max(errs) <= 1e-14, errs
        assert np.all(a_d[2] == a_o[2])
        if gravity == fvs.GRAV_SELF:
            assert np.all(a_d[:, -1] == 0.0)  # the last particle has no partners after it
            assert np.max(np.abs(a_d[:2, 0])) > 0.0


def test_gravity_self_sum_order_large(ora):
    """Self gravity over 2 full LDS tiles and a ragged one (600 particles), against the oracle, and the
    first particle's sum against a float64 direct sum in the reference's order."""
    prob = _disc(fvs.GRAV_SELF, side=25)
    osim = oracle_sim(prob)
    dev = ics.build_device_sim(prob)
    n = prob.n
    assert n > 512 and n % 256 != 0
    a_d = dev.gravity_accelerations()
    a_o = osim.gravity_accelerations().reshape(3, n)
    assert max(rel_err(a_d[k], a_o[k]) for k in range(3)) <= 1e-14
    x = prob.loc
    m = prob.mass
    acc = np.zeros(3)
    for j in range(1, n):
        d = x[j] - x[0]
        r = np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]) + prob.cfg.grav_softening
        acc += (m[j] * d) / (r * r * r)
    assert rel_err(a_d[:, 0], acc) <= 1e-14


def test_keplerian_full_run_parity(ora):
    """hfv_run with the Keplerian potential: the step takes min(hydro, Keplerian) timesteps, so the run has
    more steps than without gravity, and matches the oracle's run."""
    prob = _disc(t_end=0.2)
    osim = oracle_sim(prob)
    n_ref = osim.advance(1 << 40)
    ref = osim.get_fields()
    n_nograv = oracle_sim(_disc(fvs.GRAV_NONE, t_end=0.2)).advance(1 << 40)
    dev = ics.build_device_sim(prob)
    n_dev = dev.run(1 << 40)
    errs = compare_states(dev.get_state(), ref, prob.n)
    say(f"[kepler run] steps ref={n_ref} dev={n_dev} (no gravity {n_nograv}) errs="
        f"{ {k: f'{v:.1e}' for k, v in errs.items()} }")
    assert n_dev == n_ref and n_ref > n_nograv
    assert errs.pop("timebin_mismatch") == 0
    assert max(errs.values()) <= 1e-9, errs


def test_keplerian_phases_equal_fused_step():
    """compute_timesteps, the gravity timesteps and their minimum on the host, apply_timesteps: the fused
    step bit for bit (optimal_order.rs:45-51)."""
    prob = _disc()
    a = ics.build_device_sim(prob)
    b = ics.build_device_sim(prob)
    tb = prob.cfg.t_end / float(1 << 57)
    ti_old, ti_cur = 0, 0
    binding = 0
    for step in range(8):
        info = a.timestep_info(ti_old, ti_cur, tb)
        t1 = a.step(info)
        b.predict((ti_cur - ti_old) * tb)
        b.compute_fluxes(info)
        b.apply_fluxes()
        b.convert_conserved_to_primitive()
        b.gradients()
        dts = b.compute_timesteps()
        dtg = b.gravity_timesteps()
        binding += int(np.sum(dtg < dts))
        t2 = b.apply_timesteps(np.fmin(dts, dtg))
        assert t1 == t2, step
        sa, sb = a.get_state(), b.get_state()
        for k in sa:
            assert np.array_equal(sa[k], sb[k]), (step, k)
        ti_old, ti_cur = ti_cur, t1
    assert binding > 100


def test_no_gravity_solver_is_an_error():
    dev = ics.build_device_sim(_disc(fvs.GRAV_NONE))
    with pytest.raises(fvs.HfvError):
        dev.gravity_timesteps()
    with pytest.raises(fvs.HfvError):
        dev.gravity_accelerations()
