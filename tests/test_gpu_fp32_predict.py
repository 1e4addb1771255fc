"""FP32-flux contexts defer the engine steps' predicts: a cell's extrapolations are its stored sum plus
rate x (the predict time since its tag), added where they are read.  These tests pin that against the
summed form: the extrapolations a read returns track the FP64 context's per-step sums, flux readers
that add the pending part agree with states whose sums are settled every step, and the trait-level
phases after engine steps agree with the fused step."""
import sys

import numpy as np
import pytest

from hydro_rs_amd import fvs, ics

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _device():
    if fvs.device_count() < 1:
        pytest.skip("no HIP device")


def _kh(precision):
    # two layers of different sound speed: hierarchical time bins, and flow (non-zero rates) in every
    # cell; after steps 12-15 and 32-34 about 1700 of the 1728 cells hold a pending predict sum
    return ics.with_precision(ics.kelvin_helmholtz_3d(side=12, t_end=0.2), precision)


def _rel_l1(a, b):
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


def _ex_err(sa, sb):
    """Per-cell relative extrapolation differences (max over components / the cell's largest sum)
    over the cells with a sum that sit in the same time bin in both states (rounding can move a cell
    whose step lies on a bin boundary to the next bin: it then resets on other steps): their median
    and 90th percentile, the largest sum, and the number of moved cells.  Quantiles, not the maximum:
    a jittered mesh has cells with ill-conditioned gradient matrices (|grad v| ~ 2e4 on the 12^3
    layer, in FP64 as well), whose rates move by percents under rounding-level input changes; a
    dropped pending part puts the median at 1, a flipped one at 2."""
    same = sa["timebin"] == sb["timebin"]
    ea, eb = sa["extrap"].reshape(5, -1), sb["extrap"].reshape(5, -1)
    mag = np.abs(ea).max(axis=0)
    keep = same & (mag > 0)
    if not keep.any():
        return 0.0, 0.0, 0.0, int((~same).sum())
    rel = np.abs(ea[:, keep] - eb[:, keep]).max(axis=0) / mag[keep]
    return float(np.median(rel)), float(np.quantile(rel, 0.9)), float(mag.max()), int((~same).sum())


def test_fp32_extrapolations_track_fp64():
    """Where an FP32-flux and an FP64 context stand at the same time after the same steps, the
    extrapolations the FP32 one reports (stored sum + rate x pending time) equal the FP64 per-step
    sums to FP32 accuracy; a dropped or sign-flipped pending term is off by the sums themselves.
    (The FP32 path's signal speeds can move a cell across a time-bin boundary, after which the two
    clocks differ: only steps with equal clocks are compared.)"""
    p64 = ics.sedov_2d(side=32, t_end=0.01, seed=3)
    a = ics.build_device_sim(p64)
    b = ics.build_device_sim(ics.with_precision(p64, fvs.FP32_FLUX))
    compared = 0
    for k in range(1, 25):
        a.run(1)
        b.run(1)
        ia, ib = a.engine_info(), b.engine_info()
        if ia["ti_current"] != ib["ti_current"]:
            continue
        sa, sb = a.get_state(["extrap", "timebin"]), b.get_state(["extrap", "timebin"])
        pending = np.count_nonzero(np.abs(sa["extrap"]).reshape(5, -1).max(axis=0) > 0)
        med, p90, scale, flips = _ex_err(sa, sb)
        if pending < 10 or scale < 1e-6:
            continue
        say(f"[fp32 predict] step {k}: {pending} cells with a pending sum, max |ex| {scale:.3e}, per-cell "
            f"rel |ex32 - ex64| median {med:.2e} p90 {p90:.2e} ({flips} cells in other time bins)")
        assert med <= 1e-4 and p90 <= 1e-3 and flips <= a.n // 100, k
        compared += 1
    assert compared >= 3


def test_fp32_reads_equal_device_sums():
    """A state read sums the pending predicts on the host and leaves the device as it is; the
    trait-level predict with dt = 0 sums them on the device.  Both give the same extrapolations."""
    b = ics.build_device_sim(_kh(fvs.FP32_FLUX))
    b.run(13)
    host = b.get_state(["extrap"])["extrap"]
    again = b.get_state(["extrap"])["extrap"]
    assert np.array_equal(host, again)
    pending = np.count_nonzero(np.abs(host).reshape(5, -1).max(axis=0) > 0)
    b.predict(0.0)
    dev = b.get_state(["extrap"])["extrap"]
    err = np.abs(host - dev).max() / np.abs(host).max()
    say(f"[fp32 predict] {pending} of {b.n} cells pending after 13 steps; host vs device sums: {err:.2e}")
    assert pending > b.n // 2
    assert err <= 1e-6


def test_deferred_readers_equal_settled_sums():
    """One FP32-flux context steps with the pending predicts added by the flux kernel's readers; a
    second one writes its own state back after every step (a host write sums every cell's pending
    predicts into the stored extrapolations first), so its readers never see a pending part.  The two
    differ only by FP32 rounding of the same products."""
    p = _kh(fvs.FP32_FLUX)
    a = ics.build_device_sim(p)
    b = ics.build_device_sim(p)
    for k in (13, 34):
        a.run(k - a.engine_info()["step_count"])
        while b.engine_info()["step_count"] < k:
            b.run(1)
            b.set_state(b.get_state())
        assert a.engine_info()["ti_current"] == b.engine_info()["ti_current"]
        sa, sb = a.get_state(), b.get_state()
        e_rho = _rel_l1(sb["prim"][: a.n], sa["prim"][: a.n])
        med, p90, scale, flips = _ex_err(sa, sb)
        say(f"[fp32 predict] deferred vs settled each step, step {k}: rel L1(rho) {e_rho:.2e}, "
            f"max |ex| {scale:.2e}, per-cell rel extrap median {med:.2e} p90 {p90:.2e} "
            f"({flips} cells in other time bins)")
        assert e_rho <= 1e-6
        assert med <= 1e-5 and p90 <= 1e-3 and flips <= a.n // 100
        if k == 13:
            assert scale > 0  # pending sums at this step (1727 of 1728 cells)


def test_phases_after_engine_steps_equal_fused_step():
    """Engine steps leave predicts pending; the trait-level predict then sums them with its own dt, and
    the phase sequence matches hfv_step on the same state to FP32 rounding."""
    p = _kh(fvs.FP32_FLUX)
    a = ics.build_device_sim(p)
    b = ics.build_device_sim(p)
    a.run(13)
    b.run(13)
    tb = p.cfg.t_end / float(1 << 57)
    ei = a.engine_info()
    ti_old, ti_cur = ei["ti_current"], ei["ti_next"]
    for _ in range(3):
        info = a.timestep_info(ti_old, ti_cur, tb)
        t1 = a.step(info)
        b.predict((ti_cur - ti_old) * tb)
        b.compute_fluxes(info)
        b.apply_fluxes()
        b.convert_conserved_to_primitive()
        b.gradients()
        t2 = b.timesteps()
        assert t1 == t2
        ti_old, ti_cur = ti_cur, t1
    sa, sb = a.get_state(), b.get_state()
    e_rho = _rel_l1(sb["prim"][: a.n], sa["prim"][: a.n])
    say(f"[fp32 predict] phases vs fused after 13 engine steps: rel L1(rho) {e_rho:.2e}")
    assert e_rho <= 1e-6
