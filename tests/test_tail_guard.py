"""The FP64 rarefaction-tail sign test of the device sampler (tail_speed_for_sign, hfv_physics.hpp) first
estimates u - a * pdps^e with an FP32 pow and trusts the estimate's sign outside an error band.  This
restates that guard with numpy's FP32 pow and checks, by brute force over random and adversarial
inputs, that it never decides differently from the FP64 expression (riemann_solver.rs rarefaction tail
speed) — i.e. that the band covers the FP32 pow's error."""
import numpy as np


def guarded_sign(u, a, pdps, e):
    y32 = np.power(np.float32(pdps), np.float32(e)).astype(np.float64)
    ay = a * y32
    t = u - ay
    bound = 4e-5 * np.abs(ay) + 1e-14 * (np.abs(u) + np.abs(ay))
    usable = (pdps > 1e-30) & (pdps < 1e30)
    exact = u - a * np.power(pdps, e)
    certain = usable & ((t < -bound) | (t > bound))
    return np.where(certain, np.sign(t), np.sign(exact)), np.sign(exact), certain


def test_guard_matches_fp64_decision():
    rng = np.random.default_rng(7)
    n = 2_000_000
    for gamma in (1.4, 5.0 / 3.0, 1.1, 3.0):
        e = (gamma - 1.0) / (2.0 * gamma)
        pdps = np.exp(rng.uniform(np.log(1e-29), np.log(1e29), n))
        a = np.exp(rng.uniform(np.log(1e-6), np.log(1e6), n)) * rng.choice([-1.0, 1.0], n)
        y = np.power(pdps, e)
        # u near the root a * y (relative offsets down to 1e-9) and far from it
        u = a * y * (1.0 + rng.choice([-1.0, 1.0], n) * np.exp(rng.uniform(np.log(1e-9), np.log(10.0), n)))
        got, want, certain = guarded_sign(u, a, pdps, e)
        assert np.array_equal(got, want), gamma
        assert certain.mean() > 0.5  # the FP32 estimate decides most cases


def test_guard_edges():
    e = 0.2
    for pdps in (0.0, 1e-40, 1e40, np.inf, np.nan, -1.0, 1.0):
        for u, a in ((0.0, 1.0), (1.0, 1.0), (-1.0, 2.0), (1.0, 0.0)):
            with np.errstate(all="ignore"):
                got, want, _ = guarded_sign(np.float64(u), np.float64(a), np.float64(pdps), e)
            assert np.array_equal(got, want) or (np.isnan(got) and np.isnan(want)), (pdps, u, a)
