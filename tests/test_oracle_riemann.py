"""Pins the CPU oracle's Riemann solvers and limiter against the reference's own known answers.

Golden vectors are the reference's unit-test values, compared at the reference's margin
(float_cmp default: epsilon = f64::EPSILON, ulps = 4).
"""
import numpy as np
import pytest

G53 = 5.0 / 3.0
X = np.array([1.0, 0.0, 0.0])


def approx_eq(a, b, ulps=4, eps=np.finfo(np.float64).eps):
    """float_cmp::assert_approx_eq!(f64, a, b) default margin."""
    a = np.float64(a)
    b = np.float64(b)
    if abs(a - b) <= eps:
        return True
    ia = a.view(np.int64)
    ib = b.view(np.int64)
    return (np.sign(a) == np.sign(b)) and abs(int(ia) - int(ib)) <= ulps


def prim(rho, v, p):
    v = np.broadcast_to(np.asarray(v, dtype=np.float64), (3,)) if np.ndim(v) else np.array([v, 0.0, 0.0])
    return np.array([rho, v[0], v[1], v[2], p])


def boost(w, vel):
    w = np.array(w, dtype=np.float64)
    if w[0] > 0.0:
        w[1:4] += vel
    return w


EXACT_GOLDEN = [
    # src/riemann_solver/exact.rs:363-397 (gamma = 5/3)
    ((0.75, 0.6, 0.8, 0.55, -0.25, 0.5), (0.8927393244785966, 0.35966250129769556, 1.0709476357371466)),
    ((0.75, 0.3, 0.8, 0.55, -0.75, 0.5), (0.9114320476995261, -0.03895103710637815, 1.201228148376388)),
    ((0.75, 0.1, 0.8, 0.55, 0.5, 0.5), (0.5562398903431287, 0.4792910646431622, 0.4861363537732538)),
    ((1.75, -0.1, 0.75, 0.55, 1.5, 0.5), (0.65432665196683, 0.6088656910463873, 0.14554217677392042)),
    ((0.75, -0.5, 0.8, 0.55, -0.1, 0.5), (0.5407985848180775, -0.1207089353568378, 0.4861363537732538)),
    ((1.75, -1.5, 0.75, 1.55, 0.1, 0.5), (0.5687181105004402, -0.5249266813300747, 0.09402548983542296)),
]


@pytest.mark.parametrize("inp,exp", EXACT_GOLDEN)
def test_exact_half_state_golden(ora, inp, exp):
    rl, vl, pl, rr, vr, pr = inp
    a_l = ora.sound_speed(G53, pl, 1.0 / rl)
    a_r = ora.sound_speed(G53, pr, 1.0 / rr)
    h = ora.sample(ora.EXACT, G53, prim(rl, vl, pl), prim(rr, vr, pr), vl, vr, a_l, a_r, X)
    assert approx_eq(h[0], exp[0]), (h[0], exp[0])
    assert approx_eq(h[1], exp[1]), (h[1], exp[1])
    assert approx_eq(h[4], exp[2]), (h[4], exp[2])


def test_vacuum_half_state_golden(ora):
    # src/riemann_solver/vacuum.rs:186-220
    iv = 0.15
    rho_l, v_l, p_l = 1.0, 0.3, 0.5
    a_l = ora.sound_speed(G53, p_l, 1.0 / rho_l)
    left = boost(prim(rho_l, v_l, p_l), -iv * X)
    h = ora.sample(ora.VACUUM_SAMPLER, G53, left, prim(0.0, 0.0, 0.0), v_l - iv, 0.0, a_l, 0.0, X)
    assert approx_eq(h[0], 0.4950623323204316)
    assert approx_eq(h[1], 0.7221531968814576)
    assert approx_eq(h[4], 0.15490656018956153)


@pytest.mark.parametrize("kind", ["EXACT", "HLLC"])
def test_symmetry(ora, kind):
    # exact.rs:313-342, hllc.rs:140-166
    k = getattr(ora, kind)
    iv = -3e-1 * X
    left = prim(1.0, 0.0, 1.0)
    right = prim(1.0, -6e-1, 1.0)
    right_rev = prim(1.0, 6e-1, 1.0)
    f = ora.solve_for_flux(k, G53, boost(left, -iv), boost(right, -iv), iv, X)
    fr = ora.solve_for_flux(k, G53, boost(right_rev, iv), boost(left, iv), -iv, X)
    assert approx_eq(f[0], -fr[0])
    assert approx_eq(f[1], fr[1])
    assert approx_eq(f[4], -fr[4])
    if kind == "EXACT":
        assert approx_eq(f[2], fr[2]) and approx_eq(f[3], fr[3])


def test_vacuum_symmetry(ora):
    # vacuum.rs:152-183 (the vacuum sampler is reached through any star solver)
    iv = 0.15 * X
    vel = np.array([0.3, 0.0, 0.0])
    left = prim(1.0, vel, 0.5)
    left_rev = prim(1.0, -vel, 0.5)
    right = prim(0.0, 0.0, 0.0)
    f = ora.solve_for_flux(ora.EXACT, G53, boost(left, -iv), right, iv, X)
    fr = ora.solve_for_flux(ora.EXACT, G53, right, boost(left_rev, iv), -iv, X)
    for i, s in zip(range(5), (-1, 1, 1, 1, -1)):
        assert approx_eq(f[i], s * fr[i]), (i, f, fr)


def _invariance(ora, kind, left, right, iv, vacuum_sampler=False):
    a_l = ora.sound_speed(G53, left[4], 1.0 / left[0])
    a_r = ora.sound_speed(G53, right[4], 1.0 / right[0]) if right[0] > 0 else 0.0
    w = ora.sample(ora.VACUUM_SAMPLER if vacuum_sampler else kind, G53, left, right, left[1],
                   right[1] if right[0] > 0 else 0.0, a_l, a_r, X)
    roe = w[4] / (G53 - 1.0) + 0.5 * w[0] * np.dot(w[1:4], w[1:4])
    lab = ora.solve_for_flux(kind, G53, left, right, np.zeros(3), X)
    lab = lab - np.array([w[0] * iv[0], *(w[0] * w[1:4] * iv), roe * iv[0]])
    face = ora.solve_for_flux(kind, G53, boost(left, -iv), boost(right, -iv), iv, X)
    return lab, face


@pytest.mark.parametrize("kind,iv", [("EXACT", -0.1), ("PVRS", -0.17)])
def test_invariance(ora, kind, iv):
    # exact.rs:401-438, pvrs.rs:64-101
    lab, face = _invariance(ora, getattr(ora, kind), prim(1.5, 0.2, 1.2), prim(0.7, -0.4, 0.1), iv * X)
    assert approx_eq(lab[0], face[0])
    assert approx_eq(lab[1], face[1])
    assert approx_eq(lab[4], face[4])


@pytest.mark.xfail(strict=True, reason="vacuum.rs:223-260 samples the left rarefaction fan at x/t=0 in two frames "
                   "moving at 0.3 relative to each other; the fan state differs, so this reference test cannot "
                   "pass with the reference's own sampler (its restatement here reproduces that)")
def test_vacuum_invariance(ora):
    # vacuum.rs:223-260
    lab, face = _invariance(ora, ora.EXACT, prim(1.5, -0.2, 1.2), prim(0.0, 0.0, 0.0), 0.3 * X, True)
    assert approx_eq(lab[0], face[0])
    assert approx_eq(lab[1], face[1])
    assert approx_eq(lab[4], face[4])


def test_hllc_vs_exact(ora):
    # hllc.rs:169-195
    iv = 0.15 * X
    left = prim(1.0, 0.2, 0.5)
    right = prim(0.5, 0.1, 0.1)
    f = ora.solve_for_flux(ora.HLLC, G53, boost(left, -iv), boost(right, -iv), iv, X)
    fs = ora.solve_for_flux(ora.EXACT, G53, boost(left, -iv), boost(right, -iv), iv, X)
    for i in (0, 1, 4):
        assert abs(f[i] - fs[i]) <= 5e-2


def test_swift_near_vacuum_regression(ora):
    # exact.rs:441-460: must not panic
    left = np.array([2.1669723793e-9, -2.9266309738e00, 7.4537420273e00, 1.4240785599e01, 1.4698121886e-10])
    right = np.array([1.7172516742e-09, 1.8704223633e00, -7.5396900177e00, -1.4482626915e01, 1.1778487907e-10])
    e, star = ora.exact_star(G53, left, right, -1.5249514580e00, 4.9803251028e-01, 3.3622390032e-01,
                             3.3810544014e-01)
    assert e == 0
    assert np.all(np.isfinite(star))


def _rel_ok(a, b, tol=1e-4):
    # python/riemann_solver.py:962-966
    return a == b or abs(a - b) < tol * abs(a + b)


TORO_PSTAR = [  # python/riemann_solver.py:1094-1101, gamma = 1.4
    ((1.0, 0.0, 1.0, 0.125, 0.0, 0.1), 0.30313),
    ((1.0, -2.0, 0.4, 1.0, 2.0, 0.4), 0.001894),
    ((1.0, 0.0, 1000.0, 1.0, 0.0, 0.01), 460.894),
    ((1.0, 0.0, 0.01, 1.0, 0.0, 100.0), 46.095),
    ((5.99924, 19.5975, 460.894, 5.99242, -6.19633, 46.0950), 1691.64),
]


@pytest.mark.parametrize("inp,pstar", TORO_PSTAR)
def test_toro_pstar(ora, inp, pstar):
    rl, vl, pl, rr, vr, pr = inp
    g = 1.4
    a_l = ora.sound_speed(g, pl, 1.0 / rl)
    a_r = ora.sound_speed(g, pr, 1.0 / rr)
    e, star = ora.exact_star(g, prim(rl, vl, pl), prim(rr, vr, pr), vl, vr, a_l, a_r)
    assert e == 0
    assert _rel_ok(star[3], pstar), (star[3], pstar)


TORO_SAMPLE = [  # python/riemann_solver.py:1103-1131, gamma = 5/3, sampled at x/t = 0
    ((1.0, 0.0, 1.0, 0.125, 0.0, 0.1), (0.47969, 0.841194, 0.293945)),
    ((1.0, -2.0, 0.4, 1.0, 2.0, 0.4), (0.00617903, 0.0, 8.32249e-05)),
    ((1.0, 0.0, 1000.0, 1.0, 0.0, 0.01), (0.615719, 18.2812, 445.626)),
    ((1.0, 0.0, 0.01, 1.0, 0.0, 100.0), (0.61577, -5.78011, 44.5687)),
    ((5.99924, 19.5975, 460.894, 5.99242, -6.19633, 46.0950), (12.743, 8.56045, 1841.82)),
    ((1.0, -1.0, 1.0e-6, 1.0, 1.0, 1.0005e-6), (0.0, 0.0, 0.0)),
]


@pytest.mark.parametrize("inp,exp", TORO_SAMPLE)
def test_toro_sampled(ora, inp, exp):
    rl, vl, pl, rr, vr, pr = inp
    f = ora.solve_for_flux(ora.EXACT, G53, prim(rl, vl, pl), prim(rr, vr, pr), np.zeros(3), X)
    # recover the half state from the sampler directly
    a_l = ora.sound_speed(G53, pl, 1.0 / rl)
    a_r = ora.sound_speed(G53, pr, 1.0 / rr)
    vac = (2.0 / (G53 - 1.0)) * (a_l + a_r) <= vr - vl
    h = ora.sample(ora.VACUUM_SAMPLER if vac else ora.EXACT, G53, prim(rl, vl, pl), prim(rr, vr, pr), vl, vr,
                   a_l, a_r, X)
    # the Rust solver stops its iteration at 1e-6 relative; the Python reference at 5e-9
    assert _rel_ok(h[0], exp[0]) or abs(h[0] - exp[0]) < 1e-9
    assert _rel_ok(h[1], exp[1]) or abs(h[1] - exp[1]) < 1e-9
    assert _rel_ok(h[4], exp[2]) or abs(h[4] - exp[2]) < 1e-9
    assert np.all(np.isfinite(f))


@pytest.mark.parametrize("kind", ["EXACT", "HLLC", "PVRS", "AIRS", "TSRS", "TRRS"])
def test_consistency_equal_states(ora, kind):
    """F(W, W) is the physical Euler flux (closed form), for any solver."""
    rng = np.random.default_rng(7)
    for _ in range(20):
        rho = rng.uniform(0.1, 3.0)
        v = rng.uniform(-1, 1, 3)
        p = rng.uniform(0.1, 3.0)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        w = prim(rho, v, p)
        f = ora.solve_for_flux(getattr(ora, kind), 1.4, w, w, np.zeros(3), n)
        vn = v @ n
        E = p / 0.4 + 0.5 * rho * v @ v
        exp = np.array([rho * vn, *(rho * vn * v + p * n), (E + p) * vn])
        np.testing.assert_allclose(f, exp, rtol=1e-10, atol=1e-12)



@pytest.mark.parametrize("kind", ["EXACT", "HLLC"])
def test_isolated_contact_is_upwinded(ora, kind):
    """An isolated contact or shear layer (equal pressure and normal velocity, different densities and
    transverse velocities): the exact solver and HLLC (whose middle wave is the contact) both return the
    physical Euler flux of the upwind state, in closed form; the upwind side is the left one for u_n > 0."""
    rng = np.random.default_rng(11)
    for _ in range(20):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        un = rng.choice([-1.0, 1.0]) * rng.uniform(0.05, 1.0)
        p = rng.uniform(0.1, 3.0)
        t1, t2 = rng.normal(size=3), rng.normal(size=3)
        vl = un * n + (t1 - (t1 @ n) * n)  # transverse parts differ, normal parts equal
        vr = un * n + (t2 - (t2 @ n) * n)
        wl, wr = prim(rng.uniform(0.1, 3.0), vl, p), prim(rng.uniform(0.1, 3.0), vr, p)
        f = ora.solve_for_flux(getattr(ora, kind), 1.4, wl, wr, np.zeros(3), n)
        w = wl if un > 0 else wr
        rho, v = w[0], w[1:4]
        E = p / 0.4 + 0.5 * rho * v @ v
        exp = np.array([rho * un, *(rho * un * v + p * n), (E + p) * un])
        np.testing.assert_allclose(f, exp, rtol=1e-9, atol=1e-12)


def _two_wave_case(rng, eps, apart):
    """Nearly equal states whose normal velocities are -eps/+eps (apart: the flows separate, two
    rarefactions) or +eps/-eps (they collide, two shocks)."""
    rho, p = rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0)
    s = -1.0 if apart else 1.0
    return prim(rho, s * eps, p), prim(rho * (1.0 + 0.3 * eps), -s * eps, p * (1.0 + 0.2 * eps))


def _star_pressure(ora, wl, wr, g=1.4):
    a_l = ora.sound_speed(g, wl[4], 1.0 / wl[0])
    a_r = ora.sound_speed(g, wr[4], 1.0 / wr[0])
    e, star = ora.exact_star(g, wl, wr, wl[1], wr[1], a_l, a_r)
    assert e == 0
    return star[3]


def test_trrs_is_exact_for_two_rarefactions(ora):
    """When both waves are rarefactions the two-rarefaction estimate is the exact star pressure (trrs.rs:
    the rarefaction relations solved for p* in closed form): its flux equals the exact solver's, on
    random separating flows and on Toro's test 2 (p* = 0.001894)."""
    rng = np.random.default_rng(5)
    cases = [_two_wave_case(rng, eps, True) for eps in rng.uniform(0.05, 0.8, 20)]
    cases.append((prim(1.0, -2.0, 0.4), prim(1.0, 2.0, 0.4)))
    for wl, wr in cases:
        assert _star_pressure(ora, wl, wr) < min(wl[4], wr[4])  # two rarefactions indeed
        fe = ora.solve_for_flux(ora.EXACT, 1.4, wl, wr, np.zeros(3), X)
        ft = ora.solve_for_flux(ora.TRRS, 1.4, wl, wr, np.zeros(3), X)
        np.testing.assert_allclose(ft, fe, rtol=1e-12, atol=1e-14 * np.max(np.abs(fe)))


def test_tsrs_converges_to_exact_at_third_order(ora):
    """For colliding flows (two shocks) the two-shock estimate (tsrs.rs) differs from the exact solution
    by O(eps^3) in the wave strength eps (shock curves evaluated at the PVRS guess, which is itself
    second-order accurate): each tenfold weaker pair of shocks brings the flux at least 500x closer to
    the exact solver's.  A wrong term would leave an O(eps) or O(1) difference."""
    rng = np.random.default_rng(6)
    for _ in range(6):
        seed = int(rng.integers(1 << 30))
        d = []
        for eps in (1e-1, 1e-2, 1e-3):
            wl, wr = _two_wave_case(np.random.default_rng(seed), eps, False)
            assert _star_pressure(ora, wl, wr) > max(wl[4], wr[4])  # two shocks indeed
            fe = ora.solve_for_flux(ora.EXACT, 1.4, wl, wr, np.zeros(3), X)
            ft = ora.solve_for_flux(ora.TSRS, 1.4, wl, wr, np.zeros(3), X)
            d.append(np.max(np.abs(ft - fe)) / np.max(np.abs(fe)))
        assert d[0] / d[1] > 500 and d[1] / d[2] > 500, d


def test_linear_advection_upwinds_the_primitive_fields(ora):
    """linear_advection.rs:17-33: v = (velocity - interface velocity) . n; the flux is v times the upwind
    state's (density, velocity, pressure) placed in the conserved slots as they are (left for v > 0, right
    otherwise, v = 0 included).  Cases where a flipped interface velocity or swapped side gives another
    answer."""
    wl, wr = prim(1.3, (0.2, -0.1, 0.4), 0.7), prim(0.6, (-0.5, 0.3, 0.1), 1.9)
    n = np.array([0.6, 0.8, 0.0])
    adv = np.array([1.0, 0.5, -2.0])
    for iv, side in (((0.0, 0.0, 0.0), "L"), ((2.5, 0.0, 0.0), "R"), ((0.0, 1.5, 0.0), "R"), ((1.0, 0.5, 7.0), "R"),
                     ((-1.0, 0.0, 0.0), "L")):
        iv = np.asarray(iv)
        v = (adv - iv) @ n
        assert (v > 0) == (side == "L")
        f = ora.solve_for_flux(ora.LINADV, 1.4, wl, wr, iv, n, velocity=adv)
        exp = v * (wl if side == "L" else wr)
        np.testing.assert_array_equal(f, exp)
        np.testing.assert_array_equal(np.signbit(f), np.signbit(exp))  # at v = 0: the signed zeros of the side
