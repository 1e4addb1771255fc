"""Pins of the oracle's single functions against closed forms, invariants and brute force.

Each test states what fixes the expected value independently of the oracle's own expressions (a
textbook identity, a conservation law, an exact solution, a hand-computed case), so that a plausible
mistake in the restatement (a dropped term, a wrong sign or index, swapped operands) fails here:

  c2p round trip              physical_quantities.rs:318-339 (4 ulp), inverse of the textbook p2c
  time_extrapolations         part.rs:134-149 vs the conservative Euler flux derivative (1D), rotation
                              and moving-frame invariance (3D)
  radius                      part.rs:439-447: length / disc / sphere of the cell's volume
  Particle::timestep          part.rs:42-84: CFL R / max(msv, |v_rel| + c_s), fluid velocity at the
                              generator of a linear velocity field
  timestep_limit              part.rs:263-297: brute force over bins and times against the particle's
                              time line (its step's beginning to its new end)
  AIRS switch                 airs.rs:18-37: bitwise PVRS below the pressure-ratio threshold, exact
                              at or above it
  face velocity               flux.rs:75-78: normal part = velocity of the moving bisector plane,
                              tangential part = mean generator velocity
  LimiterData alpha           gradients.rs:58-103 with the WLS estimate (hydrodynamics.rs:67-134) on a
                              uniform 1D mesh: alpha = min(1, d_max/e_max, d_min/e_min) in closed form
  pairwise limiter            gradients.rs:105-152: hand-computed clamps, monotone, idempotent, scale
                              and shift invariant
"""
import numpy as np
import pytest

GAMMAS = (1.4, 5.0 / 3.0)


def ulps(a, b):
    a, b = float(a), float(b)
    if a == b:
        return 0
    return abs(np.int64(np.float64(a).view(np.int64)) - np.int64(np.float64(b).view(np.int64)))


def p2c(prim, volume, gamma):
    """State<Conserved>::from_primitives written from its definition: m = rho V, p = m v,
    E = 1/2 p.v + m u with u = P / ((gamma - 1) rho)."""
    rho, v, P = prim[0], np.asarray(prim[1:4]), prim[4]
    m = rho * volume
    p = m * v
    E = 0.5 * float(p @ v) + m * (P / rho / (gamma - 1.0))
    return np.array([m, *p, E])


# ----------------------------------------------------------------------------------- c2p
def test_c2p_reference_round_trip(ora):
    """test_conversions (physical_quantities.rs:318-339): rho 0.75, v (0.4, 0, 0), P 0.8, V 0.1,
    gamma 5/3; float_cmp's default margin (4 ulp)."""
    prim = np.array([0.75, 0.4, 0.0, 0.0, 0.8])
    cons = p2c(prim, 0.1, 5.0 / 3.0)
    e, back = ora.c2p(5.0 / 3.0, cons, 0.1)
    assert e == 0
    for k in range(5):
        assert ulps(back[k], prim[k]) <= 4, (k, back[k], prim[k])


def test_c2p_round_trip_random(ora):
    rng = np.random.default_rng(7)
    worst = 0
    for _ in range(3000):
        g = GAMMAS[rng.integers(2)]
        prim = np.array([10 ** rng.uniform(-3, 3), *rng.uniform(-2, 2, 3), 10 ** rng.uniform(-3, 3)])
        vol = 10 ** rng.uniform(-6, 0)
        e, back = ora.c2p(g, p2c(prim, vol, g), vol)
        assert e == 0
        np.testing.assert_allclose(back[:4], prim[:4], rtol=1e-13, atol=1e-15)
        # pressure comes from E - kinetic: its relative error grows with the kinetic/thermal ratio
        kin = 0.5 * prim[0] * float(prim[1:4] @ prim[1:4])
        worst = max(worst, abs(back[4] - prim[4]) / (prim[4] + (g - 1.0) * kin))
    assert worst < 1e-13


def test_c2p_vacuum_and_errors(ora):
    e, p = ora.c2p(1.4, np.zeros(5), 0.5)
    assert e == 0 and np.all(p == 0.0)
    # negative mass / energy are hard errors (part.rs:219-231); an energy below the kinetic energy gives
    # a negative pressure (physical_quantities.rs:220-221)
    assert ora.c2p(1.4, [-1.0, 0, 0, 0, 1.0], 1.0)[0] == -2
    assert ora.c2p(1.4, [1.0, 0, 0, 0, -1.0], 1.0)[0] == -3
    assert ora.c2p(1.4, [1.0, 2.0, 0, 0, 1.0], 1.0)[0] == -4


# ----------------------------------------------------------------------------------- time extrapolations
def euler_flux_1d(U, gamma):
    rho, mom, E = U
    u = mom / rho
    P = (gamma - 1.0) * (E - 0.5 * rho * u * u)
    return np.array([mom, mom * u + P, u * (E + P)])


def prim_to_u_1d(W, gamma):
    rho, u, P = W
    return np.array([rho, rho * u, P / (gamma - 1.0) + 0.5 * rho * u * u])


def u_to_prim_1d(U, gamma):
    rho, mom, E = U
    u = mom / rho
    return np.array([rho, u, (gamma - 1.0) * (E - 0.5 * rho * u * u)])


def test_time_extrapolations_euler_1d(ora):
    """For a 1D state with gradients along x and a fixed particle (v_rel = v), -dt x rate must equal
    dt x dW/dt of the conservative Euler equations dU/dt = -dF(U)/dx, evaluated by central differences
    of F on the linear field W(x) = W0 + W' x and mapped to primitives by the Jacobian of u_to_prim
    (also by differences)."""
    rng = np.random.default_rng(11)
    for _ in range(50):
        g = GAMMAS[rng.integers(2)]
        W0 = np.array([10 ** rng.uniform(-1, 1), rng.uniform(-1, 1), 10 ** rng.uniform(-1, 1)])
        dW = rng.uniform(-1, 1, 3)
        h = 1e-5
        Up = prim_to_u_1d(W0 + h * dW, g)
        Um = prim_to_u_1d(W0 - h * dW, g)
        dUdt = -(euler_flux_1d(Up, g) - euler_flux_1d(Um, g)) / (2 * h)
        U0 = prim_to_u_1d(W0, g)
        eps = 1e-6
        dWdt = (u_to_prim_1d(U0 + eps * dUdt, g) - u_to_prim_1d(U0 - eps * dUdt, g)) / (2 * eps)
        prim = np.array([W0[0], W0[1], 0.0, 0.0, W0[2]])
        grad = np.zeros(15)
        grad[0], grad[3], grad[12] = dW[0], dW[1], dW[2]  # d rho/dx, d vx/dx, d P/dx
        dt = 0.01
        got = ora.time_extrapolations(g, prim, grad, [W0[1], 0.0, 0.0], dt)
        exp = dt * dWdt
        np.testing.assert_allclose(got[[0, 1, 4]], exp, rtol=1e-6, atol=1e-9 * np.max(np.abs(exp)))
        assert got[2] == 0.0 and got[3] == 0.0


def _rot(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def test_time_extrapolations_rotation_and_frame(ora):
    """Rotating the state, the gradients and v_rel rotates the velocity part of the rate and leaves the
    scalar parts unchanged; the vacuum state extrapolates by zero."""
    rng = np.random.default_rng(12)
    for _ in range(200):
        g = GAMMAS[rng.integers(2)]
        prim = np.array([10 ** rng.uniform(-1, 1), *rng.uniform(-1, 1, 3), 10 ** rng.uniform(-1, 1)])
        G = rng.uniform(-1, 1, (5, 3))  # rows: grad rho, grad vx, grad vy, grad vz, grad P
        vrel = rng.uniform(-1, 1, 3)
        R = _rot(rng)
        a = ora.time_extrapolations(g, prim, G.ravel(), vrel, 0.1)
        # rotated: v' = R v, grad s' = R grad s for scalars, velocity gradient J' = R J R^T (J[i] = grad v_i)
        J = G[1:4]
        Jr = R @ J @ R.T
        Gr = np.vstack([R @ G[0], Jr, R @ G[4]])
        pr = np.array([prim[0], *(R @ prim[1:4]), prim[4]])
        b = ora.time_extrapolations(g, pr, Gr.ravel(), R @ vrel, 0.1)
        np.testing.assert_allclose(b[[0, 4]], a[[0, 4]], rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(b[1:4], R @ a[1:4], rtol=1e-12, atol=1e-14)
    assert np.all(ora.time_extrapolations(1.4, np.zeros(5), np.ones(15), np.ones(3), 0.1) == 0.0)


def test_time_extrapolations_moving_particle(ora):
    """A particle moving with w sees dW/dt|_particle = dW/dt|_fixed + (w . grad) W for the scalar
    parts (density, pressure): the rate with v_rel = v - w differs from the rate with v_rel = v by
    -dt (-w . grad s)."""
    rng = np.random.default_rng(13)
    for _ in range(100):
        g = GAMMAS[rng.integers(2)]
        prim = np.array([10 ** rng.uniform(-1, 1), *rng.uniform(-1, 1, 3), 10 ** rng.uniform(-1, 1)])
        G = rng.uniform(-1, 1, (5, 3))
        w = rng.uniform(-1, 1, 3)
        dt = 0.05
        fixed = ora.time_extrapolations(g, prim, G.ravel(), prim[1:4], dt)
        moving = ora.time_extrapolations(g, prim, G.ravel(), prim[1:4] - w, dt)
        np.testing.assert_allclose(moving[0] - fixed[0], dt * (w @ G[0]), rtol=1e-9, atol=1e-13)
        np.testing.assert_allclose(moving[4] - fixed[4], dt * (w @ G[4]), rtol=1e-9, atol=1e-13)


# ----------------------------------------------------------------------------------- radius and timestep
def test_radius_closed_forms(ora):
    assert ora.radius(0.25, 1) == 0.125  # half the cell's length
    np.testing.assert_allclose(ora.radius(np.pi, 2), 1.0, rtol=1e-15)  # disc of area V
    np.testing.assert_allclose(ora.radius(4.0 / 3.0 * np.pi * 8.0, 3), 2.0, rtol=1e-15)  # sphere of volume V
    for V in (1e-6, 0.3, 7.0):
        np.testing.assert_allclose(np.pi * ora.radius(V, 2) ** 2, V, rtol=1e-14)
        np.testing.assert_allclose(4.0 / 3.0 * np.pi * ora.radius(V, 3) ** 3, V, rtol=1e-14)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_particle_timestep_closed_form(ora, dim):
    """Fixed particles: v = 0, v_rel = fluid velocity at the generator = p/m + J (loc - gc) for a linear
    velocity field with Jacobian J, sound speed sqrt(gamma P / rho), dt = CFL R / max(msv, |v_rel| + c);
    msv is reset to 0; vacuum (m = 0) gives infinity."""
    rng = np.random.default_rng(20 + dim)
    radius_of = {1: lambda V: V / 2.0, 2: lambda V: np.sqrt(V / np.pi), 3: lambda V: np.cbrt(3.0 * V / (4.0 * np.pi))}
    for it in range(300):
        g = GAMMAS[rng.integers(2)]
        rho, P = 10 ** rng.uniform(-2, 2, 2)
        v = rng.uniform(-2, 2, 3)
        V = 10 ** rng.uniform(-6, -1)
        prim = np.array([rho, *v, P])
        cons = p2c(prim, V, g)
        J = rng.uniform(-3, 3, (3, 3))
        grad = np.zeros((5, 3))
        grad[1:4] = J
        loc = rng.uniform(0, 1, 3)
        gc = loc + rng.uniform(-0.01, 0.01, 3)
        msv = 0.0 if it % 3 else 10 ** rng.uniform(-1, 2)
        dt, vp, vrel, msv_after, e = ora.particle_timestep(g, 0.3, ora.FIXED, dim, cons, prim, grad.ravel(), loc, gc,
                                                           loc, V, msv=msv)
        assert e == 0
        fluid_v = v + J @ (loc - gc)
        c = np.sqrt(g * P / rho)
        exp = 0.3 * radius_of[dim](V) / max(msv, np.linalg.norm(fluid_v) + c)
        np.testing.assert_allclose(vrel, fluid_v, rtol=1e-12, atol=1e-14)
        assert np.all(vp == 0.0) and msv_after == 0.0
        np.testing.assert_allclose(dt, exp, rtol=1e-12)
    dt, vp, vrel, _, e = ora.particle_timestep(1.4, 0.3, ora.FIXED, 3, np.zeros(5), np.zeros(5), np.zeros(15),
                                               np.zeros(3), np.zeros(3), np.zeros(3), 1.0)
    assert np.isinf(dt) and np.all(vp == 0.0)


def test_particle_timestep_fluid_motion(ora):
    """Fluid motion: the particle moves with the fluid velocity at its generator, so v_rel = 0 and the
    signal speed is the sound speed alone."""
    rng = np.random.default_rng(30)
    for _ in range(100):
        g = 1.4
        prim = np.array([10 ** rng.uniform(-1, 1), *rng.uniform(-1, 1, 3), 10 ** rng.uniform(-1, 1)])
        V = 1e-3
        loc = rng.uniform(0, 1, 3)
        dt, vp, vrel, _, e = ora.particle_timestep(g, 0.25, ora.FLUID, 3, p2c(prim, V, g), prim, np.zeros(15), loc,
                                                   loc, loc, V)
        np.testing.assert_allclose(vp, prim[1:4], rtol=1e-14)
        assert np.all(np.abs(vrel) <= 1e-15)
        np.testing.assert_allclose(dt, 0.25 * np.cbrt(3 * V / (4 * np.pi)) / np.sqrt(g * prim[4] / prim[0]), rtol=1e-12)


# ----------------------------------------------------------------------------------- timestep_limit
def _next_end(ti, dti):
    """First multiple of dti at or after ti, found by stepping along the time line."""
    k = 0
    while k * dti < ti:
        k += 1
    return k * dti


def test_timestep_limit_brute_force(ora):
    """Every bin pair and time on a short time line: a woken particle keeps its step's beginning and
    ends at the first end of its new bin after now (a full new step when its new bin starts now), so
    its dt becomes (new end - beginning) x time_base; a particle at or below the wakeup bin is left as
    it is; a particle starting now just takes the new bin's dt.  time_base is a power of two, so every
    product is exact and the comparison is exact."""
    tb = 2.0 ** -57
    checked = 0
    for b in range(1, 10):
        dti_b = 1 << (b + 1)
        for w in range(1, 10):
            dti_w = 1 << (w + 1)
            for ti in range(2, 1 << 11, 2):
                dt0 = dti_b * tb
                nb, dt = ora.timestep_limit(dt0, b, w, ti, tb)
                if w >= b:
                    assert (nb, dt) == (b, dt0)
                    continue
                assert nb == w
                if ti % dti_b == 0:  # starting now
                    assert dt == dti_w * tb
                else:
                    end_old = _next_end(ti, dti_b)
                    begin = end_old - dti_b
                    end_new = ti + dti_w if ti % dti_w == 0 else _next_end(ti, dti_w)
                    assert dt == (end_new - begin) * tb, (b, w, ti)
                checked += 1
    assert checked > 10000


# ----------------------------------------------------------------------------------- AIRS
def test_airs_switch_bitwise(ora):
    """AIRS is PVRS where p_max / p_min < threshold and the exact solver at or above it, bit for bit."""
    rng = np.random.default_rng(40)
    n_pv = n_ex = 0
    for it in range(2000):
        thr = (2.0, 1.5, 10.0)[it % 3]
        pl = 10 ** rng.uniform(-1, 1)
        ratio = 10 ** rng.uniform(0, np.log10(4 * thr))
        if it % 50 == 0:
            ratio = thr  # exactly at the threshold: exact
        pr = pl * ratio if rng.random() < 0.5 else pl / ratio
        wl = np.array([10 ** rng.uniform(-1, 1), *rng.uniform(-0.5, 0.5, 3), pl])
        wr = np.array([10 ** rng.uniform(-1, 1), *rng.uniform(-0.5, 0.5, 3), pr])
        vf = rng.uniform(-0.2, 0.2, 3)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        a = ora.solve_for_flux(ora.AIRS, 1.4, wl, wr, vf, n, threshold=thr)
        below = max(pl, pr) / min(pl, pr) < thr
        ref = ora.solve_for_flux(ora.PVRS if below else ora.EXACT, 1.4, wl, wr, vf, n)
        assert np.array_equal(a, ref), (it, pl, pr, thr)
        n_pv += below
        n_ex += not below
    assert n_pv > 300 and n_ex > 300


# ----------------------------------------------------------------------------------- face velocity
def _signed_distance_to_bisector(f, xl, xr):
    n = (xr - xl) / np.linalg.norm(xr - xl)
    return (f - 0.5 * (xl + xr)) @ n


def test_face_velocity_moves_with_the_bisector(ora):
    """The face point f on the bisector of generators xl, xr moving with vl, vr: the normal part of
    the face velocity is the rate at which the bisector plane moves past f (numerical derivative of the
    signed distance), the tangential part the mean generator velocity; a periodic shift of the right
    generator changes nothing."""
    rng = np.random.default_rng(50)
    for _ in range(300):
        xl = rng.uniform(0, 1, 3)
        xr = xl + rng.uniform(-0.2, 0.2, 3)
        vl, vr = rng.uniform(-1, 1, (2, 3))
        n = (xr - xl) / np.linalg.norm(xr - xl)
        t1 = np.cross(n, rng.normal(size=3))
        f = 0.5 * (xl + xr) + rng.uniform(-0.05, 0.05) * t1  # a point of the face
        w = ora.face_velocity(xl, xr, np.zeros(3), vl, vr, f)
        h = 1e-6
        dsd = (_signed_distance_to_bisector(f, xl + h * vl, xr + h * vr)
               - _signed_distance_to_bisector(f, xl - h * vl, xr - h * vr)) / (2 * h)
        # the plane moves past f at -dsd along n
        np.testing.assert_allclose(w @ n, -dsd, rtol=1e-6, atol=1e-9)
        mean = 0.5 * (vl + vr)
        np.testing.assert_allclose(w - (w @ n) * n, mean - (mean @ n) * n, rtol=1e-10, atol=1e-12)
        shift = np.array([1.0, 0.0, -1.0])
        ws = ora.face_velocity(xl, xr - shift, shift, vl, vr, f)
        np.testing.assert_allclose(ws, w, rtol=1e-9, atol=1e-12)


# ----------------------------------------------------------------------------------- LimiterData alpha
def _uniform_1d_sim(ora, n):
    x = (np.arange(n) + 0.5) / n
    mesh = ora.voronoi_1d(x, 1.0, periodic=False)
    p = ora.make_params(1, ora.REFLECTIVE, (1.0, 1.0, 1.0), 1.4, 0.3, ora.HLLC, motion=ora.FIXED)
    loc = np.zeros((n, 3))
    loc[:, 0] = x
    vol = mesh["volume"]
    return ora.Sim(p, loc.ravel(), vol.copy(), np.zeros(3 * n), np.ones(n), mesh)


def _alpha(W, Wm, Wp, e):
    d_max = max(W, Wm, Wp) - W
    d_min = min(W, Wm, Wp) - W
    if e == 0.0:
        return 1.0
    return min(1.0, d_max / e, d_min / -e)


def test_limiter_alpha_uniform_1d(ora):
    """On a uniform 1D mesh of dyadic width dx, the WLS gradient of cell i is (W[i+1] - W[i-1]) / (2 dx)
    (a wall cell's mirror image carries its own density and pressure), the face extrapolations are
    +-g dx / 2, and LimiterData scales g by alpha = min(1, d_max/e_max, d_min/e_min), with alpha = 1
    when the extrapolations vanish.  Cases: linear (alpha 1), a local extremum (alpha 0), one-sided
    bounds (alpha = d_min/e_min or d_max/e_max), flat (zero gradient)."""
    n = 16
    dx = 1.0 / n
    sim = _uniform_1d_sim(ora, n)
    rho = np.array([1.0, 1.0, 1.0, 1.0, 1.5, 2.0, 2.5, 3.0, 2.0, 2.1, 3.0, 3.05, 3.1, 4.0, 1.0, 1.2])
    P = np.linspace(1.0, 2.5, n) ** 2
    f = sim.get_fields()
    f["prim"] = np.concatenate([rho, np.zeros(3 * n), P])
    sim.set_fields({"prim": f["prim"]})
    sim.gradients(np.ones(n, np.uint8))
    g = sim.get_fields()["grad"].reshape(15, n)
    kinds = set()
    for q, W, row in ((0, rho, 0), (4, P, 12)):
        for i in range(n):
            Wm = W[i - 1] if i > 0 else W[i]
            Wp = W[i + 1] if i < n - 1 else W[i]
            g0 = (Wp - Wm) / (2 * dx)
            a = _alpha(W[i], Wm, Wp, abs(g0) * dx / 2)
            kinds.add("one" if a == 1.0 else "zero" if a == 0.0 else "partial")
            np.testing.assert_allclose(g[row, i], a * g0, rtol=1e-12, atol=1e-12, err_msg=f"q={q} i={i}")
            assert g[row + 1, i] == 0.0 and g[row + 2, i] == 0.0
    assert kinds == {"one", "zero", "partial"}
    # one-sided bound, by hand: W = (0, 0.1, 1) around cell 9 gives g0 = 0.9 n / 2 ... alpha = 0.2 / 0.9
    rho2 = np.ones(n)
    rho2[8], rho2[9], rho2[10] = 1.0, 1.1, 2.0
    # a peak between equal neighbours: zero estimate, zero extrapolations, alpha 1 (not 0/0 or -1/0)
    rho2[3], rho2[4], rho2[5] = 1.0, 2.0, 1.0
    sim.set_fields({"prim": np.concatenate([rho2, np.zeros(3 * n), np.ones(n)])})
    sim.gradients(np.ones(n, np.uint8))
    g = sim.get_fields()["grad"].reshape(15, n)
    g0 = (2.0 - 1.0) / (2 * dx)
    np.testing.assert_allclose(g[0, 9], (0.1 / (g0 * dx / 2)) * g0, rtol=1e-12)  # d_min / e_min = 0.2
    assert np.all(g[:, 4] == 0.0)


# ----------------------------------------------------------------------------------- pairwise limiter
def _lim(ora, ql, qr, qd, fac):
    out = ora.pairwise_limiter([ql, 0, 0, 0, 1.0], [qr, 0, 0, 0, 1.0], [qd, 0, 0, 0, 1.0], fac)
    return out[0]


def test_pairwise_limiter_known_answers(ora):
    """Hand-computed clamps with psi1 = 1/2, psi2 = 1/4 (band q_bar -+ |dq|/4, floor/ceiling one half
    |dq| beyond the pair, positivity-preserving floor q^2/(q + |dq|/2) when that would cross zero)."""
    cases = [
        # q_l < q_r = 1, 2, face at the midpoint: clamp to [0.5, 1.75]
        ((1.0, 2.0, 0.5), [(0.0, 0.5), (0.4, 0.5), (1.0, 1.0), (1.75, 1.75), (2.0, 1.75), (9.0, 1.75)]),
        # q_l > q_r = 2, 1: clamp to [1.25, 2.5]
        ((2.0, 1.0, 0.5), [(0.0, 1.25), (1.5, 1.5), (2.4, 2.4), (3.0, 2.5)]),
        # face at a quarter: q_bar = 1.25, band top 1.5
        ((1.0, 2.0, 0.25), [(3.0, 1.5), (1.2, 1.2)]),
        # positivity: q_l, q_r = 0.1, 1: floor 0.01 / 0.55, top 0.55 + 0.225
        ((0.1, 1.0, 0.5), [(-5.0, 0.01 / 0.55), (0.3, 0.3), (5.0, 0.775)]),
    ]
    for (ql, qr, fac), pts in cases:
        for qd, exp in pts:
            np.testing.assert_allclose(_lim(ora, ql, qr, qd, fac), exp, rtol=1e-15, err_msg=str((ql, qr, fac, qd)))
    assert _lim(ora, 0.7, 0.7, 5.0, 0.5) == 0.7  # equal pair: no reconstruction


def test_pairwise_limiter_properties(ora):
    """Monotone and idempotent in q_dash (a clamp), homogeneous under power-of-two scaling, and away
    from the positivity floor invariant under a common shift."""
    rng = np.random.default_rng(3)
    for _ in range(400):
        ql, qr = rng.uniform(0.01, 2.0, 2)
        fac = rng.uniform(0.0, 1.0)
        qd = np.sort(rng.uniform(-1.0, 3.0, 12))
        out = np.array([_lim(ora, ql, qr, q, fac) for q in qd])
        assert np.all(np.diff(out) >= 0.0)
        for q in out:
            if q > 0.0:
                assert _lim(ora, ql, qr, q, fac) == q
        for q in qd[:4]:
            assert _lim(ora, 4 * ql, 4 * qr, 4 * q, fac) == 4 * _lim(ora, ql, qr, q, fac)
        c = 10.0  # far from zero: the floor is q_min - |dq|/2 > 0 on both sides
        for q in qd[:4]:
            np.testing.assert_allclose(_lim(ora, ql + c, qr + c, q + c, fac), _lim(ora, ql, qr, q, fac) + c
                                       if (min(ql, qr) - 0.5 * abs(ql - qr)) > 0 else _lim(ora, ql + c, qr + c, q + c, fac),
                                       rtol=0, atol=1e-13)


# ----------------------------------------------------------------------------------- 1D periodic rotation
def test_voronoi_1d_periodic_rotation(ora):
    """A periodic line whose generators are a rotation of the sorted order (generators that drifted through
    the box edge) has the same cells as the sorted line: every generator keeps its volume and its two
    neighbours, and the one face where the order steps down carries the periodic shift.  Known answer:
    x = (0.8, 0.1, 0.5) on [0, 1) gives the cells [0.65, 0.95], [-0.05, 0.3], [0.3, 0.65]."""
    m = ora.voronoi_1d(np.array([0.8, 0.1, 0.5]), 1.0, periodic=True)
    np.testing.assert_allclose(m["volume"], [0.3, 0.35, 0.35], rtol=1e-15)
    np.testing.assert_allclose(m["centroid"].reshape(-1, 3)[:, 0], [0.65, 0.95, 0.3], rtol=1e-15)
    assert list(m["shift"].reshape(-1, 3)[:, 0]) == [0.0, 1.0, 0.0]
    rng = np.random.default_rng(60)
    for _ in range(50):
        n = int(rng.integers(3, 40))
        x = np.sort(rng.uniform(0.0, 1.0, n))
        base = ora.voronoi_1d(x, 1.0, periodic=True)
        r = int(rng.integers(1, n))
        rot = ora.voronoi_1d(np.roll(x, -r), 1.0, periodic=True)
        np.testing.assert_allclose(np.roll(base["volume"], -r), rot["volume"], rtol=0, atol=1e-15)
        assert abs(rot["volume"].sum() - 1.0) < 1e-12
        assert int(np.count_nonzero(rot["shift"])) == 1
    with pytest.raises(RuntimeError):  # not a rotation: two downward steps
        ora.voronoi_1d(np.array([0.5, 0.1, 0.8, 0.2]), 1.0, periodic=True)


# ----------------------------------------------------------------------------------- gravity
def _gravity_sim(ora, n, gravity, soft, pos=(0.0, 0.0, 0.0), acc=(0.0, 0.0, 0.0), masses=None):
    x = (np.arange(n) + 0.5) / n
    mesh = ora.voronoi_1d(x, 1.0, periodic=True)
    p = ora.make_params(1, ora.PERIODIC, (1.0, 1.0, 1.0), 1.4, 0.3, ora.HLLC, gravity=gravity, grav_softening=soft,
                        grav_position=pos, grav_acceleration=acc)
    loc = np.zeros((n, 3))
    loc[:, 0] = x
    m = mesh["volume"].copy() if masses is None else np.asarray(masses, float)
    return ora.Sim(p, loc.ravel(), m, np.zeros(3 * n), np.ones(n), mesh), x, m


def test_self_gravity_direct_sum(ora):
    """SelfGravity::compute_accelerations (gravity.rs:24-44): particle i feels m_j (x_j - x_i) / (|x_j - x_i| +
    eps)^3 from the particles after it (j > i) only; with eps = 0 that is Newton's m/r^2 toward each of them."""
    sim, x, m = _gravity_sim(ora, 3, ora.GRAV_SELF, 0.0, masses=[1.0, 2.0, 4.0])
    a = sim.gravity_accelerations().reshape(3, 3)
    exp0 = 2.0 / (x[1] - x[0]) ** 2 + 4.0 / (x[2] - x[0]) ** 2
    np.testing.assert_allclose(a[0], [exp0, 4.0 / (x[2] - x[1]) ** 2, 0.0], rtol=1e-14)
    assert np.all(a[1:] == 0.0)
    sim, x, m = _gravity_sim(ora, 3, ora.GRAV_SELF, 0.1, masses=[1.0, 2.0, 4.0])
    a = sim.gravity_accelerations().reshape(3, 3)
    np.testing.assert_allclose(a[0, 1], 4.0 * (x[2] - x[1]) / (x[2] - x[1] + 0.1) ** 3, rtol=1e-14)
    assert np.all(np.isinf(sim.gravity_timesteps()))  # a_grav = 0 on the path: no gravity timestep


def test_keplerian_potential(ora):
    """KeplerianPotential (gravity.rs:101-148): a = -dx / |dx|^3 outside the softening length, -dx / (r (r^2 +
    eps^2)) inside; timestep 0.5 R_2D / v_circ with v_circ = sqrt(1 / r) (r + eps inside), R_2D = sqrt(V/pi)."""
    n, eps, c = 8, 0.2, (0.3, 0.0, 0.0)
    sim, x, m = _gravity_sim(ora, n, ora.GRAV_KEPLERIAN, eps, pos=c)
    a = sim.gravity_accelerations().reshape(3, n)
    dt = sim.gravity_timesteps()
    V = 1.0 / n
    for i in range(n):
        dx = x[i] - c[0]
        r = abs(dx)
        exp_a = -dx / r ** 3 if r >= eps else -dx / (r * (r * r + eps * eps))
        np.testing.assert_allclose(a[0, i], exp_a, rtol=1e-14)
        rs = r if r >= eps else r + eps
        np.testing.assert_allclose(dt[i], 0.5 * np.sqrt(V / np.pi) * np.sqrt(rs), rtol=1e-14)
    assert np.all(a[1:] == 0.0)


def test_constant_potential(ora):
    sim, x, m = _gravity_sim(ora, 4, ora.GRAV_CONSTANT, 0.0, acc=(0.0, -1.5, 0.25))
    a = sim.gravity_accelerations().reshape(3, 4)
    assert np.all(a[0] == 0.0) and np.all(a[1] == -1.5) and np.all(a[2] == 0.25)
    assert np.all(np.isinf(sim.gravity_timesteps()))
