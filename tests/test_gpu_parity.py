"""Parity of the HIP path with the CPU oracle, through the C-ABI.

FP64 path: per-face flux records and every field after one step within 1e-12 (norm-relative), integer
timebins and ti_next bit-exact; whole runs within 1e-9 (pow in the device libm may differ from glibc in
the last bit, and the step sequence compounds it).  FP32-flux path: L1 density error < 1e-5 vs FP64.
"""
import sys

import numpy as np
import pytest

from hydro_rs_amd import fvs, ics
from tests.helpers import compare_states, face_flux_rel, oracle_sim, rel_err, synced as _synced

pytestmark = pytest.mark.gpu

TOL_STEP = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _device():
    if fvs.device_count() < 1:
        pytest.skip("no HIP device")


CASES = {
    "sod1d": lambda: ics.sod_1d(n=200, t_end=0.2),
    # the other particle-motion modes of set_particle_velocity (part.rs:328-372), on a moving 1D mesh
    "sod1d_steer": lambda: ics.sod_1d(n=200, t_end=0.2, motion=fvs.STEER),
    "sod1d_steer_pakmor": lambda: ics.sod_1d(n=200, t_end=0.2, motion=fvs.STEER_PAKMOR),
    # engine and runner options: GodunovHydroRunner (godunov.rs:14-58, no gradient phases), timestep_sync
    # instead of the limiter (hydrodynamics.rs:186-191), snapshot sync points splitting the drifts
    # (engine.rs), and the isothermal equation of state (gas_law.rs)
    "sod1d_godunov": lambda: _with(ics.sod_1d(n=200, t_end=0.2), runner=fvs.GODUNOV),
    "sod1d_sync_all": lambda: _with(ics.sod_1d(n=200, t_end=0.2), sync_all=True),
    "sod1d_dt_snap": lambda: _with(ics.sod_1d(n=200, t_end=0.2), dt_snap=0.013),
    "sedov2d_sync_all": lambda: _with(ics.sedov_2d(side=32, t_end=0.05), sync_all=True),
    "tube1d_isothermal": lambda: _with(ics.riemann_tube_1d(300, 1.0, 1.0, 0.0, 0.125, 0.1, 0.0, 1.4, t_end=0.05),
                                       eos=1, isothermal_u=1.5),
    # tiny meshes: one partly filled chunk, launch grids of padding blocks
    "tube1d_n3": lambda: ics.riemann_tube_1d(3, 1.0, 1.0, 0.0, 0.125, 0.1, 0.0, 1.4, t_end=0.5),
    "sedov2d_side4": lambda: ics.sedov_2d(side=4, t_end=0.5),
    "kh3d_side3": lambda: ics.kelvin_helmholtz_3d(side=3, t_end=1.0),
    "sedov2d": lambda: ics.sedov_2d(side=32, t_end=0.05),
    "kh3d": lambda: ics.kelvin_helmholtz_3d(side=10, t_end=0.5),
    "tube2d_reflective_exact": lambda: _tube2d(),
    # the other wall kinds (flux.rs:148-169, space.rs:213-229), and walls on a 3D fixed mesh (precomputed
    # gradient geometry with LDS-staged neighbours)
    "tube2d_open_hllc": lambda: _tube2d(fvs.OPEN, fvs.HLLC),
    "tube2d_vacuum_hllc": lambda: _tube2d(fvs.VACUUM, fvs.HLLC),
    "blast3d_reflective": lambda: _blast3d(),
    # pressure ratio 100 at the interface: AIRS hands those faces to the exact-solver fallback pass
    "tube1d_airs": lambda: ics.riemann_tube_1d(300, 1.0, 10.0, 0.0, 0.125, 0.1, 0.0, 1.4, solver=fvs.AIRS,
                                               t_end=0.05),
    # the remaining solvers of the trait (tsrs.rs, trrs.rs, linear_advection.rs)
    "tube1d_tsrs": lambda: ics.riemann_tube_1d(300, 1.0, 1.0, 0.0, 0.125, 0.1, 0.0, 1.4, solver=fvs.TSRS, t_end=0.05),
    "tube1d_trrs": lambda: ics.riemann_tube_1d(300, 1.0, 1.0, 0.0, 0.125, 0.1, 0.0, 1.4, solver=fvs.TRRS, t_end=0.05),
    "tube1d_linadv": lambda: _with(ics.riemann_tube_1d(300, 1.0, 1.0, 0.0, 0.125, 0.1, 0.0, 1.4, t_end=0.05),
                                   solver=fvs.LINEAR_ADVECTION, linadv_velocity=(0.7, 0.0, 0.0)),
}


def wall_crossing_tube(precision=fvs.FP64):
    """1D moving mesh (fluid motion) with reflective walls whose outermost generators sit 0.05 dx from the
    walls and move into them at |v| = 0.5: they cross the walls at steps 6, 8, 10 and 12 and are
    reflected (Particle::box_reflect, part.rs:457-469); the reference raises a negative pressure at step
    13."""
    n = 40
    dx = 1.0 / n
    x = (np.arange(n) + 0.5) * dx
    x[0], x[-1] = 0.05 * dx, 1.0 - 0.05 * dx
    loc = np.zeros((n, 3))
    loc[:, 0] = x
    vel = np.zeros((n, 3))
    vel[:, 0] = np.where(x < 0.5, -0.5, 0.5)
    cfg = fvs.Config(dim=1, box=(1.0, 1.0, 1.0), gamma=1.4, cfl=0.3, solver=fvs.EXACT, boundary=fvs.REFLECTIVE,
                     motion=fvs.FLUID, t_end=0.05, precision=precision)
    return ics._finish("wall_crossing", cfg, loc, np.ones(n), vel, np.ones(n))


def periodic_wrap_tube(precision=fvs.FP64):
    """1D periodic moving mesh (fluid motion): a density wave advected at v = 1 for t = 0.3, so that 30 of
    the 100 generators drift through the box edge (box_wrap, part.rs:449-455) and the line's cyclic order
    is rotated; the 1D rebuild follows the rotation (the face where the order steps down carries the
    periodic shift)."""
    n = 100
    rng = np.random.default_rng(5)
    x = (np.arange(n) + 0.5) / n + 0.01 / n * rng.uniform(-1.0, 1.0, n)
    loc = np.zeros((n, 3))
    loc[:, 0] = x
    vel = np.zeros((n, 3))
    vel[:, 0] = 1.0
    cfg = fvs.Config(dim=1, box=(1.0, 1.0, 1.0), gamma=1.4, cfl=0.3, solver=fvs.HLLC, boundary=fvs.PERIODIC,
                     motion=fvs.FLUID, t_end=0.3, precision=precision)
    return ics._finish("periodic_wrap", cfg, loc, 1.0 + 0.5 * np.sin(2.0 * np.pi * x), vel, np.ones(n))


def _with(prob, **cfg):
    """The same problem with some configuration fields replaced."""
    from dataclasses import replace

    return replace(prob, cfg=replace(prob.cfg, **cfg))


def _tube2d(boundary=fvs.REFLECTIVE, solver=fvs.EXACT):
    rng = np.random.default_rng(11)
    loc = ics.jittered_grid(2, 24, (1.0, 1.0, 1.0), 0.05, rng)
    left = loc[:, 0] < 0.5
    rho = np.where(left, 1.0, 0.25)
    P = np.where(left, 1.0, 0.2)
    vel = np.zeros_like(loc)
    vel[:, 1] = 0.1
    cfg = fvs.Config(dim=2, box=(1.0, 1.0, 1.0), gamma=1.4, cfl=0.3, solver=solver, boundary=boundary,
                     motion=fvs.FIXED, t_end=0.1)
    return ics._finish("tube2d", cfg, loc, rho, vel, P)


def _blast3d(boundary=fvs.REFLECTIVE, solver=fvs.HLLC):
    """8^3 jittered cells in the unit box with walls: an overpressured sphere drifting along x."""
    rng = np.random.default_rng(12)
    loc = ics.jittered_grid(3, 8, (1.0, 1.0, 1.0), 0.1, rng)
    r = np.linalg.norm(loc - 0.5, axis=1)
    rho = np.ones(len(loc))
    P = np.where(r < 0.25, 10.0, 0.1)
    vel = np.zeros_like(loc)
    vel[:, 0] = 0.2
    cfg = fvs.Config(dim=3, box=(1.0, 1.0, 1.0), gamma=5.0 / 3.0, cfl=0.3, solver=solver, boundary=boundary,
                     motion=fvs.FIXED, t_end=0.05)
    return ics._finish("blast3d", cfg, loc, rho, vel, P)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("k", [0, 7])
def test_face_flux_parity(ora, case, k):
    say(f"[face-flux] case={case} k={k}")
    prob = CASES[case]()
    osim, dev, inf = _synced(prob, k)
    ti_old, ti_cur = inf["ti_old"], inf["ti_next"]
    tb = inf["time_base"]
    osim.set_engine_info(ti_old, ti_cur, 0)
    mab = ora.get_max_active_bin(ti_cur)
    f = osim.get_fields()
    active = (f["timebin"] <= mab).astype(np.uint8)
    ref = osim.compute_fluxes(active)
    info = dev.timestep_info(ti_old, ti_cur, tb)
    dev.compute_fluxes(info)
    got = dev.get_fluxes()
    err = face_flux_rel(got, ref)
    say(f"  faces={len(ref)} active_cells={int(active.sum())} worst face rel err={err:.3e}")
    assert err <= TOL_STEP


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("k", [0, 1, 9])
def test_step_parity(ora, case, k):
    say(f"[step] case={case} k={k}")
    prob = CASES[case]()
    osim, dev, inf = _synced(prob, k)
    ti_old, ti_cur = inf["ti_old"], inf["ti_next"]
    tb = inf["time_base"]
    osim.set_engine_info(ti_old, ti_cur, 0)
    t_ref = osim.runner_step()
    t_dev = dev.step(dev.timestep_info(ti_old, ti_cur, tb))
    ref = osim.get_fields()
    got = dev.get_state()
    errs = compare_states(got, ref, prob.n)
    say(f"  ti_next ref={t_ref} dev={t_dev} errs={ {k2: f'{v:.1e}' for k2, v in errs.items()} }")
    assert t_dev == t_ref
    assert errs.pop("timebin_mismatch") == 0
    assert max(errs.values()) <= TOL_STEP, errs
    st = dev.stats()
    assert st["error_code"] == 0


@pytest.mark.parametrize("case,steps", [("sod1d", 1 << 40), ("sod1d_steer", 1 << 40), ("sod1d_steer_pakmor", 1 << 40),
                                        ("sod1d_godunov", 1 << 40), ("sod1d_sync_all", 1 << 40),
                                        ("sod1d_dt_snap", 1 << 40), ("sedov2d_sync_all", 400),
                                        ("tube1d_isothermal", 1 << 40), ("tube1d_n3", 1 << 40),
                                        ("sedov2d_side4", 1 << 40), ("kh3d_side3", 1 << 40),
                                        ("sedov2d", 400), ("tube2d_reflective_exact", 1 << 40),
                                        ("tube2d_vacuum_hllc", 1 << 40), ("blast3d_reflective", 1 << 40),
                                        ("tube1d_tsrs", 1 << 40), ("tube1d_trrs", 1 << 40),
                                        ("tube1d_linadv", 1 << 40)])
def test_full_run_parity(ora, case, steps):
    say(f"[run] case={case}")
    prob = CASES[case]()
    osim = oracle_sim(prob)
    n_ref = osim.advance(steps)
    dev = ics.build_device_sim(prob)
    n_dev = dev.run(steps)
    ref = osim.get_fields()
    got = dev.get_state()
    errs = compare_states(got, ref, prob.n)
    say(f"  steps ref={n_ref} dev={n_dev} errs={ {k2: f'{v:.1e}' for k2, v in errs.items()} }")
    assert n_dev == n_ref
    assert errs.pop("timebin_mismatch") == 0
    assert errs["prim"] <= 1e-9 and errs["cons"] <= 1e-9, errs
    ei = dev.engine_info()
    oi = osim.info()
    assert ei["ti_current"] == oi["ti_current"]


@pytest.mark.parametrize("case", list(CASES))
def test_static_geometry_identical(case):
    """Precomputed geometry (fixed meshes) gives bit-identical states to recomputing it every step."""
    say(f"[geometry] case={case}")
    prob = CASES[case]()
    a = ics.build_device_sim(prob)
    b = ics.build_device_sim(prob)
    b.set_geometry_mode(0)
    assert not b.geometry_static()
    assert a.geometry_static() == (prob.cfg.motion == fvs.FIXED)
    na, nb = a.run(60), b.run(60)
    assert na == nb
    sa, sb = a.get_state(), b.get_state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    say(f"  {na} steps identical")


def test_conservation_on_device():
    say("[conservation] sedov2d side=48")
    prob = ics.sedov_2d(side=48, t_end=0.05)
    dev = ics.build_device_sim(prob)
    c0 = dev.conservation()
    dev.run(60)
    c1 = dev.conservation()
    assert abs(c1[0] - c0[0]) <= 1e-13 * abs(c0[0])
    assert abs(c1[4] - c0[4]) <= 1e-12 * abs(c0[4])
    assert np.all(np.abs(c1[1:4] - c0[1:4]) <= 1e-13)


@pytest.mark.parametrize("solver", ["Exact", "HLLC", "PVRS", "AIRS", "TSRS", "TRRS", "LinearAdvection"])
def test_riemann_batch_parity(ora, solver):
    say(f"[riemann] solver={solver}")
    rng = np.random.default_rng(hash(solver) % 1000)
    n = 4000
    WL = np.column_stack([10 ** rng.uniform(-1, 1, n), rng.uniform(-1, 1, (n, 3)), 10 ** rng.uniform(-2, 1, n)])
    WR = np.column_stack([10 ** rng.uniform(-1, 1, n), rng.uniform(-1, 1, (n, 3)), 10 ** rng.uniform(-2, 1, n)])
    # some vacuum and vacuum-generating pairs
    WL[:40] = 0.0
    WR[40:80] = 0.0
    WL[80:120, 1] = -8.0
    WR[80:120, 1] = 8.0
    vf = rng.uniform(-0.5, 0.5, (n, 3))
    nr = rng.normal(size=(n, 3))
    nr /= np.linalg.norm(nr, axis=1, keepdims=True)
    cfg = fvs.Config(dim=3, box=(1, 1, 1), gamma=5.0 / 3.0, cfl=0.3, solver=fvs.SOLVERS[solver],
                     linadv_velocity=(0.3, -0.2, 0.1))
    got = fvs.riemann_batch(cfg, WL, WR, vf, nr)
    ref = np.array([ora.solve_for_flux(fvs.SOLVERS[solver], 5.0 / 3.0, WL[i], WR[i], vf[i], nr[i],
                                       threshold=2.0, velocity=(0.3, -0.2, 0.1)) for i in range(n)])
    finite = np.all(np.isfinite(ref), axis=1)
    assert np.array_equal(finite, np.all(np.isfinite(got), axis=1))
    scale = np.maximum(np.max(np.abs(ref[finite]), axis=1), 1e-300)
    err = np.max(np.abs(got[finite] - ref[finite]), axis=1) / scale
    say(f"  worst rel err={err.max():.3e}, exact-equal fraction={np.mean(got[finite] == ref[finite]):.4f}")
    assert err.max() <= TOL_STEP


def test_fp32_flux_l1_density():
    """FP32 reconstruction/solve with FP64 accumulation stays within L1(rho) < 1e-5 of the FP64 path."""
    say("[fp32] sedov2d side=64, 200 steps")
    p64 = ics.sedov_2d(side=64, t_end=0.05)
    p32 = ics.sedov_2d(side=64, t_end=0.05, precision=fvs.FP32_FLUX)
    d64 = ics.build_device_sim(p64)
    d32 = ics.build_device_sim(p32)
    s64 = d64.run(200)
    s32 = d32.run(200)
    r64 = d64.get_state(["prim"])["prim"][: p64.n]
    r32 = d32.get_state(["prim"])["prim"][: p64.n]
    l1 = float(np.mean(np.abs(r32 - r64)))
    say(f"  steps {s64}/{s32}, L1(rho) = {l1:.3e}")
    assert l1 < 1e-5


FP32_CASES = {
    **{f"tube1d_{k}": (lambda k=k: ics.riemann_tube_1d(300, 1.0, 1.0, 0.0, 0.125, 0.1, 0.0, 1.4, t_end=0.05,
                                                      solver=fvs.SOLVERS[k]))
       for k in ("Exact", "HLLC", "PVRS", "AIRS", "TSRS", "TRRS")},
    "tube1d_LinearAdvection": lambda: _with(ics.riemann_tube_1d(300, 1.0, 1.0, 0.0, 0.125, 0.1, 0.0, 1.4, t_end=0.05),
                                            solver=fvs.LINEAR_ADVECTION, linadv_velocity=(0.7, 0.0, 0.0)),
    "kh3d_airs": lambda: ics.kelvin_helmholtz_3d(side=10, t_end=0.3),
    "tube2d_reflective_hllc": lambda: _tube2d(fvs.REFLECTIVE, fvs.HLLC),
}


@pytest.mark.parametrize("case", list(FP32_CASES))
def test_fp32_flux_l1_per_solver(case):
    """Every solver's FP32 instantiation (reconstruction and solve in FP32, FP64 accumulation), run to the
    end time next to the FP64 path: L1(rho) < 1e-5 at the same time."""
    say(f"[fp32] case={case}")
    p64 = FP32_CASES[case]()
    p32 = _with(p64, precision=fvs.FP32_FLUX)
    d64, d32 = ics.build_device_sim(p64), ics.build_device_sim(p32)
    s64, s32 = d64.run(1 << 40), d32.run(1 << 40)
    e64, e32 = d64.engine_info(), d32.engine_info()
    assert e64["done"] and e32["done"] and e64["error_code"] == 0 and e32["error_code"] == 0
    r64 = d64.get_state(["prim"])["prim"][: p64.n]
    r32 = d32.get_state(["prim"])["prim"][: p64.n]
    l1 = float(np.mean(np.abs(r32 - r64)))
    say(f"  steps {s64}/{s32}, L1(rho) = {l1:.3e}")
    assert l1 < 1e-5


def test_trait_phases_equal_fused_step():
    """predict/compute_fluxes/apply/c2p/gradients/timesteps through the trait entry points produce the
    fused hfv_step result bit for bit."""
    say("[phases] sedov2d side=24")
    prob = ics.sedov_2d(side=24, t_end=0.05)
    a = ics.build_device_sim(prob)
    b = ics.build_device_sim(prob)
    tb = prob.cfg.t_end / float(1 << 57)
    ti_old, ti_cur = 0, 0
    for _ in range(5):
        info = a.timestep_info(ti_old, ti_cur, tb)
        t1 = a.step(info)
        b.predict((ti_cur - ti_old) * tb)
        b.compute_fluxes(info)
        b.apply_fluxes()
        b.convert_conserved_to_primitive()
        b.gradients()
        t2 = b.timesteps()
        assert t1 == t2
        sa, sb = a.get_state(), b.get_state()
        for k in sa:
            assert np.array_equal(sa[k], sb[k]), k
        ti_old, ti_cur = ti_cur, t1
    assert rel_err(sa["prim"], sb["prim"]) == 0.0


def test_snapshot_matches_oracle(ora, tmp_path):
    """Space::dump of the device state equals the dump of the oracle's state after the same steps (the
    whole-run tolerance: the step sequence compounds last-bit differences of pow)."""
    from hydro_rs_amd import snapshot

    prob = ics.sedov_2d(side=24, t_end=0.01)
    say("[snapshot] sedov2d side=24, 40 steps")
    osim = oracle_sim(prob)
    n_ref = osim.advance(40)
    dev = ics.build_device_sim(prob)
    n_dev = dev.run(40)
    assert n_dev == n_ref
    a = snapshot.write_snapshot(tmp_path / "dev.npz", dev, prob.cfg, faces=prob.mesh)
    oi = osim.info()
    b = snapshot.snapshot_arrays(osim.get_fields(), prob.n, prob.cfg, oi["ti_current"] * oi["time_base"],
                                 faces=prob.mesh)
    assert set(a) == set(b)
    with np.load(tmp_path / "dev.npz", allow_pickle=False) as z:
        assert set(z.files) == set(a)
    errs = {k: rel_err(a[k], b[k]) for k in a}
    say(f"  worst: {max(errs.items(), key=lambda kv: kv[1])}")
    assert max(errs.values()) <= 1e-9, errs


def _crossings(osim, steps):
    """Generators reflected at a wall during `steps` oracle steps (they end up moving against their
    drift velocity)."""
    n = osim.n
    count = 0
    prev = osim.get_fields()
    for _ in range(steps):
        osim.advance(1)
        f = osim.get_fields()
        x0, x1, v = prev["loc"][:n], f["loc"][:n], prev["v"][:n]
        count += int(np.sum((v > 0) & (x1 < x0)) + np.sum((v < 0) & (x1 > x0)))
        prev = f
    return count


def test_wall_crossing_reflection_parity(ora):
    """Generators drifting through a reflective wall are reflected with their primitives and gradients
    on the device as in the oracle (box_reflect, reflect_quantities, reflect_gradients), and the next
    predict extrapolates the reflected state: 11 steps (3 reflections of each outer generator) within the
    whole-run tolerance, then the reference's negative-pressure panic at the same step with the same code."""
    prob = wall_crossing_tube()
    osim = oracle_sim(prob)
    assert _crossings(osim, 11) == 6
    dev = ics.build_device_sim(prob)
    assert dev.run(11) == 11
    ref, got = osim.get_fields(), dev.get_state()
    errs = compare_states(got, ref, prob.n)
    say(f"  11 steps: errs={ {k: f'{v:.1e}' for k, v in errs.items()} }")
    assert errs.pop("timebin_mismatch") == 0
    assert max(errs.values()) <= 1e-9, errs
    with pytest.raises(RuntimeError) as oe:
        osim.advance(1 << 40)
    with pytest.raises(fvs.HfvError) as de:
        dev.run(1 << 40)
    assert de.value.code == int(str(oe.value).split()[-1]) == -4
    assert dev.engine_info()["step_count"] == osim.info()["step_count"]


def test_wall_crossing_reflection_fp32():
    """The FP32-flux context sums its deferred predicts with the rate each reflected generator had before
    the reflection and extrapolates the reflected state afterwards: L1(rho) against the FP64 path after the
    same 11 steps (3 reflections of each outer generator) stays < 1e-5."""
    d64 = ics.build_device_sim(wall_crossing_tube(fvs.FP64))
    d32 = ics.build_device_sim(wall_crossing_tube(fvs.FP32_FLUX))
    assert d64.run(11) == 11 and d32.run(11) == 11
    assert d64.engine_info()["ti_current"] == d32.engine_info()["ti_current"]
    a = d64.get_state(["prim", "extrap"])
    b = d32.get_state(["prim", "extrap"])
    n = 40
    l1 = float(np.mean(np.abs(a["prim"][:n] - b["prim"][:n])))
    say(f"  L1(rho) fp32 vs fp64 after 11 steps: {l1:.2e}")
    assert l1 < 1e-5
    np.testing.assert_allclose(b["extrap"], a["extrap"], rtol=0, atol=1e-5 * np.max(np.abs(a["extrap"])))


def test_compute_timesteps_entry(ora):
    """hfv_compute_timesteps returns each active particle's Particle::timestep (the oracle's, from the same
    state, within 1e-12) and +inf for inactive ones; followed by hfv_apply_timesteps it equals the fused
    hfv_timesteps bit for bit."""
    prob = ics.sedov_2d(side=24, t_end=0.05)
    a = ics.build_device_sim(prob)
    b = ics.build_device_sim(prob)
    tb = prob.cfg.t_end / float(1 << 57)
    ti_old, ti_cur = 0, 0
    checked = 0
    for step in range(6):
        info = a.timestep_info(ti_old, ti_cur, tb)
        for d in (a, b):
            d.predict((ti_cur - ti_old) * tb)
            d.compute_fluxes(info)
            d.apply_fluxes()
            d.convert_conserved_to_primitive()
            d.gradients()
        t1 = a.timesteps()
        st = b.get_state()
        dts = b.compute_timesteps()
        act = st["timebin"] <= ora.get_max_active_bin(ti_cur)
        assert np.all(np.isinf(dts[~act]))
        n = prob.n
        for i in np.nonzero(act)[0][::5]:
            col = lambda k, w: st[k].reshape(w, n)[:, i]
            dt, _, vrel, _, e = ora.particle_timestep(prob.cfg.gamma, prob.cfg.cfl, prob.cfg.motion, 2,
                                                      col("cons", 5), col("prim", 5), col("grad", 15).reshape(5, 3),
                                                      col("loc", 3), col("gc", 3), col("centroid", 3),
                                                      st["volume"][i], msv=st["msv"][i], maor=st["maor"][i])
            assert e == 0
            assert abs(dts[i] - dt) <= 1e-12 * dt, (step, i, dts[i], dt)
            checked += 1
        t2 = b.apply_timesteps(dts)
        assert t1 == t2
        sa, sb = a.get_state(), b.get_state()
        for k in sa:
            assert np.array_equal(sa[k], sb[k]), (step, k)
        ti_old, ti_cur = ti_cur, t1
    assert checked > 60


def test_periodic_wrap_full_run_parity(ora):
    """Generators drifting through the periodic box edge (k_drift's box_wrap) and the rotated 1D rebuild
    (k_mesh1d_*): the whole run (1025 steps, 30 generators wrapped) within 1e-9 of the oracle, bins and
    step count exact; the FP32-flux context stays within L1(rho) < 1e-5 of it."""
    prob = periodic_wrap_tube()
    osim = oracle_sim(prob)
    n_ref = osim.advance(1 << 40)
    ref = osim.get_fields()
    xs = ref["loc"][: prob.n]
    assert int(np.argmin(xs)) == 70 and int(np.sum(np.diff(xs) <= 0)) == 1  # rotated by 30
    dev = ics.build_device_sim(prob)
    n_dev = dev.run(1 << 40)
    errs = compare_states(dev.get_state(), ref, prob.n)
    say(f"  steps ref={n_ref} dev={n_dev} errs={ {k: f'{v:.1e}' for k, v in errs.items()} }")
    assert n_dev == n_ref
    assert errs.pop("timebin_mismatch") == 0
    assert max(errs.values()) <= 1e-9, errs
    d32 = ics.build_device_sim(periodic_wrap_tube(fvs.FP32_FLUX))
    d32.run(1 << 40)
    assert d32.engine_info()["done"] and d32.engine_info()["error_code"] == 0
    l1 = float(np.mean(np.abs(d32.get_state(["prim"])["prim"][: prob.n] - ref["prim"][: prob.n])))
    say(f"  FP32-flux L1(rho) vs oracle {l1:.2e}")
    assert l1 < 1e-5
