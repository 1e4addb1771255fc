"""Seeded synthetic initial conditions for the benchmark workloads.

The reference builds preset ICs with InitialConditions::from_fn (src/initial_conditions.rs:121-186)
from an unseeded one-sided jitter; these generators are seeded and two-sided, and set each cell's mass
from its Voronoi volume so that the initial density is the stated one.  The Sod and vortex profiles are
the reference's presets (src/bin/main.rs:35-40, python/utils.py:242-256).
"""
from dataclasses import dataclass

import numpy as np

from .fvs import AIRS, EXACT, FIXED, FLUID, GRAV_KEPLERIAN, HLLC, PERIODIC, REFLECTIVE, Config
from .mesh import VoronoiMesh


@dataclass
class Problem:
    name: str
    cfg: Config
    loc: np.ndarray     # [N,3]
    mass: np.ndarray    # [N]
    vel: np.ndarray     # [N,3]
    u_int: np.ndarray   # [N]
    mesh: VoronoiMesh

    @property
    def n(self):
        return self.mass.size


def morton_order(idx):
    """Permutation sorting integer grid coordinates [m, dim] by their interleaved-bit (Z-order) code."""
    idx = np.asarray(idx, dtype=np.uint64)
    dim = idx.shape[1]
    code = np.zeros(idx.shape[0], dtype=np.uint64)
    for b in range(21):
        for a in range(dim):
            code |= ((idx[:, a] >> np.uint64(b)) & np.uint64(1)) << np.uint64(dim * b + (dim - 1 - a))
    return np.argsort(code, kind="stable")


def jittered_grid(dim, side, box, jitter, rng, order="morton"):
    """Generators at cell centres of a side^dim grid, each coordinate jittered by U(-jitter, jitter)*dx.
    order "grid": row-major (x slowest); "morton": the same generators in Z-order, so that cells close
    in space are close in index (neighbour gathers stay within a small index window)."""
    box = np.asarray(box, dtype=np.float64)
    idx = np.stack(np.meshgrid(*[np.arange(side)] * dim, indexing="ij"), -1).reshape(-1, dim)
    dx = box[:dim] / side
    pos = (idx + 0.5 + rng.uniform(-jitter, jitter, idx.shape)) * dx
    if order == "morton":
        pos = pos[morton_order(idx)]
    loc = np.zeros((pos.shape[0], 3))
    loc[:, :dim] = pos
    return loc


def _finish(name, cfg, loc, rho, vel, P, threads=0):
    mesh = VoronoiMesh(cfg.dim, cfg.box, cfg.boundary == PERIODIC, loc, n_threads=threads)
    mass = rho * mesh.volume
    u_int = P / ((cfg.gamma - 1.0) * rho)
    return Problem(name, cfg, loc, mass, vel, u_int, mesh)


def sod_1d(n=1000, seed=1, t_end=0.2, motion=FLUID, **kw):
    """1D Sod shock tube: rho/P = 1/1 left of 0.5, 0.125/0.1 right; gamma 1.4; exact solver."""
    rng = np.random.default_rng(seed)
    dx = 1.0 / n
    x = (np.arange(n) + 0.5) * dx + 0.01 * dx * rng.uniform(-1.0, 1.0, n)
    loc = np.zeros((n, 3))
    loc[:, 0] = x
    left = x < 0.5
    rho = np.where(left, 1.0, 0.125)
    P = np.where(left, 1.0, 0.1)
    vel = np.zeros((n, 3))
    cfg = Config(dim=1, box=(1.0, 1.0, 1.0), gamma=1.4, cfl=0.3, solver=EXACT, boundary=REFLECTIVE,
                 motion=motion, t_end=t_end, **kw)
    return _finish("sod1d", cfg, loc, rho, vel, P)


def sedov_2d(side=256, seed=2, t_end=0.05, solver=HLLC, threads=0, gamma=5.0 / 3.0, **kw):
    """2D Sedov blast on the unit periodic box: rho 1, P 1e-6, E = 1 shared by the cells with r < 2 dx."""
    rng = np.random.default_rng(seed)
    box = (1.0, 1.0, 1.0)
    loc = jittered_grid(2, side, box, 0.05, rng)
    n = loc.shape[0]
    dx = 1.0 / side
    rho = np.ones(n)
    P = np.full(n, 1e-6)
    vel = np.zeros((n, 3))
    cfg = Config(dim=2, box=box, gamma=gamma, cfl=0.3, solver=solver, boundary=PERIODIC, motion=FIXED,
                 t_end=t_end, **kw)
    prob = _finish(f"sedov2d_{side}", cfg, loc, rho, vel, P, threads)
    r = np.linalg.norm(loc[:, :2] - 0.5, axis=1)
    inside = r < 2.0 * dx
    n_in = int(inside.sum())
    prob.u_int = prob.u_int.copy()
    # thermal energy E/n_in per blast cell: m * u = E / n_in
    prob.u_int[inside] = (1.0 / n_in) / prob.mass[inside]
    return prob


def kelvin_helmholtz_3d(side=128, seed=3, t_end=1.0, solver=AIRS, threads=0, order="morton", **kw):
    """3D Kelvin-Helmholtz shear layer in the unit periodic box (gamma 5/3, P 2.5)."""
    rng = np.random.default_rng(seed)
    box = (1.0, 1.0, 1.0)
    loc = jittered_grid(3, side, box, 0.1, rng, order=order)
    n = loc.shape[0]
    band = np.abs(loc[:, 1] - 0.5) < 0.25
    rho = np.where(band, 2.0, 1.0)
    vel = np.zeros((n, 3))
    vel[:, 0] = np.where(band, 0.5, -0.5)
    vel[:, 1] = 0.01 * np.sin(4.0 * np.pi * loc[:, 0])
    P = np.full(n, 2.5)
    cfg = Config(dim=3, box=box, gamma=5.0 / 3.0, cfl=0.3, solver=solver, boundary=PERIODIC, motion=FIXED,
                 t_end=t_end, airs_threshold=2.0, **kw)
    return _finish(f"kh3d_{side}", cfg, loc, rho, vel, P, threads)


def isentropic_vortex_2d(side=64, seed=5, t_end=1.0, solver=HLLC, gamma=1.4, beta=5.0, threads=0, **kw):
    """Isentropic (Yee) vortex on [-5,5]^2, shifted to the box [0,10]^2 (python/utils.py:242-256)."""
    rng = np.random.default_rng(seed)
    box = (10.0, 10.0, 1.0)
    loc = jittered_grid(2, side, box, 0.05, rng)
    xy = loc[:, :2] - 5.0
    r2 = np.sum(xy * xy, axis=1)
    temp = 1.0 - (gamma - 1.0) * beta ** 2 / (8.0 * gamma * np.pi ** 2) * np.exp(1.0 - r2)
    rho = temp ** (1.0 / (gamma - 1.0))
    P = rho ** gamma
    vfac = 0.5 * beta / np.pi * np.exp(0.5 * (1.0 - r2))
    vel = np.zeros((loc.shape[0], 3))
    vel[:, 0] = -xy[:, 1] * vfac
    vel[:, 1] = xy[:, 0] * vfac
    cfg = Config(dim=2, box=box, gamma=gamma, cfl=0.3, solver=solver, boundary=PERIODIC, motion=FIXED,
                 t_end=t_end, **kw)
    return _finish(f"vortex2d_{side}", cfg, loc, rho, vel, P, threads)


def keplerian_disc_2d(side=24, seed=7, t_end=0.05, solver=HLLC, gamma=5.0 / 3.0, softening=0.05, threads=0, **kw):
    """2D gas around a Keplerian point mass at the centre of the unit box (reflective walls): rho 1, a
    pressure bump P = 0.01 + 0.5 exp(-r^2 / 0.01) and a slow solid-body rotation v = 0.5 r.  The
    Keplerian timestep (eta R / v_circ) binds in the cold outer gas, the hydro one in the bump."""
    rng = np.random.default_rng(seed)
    box = (1.0, 1.0, 1.0)
    loc = jittered_grid(2, side, box, 0.05, rng)
    xy = loc[:, :2] - 0.5
    r2 = np.sum(xy * xy, axis=1)
    rho = np.ones(loc.shape[0])
    P = 0.01 + 0.5 * np.exp(-r2 / 0.01)
    vel = np.zeros((loc.shape[0], 3))
    vel[:, 0] = -0.5 * xy[:, 1]
    vel[:, 1] = 0.5 * xy[:, 0]
    kw.setdefault("gravity", GRAV_KEPLERIAN)
    cfg = Config(dim=2, box=box, gamma=gamma, cfl=0.3, solver=solver, boundary=REFLECTIVE, motion=FIXED,
                 t_end=t_end, grav_softening=softening, grav_position=(0.5, 0.5, 0.0), **kw)
    return _finish(f"kepler2d_{side}", cfg, loc, rho, vel, P, threads)


def riemann_tube_1d(n, rho_l, p_l, v_l, rho_r, p_r, v_r, gamma, seed=0, solver=HLLC, t_end=0.1, **kw):
    """1D Riemann problem on [0,1] with reflective walls (ensemble member)."""
    rng = np.random.default_rng(seed)
    dx = 1.0 / n
    x = (np.arange(n) + 0.5) * dx + 0.01 * dx * rng.uniform(-1.0, 1.0, n)
    loc = np.zeros((n, 3))
    loc[:, 0] = x
    left = x < 0.5
    rho = np.where(left, rho_l, rho_r)
    P = np.where(left, p_l, p_r)
    vel = np.zeros((n, 3))
    vel[:, 0] = np.where(left, v_l, v_r)
    cfg = Config(dim=1, box=(1.0, 1.0, 1.0), gamma=gamma, cfl=0.3, solver=solver, boundary=REFLECTIVE,
                 motion=FIXED, t_end=t_end, **kw)
    return _finish("tube1d", cfg, loc, rho, vel, P)


def build_device_sim(prob: Problem):
    """Allocate a device simulation for a problem and load its initial conditions."""
    from .fvs import EulerEqnsFvs

    sim = EulerEqnsFvs(prob.cfg, prob.mesh)
    sim.init_from_ic(prob.loc, prob.mass, prob.vel, prob.u_int)
    return sim


def with_precision(prob: Problem, precision: int) -> Problem:
    """The same problem (shared mesh and initial arrays) with another flux precision."""
    from dataclasses import replace

    return replace(prob, cfg=replace(prob.cfg, precision=precision))
