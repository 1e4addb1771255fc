"""Shared test helpers: build the oracle simulation for a product Problem, and compare states."""
import numpy as np

from hydro_rs_amd import fvs
from hydro_rs_amd.mesh import VoronoiMesh
from oracle import ora

FLOAT_FIELDS = ["prim", "grad", "gc", "extrap", "cons", "fluxes", "volume", "loc", "centroid", "v", "v_rel",
                "maor", "msv", "dt"]


def oracle_sim(prob):
    c = prob.cfg
    p = ora.make_params(c.dim, c.boundary, c.box, c.gamma, c.cfl, c.solver, threshold=c.airs_threshold,
                        runner=c.runner, motion=c.motion, sync_all=c.sync_all, t_end=c.t_end, dt_snap=c.dt_snap,
                        eos=c.eos, isothermal_u=c.isothermal_u, velocity=c.linadv_velocity,
                        gravity=c.gravity, grav_softening=c.grav_softening, grav_position=c.grav_position,
                        grav_acceleration=c.grav_acceleration)
    sim = ora.Sim(p, prob.loc.ravel(), prob.mass, prob.vel.ravel(), prob.u_int, prob.mesh.as_dict())
    if c.dim >= 2 and c.motion != fvs.FIXED:
        # moving 2D/3D meshes: every step re-tessellates the drifted generators with the host tessellator
        # (the stand-in for meshless_voronoi that the device rebuild reproduces)
        sim.set_mesher(lambda pos, act: VoronoiMesh(c.dim, c.box, c.boundary == fvs.PERIODIC, pos,
                                                    mask=act).as_dict())
    return sim


def rel_err(a, b, floor=1e-300):
    """max |a-b| / max(|b|) over an array (norm-relative)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0:
        return 0.0
    both_nan = np.isnan(a) & np.isnan(b)
    d = np.where(both_nan, 0.0, np.abs(a - b))
    scale = max(np.max(np.abs(np.where(np.isfinite(b), b, 0.0))), floor)
    return float(np.max(d) / scale)


def per_component_rel(a, b, width, n):
    """Worst norm-relative error over the components of an SoA field."""
    a = np.asarray(a).reshape(width, n)
    b = np.asarray(b).reshape(width, n)
    return max(rel_err(a[k], b[k]) for k in range(width))


def compare_states(dev, ref, n, tol=1e-12, fields=None, ignore=()):
    """Return {field: worst relative error}; timebins compared exactly (reported as 0/1)."""
    from hydro_rs_amd.fvs import FIELD_WIDTH

    out = {}
    for k in fields or FLOAT_FIELDS:
        if k in ignore:
            continue
        out[k] = per_component_rel(dev[k], ref[k], FIELD_WIDTH[k], n)
    out["timebin_mismatch"] = int(np.sum(dev["timebin"] != ref["timebin"]))
    return out


def face_flux_rel(dev_rec, ref_rec):
    """Per-face norm-relative error of the flux records [F,7] (components 0..4 scaled by the face's flux
    magnitude, v_max and a_over_r by themselves); returns the worst face."""
    dev_rec = np.asarray(dev_rec)
    ref_rec = np.asarray(ref_rec)
    fmag = np.max(np.abs(ref_rec[:, :5]), axis=1)
    d = np.max(np.abs(dev_rec[:, :5] - ref_rec[:, :5]), axis=1)
    with np.errstate(invalid="ignore", divide="ignore"):
        r5 = np.where(fmag > 0, d / np.where(fmag > 0, fmag, 1.0), d)
        r6 = np.abs(dev_rec[:, 5] - ref_rec[:, 5]) / np.maximum(np.abs(ref_rec[:, 5]), 1e-300)
        r7 = np.abs(dev_rec[:, 6] - ref_rec[:, 6]) / np.maximum(np.abs(ref_rec[:, 6]), 1e-300)
    r6 = np.where(ref_rec[:, 5] == dev_rec[:, 5], 0.0, r6)
    r7 = np.where(ref_rec[:, 6] == dev_rec[:, 6], 0.0, r7)
    return float(max(np.max(r5, initial=0.0), np.max(r6, initial=0.0), np.max(r7, initial=0.0)))


def mesh_from_oracle(sim, fields, dim, box, periodic):
    m = sim.get_mesh()
    mesh = object.__new__(VoronoiMesh)
    mesh.dim, mesh.box, mesh.periodic = dim, np.asarray(box, float), periodic
    mesh.n_cells = sim.n
    mesh.n_faces = m["n_faces"]
    mesh.area = m["area"]
    mesh.normal = m["normal"].reshape(-1, 3)
    mesh.centroid = m["centroid"].reshape(-1, 3)
    mesh.shift = m["shift"].reshape(-1, 3)
    mesh.left, mesh.right = m["left"], m["right"]
    mesh.offset, mesh.count = m["offset"], m["count"]
    mesh.conn = m["conn"][: int(m["count"].sum())] if m["conn"].size > m["count"].sum() else m["conn"]
    n = sim.n
    mesh.volume = fields["volume"].copy()
    mesh.cell_centroid = fields["centroid"].reshape(3, n).T.copy()
    mesh.search_radius = np.zeros(n)
    return mesh


def synced(prob, k):
    """Oracle advanced k engine steps, and a device sim holding the same state and mesh."""
    osim = oracle_sim(prob)
    if k:
        osim.advance(k)
    f = osim.get_fields()
    inf = osim.info()
    dev = fvs.EulerEqnsFvs(prob.cfg)
    c = prob.cfg
    dev.set_mesh(mesh_from_oracle(osim, f, c.dim, c.box, c.boundary == fvs.PERIODIC))
    dev.set_state(f)
    return osim, dev, inf


def star_problem(base, ring, boundary, seed=5):
    """A moving-mesh problem at rest whose first generator (the box centre) has `ring` others on a sphere
    (3D, Fibonacci points) or circle (2D) of radius 0.08 around it, in a jittered 5^dim lattice (points
    within 0.15 of the centre dropped): the centre cell has exactly `ring` faces, one per ring point.
    base: a Problem of the same dimension whose configuration is kept (motion, solver, ...)."""
    import dataclasses

    dim = base.cfg.dim
    rng = np.random.default_rng(seed)
    g = (np.arange(5) + 0.5) / 5
    bg = np.stack(np.meshgrid(*[g] * dim, indexing="ij"), -1).reshape(-1, dim)
    bg = bg + rng.uniform(-0.02, 0.02, bg.shape)
    c = np.full(dim, 0.5)
    bg = bg[np.linalg.norm(bg - c, axis=1) > 0.15]
    if dim == 3:
        k = np.arange(ring) + 0.5
        th, ph = np.arccos(1 - 2 * k / ring), np.pi * (1 + 5 ** 0.5) * k
        rim = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1)
    else:
        a = 2 * np.pi * (np.arange(ring) + 0.25) / ring
        rim = np.stack([np.cos(a), np.sin(a)], -1)
    pts = np.concatenate([c[None], c + 0.08 * rim, bg])
    n = len(pts)
    loc = np.zeros((n, 3))
    loc[:, :dim] = pts
    cfg = dataclasses.replace(base.cfg, boundary=boundary)
    mesh = VoronoiMesh(dim, cfg.box, boundary == fvs.PERIODIC, loc)
    return dataclasses.replace(base, cfg=cfg, loc=loc, mass=np.full(n, 1.0 / n), vel=np.zeros((n, 3)),
                               u_int=np.full(n, 1.5), mesh=mesh)


def advected_wave(dim, side, vel, t_end, seed=7, precision=fvs.FP64):
    """A density wave rho = 1 + 0.3 prod_a sin(2 pi x_a) carried by a uniform flow `vel` at uniform pressure
    through the unit periodic box, generators moving with the fluid (HLLC, gamma 1.4): over t_end a
    fraction vel_a * t_end of the generators drifts through each box face (box_wrap, part.rs:449-455)."""
    from hydro_rs_amd import ics

    rng = np.random.default_rng(seed)
    loc = ics.jittered_grid(dim, side, (1.0, 1.0, 1.0), 0.1, rng)
    n = loc.shape[0]
    rho = 1.0 + 0.3 * np.prod(np.sin(2 * np.pi * loc[:, :dim]), axis=1)
    v = np.zeros((n, 3))
    v[:, :dim] = vel
    cfg = fvs.Config(dim=dim, box=(1.0, 1.0, 1.0), gamma=1.4, cfl=0.3, solver=fvs.HLLC, boundary=fvs.PERIODIC,
                     motion=fvs.FLUID, t_end=t_end, precision=precision)
    return ics._finish(f"advected_wave_{dim}d", cfg, loc, rho, v, np.ones(n))


def wrapped_fraction(loc0, loc1, vel):
    """Per axis with a flow, the fraction of generators that drifted through the box face downstream."""
    return [float(np.mean(loc1[:, a] < loc0[:, a]) if u > 0 else np.mean(loc1[:, a] > loc0[:, a]))
            for a, u in enumerate(vel) if u != 0]


def wall_box(dim, side, precision=fvs.FP64):
    """Gas at rest pressure and density (1, 1) in the unit box with reflective walls, flowing out along
    every axis at 0.5 (v_a = 0.5 sign(x_a - 0.5)), generators moving with it: a side^dim lattice (inner
    coordinates jittered by 0.05 dx) whose outer layers sit 0.05 dx from the walls, so that they drift
    through them within a few steps and are reflected (Particle::box_reflect, part.rs:457-469)."""
    from hydro_rs_amd import ics

    dx = 1.0 / side
    g = (np.arange(side) + 0.5) * dx
    g[0], g[-1] = 0.05 * dx, 1.0 - 0.05 * dx
    rng = np.random.default_rng(9)
    pts = np.stack(np.meshgrid(*[g] * dim, indexing="ij"), -1).reshape(-1, dim)
    inner = (pts > 0.06 * dx) & (pts < 1 - 0.06 * dx)
    pts = np.where(inner, pts + rng.uniform(-0.05, 0.05, pts.shape) * dx, pts)
    n = len(pts)
    loc = np.zeros((n, 3))
    loc[:, :dim] = pts
    vel = np.zeros((n, 3))
    vel[:, :dim] = 0.5 * np.sign(pts - 0.5)
    cfg = fvs.Config(dim=dim, box=(1.0, 1.0, 1.0), gamma=1.4, cfl=0.3, solver=fvs.HLLC, boundary=fvs.REFLECTIVE,
                     motion=fvs.FLUID, t_end=0.05, precision=precision)
    return ics._finish(f"wall_box_{dim}d", cfg, loc, np.ones(n), vel, np.ones(n))


def reflections(osim, steps, dim):
    """Generator coordinates reflected at a wall during `steps` oracle steps (they moved against the
    velocity they drifted with)."""
    n = osim.n
    count = 0
    prev = osim.get_fields()
    for _ in range(steps):
        osim.advance(1)
        f = osim.get_fields()
        x0 = np.asarray(prev["loc"]).reshape(3, -1)[:dim, :n]
        x1 = np.asarray(f["loc"]).reshape(3, -1)[:dim, :n]
        v = np.asarray(prev["v"]).reshape(3, -1)[:dim, :n]
        count += int(np.sum((v > 0) & (x1 < x0)) + np.sum((v < 0) & (x1 > x0)))
        prev = f
    return count
