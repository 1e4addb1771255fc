"""Python surface of the device path, mirroring the reference's plugin interface.

  EulerEqnsFvs          FiniteVolumeSolver + Engine state of one simulation on one GPU
                        (src/finite_volume_solver.rs:9-17, euler_eqns_fvs.rs, engine.rs)
  riemann_batch         RiemannFluxSolver::solve_for_flux over a batch (src/riemann_solver.rs:78-87)

Every call goes through the C-ABI of include/hfv.h into lib/libhfv.so; there is no CPU fallback.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib

HfvError = _lib.HfvError

EXACT, HLLC, PVRS, AIRS, TSRS, TRRS, LINEAR_ADVECTION = range(7)
SOLVERS = {"Exact": EXACT, "HLLC": HLLC, "PVRS": PVRS, "AIRS": AIRS, "TSRS": TSRS, "TRRS": TRRS,
           "LinearAdvection": LINEAR_ADVECTION}
PERIODIC, REFLECTIVE, OPEN, VACUUM = range(4)
BOUNDARIES = {"periodic": PERIODIC, "reflective": REFLECTIVE, "open": OPEN, "vacuum": VACUUM}
FIXED, STEER, STEER_PAKMOR, FLUID = range(4)
MOTIONS = {"fixed": FIXED, "steer": STEER, "steer_pakmor": STEER_PAKMOR, "fluid": FLUID}
OPTIMAL_ORDER, GODUNOV = range(2)
RUNNERS = {"OptimalOrder": OPTIMAL_ORDER, "Godunov": GODUNOV}
FP64, FP32_FLUX = range(2)
# gravity solvers (src/bin/main.rs:643-666)
GRAV_NONE, GRAV_SELF, GRAV_CONSTANT, GRAV_KEPLERIAN = range(4)

FIELD_WIDTH = {"prim": 5, "grad": 15, "gc": 3, "extrap": 5, "cons": 5, "fluxes": 5, "volume": 1, "loc": 3,
               "centroid": 3, "v": 3, "v_rel": 3, "maor": 1, "msv": 1, "dt": 1}


@dataclass
class Config:
    """Solver and engine configuration (the reference's YAML sections hydrodynamics, engine,
    time_integration, space; src/bin/main.rs:253-606)."""
    dim: int
    box: tuple
    gamma: float
    cfl: float
    solver: int = HLLC
    boundary: int = REFLECTIVE
    motion: int = FIXED
    runner: int = OPTIMAL_ORDER
    sync_all: bool = False
    t_end: float = 1.0
    dt_snap: float = 0.0
    airs_threshold: float = 2.0
    linadv_velocity: tuple = (0.0, 0.0, 0.0)
    eos: int = 0
    isothermal_u: float = 0.0
    precision: int = FP64
    gravity: int = GRAV_NONE
    grav_softening: float = 0.0
    grav_position: tuple = (0.0, 0.0, 0.0)
    grav_acceleration: tuple = (0.0, 0.0, 0.0)
    extra: dict = field(default_factory=dict)

    def to_c(self):
        p = _lib.Params()
        p.dim = self.dim
        p.boundary = self.boundary
        for a in range(3):
            p.box[a] = float(self.box[a])
            p.linadv_velocity[a] = float(self.linadv_velocity[a])
        p.gamma = self.gamma
        p.eos = self.eos
        p.isothermal_u = self.isothermal_u
        p.cfl = self.cfl
        p.solver = self.solver
        p.airs_threshold = self.airs_threshold
        p.runner = self.runner
        p.motion = self.motion
        p.sync_all = int(self.sync_all)
        p.t_end = self.t_end
        p.dt_snap = self.dt_snap
        p.precision = self.precision
        p.gravity = self.gravity
        p.grav_softening = self.grav_softening
        for a in range(3):
            p.grav_position[a] = float(self.grav_position[a])
            p.grav_acceleration[a] = float(self.grav_acceleration[a])
        return p


def _dp(a):
    return a.ctypes.data_as(_lib.DP)


class EulerEqnsFvs:
    """One simulation's device state: the FiniteVolumeSolver trait surface, the fused runner step and
    the engine loop."""

    def __init__(self, cfg: Config, mesh=None):
        self.cfg = cfg
        self.lib = _lib.hfv()
        self._params = cfg.to_c()
        self.ptr = C.c_void_p()
        _lib.check(self.lib.hfv_create(C.byref(self._params), C.byref(self.ptr)))
        self.n = 0
        if mesh is not None:
            self.set_mesh(mesh)

    def close(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            self.lib.hfv_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    # ------------------------------------------------------------ setup
    def set_mesh(self, mesh, update_mask=None):
        m = mesh.arrays_struct()
        mask = None
        if update_mask is not None:
            mask = np.ascontiguousarray(update_mask, dtype=np.uint8)
        _lib.check(self.lib.hfv_set_mesh(self.ptr, C.byref(m),
                                         mask.ctypes.data_as(_lib.U8P) if mask is not None else None))
        self.n = mesh.n_cells
        self.n_faces = mesh.n_faces

    def init_from_ic(self, loc, mass, vel, u_int):
        loc = np.ascontiguousarray(loc, dtype=np.float64).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        u_int = np.ascontiguousarray(u_int, dtype=np.float64)
        _lib.check(self.lib.hfv_init_from_ic(self.ptr, loc.shape[0], _dp(loc), _dp(mass), _dp(vel), _dp(u_int)))

    def get_state(self, fields=None):
        n = self.n
        fields = fields or list(FIELD_WIDTH) + ["timebin"]
        out = {}
        st = _lib.State()
        for k in fields:
            if k == "timebin":
                out[k] = np.zeros(n, np.int8)
                st.timebin = out[k].ctypes.data_as(_lib.I8P)
            else:
                out[k] = np.zeros(FIELD_WIDTH[k] * n)
                setattr(st, k, _dp(out[k]))
        _lib.check(self.lib.hfv_get_state(self.ptr, C.byref(st)))
        return out

    def set_state(self, fields):
        st = _lib.State()
        keep = {}
        for k, v in fields.items():
            if k == "timebin":
                keep[k] = np.ascontiguousarray(v, dtype=np.int8)
                st.timebin = keep[k].ctypes.data_as(_lib.I8P)
            elif k in FIELD_WIDTH:
                keep[k] = np.ascontiguousarray(v, dtype=np.float64)
                assert keep[k].size == FIELD_WIDTH[k] * self.n, k
                setattr(st, k, _dp(keep[k]))
        _lib.check(self.lib.hfv_set_state(self.ptr, C.byref(st)))

    # ------------------------------------------------------------ trait surface
    @staticmethod
    def timestep_info(ti_old, ti_current, time_base, ti_next=0):
        t = _lib.TimestepInfo()
        t.ti_old = ti_old
        t.ti_current = ti_current
        t.ti_next = ti_next
        t.time_base = time_base
        t.time_base_inv = 1.0 / time_base
        return t

    def predict(self, dt):
        _lib.check(self.lib.hfv_predict(self.ptr, dt))

    def compute_fluxes(self, info):
        _lib.check(self.lib.hfv_compute_fluxes(self.ptr, C.byref(info)))

    def get_fluxes(self):
        out = np.zeros(7 * self.n_faces)
        _lib.check(self.lib.hfv_get_fluxes(self.ptr, _dp(out)))
        return out.reshape(-1, 7)

    def apply_fluxes(self):
        _lib.check(self.lib.hfv_apply_fluxes(self.ptr))

    def convert_conserved_to_primitive(self):
        _lib.check(self.lib.hfv_conserved_to_primitive(self.ptr))

    def gradients(self):
        _lib.check(self.lib.hfv_gradients(self.ptr))

    def timesteps(self):
        t = _lib.U64(0)
        _lib.check(self.lib.hfv_timesteps(self.ptr, C.byref(t)))
        return t.value

    def compute_timesteps(self):
        """FiniteVolumeSolver::compute_timesteps: each particle's requested dt (+inf if inactive)."""
        out = np.zeros(self.n)
        _lib.check(self.lib.hfv_compute_timesteps(self.ptr, _dp(out)))
        return out

    def apply_timesteps(self, dt):
        """timesteps_apply + limiter/sync + reset + kick1 of requested timesteps; returns ti_next."""
        d = np.ascontiguousarray(dt, dtype=np.float64)
        assert d.size == self.n
        t = _lib.U64(0)
        _lib.check(self.lib.hfv_apply_timesteps(self.ptr, _dp(d), C.byref(t)))
        return t.value

    # ------------------------------------------------------------ GravitySolver
    def gravity_timesteps(self):
        """GravitySolver::compute_timesteps of every particle (a_grav = 0, as on the step path)."""
        out = np.zeros(self.n)
        _lib.check(self.lib.hfv_gravity_timesteps(self.ptr, _dp(out)))
        return out

    def gravity_accelerations(self):
        """GravitySolver::compute_accelerations at the current positions and masses, [3, N]."""
        out = np.zeros(3 * self.n)
        _lib.check(self.lib.hfv_gravity_accelerations(self.ptr, _dp(out)))
        return out.reshape(3, self.n)

    # ------------------------------------------------------------ step / engine
    def step(self, info):
        t = _lib.U64(0)
        _lib.check(self.lib.hfv_step(self.ptr, C.byref(info), C.byref(t)))
        return t.value

    def run(self, max_steps=1 << 62):
        done = _lib.I64(0)
        _lib.check(self.lib.hfv_run(self.ptr, int(max_steps), C.byref(done)))
        return done.value

    def engine_info(self):
        e = _lib.EngineInfo()
        _lib.check(self.lib.hfv_get_engine_info(self.ptr, C.byref(e)))
        d = {k: getattr(e, k) for k, _ in _lib.EngineInfo._fields_ if k != "ts"}
        d.update({k: getattr(e.ts, k) for k, _ in _lib.TimestepInfo._fields_})
        return d

    def stats(self):
        s = _lib.Stats()
        _lib.check(self.lib.hfv_get_stats(self.ptr, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _lib.Stats._fields_}

    def conservation(self):
        out = np.zeros(5)
        _lib.check(self.lib.hfv_conservation(self.ptr, _dp(out)))
        return out

    def get_mesh(self):
        """The current mesh (for moving meshes the device's latest rebuild) as a dict of host arrays in
        the tessellator's layout."""
        m = _lib.MeshArrays()
        _lib.check(self.lib.hfv_get_mesh(self.ptr, C.byref(m)))
        n, nf, nk = m.n_cells, m.n_faces, m.n_conn
        out = {"n_cells": n, "n_faces": nf, "n_conn": nk}
        keep = {}
        for k, size, dt in (("area", nf, np.float64), ("normal", 3 * nf, np.float64), ("centroid", 3 * nf, np.float64),
                            ("shift", 3 * nf, np.float64), ("left", nf, np.int64), ("right", nf, np.int64),
                            ("volume", n, np.float64), ("cell_centroid", 3 * n, np.float64),
                            ("search_radius", n, np.float64), ("offset", n, np.int64), ("count", n, np.int64),
                            ("conn", nk, np.int64)):
            keep[k] = np.zeros(max(size, 1), dtype=dt)
            setattr(m, k, keep[k].ctypes.data_as(_lib.DP if dt == np.float64 else _lib.I64P))
        _lib.check(self.lib.hfv_get_mesh(self.ptr, C.byref(m)))
        for k, a in keep.items():
            size = {"normal": 3 * nf, "centroid": 3 * nf, "shift": 3 * nf, "cell_centroid": 3 * n}.get(
                k, nf if k in ("area", "left", "right") else nk if k == "conn" else n)
            out[k] = a[:size]
        return out

    def profile(self, enable=True):
        _lib.check(self.lib.hfv_profile_enable(self.ptr, int(bool(enable))))

    def kernel_times(self):
        t = _lib.KernelTimes()
        _lib.check(self.lib.hfv_profile_read(self.ptr, C.byref(t)))
        return {name: {"launches": t.launches[k], "ms": t.ms[k]} for k, name in enumerate(_lib.KERNEL_SLOTS)}

    def synchronize(self):
        _lib.check(self.lib.hfv_synchronize(self.ptr))

    def set_geometry_mode(self, use_static):
        """1 (default): fixed meshes use precomputed geometry; 0: recompute it every step.  Bit-identical."""
        _lib.check(self.lib.hfv_set_geometry_mode(self.ptr, int(use_static)))

    def geometry_static(self, slot=0):
        v = C.c_int(0)
        _lib.check(self.lib.hfv_get_geometry_mode(self.ptr, int(slot), C.byref(v)))
        return bool(v.value)

    def autotune(self, launches=8, rounds=2):
        """Choose the flux tiling and gradient geometry per size class by timing real launches
        (hfv_autotune; bit-identical candidates, so the launches are ordinary steps)."""
        return _autotune(self.lib, self.ptr, launches, rounds)

    def set_tuning(self, slot=0, staged=-1, static_grad=-1, gstaged=-1):
        """Flux tiling (staged), gradient geometry (static_grad) and gradient neighbour staging
        (gstaged) of one slot: 1/0, -1 = default."""
        _lib.check(self.lib.hfv_set_tuning(self.ptr, int(slot), int(staged), int(static_grad), int(gstaged)))

    def tuning(self, slot=0):
        a, b, g = C.c_int(0), C.c_int(0), C.c_int(0)
        _lib.check(self.lib.hfv_get_tuning(self.ptr, int(slot), C.byref(a), C.byref(b), C.byref(g)))
        return {"staged": a.value, "static_grad": b.value, "gstaged": g.value}

    def set_launch_shape(self, flux_faces_per_thread=1):
        """The flux kernel's face chunks per block (faces per thread: 1, 2 or 4) for the whole context."""
        _lib.check(self.lib.hfv_set_launch_shape(self.ptr, int(flux_faces_per_thread)))

    def launch_shape(self):
        """{"flux_faces_per_thread": current choice, "ns_per_update": the tuner's cost of 1, 2, 4}."""
        return _launch_shape(self.lib, self.ptr)


def _engine_dict(e):
    d = {k: getattr(e, k) for k, _ in _lib.EngineInfo._fields_ if k != "ts"}
    d.update({k: getattr(e.ts, k) for k, _ in _lib.TimestepInfo._fields_})
    return d


class Batch:
    """Many independent simulations in one device context (one launch sequence steps them all).

    Slots have fixed HBM capacities (cells, faces, CSR entries); load() puts a simulation into a slot
    and can be called again on a finished slot (continuous batching)."""

    def __init__(self, cell_caps, face_caps, conn_caps):
        self.lib = _lib.hfv()
        cc = np.ascontiguousarray(cell_caps, dtype=np.int64)
        fc = np.ascontiguousarray(face_caps, dtype=np.int64)
        kc = np.ascontiguousarray(conn_caps, dtype=np.int64)
        self.nslots = cc.size
        self.caps = (cc.copy(), fc.copy(), kc.copy())
        self.ptr = C.c_void_p()
        _lib.check(self.lib.hfv_batch_create(self.nslots, cc.ctypes.data_as(_lib.I64P), fc.ctypes.data_as(_lib.I64P),
                                             kc.ctypes.data_as(_lib.I64P), C.byref(self.ptr)))
        self.n_cells = [0] * self.nslots

    @classmethod
    def for_problems(cls, probs, headroom=1.0):
        caps = [(p.n, p.mesh.n_faces, p.mesh.conn.size) for p in probs]
        b = cls(*[np.ceil(np.array(c) * headroom).astype(np.int64) for c in zip(*caps)])
        for s, p in enumerate(probs):
            b.load(s, p)
        return b

    def close(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            self.lib.hfv_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    def load(self, slot, prob):
        m = prob.mesh.arrays_struct()
        p = prob.cfg.to_c()
        loc = np.ascontiguousarray(prob.loc, dtype=np.float64).reshape(-1, 3)
        vel = np.ascontiguousarray(prob.vel, dtype=np.float64).reshape(-1, 3)
        mass = np.ascontiguousarray(prob.mass, dtype=np.float64)
        u = np.ascontiguousarray(prob.u_int, dtype=np.float64)
        _lib.check(self.lib.hfv_batch_load(self.ptr, slot, C.byref(p), C.byref(m), _dp(loc), _dp(mass), _dp(vel),
                                           _dp(u)))
        self.n_cells[slot] = prob.n

    def unload(self, slot):
        _lib.check(self.lib.hfv_batch_unload(self.ptr, slot))
        self.n_cells[slot] = 0

    def run(self, max_launches=1 << 62):
        k = _lib.I64(0)
        _lib.check(self.lib.hfv_batch_run(self.ptr, int(max_launches), C.byref(k)))
        return k.value

    def engine_info(self, slot):
        e = _lib.EngineInfo()
        _lib.check(self.lib.hfv_batch_get_engine_info(self.ptr, slot, C.byref(e)))
        return _engine_dict(e)

    def get_state(self, slot, fields=None):
        n = self.n_cells[slot]
        fields = fields or list(FIELD_WIDTH) + ["timebin"]
        out = {}
        st = _lib.State()
        for k in fields:
            if k == "timebin":
                out[k] = np.zeros(n, np.int8)
                st.timebin = out[k].ctypes.data_as(_lib.I8P)
            else:
                out[k] = np.zeros(FIELD_WIDTH[k] * n)
                setattr(st, k, _dp(out[k]))
        _lib.check(self.lib.hfv_batch_get_state(self.ptr, slot, C.byref(st)))
        return out

    def stats(self):
        s = _lib.Stats()
        _lib.check(self.lib.hfv_get_stats(self.ptr, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _lib.Stats._fields_}

    def profile(self, enable=True):
        _lib.check(self.lib.hfv_profile_enable(self.ptr, int(bool(enable))))

    def kernel_times(self):
        t = _lib.KernelTimes()
        _lib.check(self.lib.hfv_profile_read(self.ptr, C.byref(t)))
        return {name: {"launches": t.launches[k], "ms": t.ms[k]} for k, name in enumerate(_lib.KERNEL_SLOTS)}

    def synchronize(self):
        _lib.check(self.lib.hfv_synchronize(self.ptr))

    def set_geometry_mode(self, use_static):
        """1 (default): fixed meshes use precomputed geometry; 0: recompute it every step.  Bit-identical."""
        _lib.check(self.lib.hfv_set_geometry_mode(self.ptr, int(use_static)))

    def geometry_static(self, slot=0):
        v = C.c_int(0)
        _lib.check(self.lib.hfv_get_geometry_mode(self.ptr, int(slot), C.byref(v)))
        return bool(v.value)

    def autotune(self, launches=8, rounds=2):
        """Choose the flux tiling and gradient geometry per size class by timing real launches
        (hfv_autotune; bit-identical candidates, so the launches are ordinary steps)."""
        return _autotune(self.lib, self.ptr, launches, rounds)

    def set_tuning(self, slot=0, staged=-1, static_grad=-1, gstaged=-1):
        """Flux tiling (staged), gradient geometry (static_grad) and gradient neighbour staging
        (gstaged) of one slot: 1/0, -1 = default."""
        _lib.check(self.lib.hfv_set_tuning(self.ptr, int(slot), int(staged), int(static_grad), int(gstaged)))

    def tuning(self, slot=0):
        a, b, g = C.c_int(0), C.c_int(0), C.c_int(0)
        _lib.check(self.lib.hfv_get_tuning(self.ptr, int(slot), C.byref(a), C.byref(b), C.byref(g)))
        return {"staged": a.value, "static_grad": b.value, "gstaged": g.value}

    def set_launch_shape(self, flux_faces_per_thread=1):
        """The flux kernel's face chunks per block (faces per thread: 1, 2 or 4) for the whole context."""
        _lib.check(self.lib.hfv_set_launch_shape(self.ptr, int(flux_faces_per_thread)))

    def launch_shape(self):
        """{"flux_faces_per_thread": current choice, "ns_per_update": the tuner's cost of 1, 2, 4}."""
        return _launch_shape(self.lib, self.ptr)


def _launch_shape(lib, ptr):
    f = C.c_int(0)
    ns = (C.c_double * 3)()
    _lib.check(lib.hfv_get_launch_shape(ptr, C.byref(f), ns))
    return {"flux_faces_per_thread": f.value, "ns_per_update": [round(x, 4) for x in ns]}


def _autotune(lib, ptr, launches, rounds):
    out = (_lib.TuneClass * 64)()
    n = C.c_int(0)
    _lib.check(lib.hfv_autotune(ptr, int(launches), int(rounds), out, 64, C.byref(n)))
    return [{"dim": o.dim, "n_sims": o.n_sims, "cells": [o.cells_lo, o.cells_hi], "staged": o.staged,
             "static_grad": o.static_grad, "gstaged": o.gstaged,
             "ns_per_update": [round(x, 4) for x in o.ns_per_update]}
            for o in out[: min(n.value, 64)]]


NOALLOC_MAX_FACES = 65536  # HFV_NOALLOC_MAX_FACES


def riemann_batch(cfg: Config, WL, WR, vface, nrm, noalloc=False):
    """RiemannFluxSolver::solve_for_flux for n face problems; arrays are [n,5] / [n,3].  noalloc: through
    hfv_riemann_batch_noalloc (no device allocation, at most NOALLOC_MAX_FACES problems)."""
    WL = np.ascontiguousarray(np.asarray(WL, dtype=np.float64).T)
    WR = np.ascontiguousarray(np.asarray(WR, dtype=np.float64).T)
    vf = np.ascontiguousarray(np.asarray(vface, dtype=np.float64).T)
    nr = np.ascontiguousarray(np.asarray(nrm, dtype=np.float64).T)
    n = WL.shape[1]
    out = np.zeros((5, n))
    p = cfg.to_c()
    fn = _lib.hfv().hfv_riemann_batch_noalloc if noalloc else _lib.hfv().hfv_riemann_batch
    _lib.check(fn(C.byref(p), n, _dp(WL), _dp(WR), _dp(vf), _dp(nr), _dp(out)))
    return out.T.copy()


def device_count():
    n = C.c_int(0)
    code = _lib.hfv().hfv_device_count(C.byref(n))
    return n.value if code == 0 else 0
