"""ctypes front-end of the CPU FP64 oracle (oracle/*.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the reference that results are compared against or as the
timed CPU baseline, never by the product package hydro_rs_amd.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORA_LIB: a build of the same sources elsewhere (the sanitizer test's instrumented copy)
_LIB_PATH = os.environ.get("ORA_LIB") or os.path.join(_HERE, "_build", "liboracle.so")

EXACT, HLLC, PVRS, AIRS, TSRS, TRRS, LINADV = range(7)
VACUUM_SAMPLER = 100
PERIODIC, REFLECTIVE, OPEN, VACUUM = range(4)
FIXED, STEER, STEER_PAKMOR, FLUID = range(4)
OPTIMAL_ORDER, GODUNOV = range(2)
GRAV_NONE, GRAV_SELF, GRAV_CONSTANT, GRAV_KEPLERIAN = range(4)

I64 = C.c_int64
DP = C.POINTER(C.c_double)
I64P = C.POINTER(C.c_int64)
I8P = C.POINTER(C.c_int8)


def build(force=False):
    if os.environ.get("ORA_LIB"):
        return _LIB_PATH
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE])
    return _LIB_PATH


class Riemann(C.Structure):
    _fields_ = [("kind", C.c_int), ("threshold", C.c_double), ("velocity", C.c_double * 3)]


class Params(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("boundary", C.c_int), ("box", C.c_double * 3), ("gamma", C.c_double),
        ("eos", C.c_int), ("isothermal_u", C.c_double), ("cfl", C.c_double), ("riemann", Riemann),
        ("runner", C.c_int), ("motion", C.c_int), ("sync_all", C.c_int), ("t_end", C.c_double),
        ("dt_snap", C.c_double), ("gravity", C.c_int), ("grav_softening", C.c_double),
        ("grav_position", C.c_double * 3), ("grav_acceleration", C.c_double * 3),
    ]


class Mesh(C.Structure):
    _fields_ = [
        ("n_faces", I64), ("area", DP), ("normal", DP), ("centroid", DP), ("left", I64P), ("right", I64P),
        ("shift", DP), ("n_conn", I64), ("conn", I64P), ("offset", I64P), ("count", I64P), ("volume", DP),
        ("cell_centroid", DP), ("search_radius", DP),
    ]


class Counters(C.Structure):
    _fields_ = [("check_physical_resets", I64), ("apply_flux_resets", I64), ("gradient_fallbacks", I64),
                ("brent_calls", I64)]


class EngineInfo(C.Structure):
    _fields_ = [("ti_old", C.c_uint64), ("ti_current", C.c_uint64), ("ti_next", C.c_uint64),
                ("time_base", C.c_double), ("time_base_inv", C.c_double), ("t_current", C.c_double),
                ("t_end", C.c_double), ("step_count", I64), ("error", C.c_int), ("last_n_active", I64),
                ("last_n_face_active", I64)]


class Fields(C.Structure):
    _fields_ = [(k, DP) for k in ("prim", "grad", "gc", "extrap", "cons", "fluxes", "volume", "loc", "centroid",
                                  "v", "v_rel", "maor", "msv", "dt")] + [("timebin", I8P)]


FIELD_SHAPES = {"prim": 5, "grad": 15, "gc": 3, "extrap": 5, "cons": 5, "fluxes": 5, "volume": 1, "loc": 3,
                "centroid": 3, "v": 3, "v_rel": 3, "maor": 1, "msv": 1, "dt": 1}

_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        L.ora_solve_for_flux.argtypes = [C.POINTER(Riemann), C.c_double, DP, DP, DP, DP, DP]
        L.ora_sample.argtypes = [C.c_int, C.c_double, C.c_double, DP, DP, C.c_double, C.c_double, C.c_double,
                                 C.c_double, DP, DP]
        L.ora_exact_star.argtypes = [C.c_double, DP, DP, C.c_double, C.c_double, C.c_double, C.c_double, DP]
        L.ora_pairwise_limiter.argtypes = [DP, DP, DP, C.c_double, DP]
        L.ora_pairwise_limiter.restype = None
        L.ora_sound_speed.argtypes = [C.c_double, C.c_double, C.c_double]
        L.ora_sound_speed.restype = C.c_double
        L.ora_get_integer_timestep.argtypes = [C.c_int8]
        L.ora_get_integer_timestep.restype = C.c_uint64
        L.ora_get_time_bin.argtypes = [C.c_uint64]
        L.ora_get_time_bin.restype = C.c_int8
        L.ora_get_integer_time_end.argtypes = [C.c_uint64, C.c_int8]
        L.ora_get_integer_time_end.restype = C.c_uint64
        L.ora_get_max_active_bin.argtypes = [C.c_uint64]
        L.ora_get_max_active_bin.restype = C.c_int8
        L.ora_make_integer_timestep.argtypes = [C.c_double, C.c_int8, C.c_int8, C.c_uint64, C.c_double]
        L.ora_make_integer_timestep.restype = C.c_uint64
        L.ora_saturating_f64_to_u64.argtypes = [C.c_double]
        L.ora_saturating_f64_to_u64.restype = C.c_uint64
        L.ora_sim_new.argtypes = [C.POINTER(Params), I64, DP, DP, DP, DP, C.POINTER(Mesh)]
        L.ora_sim_new.restype = C.c_void_p
        L.ora_sim_free.argtypes = [C.c_void_p]
        L.ora_sim_free.restype = None
        L.ora_sim_set_threads.argtypes = [C.c_int]
        L.ora_sim_advance.argtypes = [C.c_void_p, I64]
        L.ora_sim_advance.restype = I64
        L.ora_sim_runner_step.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
        L.ora_sim_engine_info.argtypes = [C.c_void_p, C.POINTER(EngineInfo)]
        L.ora_sim_engine_info.restype = None
        L.ora_sim_set_engine_info.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64]
        L.ora_sim_set_engine_info.restype = None
        L.ora_sim_counters.argtypes = [C.c_void_p, C.POINTER(Counters)]
        L.ora_sim_counters.restype = None
        L.ora_sim_get_fields.argtypes = [C.c_void_p, C.POINTER(Fields)]
        L.ora_sim_get_fields.restype = None
        L.ora_sim_set_fields.argtypes = [C.c_void_p, C.POINTER(Fields)]
        L.ora_sim_set_fields.restype = None
        L.ora_sim_n_cells.argtypes = [C.c_void_p]
        L.ora_sim_n_cells.restype = I64
        L.ora_sim_n_faces.argtypes = [C.c_void_p]
        L.ora_sim_n_faces.restype = I64
        L.ora_sim_get_mesh.argtypes = [C.c_void_p, DP, DP, DP, I64P, I64P, DP, I64P, I64P, I64P]
        L.ora_sim_get_mesh.restype = None
        L.ora_phase_predict.argtypes = [C.c_void_p, C.c_double]
        L.ora_phase_predict.restype = None
        L.ora_phase_compute_fluxes.argtypes = [C.c_void_p, C.POINTER(C.c_uint8), DP]
        L.ora_phase_compute_fluxes.restype = None
        L.ora_sim_conservation.argtypes = [C.c_void_p, DP]
        L.ora_sim_conservation.restype = None
        L.ora_sim_set_mesher.argtypes = [C.c_void_p, MESHER, C.c_void_p]
        L.ora_sim_set_mesher.restype = None
        L.ora_sim_apply_mesh.argtypes = [C.c_void_p, C.POINTER(Mesh)]
        L.ora_sim_apply_mesh.restype = None
        L.ora_gravity_accelerations.argtypes = [C.c_void_p, DP]
        L.ora_gravity_accelerations.restype = None
        L.ora_gravity_timesteps.argtypes = [C.c_void_p, DP]
        L.ora_gravity_timesteps.restype = None
        L.ora_phase_gradients.argtypes = [C.c_void_p, C.POINTER(C.c_uint8)]
        L.ora_phase_gradients.restype = None
        L.ora_t_c2p.argtypes = [C.c_double, DP, C.c_double, DP]
        L.ora_t_time_extrapolations.argtypes = [C.c_double, DP, DP, DP, C.c_double, DP]
        L.ora_t_time_extrapolations.restype = None
        L.ora_t_radius.argtypes = [C.c_double, C.c_int]
        L.ora_t_radius.restype = C.c_double
        L.ora_t_particle_timestep.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, DP, DP]
        L.ora_t_timestep_limit.argtypes = [C.c_double, C.c_int8, C.c_int8, C.c_uint64, C.c_double, DP]
        L.ora_t_timestep_limit.restype = C.c_int8
        L.ora_t_face_velocity.argtypes = [DP, DP, DP, DP, DP, DP, DP]
        L.ora_t_face_velocity.restype = None
        L.ora_voronoi_1d.argtypes = [I64, DP, C.c_double, C.c_int, DP, DP, DP, I64P, I64P, DP, I64P, I64P, I64P,
                                     DP, DP, DP, I64P]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(DP)


def _i64p(a):
    return a.ctypes.data_as(I64P)


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None:
        assert a.size == n
    return a


def riemann(kind, threshold=2.0, velocity=(0.0, 0.0, 0.0)):
    r = Riemann()
    r.kind = kind
    r.threshold = threshold
    for i in range(3):
        r.velocity[i] = velocity[i]
    return r


def solve_for_flux(kind, gamma, left, right, iface_v, n_unit, threshold=2.0, velocity=(0.0, 0.0, 0.0)):
    rs = riemann(kind, threshold, velocity)
    out = np.zeros(5)
    e = lib().ora_solve_for_flux(C.byref(rs), gamma, _dp(_f64(left, 5)), _dp(_f64(right, 5)),
                                 _dp(_f64(iface_v, 3)), _dp(_f64(n_unit, 3)), _dp(out))
    if e:
        raise RuntimeError(f"oracle error {e}")
    return out


def sample(kind, gamma, left, right, v_l, v_r, a_l, a_r, n_unit, threshold=2.0):
    out = np.zeros(5)
    e = lib().ora_sample(kind, threshold, gamma, _dp(_f64(left, 5)), _dp(_f64(right, 5)), v_l, v_r, a_l, a_r,
                         _dp(_f64(n_unit, 3)), _dp(out))
    if e:
        raise RuntimeError(f"oracle error {e}")
    return out


def exact_star(gamma, left, right, v_l, v_r, a_l, a_r):
    out = np.zeros(4)
    e = lib().ora_exact_star(gamma, _dp(_f64(left, 5)), _dp(_f64(right, 5)), v_l, v_r, a_l, a_r, _dp(out))
    return e, out


def pairwise_limiter(pl, pr, pdash, dx_fac):
    out = np.zeros(5)
    lib().ora_pairwise_limiter(_dp(_f64(pl, 5)), _dp(_f64(pr, 5)), _dp(_f64(pdash, 5)), dx_fac, _dp(out))
    return out


def sound_speed(gamma, p, rho_inv):
    return lib().ora_sound_speed(gamma, p, rho_inv)


def voronoi_1d(x, box_x, periodic=False):
    x = _f64(x)
    n = x.size
    nf = n if periodic else n + 1
    m = {
        "area": np.zeros(nf), "normal": np.zeros(3 * nf), "centroid": np.zeros(3 * nf),
        "left": np.zeros(nf, np.int64), "right": np.zeros(nf, np.int64), "shift": np.zeros(3 * nf),
        "conn": np.zeros(2 * n, np.int64), "offset": np.zeros(n, np.int64), "count": np.zeros(n, np.int64),
        "volume": np.zeros(n), "cell_centroid": np.zeros(3 * n), "search_radius": np.zeros(n),
    }
    nfo = I64(0)
    e = lib().ora_voronoi_1d(n, _dp(x), box_x, int(periodic), _dp(m["area"]), _dp(m["normal"]),
                             _dp(m["centroid"]), _i64p(m["left"]), _i64p(m["right"]), _dp(m["shift"]),
                             _i64p(m["conn"]), _i64p(m["offset"]), _i64p(m["count"]), _dp(m["volume"]),
                             _dp(m["cell_centroid"]), _dp(m["search_radius"]), C.byref(nfo))
    if e:
        raise RuntimeError(f"oracle 1D voronoi error {e}")
    m["n_faces"] = int(nfo.value)
    return m


def make_params(dim, boundary, box, gamma, cfl, solver, threshold=2.0, runner=OPTIMAL_ORDER, motion=FIXED,
                sync_all=False, t_end=1.0, dt_snap=0.0, eos=0, isothermal_u=0.0, velocity=(0.0, 0.0, 0.0),
                gravity=GRAV_NONE, grav_softening=0.0, grav_position=(0.0, 0.0, 0.0),
                grav_acceleration=(0.0, 0.0, 0.0)):
    p = Params()
    p.gravity = gravity
    p.grav_softening = grav_softening
    for i in range(3):
        p.grav_position[i] = grav_position[i]
        p.grav_acceleration[i] = grav_acceleration[i]
    p.dim = dim
    p.boundary = boundary
    for i in range(3):
        p.box[i] = box[i]
    p.gamma = gamma
    p.eos = eos
    p.isothermal_u = isothermal_u
    p.cfl = cfl
    p.riemann = riemann(solver, threshold, velocity)
    p.runner = runner
    p.motion = motion
    p.sync_all = int(sync_all)
    p.t_end = t_end
    p.dt_snap = dt_snap
    return p


def _mesh_struct(mesh):
    """An ora_mesh over a mesh dict (the tessellator's layout) and the arrays it points into."""
    mk = {k: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v) for k, v in mesh.items()}
    m = Mesh()
    m.n_faces = int(mk["n_faces"])
    for k in ("area", "normal", "centroid", "shift", "volume", "cell_centroid", "search_radius"):
        mk[k] = np.ascontiguousarray(mk[k], dtype=np.float64)
        setattr(m, k, _dp(mk[k]))
    for k in ("left", "right", "conn", "offset", "count"):
        mk[k] = np.ascontiguousarray(mk[k], dtype=np.int64)
        setattr(m, k, _i64p(mk[k]))
    m.n_conn = int(mk["conn"].size)
    return m, mk


MESHER = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_uint8))


class Sim:
    """The oracle's Space + Engine for one simulation."""

    def __init__(self, params, loc, mass, vel, u_int, mesh):
        self.lib = lib()
        n = len(mass)
        self.n = n
        m, self._keep = _mesh_struct(mesh)
        self.params = params
        self.ptr = self.lib.ora_sim_new(C.byref(params), n, _dp(_f64(loc, 3 * n)), _dp(_f64(mass, n)),
                                        _dp(_f64(vel, 3 * n)), _dp(_f64(u_int, n)), C.byref(m))

    def __del__(self):
        if getattr(self, "ptr", None):
            self.lib.ora_sim_free(self.ptr)
            self.ptr = None

    def set_mesher(self, fn):
        """2D/3D moving meshes: fn(loc (n, 3) array, active mask (n,) bool) -> the mesh dict of those
        generators with the active ones' cells built (the tessellator standing in for meshless_voronoi's
        build_partial), called by every step after the predict."""
        import traceback

        def cb(_user, sim, n, loc, mask):
            try:
                pos = np.ctypeslib.as_array(loc, shape=(3 * n,)).reshape(n, 3).copy()
                act = np.ctypeslib.as_array(mask, shape=(n,)).astype(bool)
                m, _keep = _mesh_struct(fn(pos, act))
                self.lib.ora_sim_apply_mesh(sim, C.byref(m))
                return 0
            except Exception:  # noqa: BLE001 - reported, and the step stops with ORA_ERR_MESH
                traceback.print_exc()
                return 1

        self._mesher = MESHER(cb)
        self.lib.ora_sim_set_mesher(self.ptr, self._mesher, None)

    def advance(self, max_steps):
        r = self.lib.ora_sim_advance(self.ptr, max_steps)
        if r < 0:
            raise RuntimeError(f"oracle error {r}")
        return r

    def runner_step(self):
        t = C.c_uint64(0)
        e = self.lib.ora_sim_runner_step(self.ptr, C.byref(t))
        if e:
            raise RuntimeError(f"oracle error {e}")
        return t.value

    def info(self):
        inf = EngineInfo()
        self.lib.ora_sim_engine_info(self.ptr, C.byref(inf))
        return {k: getattr(inf, k) for k, _ in EngineInfo._fields_}

    def set_engine_info(self, ti_old, ti_current, ti_next):
        self.lib.ora_sim_set_engine_info(self.ptr, ti_old, ti_current, ti_next)

    def counters(self):
        c = Counters()
        self.lib.ora_sim_counters(self.ptr, C.byref(c))
        return {k: getattr(c, k) for k, _ in Counters._fields_}

    def get_fields(self):
        n = self.n
        out = {k: np.zeros(w * n) for k, w in FIELD_SHAPES.items()}
        out["timebin"] = np.zeros(n, np.int8)
        f = Fields()
        for k in FIELD_SHAPES:
            setattr(f, k, _dp(out[k]))
        f.timebin = out["timebin"].ctypes.data_as(I8P)
        self.lib.ora_sim_get_fields(self.ptr, C.byref(f))
        return out

    def set_fields(self, fields):
        f = Fields()
        keep = {}
        for k in FIELD_SHAPES:
            if k in fields:
                keep[k] = np.ascontiguousarray(fields[k], dtype=np.float64)
                setattr(f, k, _dp(keep[k]))
        if "timebin" in fields:
            keep["timebin"] = np.ascontiguousarray(fields["timebin"], dtype=np.int8)
            f.timebin = keep["timebin"].ctypes.data_as(I8P)
        self.lib.ora_sim_set_fields(self.ptr, C.byref(f))

    def n_faces(self):
        return self.lib.ora_sim_n_faces(self.ptr)

    def get_mesh(self):
        nf = self.n_faces()
        n = self.n
        m = {"area": np.zeros(nf), "normal": np.zeros(3 * nf), "centroid": np.zeros(3 * nf),
             "left": np.zeros(nf, np.int64), "right": np.zeros(nf, np.int64), "shift": np.zeros(3 * nf),
             "conn": np.zeros(2 * nf, np.int64), "offset": np.zeros(n, np.int64), "count": np.zeros(n, np.int64)}
        self.lib.ora_sim_get_mesh(self.ptr, _dp(m["area"]), _dp(m["normal"]), _dp(m["centroid"]),
                                  _i64p(m["left"]), _i64p(m["right"]), _dp(m["shift"]), _i64p(m["conn"]),
                                  _i64p(m["offset"]), _i64p(m["count"]))
        m["n_faces"] = nf
        return m

    def predict(self, dt):
        self.lib.ora_phase_predict(self.ptr, dt)

    def compute_fluxes(self, active):
        act = np.ascontiguousarray(active, dtype=np.uint8)
        out = np.zeros(7 * self.n_faces())
        self.lib.ora_phase_compute_fluxes(self.ptr, act.ctypes.data_as(C.POINTER(C.c_uint8)), _dp(out))
        return out.reshape(-1, 7)

    def conservation(self):
        out = np.zeros(5)
        self.lib.ora_sim_conservation(self.ptr, _dp(out))
        return out

    def gravity_accelerations(self):
        out = np.zeros(3 * self.n)
        self.lib.ora_gravity_accelerations(self.ptr, _dp(out))
        return out

    def gravity_timesteps(self):
        out = np.zeros(self.n)
        self.lib.ora_gravity_timesteps(self.ptr, _dp(out))
        return out

    def gradients(self, active):
        act = np.ascontiguousarray(active, dtype=np.uint8)
        self.lib.ora_phase_gradients(self.ptr, act.ctypes.data_as(C.POINTER(C.c_uint8)))


# single functions of the path (ideal gas), for the tests that pin them
def c2p(gamma, cons, volume):
    out = np.zeros(5)
    e = lib().ora_t_c2p(gamma, _dp(_f64(cons, 5)), volume, _dp(out))
    return e, out


def time_extrapolations(gamma, prim, grad, v_rel, dt):
    out = np.zeros(5)
    lib().ora_t_time_extrapolations(gamma, _dp(_f64(prim, 5)), _dp(_f64(grad, 15)), _dp(_f64(v_rel, 3)), dt,
                                    _dp(out))
    return out


def radius(volume, dim):
    return lib().ora_t_radius(volume, dim)


def particle_timestep(gamma, cfl, motion, dim, cons, prim, grad, loc, gc, centroid, volume, msv=0.0, maor=0.0):
    """Particle::timestep: returns (dt, v, v_rel, msv after the call, error flag)."""
    inp = np.concatenate([_f64(cons, 5).ravel(), _f64(prim, 5).ravel(), _f64(grad, 15).ravel(), _f64(loc, 3).ravel(),
                          _f64(gc, 3).ravel(), _f64(centroid, 3).ravel(), [volume, msv, maor]])
    out = np.zeros(8)
    e = lib().ora_t_particle_timestep(gamma, cfl, motion, dim, _dp(inp), _dp(out))
    return out[0], out[1:4].copy(), out[4:7].copy(), out[7], e


def timestep_limit(dt, timebin, wakeup, ti_current, time_base):
    """Particle::timestep_limit: returns (timebin, dt)."""
    d = C.c_double(0.0)
    b = lib().ora_t_timestep_limit(dt, timebin, wakeup, ti_current, time_base, C.byref(d))
    return b, d.value


def face_velocity(loc_l, loc_r, shift, v_l, v_r, fc):
    out = np.zeros(3)
    lib().ora_t_face_velocity(_dp(_f64(loc_l, 3)), _dp(_f64(loc_r, 3)), _dp(_f64(shift, 3)), _dp(_f64(v_l, 3)),
                              _dp(_f64(v_r, 3)), _dp(_f64(fc, 3)), _dp(out))
    return out


def set_threads(n):
    """OpenMP threads for the oracle (n > 0), or the process's default (n <= 0); returns the count in effect."""
    return lib().ora_sim_set_threads(n)


# timeline
def get_integer_timestep(b):
    return lib().ora_get_integer_timestep(b)


def get_time_bin(dti):
    return lib().ora_get_time_bin(dti)


def get_integer_time_end(ti, b):
    return lib().ora_get_integer_time_end(ti, b)


def get_max_active_bin(ti):
    return lib().ora_get_max_active_bin(ti)


def make_integer_timestep(dt, old_bin, min_ngb_bin, ti_current, time_base_inv):
    return lib().ora_make_integer_timestep(dt, old_bin, min_ngb_bin, ti_current, time_base_inv)


def saturating_cast(x):
    return lib().ora_saturating_f64_to_u64(x)
