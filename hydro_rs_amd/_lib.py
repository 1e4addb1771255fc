"""ctypes bindings of the hfv C-ABI (include/hfv.h, include/hfv_host.h).

The product path is the HIP library lib/libhfv.so; there is no CPU fallback.  Loading fails loudly
when a library has not been built (python -c "import __graft_entry__; __graft_entry__.build()").
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
HFV_SO = os.path.join(LIB_DIR, "libhfv.so")
# HFV_HOST_LIB: another build of the host tessellator (e.g. one instrumented with sanitizers)
HOST_SO = os.environ.get("HFV_HOST_LIB") or os.path.join(LIB_DIR, "libhfv_host.so")

I64 = C.c_int64
U64 = C.c_uint64
DP = C.POINTER(C.c_double)
I64P = C.POINTER(C.c_int64)
I8P = C.POINTER(C.c_int8)
U8P = C.POINTER(C.c_uint8)

OK = 0
ERRORS = {
    -1: "HFV_ERR_NONFINITE", -2: "HFV_ERR_NEG_MASS", -3: "HFV_ERR_NEG_ENERGY", -4: "HFV_ERR_NEG_PRESSURE",
    -5: "HFV_ERR_NEWTON", -6: "HFV_ERR_BRENT", -7: "HFV_ERR_UNSUPPORTED", -8: "HFV_ERR_DRIFT",
    -20: "HFV_ERR_INVALID", -21: "HFV_ERR_HIP", -22: "HFV_ERR_MESH", -23: "HFV_ERR_NO_DEVICE",
}


class HfvError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__(f"{ERRORS.get(code, code)}: {what}")


class Params(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("boundary", C.c_int), ("box", C.c_double * 3), ("gamma", C.c_double),
        ("eos", C.c_int), ("isothermal_u", C.c_double), ("cfl", C.c_double), ("solver", C.c_int),
        ("airs_threshold", C.c_double), ("linadv_velocity", C.c_double * 3), ("runner", C.c_int),
        ("motion", C.c_int), ("sync_all", C.c_int), ("t_end", C.c_double), ("dt_snap", C.c_double),
        ("precision", C.c_int), ("gravity", C.c_int), ("grav_softening", C.c_double),
        ("grav_position", C.c_double * 3), ("grav_acceleration", C.c_double * 3),
    ]


class MeshArrays(C.Structure):
    _fields_ = [
        ("n_cells", I64), ("n_faces", I64), ("n_conn", I64), ("area", DP), ("normal", DP), ("centroid", DP),
        ("shift", DP), ("left", I64P), ("right", I64P), ("volume", DP), ("cell_centroid", DP),
        ("search_radius", DP), ("offset", I64P), ("count", I64P), ("conn", I64P),
    ]


class State(C.Structure):
    _fields_ = [(k, DP) for k in ("prim", "grad", "gc", "extrap", "cons", "fluxes", "volume", "loc", "centroid",
                                  "v", "v_rel", "maor", "msv", "dt")] + [("timebin", I8P)]


class TimestepInfo(C.Structure):
    _fields_ = [("ti_old", U64), ("ti_current", U64), ("ti_next", U64), ("time_base", C.c_double),
                ("time_base_inv", C.c_double)]


class EngineInfo(C.Structure):
    _fields_ = [("ts", TimestepInfo), ("t_current", C.c_double), ("t_end", C.c_double), ("step_count", I64),
                ("cell_updates", I64), ("face_fluxes", I64), ("last_n_active", I64),
                ("last_n_face_active", I64), ("done", C.c_int), ("error_code", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("check_physical_resets", I64), ("apply_flux_resets", I64), ("gradient_fallbacks", I64),
                ("brent_calls", I64), ("error_code", I64)]


class KernelTimes(C.Structure):
    _fields_ = [("launches", I64 * 8), ("ms", C.c_double * 8)]


class TuneClass(C.Structure):
    _fields_ = [("dim", C.c_int), ("n_sims", C.c_int), ("cells_lo", I64), ("cells_hi", I64), ("staged", C.c_int),
                ("static_grad", C.c_int), ("gstaged", C.c_int), ("pad_", C.c_int), ("ns_per_update", C.c_double * 8)]


KERNEL_SLOTS = ["mask_predict", "mesh_rebuild", "flux", "apply_c2p", "grad_timestep", "wakeup", "limit_kick", "drift"]


# Every symbol include/hfv.h and include/hfv_host.h declare, with its ctypes signature.
HFV_SYMBOLS = {
    "hfv_last_error": (C.c_char_p, []),
    "hfv_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "hfv_create": (C.c_int, [C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "hfv_destroy": (C.c_int, [C.c_void_p]),
    "hfv_set_mesh": (C.c_int, [C.c_void_p, C.POINTER(MeshArrays), U8P]),
    "hfv_init_from_ic": (C.c_int, [C.c_void_p, I64, DP, DP, DP, DP]),
    "hfv_set_state": (C.c_int, [C.c_void_p, C.POINTER(State)]),
    "hfv_get_state": (C.c_int, [C.c_void_p, C.POINTER(State)]),
    "hfv_get_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "hfv_synchronize": (C.c_int, [C.c_void_p]),
    "hfv_predict": (C.c_int, [C.c_void_p, C.c_double]),
    "hfv_compute_fluxes": (C.c_int, [C.c_void_p, C.POINTER(TimestepInfo)]),
    "hfv_get_fluxes": (C.c_int, [C.c_void_p, DP]),
    "hfv_apply_fluxes": (C.c_int, [C.c_void_p]),
    "hfv_conserved_to_primitive": (C.c_int, [C.c_void_p]),
    "hfv_gradients": (C.c_int, [C.c_void_p]),
    "hfv_timesteps": (C.c_int, [C.c_void_p, C.POINTER(U64)]),
    "hfv_compute_timesteps": (C.c_int, [C.c_void_p, DP]),
    "hfv_apply_timesteps": (C.c_int, [C.c_void_p, DP, C.POINTER(U64)]),
    "hfv_gravity_timesteps": (C.c_int, [C.c_void_p, DP]),
    "hfv_gravity_accelerations": (C.c_int, [C.c_void_p, DP]),
    "hfv_step": (C.c_int, [C.c_void_p, C.POINTER(TimestepInfo), C.POINTER(U64)]),
    "hfv_run": (C.c_int, [C.c_void_p, I64, C.POINTER(I64)]),
    "hfv_get_engine_info": (C.c_int, [C.c_void_p, C.POINTER(EngineInfo)]),
    "hfv_conservation": (C.c_int, [C.c_void_p, DP]),
    "hfv_get_mesh": (C.c_int, [C.c_void_p, C.POINTER(MeshArrays)]),
    "hfv_riemann_batch": (C.c_int, [C.POINTER(Params), I64, DP, DP, DP, DP, DP]),
    "hfv_riemann_batch_noalloc": (C.c_int, [C.POINTER(Params), I64, DP, DP, DP, DP, DP]),
    "hfv_batch_create": (C.c_int, [C.c_int, I64P, I64P, I64P, C.POINTER(C.c_void_p)]),
    "hfv_batch_load": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(Params), C.POINTER(MeshArrays), DP, DP, DP, DP]),
    "hfv_batch_unload": (C.c_int, [C.c_void_p, C.c_int]),
    "hfv_batch_run": (C.c_int, [C.c_void_p, I64, C.POINTER(I64)]),
    "hfv_batch_get_engine_info": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(EngineInfo)]),
    "hfv_batch_get_state": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(State)]),
    "hfv_batch_num_slots": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "hfv_set_geometry_mode": (C.c_int, [C.c_void_p, C.c_int]),
    "hfv_get_geometry_mode": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "hfv_autotune": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(TuneClass), C.c_int, C.POINTER(C.c_int)]),
    "hfv_set_tuning": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "hfv_get_tuning": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "hfv_set_launch_shape": (C.c_int, [C.c_void_p, C.c_int]),
    "hfv_get_launch_shape": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), DP]),
    "hfv_profile_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "hfv_profile_read": (C.c_int, [C.c_void_p, C.POINTER(KernelTimes)]),
}
HOST_SYMBOLS = {
    "hfv_voronoi_build": (C.c_int, [C.c_int, DP, C.c_int, I64, DP, C.c_int, C.POINTER(C.c_void_p)]),
    "hfv_voronoi_build_partial": (C.c_int, [C.c_int, DP, C.c_int, I64, DP, U8P, C.c_int, C.POINTER(C.c_void_p)]),
    "hfv_mesh_view": (C.c_int, [C.c_void_p, C.POINTER(MeshArrays)]),
    "hfv_mesh_free": (None, [C.c_void_p]),
}

_hfv = None
_host = None


def _bind(path, symbols):
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build() first (no CPU fallback exists)")
    lib = C.CDLL(path)
    for name, (res, args) in symbols.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def hfv():
    """The HIP library (device path)."""
    global _hfv
    if _hfv is None:
        _hfv = _bind(HFV_SO, HFV_SYMBOLS)
    return _hfv


def host():
    """The host tessellation library."""
    global _host
    if _host is None:
        _host = _bind(HOST_SO, HOST_SYMBOLS)
    return _host


def check(code, lib=None):
    if code != OK:
        what = ""
        try:
            what = (lib or hfv()).hfv_last_error().decode()
        except Exception:  # noqa: BLE001 - error text is best effort
            pass
        raise HfvError(code, what)
    return code
