"""Argument checks of the C-ABI that run before any device work (CPU): invalid parameters and arrays
return HFV_ERR_INVALID with a message, as the reference's configuration parsing rejects them
(src/bin/main.rs:253-642)."""
import ctypes as C

import numpy as np
import pytest

from hydro_rs_amd import _lib, fvs

HFV_ERR_INVALID = -20


def _cfg(**kw):
    base = dict(dim=1, box=(1.0, 1.0, 1.0), gamma=1.4, cfl=0.3)
    base.update(kw)
    return fvs.Config(**base)


def _create(params):
    lib = _lib.hfv()
    ptr = C.c_void_p()
    rc = lib.hfv_create(C.byref(params), C.byref(ptr))
    msg = lib.hfv_last_error()
    if rc == 0:
        lib.hfv_destroy(ptr)
    return rc, (msg.decode() if msg else "")


@pytest.mark.parametrize("field,value,text", [
    ("dim", 0, "dim"), ("dim", 4, "dim"),
    ("boundary", 4, "boundary"), ("solver", 7, "riemann solver"), ("solver", -1, "riemann solver"),
    ("motion", 4, "particle motion"), ("runner", 2, "runner"), ("precision", 2, "precision"),
    ("eos", 2, "equation of state"), ("gamma", 1.0, "gamma"), ("gamma", float("nan"), "gamma"),
    ("t_end", 0.0, "t_end"), ("t_end", -1.0, "t_end"),
])
def test_create_rejects_invalid_params(field, value, text):
    p = _cfg().to_c()
    setattr(p, field, value)
    rc, msg = _create(p)
    assert rc == HFV_ERR_INVALID
    assert text in msg


def test_create_rejects_null():
    lib = _lib.hfv()
    ptr = C.c_void_p()
    assert lib.hfv_create(None, C.byref(ptr)) == HFV_ERR_INVALID
    assert lib.hfv_create(C.byref(_cfg().to_c()), None) == HFV_ERR_INVALID


def test_riemann_batch_rejects_bad_arrays():
    lib = _lib.hfv()
    p = _cfg().to_c()
    a = np.zeros(5)
    d = a.ctypes.data_as(_lib.DP)
    assert lib.hfv_riemann_batch(C.byref(p), 0, d, d, d, d, d) == HFV_ERR_INVALID
    assert lib.hfv_riemann_batch(C.byref(p), 1, None, d, d, d, d) == HFV_ERR_INVALID
    bad = _cfg(solver=9).to_c()
    assert lib.hfv_riemann_batch(C.byref(bad), 1, d, d, d, d, d) == HFV_ERR_INVALID


@pytest.mark.parametrize("caps", [([0], [4], [8]), ([-3], [4], [8]), ([10], [-1], [8]), ([10], [4], [-8])])
def test_batch_create_rejects_bad_capacities(caps):
    lib = _lib.hfv()
    arrs = [np.asarray(c, dtype=np.int64) for c in caps]
    ptr = C.c_void_p()
    rc = lib.hfv_batch_create(1, *[a.ctypes.data_as(C.POINTER(C.c_int64)) for a in arrs], C.byref(ptr))
    assert rc == HFV_ERR_INVALID
    assert lib.hfv_batch_create(0, *[a.ctypes.data_as(C.POINTER(C.c_int64)) for a in arrs], C.byref(ptr)) == \
        HFV_ERR_INVALID
