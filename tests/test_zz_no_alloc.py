"""hfv_riemann_batch_noalloc: the allocation-free variant of hfv_riemann_batch gives the same fluxes bit for
bit, on consecutive calls whose second is smaller (the reserved device array keeps the first call's
contents), and refuses batches beyond HFV_NOALLOC_MAX_FACES before touching a device."""
import sys

import numpy as np
import pytest

from hydro_rs_amd import fvs


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _problems(n, seed):
    rng = np.random.default_rng(seed)
    WL = np.column_stack([10 ** rng.uniform(-1, 1, n), rng.uniform(-1, 1, (n, 3)), 10 ** rng.uniform(-2, 1, n)])
    WR = np.column_stack([10 ** rng.uniform(-1, 1, n), rng.uniform(-1, 1, (n, 3)), 10 ** rng.uniform(-2, 1, n)])
    WL[: n // 50] = 0.0  # vacuum and vacuum-generating pairs
    WR[n // 50: n // 25] = 0.0
    WL[n // 25: 3 * n // 50, 1] = -8.0
    WR[n // 25: 3 * n // 50, 1] = 8.0
    vf = rng.uniform(-0.5, 0.5, (n, 3))
    nr = rng.normal(size=(n, 3))
    nr /= np.linalg.norm(nr, axis=1, keepdims=True)
    return WL, WR, vf, nr


def test_noalloc_rejects_oversized_batch():
    """Beyond HFV_NOALLOC_MAX_FACES: HFV_ERR_INVALID with a message, checked before any device call."""
    cfg = fvs.Config(dim=3, box=(1, 1, 1), gamma=1.4, cfl=0.3, solver=fvs.HLLC)
    n = fvs.NOALLOC_MAX_FACES + 1
    z = np.zeros((n, 5))
    with pytest.raises(fvs.HfvError) as e:
        fvs.riemann_batch(cfg, z, z, np.zeros((n, 3)), np.zeros((n, 3)), noalloc=True)
    assert e.value.code == -20  # HFV_ERR_INVALID
    assert "HFV_NOALLOC_MAX_FACES" in str(e.value)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [fvs.FP64, fvs.FP32_FLUX])
@pytest.mark.parametrize("solver", ["Exact", "HLLC", "AIRS", "TSRS"])
def test_noalloc_matches_riemann_batch(solver, precision):
    if fvs.device_count() < 1:
        pytest.skip("no HIP device")
    cfg = fvs.Config(dim=3, box=(1, 1, 1), gamma=5.0 / 3.0, cfl=0.3, solver=fvs.SOLVERS[solver], precision=precision)
    for n, seed in ((fvs.NOALLOC_MAX_FACES, 1), (3001, 2)):  # the second call smaller than the first
        say(f"[noalloc] solver={solver} precision={precision} n={n}")
        WL, WR, vf, nr = _problems(n, seed)
        ref = fvs.riemann_batch(cfg, WL, WR, vf, nr)
        got = fvs.riemann_batch(cfg, WL, WR, vf, nr, noalloc=True)
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        assert np.array_equal(np.nan_to_num(got), np.nan_to_num(ref))
