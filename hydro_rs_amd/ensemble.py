"""Batched ensembles of independent simulations advanced by one graph-captured batch step.

ensemble_problems() draws the ensemble workload: `count` 1D/2D Riemann problems (half of each kind),
cell counts log-uniform in [n_min, n_max], rho_L, rho_R log-uniform in [0.1, 10], P log-uniform in
[0.01, 10], v uniform in [-1, 1], gamma in {1.4, 5/3}, 2D generators jittered by +-0.05 dx, HLLC,
reflective walls, fixed mesh.  run_ensemble() advances all of them in one Batch.
"""
import time

import numpy as np

from . import fvs, ics


def riemann_tube_2d(side, rho_l, p_l, v_l, rho_r, p_r, v_r, gamma, seed=0, t_end=0.05, **kw):
    """Planar 2D Riemann problem on the unit box (interface at x = 0.5, velocities along x)."""
    rng = np.random.default_rng(seed)
    loc = ics.jittered_grid(2, side, (1.0, 1.0, 1.0), 0.05, rng)
    left = loc[:, 0] < 0.5
    rho = np.where(left, rho_l, rho_r)
    P = np.where(left, p_l, p_r)
    vel = np.zeros_like(loc)
    vel[:, 0] = np.where(left, v_l, v_r)
    cfg = fvs.Config(dim=2, box=(1.0, 1.0, 1.0), gamma=gamma, cfl=0.3, solver=fvs.HLLC, boundary=fvs.REFLECTIVE,
                     motion=fvs.FIXED, t_end=t_end, **kw)
    return ics._finish("tube2d", cfg, loc, rho, vel, P)


def ensemble_problems(count=512, seed=4, n_min=200, n_max=65536, t_end=0.05, precision=fvs.FP32_FLUX):
    rng = np.random.default_rng(seed)
    probs = []
    for k in range(count):
        n = int(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        rho_l, rho_r = np.exp(rng.uniform(np.log(0.1), np.log(10.0), 2))
        p_l, p_r = np.exp(rng.uniform(np.log(0.01), np.log(10.0), 2))
        v_l, v_r = rng.uniform(-1.0, 1.0, 2)
        gamma = 1.4 if rng.random() < 0.5 else 5.0 / 3.0
        sub = int(rng.integers(0, 1 << 30))
        if k % 2 == 0:
            probs.append(ics.riemann_tube_1d(n, rho_l, p_l, v_l, rho_r, p_r, v_r, gamma, seed=sub, t_end=t_end,
                                             precision=precision))
        else:
            side = max(4, int(round(np.sqrt(n))))
            probs.append(riemann_tube_2d(side, rho_l, p_l, v_l, rho_r, p_r, v_r, gamma, seed=sub, t_end=t_end,
                                         precision=precision))
    return probs


def run_ensemble(probs, max_launches=1 << 40, profile=False):
    """Advance every problem to its t_end in one Batch; returns (batch, stats)."""
    b = fvs.Batch.for_problems(probs)
    b.synchronize()
    if profile:
        b.profile(True)
    t0 = time.perf_counter()
    launches = b.run(max_launches)
    b.synchronize()
    t1 = time.perf_counter()
    infos = [b.engine_info(s) for s in range(len(probs))]
    return b, {
        "seconds": t1 - t0, "launches": launches,
        "cell_updates": int(sum(i["cell_updates"] for i in infos)),
        "face_fluxes": int(sum(i["face_fluxes"] for i in infos)),
        "steps": int(sum(i["step_count"] for i in infos)),
        "all_done": all(i["done"] for i in infos),
        "failed": [s for s, i in enumerate(infos) if i["error_code"]],
        "kernel_times": b.kernel_times() if profile else None,
    }


def run_plain(probs, max_steps=1 << 40, warmup=0):
    """Each problem alone through the single-simulation path (FP64 unless the problem says otherwise):
    `warmup` untimed steps, then up to `max_steps` timed steps.  A simulation that raises a hard error
    stops there (its completed steps count)."""
    cu = 0
    steps = 0
    secs = 0.0
    failed = []
    for k, p in enumerate(probs):
        d = ics.build_device_sim(p)
        try:
            if warmup:
                d.run(warmup)
            d.synchronize()
            e0 = d.engine_info()
            t0 = time.perf_counter()
            try:
                d.run(max_steps)
            except fvs.HfvError as e:
                failed.append((k, e.code))
            d.synchronize()
            secs += time.perf_counter() - t0
            e1 = d.engine_info()
            cu += e1["cell_updates"] - e0["cell_updates"]
            steps += e1["step_count"] - e0["step_count"]
        except fvs.HfvError as e:
            failed.append((k, e.code))
        finally:
            d.close()
    return {"seconds": secs, "cell_updates": cu, "steps": steps, "failed": failed}
