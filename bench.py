"""Benchmark of the MI355X finite-volume path.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--side 256] [--cpu-seconds S] [--no-cpu]

Workload: 2D Sedov blast, 256x256 generators on the unit periodic box (jitter +-0.05 dx), gamma 5/3,
HLLC, hierarchical timebins, fixed mesh.  One step is one OptimalOrderRunner::step (one engine sync
point) over the device-resident state; W untimed steps, then K timed steps bracketed by device
synchronisation.  value = cell updates (active cells summed over the K steps) / wall time.

Printed JSON line: the contract keys plus
  roofline     dominant kernel (by device time in the timed region, from HIP events on the context's
               stream): algorithmic bytes per launch / average launch time vs 8 TB/s
  cpu_baseline the oracle (C restatement of the reference path, OpenMP at the reference's rayon sites)
               on the host cores, over a bounded sample of the same steps
  fp32         the FP32-flux / FP64-accumulate path on the same steps, with L1(rho) vs FP64
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "cell-updates/s vs plain FP64 GPU path; p50/p99 job latency at L1 err<1e-5"

# Algorithmic HBM bytes per unit of work, FP64 layout (one read of every operand a phase needs and one
# write of every output, ideal caching within a phase):
#   flux kernel: per active cell 329 B (prim 40, grad 120, gc 24, extrap 40, centroid 24, loc 24, v 24,
#                v_rel 24, dt 8, active flag 1) + per computed face 121 B (area 8, normal 24, centroid 24,
#                left/right 8, shift code 1, flux record 56)
#   accumulate+c2p: per cell 89 B (flag 1, accumulator r/w 80, CSR offset 8) + per active cell 208 B
#                (conserved r/w 80, msv/maor r/w 32, volume 8, prim write 40, + 48 spare of the record's
#                v_max/a_over_r), per face record read 56 B x 2 sides + CSR entry 4 B x 2
#   gradients+timestep: per active cell 600 B, per face of an active cell 60 B
ALGO = {
    "flux": lambda n_act, f_act, n_all: 329 * n_act + 121 * f_act,
    "apply_c2p": lambda n_act, f_act, n_all: 89 * n_all + 208 * n_act + 120 * f_act,
    "grad_timestep": lambda n_act, f_act, n_all: 600 * n_act + 120 * f_act,
    "mask_predict": lambda n_act, f_act, n_all: 193 * n_all,
    "wakeup": lambda n_act, f_act, n_all: 18 * n_all + 8 * f_act,
    "limit_kick": lambda n_act, f_act, n_all: 19 * n_all + 80 * n_act,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_device(prob, warmup, steps, profile=True):
    from hydro_rs_amd import ics

    dev = ics.build_device_sim(prob)
    dev.run(warmup)
    dev.synchronize()
    e0 = dev.engine_info()
    if profile:
        dev.profile(True)
    t0 = time.perf_counter()
    done = dev.run(steps)
    dev.synchronize()
    t1 = time.perf_counter()
    e1 = dev.engine_info()
    kt = dev.kernel_times() if profile else None
    return dev, {
        "steps": done, "seconds": t1 - t0, "cell_updates": e1["cell_updates"] - e0["cell_updates"],
        "face_fluxes": e1["face_fluxes"] - e0["face_fluxes"], "kernel_times": kt, "t_end_reached": e1["done"],
        "step0": e0["step_count"],
    }


def roofline(stats, n_all):
    kt = stats["kernel_times"]
    name, rec = max(((k, v) for k, v in kt.items() if v["launches"] > 0), key=lambda kv: kv[1]["ms"])
    launches = rec["launches"]
    avg_s = rec["ms"] / 1e3 / launches
    steps = max(stats["steps"], 1)
    n_act = stats["cell_updates"] / steps
    f_act = stats["face_fluxes"] / steps
    algo = ALGO.get(name, lambda *_: 0)(n_act, f_act, n_all)
    achieved = algo / avg_s / 1e9
    return {
        "kernel": name, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "avg_launch_us": round(avg_s * 1e6, 3),
        "algorithmic_bytes_per_launch": int(algo), "launches": launches,
        "per_kernel_ms": {k: round(v["ms"], 3) for k, v in kt.items() if v["launches"]},
    }


def cpu_baseline(prob, warmup, budget_s):
    from oracle import ora  # checker / CPU baseline only
    from tests.helpers import oracle_sim

    threads = ora.set_threads(0)
    sim = oracle_sim(prob)
    sim.advance(warmup)
    i0 = sim.info()
    cu = 0
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < budget_s:
        if sim.advance(1) == 0:
            break
        cu += sim.info()["last_n_active"]
        steps += 1
    dt = time.perf_counter() - t0
    return {"value": cu / dt if dt > 0 else 0.0, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of the same run after step {i0['step_count']} ({cu} cell updates, "
                      f"{dt:.1f} s, OpenMP threads={threads})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--side", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only run a plain device copy of this many MiB (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.copy_probe > 0:
        import torch

        nbytes = int(args.copy_probe * 2 ** 20)
        x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": nbytes, "copy_probe_bytes_written": nbytes, "launches": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (each job runs on one GPU, "
              "see DESIGN.md for how the path shards)", file=sys.stderr)
        sys.exit(2)

    from hydro_rs_amd import fvs, ics

    if fvs.device_count() < 1:
        raise SystemExit("no HIP device visible")
    log(f"[bench] building sedov2d side={args.side}")
    prob = ics.sedov_2d(side=args.side, t_end=0.05)
    n = prob.n
    log(f"[bench] N={n} F={prob.mesh.n_faces}; FP64 warmup={args.warmup} steps={args.steps}")
    dev, s64 = run_device(prob, args.warmup, args.steps)
    rl = roofline(s64, n)
    value = s64["cell_updates"] / s64["seconds"]
    log(f"[bench] fp64: {s64['steps']} steps in {s64['seconds']:.3f}s -> {value:.4e} cell-updates/s; "
        f"dominant {rl['kernel']} {rl['achieved']} GB/s")
    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": s64["steps"],
        "warmup": args.warmup, "ms_per_step": 1e3 * s64["seconds"] / max(s64["steps"], 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"sedov2d_{args.side}x{args.side}_periodic_hllc_fp64", "cells": n,
                   "faces": prob.mesh.n_faces, "gamma": 5.0 / 3.0, "t_end": 0.05, "timebins": "hierarchical",
                   "parallelism": "single GPU"},
        "face_fluxes_per_s": s64["face_fluxes"] / s64["seconds"],
        "cell_updates": s64["cell_updates"],
        "roofline": rl,
    }
    if not args.no_fp32:
        p32 = ics.sedov_2d(side=args.side, t_end=0.05, precision=fvs.FP32_FLUX)
        d32, s32 = run_device(p32, args.warmup, args.steps, profile=False)
        r64 = dev.get_state(["prim"])["prim"][:n]
        r32 = d32.get_state(["prim"])["prim"][:n]
        out["fp32_flux"] = {"value": s32["cell_updates"] / s32["seconds"], "unit": "cell-updates/s",
                            "steps": s32["steps"], "l1_rho_vs_fp64": float(np.mean(np.abs(r32 - r64)))}
        log(f"[bench] fp32: {out['fp32_flux']}")
    if not args.no_cpu:
        log(f"[bench] cpu baseline (oracle), budget {args.cpu_seconds}s")
        out["cpu_baseline"] = cpu_baseline(prob, args.warmup, args.cpu_seconds)
        log(f"[bench] cpu: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
