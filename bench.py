"""Benchmark of the MI355X finite-volume path (one GPU).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload all|ensemble|sedov|kh3d|serve] ...

Headline (value): the batched ensemble -- 512 independent 1D/2D Riemann-problem simulations (cell
counts log-uniform in [200, 65536], HLLC, FP32 flux / FP64 accumulate) in one device context.  One
step is one batch launch: every unfinished member advances by one engine step (one
OptimalOrderRunner::step with its own hierarchical timebins).  The autotuner first times a few real
launches per size class and candidate (flux tiling, gradient geometry; bit-identical candidates), then
W untimed launches (more if needed for the window to start at launch 260, whatever the tuner's
length; `warmup` reports the window's start, `warmup_requested` W), then K timed launches, replayed
from the captured step graph with no timing events, bracketed by device synchronisation; value = cell
updates (active cells summed over members and launches) / wall time of the K launches.

Printed JSON line: the contract keys plus
  speedup_vs_plain_fp64  value / the plain single-simulation FP64 device path on the same members and
                         step windows (members 0, 16, 32, ... each run alone, graph replays)
  l1_vs_oracle           L1(rho) of sampled FP32-flux members against the FP64 oracle advanced to the
                         member's step count (at equal time)
  l1_gate                L1(rho) of every FP32-flux member against the FP64 batch on the device
  roofline               dominant kernel of the timed window: the step byte model's algorithmic bytes of that
                         kernel per launch (ALGO below) / its mean duration from HIP events on the
                         context's stream, measured on a replica context replaying the same window
  cpu_baseline           the oracle (C restatement of the reference path, OpenMP at the reference's
                         rayon sites) on all host cores, over a bounded sample of the same members,
                         with the host CPU model and a shorter one-thread sample (single_thread)
  sedov2d / kh3d / serve the 2D Sedov blast (plain FP64 and FP32, L1 against the oracle at t_end), the
                         3D Kelvin-Helmholtz layer (plain FP64 with its step roofline and FP32; L1
                         against the oracle at a fixed time) and the served request stream over a sweep
                         of arrival rates up to saturation (p50/p99 job latency, cell updates/s, L1
                         against the oracle)
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

# Algorithmic HBM bytes of one OptimalOrderRunner::step on the FP64 SoA (the step byte model,
#   B_step = 192 N_all + 1247 N_act + 248 F_act (+8 F_act periodic)),
# attributed to the kernels that do each part of the step:
#   mask_predict   192 / cell      predict of every cell: rho, P, 9 gradient components, v_rel; extrap r/w
#   flux           329 / act cell  phase-1 gather (prim 40, grad 120, gc 24, extrap 40, centroid 24, loc 24,
#                                  v 24, v_rel 24, dt 8, bin 1)
#                  120 / act face  geometry 64 (area, normal, centroid, left/right) + flux record write 56
#                                  (+8 periodic shift)
#   apply_c2p      281 / act cell  accumulate 193 (accumulator r/w 80, conserved r/w 80, msv/maor r/w 32,
#                                  flag 1) + conserved -> primitive 88
#                   64 / act face  flux record read 56 + CSR 8
#   grad_timestep  557 / act cell  gradients 208 (neighbour prim + centroid 64, gradient write 120, gc write
#                                  24) + the timestep share of the 429 B timestep/limiter/reset/kick1 part
#                                  that k_limit does not carry (349)
#                   48 / act face  CSR 8 + gradient-phase geometry re-read 40
#   wakeup          16 / act face  CSR 8 + limiter neighbour ids 8
#   limit_kick      80 / act cell  reset of the extrapolations and kick1
# Sum: 192 N + 1247 N_act + 248 F_act.  FP32-flux contexts store gradients, extrapolations and face
# records as floats: the same terms at 4-byte widths for those fields (gradients 120 -> 60,
# extrapolations 40 -> 20, records 56 -> 28), i.e. 116 N + 1087 N_act + 192 F_act (+8 F_act periodic).
ALGO = {
    "mask_predict": lambda n_act, f_act, n_all: 192 * n_all,
    "flux": lambda n_act, f_act, n_all: 329 * n_act + 120 * f_act,
    "apply_c2p": lambda n_act, f_act, n_all: 281 * n_act + 64 * f_act,
    "grad_timestep": lambda n_act, f_act, n_all: 557 * n_act + 48 * f_act,
    "wakeup": lambda n_act, f_act, n_all: 16 * f_act,
    "limit_kick": lambda n_act, f_act, n_all: 80 * n_act,
}
ALGO_FP32 = {
    "mask_predict": lambda n_act, f_act, n_all: 116 * n_all,
    "flux": lambda n_act, f_act, n_all: 249 * n_act + 92 * f_act,
    "apply_c2p": lambda n_act, f_act, n_all: 281 * n_act + 36 * f_act,
    "grad_timestep": lambda n_act, f_act, n_all: 497 * n_act + 48 * f_act,
    "wakeup": ALGO["wakeup"],
    "limit_kick": lambda n_act, f_act, n_all: 60 * n_act,
}
PERIODIC_SHIFT = 8  # bytes per active face of a periodic mesh (the flux phase's shift read)


def step_bytes(n_all, n_act, f_act, fp32=False, periodic=False):
    """B_step of the step byte model at the context's storage widths."""
    t = ALGO_FP32 if fp32 else ALGO
    return sum(f(n_act, f_act, n_all) for f in t.values()) + (PERIODIC_SHIFT * f_act if periodic else 0)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def roofline(kt, launches_timed, cell_updates, face_fluxes, n_all, fp32=False, periodic=False):
    """Dominant kernel (most device time) of a timed window; bytes from ALGO (ALGO_FP32 for FP32-flux
    contexts) at the window's mean active counts per launch."""
    name, rec = max(((k, v) for k, v in kt.items() if v["launches"] > 0), key=lambda kv: kv[1]["ms"])
    launches = rec["launches"]
    avg_s = rec["ms"] / 1e3 / launches
    steps = max(launches_timed, 1)
    n_act = cell_updates / steps
    f_act = face_fluxes / steps
    algo = (ALGO_FP32 if fp32 else ALGO).get(name, lambda *_: 0)(n_act, f_act, n_all)
    if name == "flux" and periodic:
        algo += PERIODIC_SHIFT * f_act
    achieved = algo / avg_s / 1e9
    return {
        "kernel": name, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "avg_launch_us": round(avg_s * 1e6, 3),
        "algorithmic_bytes_per_launch": int(algo), "launches": launches,
        "mean_active_cells": round(n_act, 1), "mean_active_faces": round(f_act, 1), "cells": int(n_all),
        "byte_model": "step byte model B_step per kernel (bench.py ALGO%s)" % ("_FP32" if fp32 else ""),
        "per_kernel_ms": {k: round(v["ms"], 3) for k, v in kt.items() if v["launches"]},
    }


def _l1(a, b):
    return float(np.mean(np.abs(np.asarray(a) - np.asarray(b))))


# ------------------------------------------------------------------ rocprofv3 summaries
PROFILES = os.path.join(ROOT, "profiles")


def _counter_rows(d):
    import csv
    import glob

    rows = []
    for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        rows += list(csv.DictReader(open(path)))
    return rows


def summarize_pmc(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, label, command, copy_mib, window=0):
    """profiles/<label>_pmc.md: HBM bytes per dispatch of every kernel from a FETCH_SIZE run and a
    WRITE_SIZE run (KiB per dispatch), each scaled by what the counter reads for a plain device copy of
    copy_mib MiB (true bytes / counted bytes), plus a TRAFFIC_JSON line that bench.py reads back.  With
    a window, each kernel's figures are the mean over its last `window` dispatches (the timed window of
    a run that ends with it)."""
    def per_kernel(d, counter, win=0):
        rows = [r for r in _counter_rows(d) if r["Counter_Name"] == counter]
        if win > 0:
            return _last_dispatches(rows, win, lambda r: float(r["Counter_Value"]) * 1024.0)
        acc = {}
        for r in rows:
            acc.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]) * 1024.0)
        return acc

    true_bytes = copy_mib * 2 ** 20
    cf = per_kernel(copy_fetch_dir, "FETCH_SIZE")
    cw = per_kernel(copy_write_dir, "WRITE_SIZE")
    big = lambda acc: max(v for vals in acc.values() for v in vals)  # the copy dispatch of copy_mib MiB
    f_scale = true_bytes / big(cf)
    w_scale = true_bytes / big(cw)
    fetch = per_kernel(fetch_dir, "FETCH_SIZE", window)
    write = per_kernel(write_dir, "WRITE_SIZE", window)
    table = {}
    for k in fetch:
        if k not in write or not k.startswith(("void hfv::", "hfv::")):
            continue
        fb = float(np.mean(fetch[k])) * f_scale
        wb = float(np.mean(write[k])) * w_scale
        table[k] = {"dispatches": len(fetch[k]), "fetch_bytes": fb, "write_bytes": wb, "traffic_bytes": fb + wb}
    lines = [f"# HBM traffic per dispatch: {label}", "", "Commands (run from /tmp with TMPDIR=/tmp on the MI355X host, R = repository root):", ""]
    lines += [f"    {c}" for c in command]
    lines += ["", f"Calibration: a {copy_mib} MiB device copy reads FETCH_SIZE x {f_scale:.3f} and WRITE_SIZE x "
                  f"{w_scale:.3f} of its true bytes; every kernel's counters below are scaled by these factors."]
    if window > 0:
        lines += ["", f"Per kernel: the mean over its last {window} dispatches (the run's final timed window)."]
    lines += ["", "| kernel | dispatches | HBM read MB | HBM write MB | total MB |", "|---|---|---|---|---|"]
    for k, v in sorted(table.items(), key=lambda kv: -kv[1]["traffic_bytes"]):
        lines.append(f"| `{k}` | {v['dispatches']} | {v['fetch_bytes'] / 1e6:.1f} | {v['write_bytes'] / 1e6:.1f} | "
                     f"{v['traffic_bytes'] / 1e6:.1f} |")
    lines += ["", "TRAFFIC_JSON " + json.dumps({"label": label, "kernels": table})]
    out = os.path.join(PROFILES, f"{label}_pmc.md")
    open(out, "w").write("\n".join(lines) + "\n")
    return out


def _last_dispatches(rows, window, key):
    """Per kernel, the `key` values of its last `window` dispatches (by dispatch id): in a run that ends
    with its timed window, the dispatches of that window."""
    per = {}
    for r in sorted(rows, key=lambda r: int(r["Dispatch_Id"])):
        per.setdefault(r["Kernel_Name"], []).append(key(r))
    return {k: v[-window:] for k, v in per.items()}


def summarize_trace(stats_csv, label, command, note="", trace_csv=None, window=0):
    """profiles/<label>_kernel_stats.md from rocprofv3 --kernel-trace --stats (kernel_stats.csv); with
    the per-dispatch trace and a window, also each kernel's mean duration over its last `window`
    dispatches (the timed window of a run that ends with it)."""
    import csv

    rows = list(csv.DictReader(open(stats_csv)))
    win = {}
    if trace_csv and window > 0:
        win = _last_dispatches(list(csv.DictReader(open(trace_csv))), window,
                               lambda r: int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    wcol = f" last {window} avg us |" if win else ""
    lines = [f"# Kernel stats: {label}", "", "Command (run from /tmp with TMPDIR=/tmp on the MI355X host, R = repository root):", "",
             f"    {command}", "", "| kernel | calls | total us | avg us | min us | max us | % |" + wcol,
             "|---|---|---|---|---|---|---|" + ("---|" if win else "")]
    for r in rows:
        w = win.get(r["Name"])
        wv = (f" {np.mean(w) / 1e3:.2f} |" if w else " |") if win else ""
        lines.append(f"| `{r['Name']}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e3:.1f} | "
                     f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | "
                     f"{float(r['Percentage']):.2f} |" + wv)
    if win:
        lines += ["", f"`last {window} avg us`: mean duration of each kernel's last {window} dispatches, i.e. over "
                      "the run's final timed window (the launches bench.py times with HIP events)."]
    if note:
        lines += ["", note]
    out = os.path.join(PROFILES, f"{label}_kernel_stats.md")
    open(out, "w").write("\n".join(lines) + "\n")
    return out


def traffic_for(label, kernel_name):
    """HBM bytes per dispatch of kernel_name from profiles/<label>_pmc.md (None when not profiled)."""
    path = os.path.join(PROFILES, f"{label}_pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("TRAFFIC_JSON "):
            rec = json.loads(line[len("TRAFFIC_JSON "):]).get("kernels", {}).get(kernel_name)
            return rec["traffic_bytes"] if rec else None
    return None


def flux_kernel_name(fp32, solver, static_geometry=True, fpt=1):
    """The flux kernel's name as the profiler reports it (fpt: face chunks per block, the launch shape)."""
    return (f"void hfv::k_flux<{'float' if fp32 else 'double'}, {solver}, "
            f"{'true' if static_geometry else 'false'}, {fpt or 1}>(hfv::Batch)")


SLOT_KERNELS = {"apply_c2p": ["hfv::k_apply(hfv::Batch, int)"], "mask_predict": ["hfv::k_mask(hfv::Batch, int)"],
                "wakeup": ["hfv::k_wakeup(hfv::Batch, int)"], "limit_kick": ["hfv::k_limit(hfv::Batch, int)"]}


def slot_kernels(slot, fp32, solver, tuning, fpt=1):
    """Kernel names that a timing slot's launches consist of (static geometry; the gradient variants
    the autotuner chose)."""
    if slot == "flux":
        return [flux_kernel_name(fp32, solver, True, fpt)]
    if slot == "grad_timestep":
        chosen = {c["static_grad"] for c in (tuning or [])} or {0}
        return [f"void hfv::k_grad<{'true' if g else 'false'}>(hfv::Batch, int, int)" for g in sorted(chosen)]
    return SLOT_KERNELS.get(slot, [])


def attach_traffic(rl, label, fp32, solver, tuning, fpt=1):
    """roofline["traffic"]: HBM bytes per launch of the dominant slot's kernels from profiles/."""
    names = slot_kernels(rl["kernel"], fp32, solver, tuning, fpt)
    vals = [traffic_for(label, k) for k in names]
    rl["kernel_name"] = " + ".join(names)
    rl["traffic"] = sum(vals) if vals and all(v is not None for v in vals) else None


# ------------------------------------------------------------------ tuning taken from an earlier run
def load_report(path):
    """The JSON line of an earlier bench.py run (the file's last non-empty line)."""
    lines = [ln for ln in open(path).read().splitlines() if ln.strip()]
    return json.loads(lines[-1])


def apply_tuning(sim, members, classes, shape=None):
    """Set each slot's tuning to the choice the autotuner made for its size class (dimension and cell
    count in [lo, hi)); members: (slot, dim, cells).  shape: the tuner's flux launch shape (faces per
    thread), if it chose one."""
    if shape:
        sim.set_launch_shape(shape)
    for slot, dim, n in members:
        for c in classes or []:
            if c["dim"] == dim and c["cells"][0] <= n < c["cells"][1]:
                sim.set_tuning(slot, c["staged"], c["static_grad"], c["gstaged"])
                break


# ------------------------------------------------------------------ ensemble (headline)
def window_warmup(window_start, warmup, tune_launches, replay):
    """Untimed launches to run after tuning (or, replaying an earlier run's tuning, from launch 0) so
    that the timed window starts at launch max(window_start, tune_launches + warmup); returns
    (launches to run, window start)."""
    start = max(window_start, tune_launches + warmup)
    return (start if replay else start - tune_launches), start


def bench_ensemble(args):
    from hydro_rs_amd import ensemble, fvs, ics

    t = time.perf_counter()
    probs = ensemble.ensemble_problems(count=args.members, seed=4, precision=fvs.FP32_FLUX)
    n_all = sum(p.n for p in probs)
    members = [(s, p.cfg.dim, p.n) for s, p in enumerate(probs)]
    log(f"[ensemble] {len(probs)} members, {n_all} cells, generated in {time.perf_counter() - t:.1f}s")
    b = fvs.Batch.for_problems(probs)
    # autotuner (per size class: flux tiling, gradient geometry); its launches are ordinary steps
    tune_launches = 0
    tuning = None
    shape = None
    # the timed window starts at launch max(window_start, tuning + W), whatever the tuner's length
    if args.tuning_from:
        # the earlier run's choices, and its tuner's launches as warmup: every candidate gives the same
        # bits, so the timed window holds the same steps as that run's
        rep = load_report(args.tuning_from)["autotune"]
        shape = (rep.get("launch_shape") or {}).get("flux_faces_per_thread")
        apply_tuning(b, members, rep["classes"], shape)
        tune_launches = rep["launches"]
        tuning = rep["classes"]
        warmup, start = window_warmup(args.window_start, args.warmup, tune_launches, True)
        log(f"[ensemble] tuning from {args.tuning_from}; {warmup} warmup launches")
    elif args.tune_launches > 0:
        s0 = [b.engine_info(s)["step_count"] for s in range(len(probs))]
        t = time.perf_counter()
        tuning = b.autotune(launches=args.tune_launches, rounds=args.tune_rounds)
        shape = b.launch_shape()["flux_faces_per_thread"]
        tune_launches = max(b.engine_info(s)["step_count"] - s0[s] for s in range(len(probs)))
        log(f"[ensemble] autotune: {tune_launches} launches in {time.perf_counter() - t:.1f}s: "
            f"{[(c['dim'], c['cells'], c['staged'], c['static_grad'], c['gstaged']) for c in tuning]}")
        warmup, start = window_warmup(args.window_start, args.warmup, tune_launches, False)
    else:
        warmup, start = window_warmup(args.window_start, args.warmup, 0, False)
    b.run(warmup)
    b.synchronize()
    i0 = [b.engine_info(s) for s in range(len(probs))]
    # the timed window: graph replays, no timing events (as a production run executes)
    t0 = time.perf_counter()
    launches = b.run(args.steps)
    b.synchronize()
    secs = time.perf_counter() - t0
    i1 = [b.engine_info(s) for s in range(len(probs))]
    cu = sum(a["cell_updates"] - z["cell_updates"] for a, z in zip(i1, i0))
    ff = sum(a["face_fluxes"] - z["face_fluxes"] for a, z in zip(i1, i0))
    value = cu / secs
    log(f"[ensemble] fp32: {launches} launches in {secs:.3f}s -> {value:.4e} cell-updates/s")
    # per-kernel device times over the same window: a replica context with the same tuning replays the
    # first `start` launches (bit-identical states), then the window's launches with HIP events on its
    # stream around every kernel
    kt, rep_cu = kernel_times_replay(lambda: fvs.Batch.for_problems(probs), members, tuning, start, launches,
                                     shape)
    assert rep_cu == cu, "the replica window must hold the timed window's steps"
    res = {"launches": launches, "seconds": secs, "cell_updates": cu, "face_fluxes": ff, "value": value,
           "window_start": start,
           "autotune": {"launches": tune_launches, "classes": tuning, "launch_shape": b.launch_shape()},
           "n_all": n_all, "members_done": sum(1 for i in i1 if i["done"]),
           "members_failed": [s for s, i in enumerate(i1) if i["error_code"]],
           "roofline": roofline(kt, launches, cu, ff, n_all, fp32=True)}
    attach_traffic(res["roofline"], "ensemble512", True, fvs.HLLC, tuning,
                   b.launch_shape()["flux_faces_per_thread"])
    if args.no_gate:  # a profiling run: the timed window only
        res.update({"fp64_batched_value": None, "l1_gate": None, "plain_fp64": None, "speedup_vs_plain_fp64": None,
                    "l1_vs_oracle": None})
        b.close()
        return res, probs
    # L1(rho) against the FP64 oracle: sampled members advanced by the oracle to the batch member's own
    # step count, compared where both stand at the same time
    res["l1_vs_oracle"] = None if args.no_oracle else \
        l1_members_vs_oracle(b, probs, i1, range(0, len(probs), args.oracle_stride))
    log(f"[ensemble] L1 vs oracle: {res['l1_vs_oracle']}")
    # the same window in FP64 (batched) for the L1 gate over every member
    p64 = [ics.with_precision(p, fvs.FP64) for p in probs]
    b64 = fvs.Batch.for_problems(p64)
    t1 = time.perf_counter()
    b64.run(start + launches)
    b64.synchronize()
    res["fp64_batched_value"] = sum(b64.engine_info(s)["cell_updates"] for s in range(len(p64))) / (time.perf_counter() - t1)
    l1s, same_t = [], 0
    for s, p in enumerate(probs):
        e32, e64 = i1[s], b64.engine_info(s)
        if e32["error_code"] or e64["error_code"]:
            continue
        if e32["ti_current"] != e64["ti_current"]:
            continue
        same_t += 1
        r32 = b.get_state(s, ["prim"])["prim"][: p.n]
        r64 = b64.get_state(s, ["prim"])["prim"][: p.n]
        l1s.append(_l1(r32, r64))
    res["l1_gate"] = {"members_compared": same_t, "max_l1_rho": max(l1s) if l1s else None,
                      "mean_l1_rho": float(np.mean(l1s)) if l1s else None,
                      "pass": bool(l1s) and max(l1s) < 1e-5, "against": "FP64 batch on the device, every member"}
    log(f"[ensemble] L1 gate: {res['l1_gate']}")
    b64.close()
    b.close()
    # plain single-simulation FP64 path on a sample of members, same windows (graph replays)
    idx = list(range(0, len(probs), args.plain_stride))
    st = ensemble.run_plain([p64[i] for i in idx], max_steps=args.steps, warmup=start)
    res["plain_fp64"] = {"value": st["cell_updates"] / st["seconds"], "members": idx, "steps": st["steps"],
                         "seconds": st["seconds"], "failed": st["failed"]}
    res["speedup_vs_plain_fp64"] = value / res["plain_fp64"]["value"]
    log(f"[ensemble] plain fp64 sample: {res['plain_fp64']['value']:.4e} cell-updates/s -> "
        f"speedup {res['speedup_vs_plain_fp64']:.1f}x")
    return res, probs


def kernel_times_replay(make, members, tuning, start, launches, shape=None):
    """Per-kernel device times (HIP events on the context's stream) of launches [start, start + launches)
    of a fresh replica with the given tuning; returns (kernel_times, cell updates of that window)."""
    r = make()
    if tuning or shape:
        apply_tuning(r, members, tuning, shape)
    r.run(start)
    r.synchronize()
    u0 = sum(r.engine_info(s)["cell_updates"] for s, _, _ in members)
    r.profile(True)
    r.run(launches)
    r.synchronize()
    kt = r.kernel_times()
    r.profile(False)
    u1 = sum(r.engine_info(s)["cell_updates"] for s, _, _ in members)
    r.close()
    return kt, u1 - u0


def l1_members_vs_oracle(batch, probs, infos, idx):
    """L1(rho) of FP32-flux batch members against the FP64 oracle advanced to the member's step count;
    members that stopped on a hard error or stand at another time than the oracle are skipped."""
    from hydro_rs_amd import fvs, ics
    from tests.helpers import oracle_sim  # the oracle is the reference here, not the measured path

    l1s, skipped = {}, []
    for s in idx:
        e = infos[s]
        if e["error_code"]:
            skipped.append(s)
            continue
        osim = oracle_sim(ics.with_precision(probs[s], fvs.FP64))
        try:
            osim.advance(e["step_count"])
        except RuntimeError:
            skipped.append(s)
            continue
        if osim.info()["ti_current"] != e["ti_current"]:
            skipped.append(s)
            continue
        ref = osim.get_fields()["prim"][: probs[s].n]
        l1s[s] = _l1(batch.get_state(s, ["prim"])["prim"][: probs[s].n], ref)
    vals = list(l1s.values())
    return {"members_compared": len(vals), "members": sorted(l1s), "skipped": skipped,
            "max_l1_rho": max(vals) if vals else None, "mean_l1_rho": float(np.mean(vals)) if vals else None,
            "pass": bool(vals) and max(vals) < 1e-5}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_ensemble(probs, stride, warmup, budget_s):
    """The oracle over members 0, stride, 2*stride, ... (FP64, as the reference), each for up to
    budget/len steps' worth of time after `warmup` untimed steps, on all host threads (OpenMP at the
    reference's rayon sites, as rayon uses every logical CPU); then a shorter sample on one thread."""
    from oracle import ora  # timed CPU baseline only
    from tests.helpers import oracle_sim
    from hydro_rs_amd import fvs, ics

    def sample(idx, per, threads):
        cu, secs, steps = 0, 0.0, 0
        for i in idx:
            sim = oracle_sim(ics.with_precision(probs[i], fvs.FP64))
            ora.set_threads(0)  # warmup on every thread, whatever the timed count
            try:
                sim.advance(warmup)
            except RuntimeError:
                continue
            ora.set_threads(threads)
            t0 = time.perf_counter()
            while time.perf_counter() - t0 < per:
                try:
                    if sim.advance(1) == 0:
                        break
                except RuntimeError:
                    break
                cu += sim.info()["last_n_active"]
                steps += 1
            secs += time.perf_counter() - t0
        ora.set_threads(0)
        return cu, secs, steps

    threads = ora.set_threads(0)
    idx = list(range(0, len(probs), stride))
    per = budget_s / len(idx)
    cu, secs, steps = sample(idx, per, 0)
    idx1 = idx[::4]
    per1 = budget_s / 3 / len(idx1)
    cu1, secs1, steps1 = sample(idx1, per1, 1)
    return {"value": cu / secs if secs > 0 else 0.0, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"members {idx[0]}..{idx[-1]} every {stride} ({len(idx)}) of the ensemble, each for "
                      f"{per:.2f}s after {warmup} untimed steps: {steps} steps, {cu} cell updates in {secs:.1f}s, "
                      f"OpenMP threads={threads}",
            "cpu_model": _cpu_model(),
            "single_thread": {"value": cu1 / secs1 if secs1 > 0 else 0.0, "cores": 1,
                              "sample": f"members {idx1[0]}..{idx1[-1]} every {4 * stride} ({len(idx1)}), each for "
                                        f"{per1:.2f}s after the same warmup: {steps1} steps, {cu1} cell updates "
                                        f"in {secs1:.1f}s"}}


# ------------------------------------------------------------------ single simulations
def run_device(prob, warmup, steps, profile=False, tune=0, tuning_from=None, tune_rounds=1):
    """W untimed steps, then K timed ones (graph replays, no timing events), of a fresh context.  tune > 0:
    the autotuner runs first on a separate context, after the same W steps (so it samples the phase of the
    run that is timed), and its choice is applied, so the timed window is the same steps whatever the
    tuner's length; tuning_from: (classes, launches) of an earlier run's tuner, applied the same way.
    profile: per-kernel device times of the same window from a replica context (same tuning, the same W
    steps first, then the window with HIP events around every kernel)."""
    from hydro_rs_amd import ics

    tuning, tune_launches, shape = None, 0, None
    if tuning_from:
        tuning, tune_launches, shape = tuning_from
    elif tune > 0:
        probe = ics.build_device_sim(prob)
        probe.run(warmup)
        tuning = probe.autotune(launches=tune, rounds=tune_rounds)
        shape = probe.launch_shape()["flux_faces_per_thread"]
        tune_launches = probe.engine_info()["step_count"] - warmup
        probe.close()

    def fresh():
        d = ics.build_device_sim(prob)
        if tuning or shape:
            apply_tuning(d, [(0, prob.cfg.dim, prob.n)], tuning, shape)
        d.run(warmup)
        d.synchronize()
        return d

    dev = fresh()
    e0 = dev.engine_info()
    t0 = time.perf_counter()
    done = dev.run(steps)
    dev.synchronize()
    t1 = time.perf_counter()
    e1 = dev.engine_info()
    kt = None
    if profile:
        r = fresh()
        r.profile(True)
        r.run(done)
        r.synchronize()
        kt = r.kernel_times()
        assert r.engine_info()["cell_updates"] == e1["cell_updates"], "replica window differs"
        r.close()
    return dev, {
        "steps": done, "seconds": t1 - t0, "cell_updates": e1["cell_updates"] - e0["cell_updates"],
        "face_fluxes": e1["face_fluxes"] - e0["face_fluxes"], "kernel_times": kt, "done": e1["done"],
        "ti_current": e1["ti_current"], "autotune": tuning, "tune_launches": tune_launches,
        "launch_shape": shape,
    }


def oracle_run(prob, steps=1 << 40):
    """The FP64 oracle (OpenMP on every host thread) advanced by `steps` engine steps: (density, info)."""
    from hydro_rs_amd import fvs, ics
    from tests.helpers import oracle_sim  # the oracle is the reference here, not the measured path

    osim = oracle_sim(ics.with_precision(prob, fvs.FP64))
    osim.advance(steps)
    return osim.get_fields()["prim"][: prob.n], osim.info()


def bench_sedov(args):
    from hydro_rs_amd import fvs, ics

    prob = ics.sedov_2d(side=args.sedov_side, t_end=0.05)
    log(f"[sedov] {prob.n} cells, {prob.mesh.n_faces} faces; FP64 and FP32 flux to t_end")
    d64, s64 = run_device(prob, 0, 1 << 40)
    d32, s32 = run_device(ics.with_precision(prob, fvs.FP32_FLUX), 0, 1 << 40)
    r64 = d64.get_state(["prim"])["prim"][: prob.n]
    r32 = d32.get_state(["prim"])["prim"][: prob.n]
    out = {"workload": f"sedov2d_{args.sedov_side}x{args.sedov_side}_periodic_hllc_t0.05", "cells": prob.n,
           "faces": prob.mesh.n_faces}
    for k, s in (("fp64", s64), ("fp32_flux", s32)):
        out[k] = {"cell_updates_per_s": s["cell_updates"] / s["seconds"],
                  "face_fluxes_per_s": s["face_fluxes"] / s["seconds"], "steps": s["steps"],
                  "us_per_step": 1e6 * s["seconds"] / max(s["steps"], 1), "reached_t_end": bool(s["done"])}
    out["l1_rho_fp32_vs_fp64_at_t_end"] = _l1(r32, r64)
    if not args.no_oracle:
        log("[sedov] the FP64 oracle to t_end (L1 check)")
        t = time.perf_counter()
        ref, oi = oracle_run(prob)
        out["l1_vs_oracle"] = {"fp32_flux": _l1(r32, ref), "fp64": _l1(r64, ref), "at": "t_end 0.05",
                               "oracle_steps": oi["step_count"], "oracle_seconds": time.perf_counter() - t,
                               "pass": _l1(r32, ref) < 1e-5}
    log(f"[sedov] {out}")
    d64.close()
    d32.close()
    return out, prob


def bench_moving(args, dim):
    """A moving mesh (steered generators, the reference's default particle motion): drift, device Voronoi
    rebuild and step per engine step.  2D: the Sedov blast; 3D: the Kelvin-Helmholtz shear layer.  The
    rebuild's device time against the host tessellator's (OpenMP, every host thread) on the same
    generators; L1(rho) against the FP64 oracle (which re-tessellates with the host tessellator) at
    t = --moving-l1-t."""
    from dataclasses import replace

    from hydro_rs_amd import fvs, ics
    from hydro_rs_amd.mesh import VoronoiMesh

    tag = f"moving{dim}d"
    if dim == 2:
        side, steps, warmup = args.moving_side, args.moving_steps, args.moving_warmup
        base = ics.sedov_2d(side=side, t_end=0.05)
        workload = f"sedov2d_{side}x{side}_periodic_hllc_steer"
    else:
        side, steps, warmup = args.moving3d_side, args.moving3d_steps, args.moving3d_warmup
        base = ics.kelvin_helmholtz_3d(side=side, t_end=1.0, threads=args.threads)
        workload = f"kh3d_{side}^3_periodic_airs_steer"
    prob = replace(base, cfg=replace(base.cfg, motion=fvs.STEER))
    log(f"[{tag}] {prob.n} cells, steered generators, {warmup} + {steps} steps")
    out = {"workload": workload, "cells": prob.n, "warmup": warmup, "steps": steps}
    for k, prec in (("fp64", fvs.FP64), ("fp32_flux", fvs.FP32_FLUX)):
        d, st = run_device(ics.with_precision(prob, prec), warmup, steps, profile=True)
        kt = st["kernel_times"]
        per = {name: round(v["ms"] / max(v["launches"], 1) * 1e3, 2) for name, v in kt.items() if v["launches"]}
        out[k] = {"cell_updates_per_s": st["cell_updates"] / st["seconds"],
                  "us_per_step": 1e6 * st["seconds"] / max(st["steps"], 1),
                  "mean_active_cells": st["cell_updates"] / max(st["steps"], 1),
                  "per_kernel_us": per}
        if k == "fp64":
            faces = d.get_mesh()["n_faces"]
        d.close()
        log(f"[{tag}] {k}: {out[k]}")
    # a full rebuild (every cell active: timestep_sync) on the device, against the host tessellator's full
    # build of the same generator positions (the device rebuild's CPU counterpart)
    dfull, sfull = run_device(replace(prob, cfg=replace(prob.cfg, sync_all=True)), 2, 10 if dim == 2 else 4,
                              profile=True)
    loc_full = dfull.get_state(["loc"])["loc"].reshape(3, prob.n).T.copy()
    dfull.close()
    full_us = sfull["kernel_times"]["mesh_rebuild"]["ms"] / max(sfull["kernel_times"]["mesh_rebuild"]["launches"], 1) * 1e3
    host = []
    for _ in range(3):  # best of 3: a single call can catch the host busy
        t = time.perf_counter()
        VoronoiMesh(dim, prob.cfg.box, True, loc_full, n_threads=args.threads)
        host.append(time.perf_counter() - t)
    host_ms = 1e3 * min(host)
    out["rebuild"] = {"partial_device_us": out["fp64"]["per_kernel_us"].get("mesh_rebuild"),
                      "partial_faces": int(faces), "full_device_us": full_us, "full_host_tessellator_ms": host_ms, "host_timing": "best of 3 calls",
                      "host_threads": args.threads or int(os.environ.get("OMP_NUM_THREADS", os.cpu_count())),
                      "full_speedup_vs_host": 1e3 * host_ms / full_us if full_us else None,
                      "device_time_from": "HIP events around the rebuild's kernels (eager launches)"}
    if not args.no_oracle:
        l1_t = args.moving_l1_t if dim == 2 else args.moving3d_l1_t
        short = replace(prob, cfg=replace(prob.cfg, t_end=l1_t))
        log(f"[{tag}] FP64, FP32 flux and the FP64 oracle to t = {l1_t} (L1 check)")
        res = {}
        for k, prec in (("fp64", fvs.FP64), ("fp32_flux", fvs.FP32_FLUX)):
            dd = ics.build_device_sim(ics.with_precision(short, prec))
            dd.run()
            res[k] = dd.get_state(["prim"])["prim"][: prob.n]
            dd.close()
        t = time.perf_counter()
        ref, oi = oracle_run(short)
        out["l1_vs_oracle"] = {k: _l1(v, ref) for k, v in res.items()}
        out["l1_vs_oracle"].update({"at": f"t_end {l1_t}", "oracle_steps": oi["step_count"],
                                    "oracle_seconds": time.perf_counter() - t,
                                    "pass": _l1(res["fp32_flux"], ref) < 1e-5})
    log(f"[{tag}] {out}")
    return out


def bench_kh3d(args):
    from hydro_rs_amd import fvs, ics

    t = time.perf_counter()
    prob = ics.kelvin_helmholtz_3d(side=args.kh_side, t_end=1.0, threads=args.threads)
    log(f"[kh3d] {prob.n} cells, {prob.mesh.n_faces} faces, tessellated in {time.perf_counter() - t:.1f}s")
    out = {"workload": f"kh3d_{args.kh_side}^3_periodic_airs", "cells": prob.n, "faces": prob.mesh.n_faces,
           "warmup": args.kh_warmup, "steps": args.kh_steps}
    devs = {}
    rep = load_report(args.tuning_from).get("kh3d") if args.tuning_from else None
    precs = [("fp64", fvs.FP64), ("fp32_flux", fvs.FP32_FLUX)]
    if args.kh_precision != "both":  # profiling runs: one precision, so that a kernel's last dispatches are its window
        precs = [kp for kp in precs if kp[0].startswith(args.kh_precision)]
    for k, prec in precs:
        tf = (rep[k]["autotune"], rep[k]["tune_launches"], rep[k].get("flux_faces_per_thread")) if rep else None
        fp32 = prec == fvs.FP32_FLUX
        d, s = run_device(ics.with_precision(prob, prec), args.kh_warmup, args.kh_steps, profile=True,
                          tune=args.kh_tune, tuning_from=tf, tune_rounds=args.kh_tune_rounds)
        rl = roofline(s["kernel_times"], s["steps"], s["cell_updates"], s["face_fluxes"], prob.n, fp32=fp32,
                      periodic=True)
        attach_traffic(rl, f"kh3d_{args.kh_side}_{'fp32' if fp32 else 'fp64'}", fp32, fvs.AIRS, s["autotune"],
                       s["launch_shape"])
        n_act = s["cell_updates"] / max(s["steps"], 1)
        f_act = s["face_fluxes"] / max(s["steps"], 1)
        sec_step = s["seconds"] / max(s["steps"], 1)
        out[k] = {"cell_updates_per_s": s["cell_updates"] / s["seconds"],
                  "face_fluxes_per_s": s["face_fluxes"] / s["seconds"],
                  "ms_per_step": 1e3 * sec_step,
                  "step_bytes": step_bytes(prob.n, n_act, f_act, fp32=fp32, periodic=True),
                  "step_roofline_frac": step_bytes(prob.n, n_act, f_act, fp32=fp32, periodic=True) / sec_step
                  / (HBM_PEAK_GBS * 1e9),
                  "mean_active_cells": n_act, "mean_active_faces": f_act,
                  "roofline": rl, "ti_current": s["ti_current"], "autotune": s["autotune"],
                  "tune_launches": s["tune_launches"], "flux_faces_per_thread": s["launch_shape"]}
        devs[k] = d
        log(f"[kh3d] {k}: {out[k]['ms_per_step']:.2f} ms/step, {out[k]['cell_updates_per_s']:.3e} cell-updates/s, "
            f"step {out[k]['step_roofline_frac']:.3f} of HBM, dominant {rl['kernel']} {rl['achieved']} GB/s")
    for d in devs.values():
        d.close()
    if not args.no_oracle:
        # FP32-flux result against the FP64 oracle at a common time: both run to t_end = kh_l1_t
        t = time.perf_counter()
        short = ics.with_precision(prob, fvs.FP32_FLUX)
        from dataclasses import replace

        short = replace(short, cfg=replace(short.cfg, t_end=args.kh_l1_t))
        log(f"[kh3d] FP64, FP32 flux and the FP64 oracle to t = {args.kh_l1_t} (L1 check)")
        d32 = ics.build_device_sim(short)
        n32 = d32.run()
        r32 = d32.get_state(["prim"])["prim"][: prob.n]
        d32.close()
        d64 = ics.build_device_sim(ics.with_precision(short, fvs.FP64))
        n64 = d64.run()
        r64 = d64.get_state(["prim"])["prim"][: prob.n]
        d64.close()
        ref, oi = oracle_run(short)
        out["l1_vs_oracle"] = {"fp32_flux": _l1(r32, ref), "fp64": _l1(r64, ref), "at": f"t_end {args.kh_l1_t}",
                               "device_steps": {"fp32_flux": n32, "fp64": n64}, "oracle_steps": oi["step_count"],
                               "oracle_seconds": time.perf_counter() - t, "pass": _l1(r32, ref) < 1e-5}
        log(f"[kh3d] L1 vs oracle: {out['l1_vs_oracle']}")
    return out


def serve_once(args, rate, count, hbm_gb, seed=5):
    """One served stream: Poisson arrivals at `rate` jobs/s (rate <= 0: every job arrives at t = 0)."""
    from hydro_rs_amd import serve

    jobs = serve.served_jobs(count=count, rate=rate if rate > 0 else 1e12, seed=seed, side_min=32,
                             side_max=args.serve_side_max, sedov_t_end=args.serve_sedov_t_end,
                             vortex_t_end=args.serve_vortex_t_end, threads=args.threads)
    srv = serve.Server(hbm_budget_bytes=int(hbm_gb * 2 ** 30), launches_per_poll=args.serve_poll)
    done = srv.serve(jobs, fetch=("prim",), progress=lambda t: log(f"[serve] rate {rate}: {t}"))
    summ = serve.latency_summary(done, srv.wall)
    summ.update({"rate_jobs_per_s": rate if rate > 0 else "all at t=0", "offered_cells": int(sum(j.prob.n for j in jobs)),
                 "slots": len(srv.caps[0]), "hbm_budget_gb": hbm_gb,
                 "cell_updates_per_s_while_queued": srv.saturated_rate()})
    return summ, done


def bench_serve(args):
    """The served stream over a sweep of arrival rates up to saturation (every job queued at t = 0): p50/p99
    latency and cell-updates/s per rate; L1(rho) of sampled jobs against the FP64 oracle."""
    from hydro_rs_amd import fvs, ics

    rates = [float(r) for r in args.serve_rates.split(",")]
    table, last = [], None
    for r in rates:
        summ, done = serve_once(args, r, args.serve_jobs, args.serve_hbm_gb)
        log(f"[serve] rate {r}: {summ}")
        table.append(summ)
        last = done
    # saturation: more jobs than the slots of a smaller budget, all queued at t = 0, so that every slot a job
    # fits is busy until the queue drains; throughput while jobs wait in the queue
    sat, sat_done = serve_once(args, 0.0, args.serve_sat_jobs, args.serve_sat_hbm_gb, seed=6)
    log(f"[serve] saturation: {sat}")
    # the plain FP64 path on the same jobs: a size-stratified sample of the saturation stream, each job alone
    # in its own single-simulation context (hfv_run, graph replays) from t = 0 to t_end
    from hydro_rs_amd import ensemble

    sample = sorted(sat_done, key=lambda j: j.prob.n)[::args.serve_plain_stride]
    log(f"[serve] plain FP64 path on {len(sample)} of the saturation stream's jobs")
    plain = ensemble.run_plain([ics.with_precision(j.prob, fvs.FP64) for j in sample])
    plain_rate = plain["cell_updates"] / plain["seconds"] if plain["seconds"] > 0 else None
    sat_rate = sat["cell_updates_per_s_while_queued"]
    out = {"rates": table, "saturation": sat,
           "saturated_cell_updates_per_s": sat_rate,
           "plain_fp64_same_jobs": {"cell_updates_per_s": plain_rate, "jobs": len(sample),
                                    "cells": [int(j.prob.n) for j in sample], "steps": plain["steps"],
                                    "seconds": plain["seconds"], "failed": plain["failed"]},
           "saturated_vs_plain_fp64_same_jobs": (sat_rate / plain_rate) if sat_rate and plain_rate else None,
           "jobs_per_rate": args.serve_jobs, "hbm_budget_gb": args.serve_hbm_gb,
           "jobs_desc": f"Poisson arrivals, 50/50 Sedov (t_end {args.serve_sedov_t_end}) / isentropic vortex "
                        f"(t_end {args.serve_vortex_t_end}), side log-uniform in [32, {args.serve_side_max}], "
                        "gamma 1.4, FP32 flux / FP64 accumulate; the same job draw at every rate"}
    # L1 gate: the smallest jobs of each kind of the last stream, against the FP64 oracle at t_end
    if not args.no_oracle:
        l1s = {}
        for kind in ("sedov", "vortex"):
            for j in sorted((j for j in last if j.kind == kind and not j.error_code), key=lambda j: j.prob.n)[:args.serve_l1_jobs]:
                log(f"[serve] the FP64 oracle on job {j.jid} ({kind}, {j.prob.n} cells)")
                ref, _ = oracle_run(j.prob)
                l1s[j.jid] = {"kind": kind, "side": int(round(np.sqrt(j.prob.n))),
                              "l1_rho": _l1(j.result["prim"][: j.prob.n], ref)}
        vals = [v["l1_rho"] for v in l1s.values()]
        out["l1_vs_oracle"] = {"jobs": l1s, "max": max(vals) if vals else None,
                               "pass": bool(vals) and max(vals) < 1e-5}
    log(f"[serve] {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed batch launches of the ensemble")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--window-start", type=int, default=260,
                    help="ensemble: untimed launches (tuning included) before the timed window, at least")
    ap.add_argument("--workload", default="all",
                    choices=["all", "ensemble", "sedov", "moving2d", "moving3d", "kh3d", "serve"])
    ap.add_argument("--moving-side", type=int, default=256)
    ap.add_argument("--moving-steps", type=int, default=100)
    ap.add_argument("--moving-warmup", type=int, default=20)
    ap.add_argument("--moving-l1-t", type=float, default=0.002,
                    help="moving 2D workload: end time of the L1 check against the oracle")
    ap.add_argument("--moving3d-side", type=int, default=48)
    ap.add_argument("--moving3d-steps", type=int, default=20)
    ap.add_argument("--moving3d-warmup", type=int, default=5)
    ap.add_argument("--moving3d-l1-t", type=float, default=0.002,
                    help="moving 3D workload: end time of the L1 check against the oracle")
    ap.add_argument("--members", type=int, default=512)
    ap.add_argument("--plain-stride", type=int, default=16)
    ap.add_argument("--oracle-stride", type=int, default=64, help="ensemble members compared with the oracle")
    ap.add_argument("--tune-launches", type=int, default=2,
                    help="ensemble autotuner: launches per timed group of a candidate trial (0: no tuning)")
    ap.add_argument("--tune-rounds", type=int, default=1)
    ap.add_argument("--kh-tune", type=int, default=8, help="3D workload autotuner: launches per timed group of a trial (0: none)")
    ap.add_argument("--kh-tune-rounds", type=int, default=2,
                    help="3D workload autotuner: rounds over the candidates (each samples another part of the time-bin cycle)")
    ap.add_argument("--serve-plain-stride", type=int, default=8,
                    help="serving: every k-th job of the saturation stream (by size) run alone on the plain FP64 path")
    ap.add_argument("--sedov-side", type=int, default=256)
    ap.add_argument("--kh-side", type=int, default=128)
    ap.add_argument("--kh-steps", type=int, default=20)
    ap.add_argument("--kh-warmup", type=int, default=80)
    ap.add_argument("--kh-precision", default="both", choices=["both", "fp64", "fp32"],
                    help="3D workload: the precisions run (profiling runs take one)")
    ap.add_argument("--kh-l1-t", type=float, default=0.002,
                    help="3D workload: end time of the L1 check against the oracle (12 engine steps at 0.002)")
    ap.add_argument("--no-oracle", action="store_true", help="skip the L1 checks against the oracle")
    ap.add_argument("--serve-jobs", type=int, default=32)
    ap.add_argument("--serve-rates", default="4,16,0", help="arrival rates in jobs/s (0: all jobs at t = 0)")
    ap.add_argument("--serve-side-max", type=int, default=512)
    ap.add_argument("--serve-sedov-t-end", type=float, default=0.05)
    ap.add_argument("--serve-vortex-t-end", type=float, default=0.5)
    ap.add_argument("--serve-hbm-gb", type=float, default=24.0)
    ap.add_argument("--serve-sat-jobs", type=int, default=96, help="jobs of the saturation stream")
    ap.add_argument("--serve-sat-hbm-gb", type=float, default=3.0, help="HBM budget of the saturation stream")
    ap.add_argument("--serve-poll", type=int, default=8)
    ap.add_argument("--serve-l1-jobs", type=int, default=2, help="smallest jobs of each kind checked against the oracle")
    ap.add_argument("--threads", type=int, default=0, help="host tessellation threads (0: all)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only run a plain device copy of this many MiB (PMC byte-counter calibration)")
    ap.add_argument("--summarize-pmc", nargs=5, metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR", "LABEL"),
                    help="write profiles/LABEL_pmc.md from rocprofv3 --pmc output folders (no GPU needed)")
    ap.add_argument("--summarize-trace", nargs=2, metavar=("KERNEL_STATS_CSV", "LABEL"),
                    help="write profiles/LABEL_kernel_stats.md from rocprofv3 --stats output (no GPU needed)")
    ap.add_argument("--command", action="append", default=[], help="command line(s) recorded in a summary")
    ap.add_argument("--note", default="")
    ap.add_argument("--trace-csv", default=None, help="per-dispatch kernel_trace.csv for --summarize-trace")
    ap.add_argument("--window", type=int, default=0,
                    help="summaries: also/only each kernel's last N dispatches (the run's timed window)")
    ap.add_argument("--tuning-from", default=None,
                    help="apply the autotuner choices of an earlier run's JSON line instead of tuning (the "
                         "ensemble also runs its tuning launches as warmup): same timed window, no tuning phase "
                         "(profiling runs)")
    ap.add_argument("--no-gate", action="store_true",
                    help="ensemble: time the window only (no FP64 gate batch, no plain path); for profiling")
    args = ap.parse_args()
    if args.summarize_pmc:
        fd, wd, cfd, cwd, label = args.summarize_pmc
        print(summarize_pmc(fd, wd, cfd, cwd, label, args.command, args.copy_probe or 1024.0, args.window))
        return
    if args.summarize_trace:
        print(summarize_trace(args.summarize_trace[0], args.summarize_trace[1], " ".join(args.command), args.note,
                              args.trace_csv, args.window))
        return
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
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: every simulation and every "
              "batch runs on one GPU here (the multi-GPU layout is design only)", file=sys.stderr)
        sys.exit(2)

    from hydro_rs_amd import fvs

    if fvs.device_count() < 1:
        raise SystemExit("no HIP device visible")
    w = args.workload
    out = {"metric": METRIC, "unit": "cell-updates/s", "n_gpus": 1, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "data": "synthetic"}
    if w in ("all", "ensemble"):
        ens, probs = bench_ensemble(args)
        out.update({
            "value": ens["value"], "steps": ens["launches"], "warmup": ens["window_start"],
            "ms_per_step": 1e3 * ens["seconds"] / max(ens["launches"], 1), "dtype": "f32 flux / f64 accumulate",
            "config": {"workload": f"ensemble{args.members}_riemann_1d2d_hllc_fp32flux", "members": args.members,
                       "cells": ens["n_all"], "cell_range": [200, 65536], "t_end": 0.05,
                       "step": "one batch launch = one engine step of every unfinished member",
                       "parallelism": "single GPU, one batch context"},
            "face_fluxes_per_s": ens["face_fluxes"] / ens["seconds"],
            "speedup_vs_plain_fp64": ens["speedup_vs_plain_fp64"], "plain_fp64": ens["plain_fp64"],
            "fp64_batched_value": ens["fp64_batched_value"], "l1_gate": ens["l1_gate"],
            "l1_vs_oracle": ens["l1_vs_oracle"], "warmup_requested": args.warmup,
            "members_done": ens["members_done"], "members_failed": ens["members_failed"],
            "roofline": ens["roofline"], "autotune": ens["autotune"],
        })
        if not args.no_cpu:
            log(f"[bench] cpu baseline (oracle), budget {args.cpu_seconds}s")
            out["cpu_baseline"] = cpu_baseline_ensemble(probs, args.plain_stride, ens["window_start"],
                                                        args.cpu_seconds)
            log(f"[bench] cpu: {out['cpu_baseline']}")
        del probs
    if w in ("all", "sedov"):
        out["sedov2d"], _ = bench_sedov(args)
    if w in ("all", "moving2d"):
        out["moving2d"] = bench_moving(args, 2)
    if w in ("all", "moving3d"):
        out["moving3d"] = bench_moving(args, 3)
    if w in ("all", "kh3d"):
        out["kh3d"] = bench_kh3d(args)
    if w in ("all", "serve"):
        out["serve"] = bench_serve(args)
    if "value" not in out:  # a single secondary workload: its own headline numbers
        if w == "sedov":
            s = out["sedov2d"]["fp64"]
            out.update({"value": s["cell_updates_per_s"], "steps": s["steps"], "warmup": 0,
                        "ms_per_step": s["us_per_step"] / 1e3, "dtype": "f64",
                        "config": {"workload": out["sedov2d"]["workload"]}})
        elif w in ("moving2d", "moving3d"):
            s = out[w]["fp64"]
            out.update({"value": s["cell_updates_per_s"], "steps": out[w]["steps"], "warmup": out[w]["warmup"],
                        "ms_per_step": s["us_per_step"] / 1e3, "dtype": "f64",
                        "config": {"workload": out[w]["workload"]}})
        elif w == "kh3d":
            key = "fp64" if "fp64" in out["kh3d"] else "fp32_flux"
            s = out["kh3d"][key]
            out.update({"value": s["cell_updates_per_s"], "steps": args.kh_steps, "warmup": args.kh_warmup,
                        "ms_per_step": s["ms_per_step"], "dtype": "f64" if key == "fp64" else "f32 flux / f64 accumulate",
                        "roofline": s["roofline"], "config": {"workload": out["kh3d"]["workload"]}})
        else:
            s = out["serve"]["saturation"]
            out.update({"value": s["cell_updates_per_s_while_queued"] or s["cell_updates_per_s"], "steps": s["jobs"],
                        "warmup": 0,
                        "ms_per_step": 1e3 * s["wall_s"] / max(s["jobs"], 1), "dtype": "f32 flux / f64 accumulate",
                        "config": {"workload": "served_stream", "jobs": s["jobs"]}})
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
