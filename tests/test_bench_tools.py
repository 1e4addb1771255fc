"""bench.py's measurement helpers on synthetic profiler output (CPU only): the timed-window selection of
per-dispatch rows, the calibrated traffic summary, the byte models and tuning taken from an earlier
run's report."""
import csv
import json
import os

import numpy as np
import pytest

import bench

COUNTER_COLS = ["Correlation_Id", "Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]


def _write_counters(d, counter, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "x_counter_collection.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(COUNTER_COLS)
        for i, (name, value) in enumerate(rows):
            w.writerow([i + 1, i + 1, name, counter, value])


@pytest.fixture
def profiles(tmp_path, monkeypatch):
    out = tmp_path / "profiles"
    out.mkdir()
    monkeypatch.setattr(bench, "PROFILES", str(out))
    return out


def test_last_dispatches_follow_dispatch_order():
    rows = [{"Dispatch_Id": str(i), "Kernel_Name": "k" if i % 2 else "j", "v": i} for i in (5, 1, 3, 2, 4, 6)]
    got = bench._last_dispatches(rows, 2, lambda r: r["v"])
    assert got == {"k": [3, 5], "j": [4, 6]}


def test_pmc_summary_window_and_calibration(tmp_path, profiles):
    # a 1 MiB copy whose FETCH_SIZE reads half its bytes and WRITE_SIZE all of them (KiB per dispatch)
    _write_counters(tmp_path / "cf", "FETCH_SIZE", [("copy", 512.0)])
    _write_counters(tmp_path / "cw", "WRITE_SIZE", [("copy", 1024.0)])
    # kernel A: 5 dispatches, the last 2 are the window; kernel B: 3 dispatches
    fetch = [("hfv::A", v) for v in (1.0, 1.0, 1.0, 10.0, 20.0)] + [("hfv::B", v) for v in (4.0, 4.0, 4.0)]
    write = [("hfv::A", v) for v in (2.0, 2.0, 2.0, 6.0, 8.0)] + [("hfv::B", v) for v in (1.0, 1.0, 1.0)]
    _write_counters(tmp_path / "f", "FETCH_SIZE", fetch)
    _write_counters(tmp_path / "w", "WRITE_SIZE", write)
    bench.summarize_pmc(str(tmp_path / "f"), str(tmp_path / "w"), str(tmp_path / "cf"), str(tmp_path / "cw"),
                        "synthetic", ["cmd"], 1.0, window=2)
    text = (profiles / "synthetic_pmc.md").read_text()
    rec = json.loads(text.split("TRAFFIC_JSON ", 1)[1])["kernels"]
    kib = 1024.0
    assert rec["hfv::A"]["dispatches"] == 2
    assert rec["hfv::A"]["fetch_bytes"] == pytest.approx(2.0 * kib * (10.0 + 20.0) / 2)  # reads scaled by 2
    assert rec["hfv::A"]["write_bytes"] == pytest.approx(1.0 * kib * (6.0 + 8.0) / 2)
    assert rec["hfv::B"]["traffic_bytes"] == pytest.approx(kib * (2.0 * 4.0 + 1.0))
    assert bench.traffic_for("synthetic", "hfv::A") == pytest.approx(kib * (30.0 + 7.0))
    assert bench.traffic_for("synthetic", "hfv::missing") is None


def test_trace_summary_window_column(tmp_path, profiles):
    stats = tmp_path / "stats.csv"
    with open(stats, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Name", "Calls", "TotalDurationNs", "AverageNs", "MinNs", "MaxNs", "Percentage"])
        w.writerow(["hfv::A", 4, 10000, 2500, 1000, 4000, 100.0])
    trace = tmp_path / "trace.csv"
    with open(trace, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        for i, dur in enumerate((1000, 2000, 3000, 4000)):
            w.writerow([i + 1, "hfv::A", 100 * i, 100 * i + dur])
    bench.summarize_trace(str(stats), "synth", "cmd", trace_csv=str(trace), window=2)
    line = [ln for ln in (profiles / "synth_kernel_stats.md").read_text().splitlines() if "`hfv::A`" in ln][0]
    assert line.rstrip().endswith("| 3.50 |")  # (3000 + 4000) / 2 ns = 3.5 us


def test_byte_model_is_the_step_model():
    """The per-kernel attribution sums to B_step = 192 N + 1247 N_act + 248 F_act (+8 F_act periodic) in
    FP64 contexts, and the FP32-flux widths remove exactly the stored-FP32 fields: gradients 120 -> 60 (read
    by predict 72 -> 36 and the flux gather, written by the gradient phase), extrapolations 40 -> 20 (predict
    r/w, flux gather, reset) and face records 56 -> 28 (written and read back)."""
    n_act, f_act, n_all = 1000.0, 3000.0, 5000.0
    assert bench.step_bytes(n_all, n_act, f_act) == 192 * n_all + 1247 * n_act + 248 * f_act
    assert bench.step_bytes(n_all, n_act, f_act, periodic=True) == 192 * n_all + 1247 * n_act + 256 * f_act
    d = {k: bench.ALGO[k](n_act, f_act, n_all) - bench.ALGO_FP32[k](n_act, f_act, n_all) for k in bench.ALGO}
    assert d["mask_predict"] == (36 + 40) * n_all
    assert d["flux"] == (60 + 20) * n_act + 28 * f_act
    assert d["apply_c2p"] == 28 * f_act
    assert d["grad_timestep"] == 60 * n_act
    assert d["limit_kick"] == 20 * n_act
    assert d["wakeup"] == 0
    assert bench.step_bytes(n_all, n_act, f_act, fp32=True) == 116 * n_all + 1087 * n_act + 192 * f_act


def test_tuning_from_report(tmp_path):
    classes = [{"dim": 1, "cells": [0, 1024], "staged": 0, "static_grad": 0, "gstaged": 1},
               {"dim": 2, "cells": [1024, 8192], "staged": 1, "static_grad": 0, "gstaged": 0},
               {"dim": 3, "cells": [65536, 2 ** 63 - 1], "staged": 1, "static_grad": 1, "gstaged": 1}]
    path = tmp_path / "bench.json"
    path.write_text("[log line]\n" + json.dumps({"autotune": {"launches": 7, "classes": classes}}) + "\n\n")
    rep = bench.load_report(str(path))
    assert rep["autotune"]["launches"] == 7

    class Sim:
        def __init__(self):
            self.calls = []

        def set_tuning(self, slot, staged, static_grad, gstaged):
            self.calls.append((slot, staged, static_grad, gstaged))

    sim = Sim()
    # slot 2 (2D, 900 cells) falls in no class and keeps its defaults
    bench.apply_tuning(sim, [(0, 1, 1023), (1, 2, 1024), (2, 2, 900), (3, 3, 2097152)], rep["autotune"]["classes"])
    assert sim.calls == [(0, 0, 0, 1), (1, 1, 0, 0), (3, 1, 1, 1)]


def test_ensemble_window_start_does_not_depend_on_the_tuner():
    # the tuner ran 120 or 240 launches: the window still starts at launch 260
    assert bench.window_warmup(260, 20, 120, False) == (140, 260)
    assert bench.window_warmup(260, 20, 240, False) == (20, 260)
    # a replay of that run's tuning runs no tuner: every launch before the window is warmup
    assert bench.window_warmup(260, 20, 120, True) == (260, 260)
    # a long tuner pushes the window back, with W launches after it
    assert bench.window_warmup(260, 20, 300, False) == (20, 320)
    assert bench.window_warmup(260, 20, 300, True) == (320, 320)
