"""bench.py --mode summary on rocprofv3-style databases built here: the
window/replay split of each kernel's launches, the copy calibration of the
HBM counters, and the per-kernel figures the bench line reads back."""
import json
import os
import sqlite3
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def make_db(path, kernels=(), counters=()):
    os.makedirs(path, exist_ok=True)
    con = sqlite3.connect(os.path.join(path, "run_results.db"))
    con.execute("create table kernels (name text, duration real, start real)")
    con.execute("create table counters_collection (kernel_name text, counter_name text, value real, dispatch_id int)")
    for i, (n, d) in enumerate(kernels):
        con.execute("insert into kernels values (?, ?, ?)", (n, d, i))
    for i, (n, c, v) in enumerate(counters):
        con.execute("insert into counters_collection values (?, ?, ?, ?)", (n, c, v, i))
    con.commit()
    con.close()


def test_summary_splits_window_and_replay_and_calibrates(tmp_path):
    steps = 3
    force = "void (anonymous namespace)::k_force_f32<1, 4>(DevState, int, int)"
    # two launches before the window (list growth), then 3 window and 3 replay launches
    durs_ns = [9e6, 9e6] + [5e6] * steps + [4e6] * steps
    make_db(tmp_path / "prof_trace", kernels=[(force, d) for d in durs_ns])
    n = len(durs_ns)
    make_db(tmp_path / "prof_FETCH_SIZE", counters=[(force, "FETCH_SIZE", 999.0)] * 2 + [(force, "FETCH_SIZE", 100.0)] * (n - 2))
    make_db(tmp_path / "prof_WRITE_SIZE", counters=[(force, "WRITE_SIZE", 50.0)] * n)
    make_db(tmp_path / "prof_SQ", counters=[(force, "SQ_INSTS_VALU", 1e9), (force, "SQ_WAVES", 1e5)] * n)
    make_db(tmp_path / "prof_STALL", counters=[(force, "SQ_WAVE_CYCLES", 2e9)] * n)
    copy = "k_copy(float4 const*, float4*, long)"
    make_db(tmp_path / "prof_copy_FETCH_SIZE", counters=[(copy, "FETCH_SIZE", 524288.0)])
    make_db(tmp_path / "prof_copy_WRITE_SIZE", counters=[(copy, "WRITE_SIZE", 1048576.0)])
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--mode", "summary", "--prof-dir",
                        str(tmp_path), "--workload", "cfgX", "--steps", str(steps)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.load(open(tmp_path / "prof_counters.json"))["cfgX"]["force"]
    assert out["avg_us_window"] == 5000.0 and out["avg_us_replay"] == 4000.0
    # FETCH_SIZE reads half the bytes of the calibration copy: x2; WRITE_SIZE x1 (KiB -> bytes)
    assert out["hbm_bytes"] == (100.0 * 2 + 50.0) * 1024
    assert out["valu_insts"] == 1e9 and out["waves"] == 1e5
    assert out["stall"]["SQ_WAVE_CYCLES"] == 2e9
    assert "| k_force_f32 | 8 |" in open(tmp_path / "prof_summary.md").read()
