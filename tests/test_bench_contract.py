"""bench.py's measurement plumbing, on the CPU: the algorithmic bytes per pass, the
HBM traffic it reads from the committed rocprofv3 summary for the default workload,
and the refusal of more than one GPU."""
import importlib.util
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_bytes_per_pass_counts_state_velocity_and_source(bench):
    # FP32: read d, x, y, z, u, v, w and write x, y, z, u, v, w (13 x 4 B) per parcel;
    # per cell the velocity (3 x 4 B) once and the FP64 source read and written (2 x 24 B)
    assert bench.bytes_per_pass(1, 0, True) == 13 * 4
    assert bench.bytes_per_pass(0, 1, True) == 3 * 4 + 2 * 24
    assert bench.bytes_per_pass(1, 0, False) == 13 * 8
    assert bench.bytes_per_pass(0, 1, False) == 3 * 8 + 2 * 24
    assert bench.bytes_per_pass(1 << 24, 128 ** 3, True) == 998_244_352
    assert bench.bytes_per_pass(1 << 24, 128 ** 3, False) == 1_895_825_408


def default_workload(bench):
    saved = sys.argv
    sys.argv = ["bench.py"]
    try:
        args = bench.parse()
    finally:
        sys.argv = saved
    return {"precision": args.precision, "method": args.method, "parcels": args.parcels, "cells": args.cells,
            "brick": args.brick, "block": args.block, "substeps_per_pass": args.substeps_per_pass,
            "flush": args.flush}


def test_default_run_finds_its_profiled_traffic(bench):
    path = os.path.join(ROOT, "profiles", "k_pass_sorted_fp32_16M_128c.md")
    want = default_workload(bench)
    traffic = bench.profiled_traffic(path, want)
    assert traffic is not None
    # the measured traffic is the algorithmic bytes plus a few per cent of halo re-reads
    algo = bench.bytes_per_pass(want["parcels"], want["cells"] ** 3, True)
    assert algo <= traffic <= 1.15 * algo
    # a summary of another workload is not used
    assert bench.profiled_traffic(path, dict(want, block=256)) is None
    assert bench.profiled_traffic(os.path.join(ROOT, "profiles", "missing.md"), want) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stdout + r.stderr


def _csv(path, header, rows):
    import csv

    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        for r in rows:
            w.writerow(r)
    return str(path)


def test_rocprof_summary_calibrates_and_sums_counter_instances(tmp_path, capsys):
    """Synthetic rocprofv3 CSVs: the kernel's FETCH_SIZE comes in two instance rows per
    dispatch (summed), the copy of 1 GiB reads half its bytes in FETCH_SIZE (factor 2)
    and all of them in WRITE_SIZE (factor 1); the traffic is each kernel counter times
    its factor, and the VALU issue rate is instructions / SIMDs / duration against one
    wave64 instruction per 4 cycles."""
    from scaletrack import rocprof_summary as rs

    kern, copy = "st::k_pass_sorted<float, true, false, false>(x)", "st::k_copy(a, b, n)"
    stats = _csv(tmp_path / "stats.csv", ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage"],
                 [[kern, 2, 400000, 200000, 80.0], [copy, 1, 100000, 100000, 20.0]])
    head = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    fetch = _csv(tmp_path / "fetch.csv", head,
                 [[1, kern, "FETCH_SIZE", 100000.0], [1, kern, "FETCH_SIZE", 184289.0],
                  [2, kern, "FETCH_SIZE", 150000.0], [2, kern, "FETCH_SIZE", 134289.0],
                  [3, copy, "FETCH_SIZE", 524288.0]])
    write = _csv(tmp_path / "write.csv", head,
                 [[1, kern, "WRITE_SIZE", 463318.0], [2, kern, "WRITE_SIZE", 463318.0],
                  [3, copy, "WRITE_SIZE", 1048576.0]])
    pmc = _csv(tmp_path / "pmc.csv", head,
               [[1, kern, "SQ_INSTS_VALU", 6.0e7], [1, kern, "SQ_WAVES", 3.0e4], [1, kern, "GRBM_GUI_ACTIVE", 1e6]])
    rs.main(["--stats", stats, "--fetch", fetch, "--write", write, "--pmc", pmc, "--simds", "1000",
             "--kernel", "k_pass_sorted", "--algorithmic-bytes", "1000000000", "--set", "precision=fp32"])
    out = capsys.readouterr().out
    summary = tmp_path / "summary.md"
    summary.write_text(out)
    fig = rs.read_figures(str(summary))
    assert fig["fetch_factor"] == "2.0000" and fig["write_factor"] == "1.0000"
    assert int(fig["traffic_bytes"]) == 284289 * 1024 * 2 + 463318 * 1024
    assert fig["kernel_avg_ns"] == "200000" and fig["precision"] == "fp32"
    assert fig["valu_per_wave"] == "2000"
    # 6e7 / 1000 SIMDs / 200 us = 3e8 per s, against 2.4e9 / 4 = 6e8 per s
    assert fig["valu_issue_fraction"] == "0.500"
