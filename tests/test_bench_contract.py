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
