"""Throughput of the parcel step on one MI355X: parcel-updates/s and HBM roofline.

Workload (the forward configuration the metric is quoted on): periodic box of 128^3
uniform hex cells on [0,1]^3 with a frozen Taylor-Green carrier (U0 = 1 m/s), 2^24
parcels at seeded uniform positions, log-normal diameters (median 30 um, sigma 0.5),
rho_p = 1000 kg/m^3, zero initial velocity.  One step = one fluid step of
dt = 1e-3 s: UTrans zeroed, then nsub = 10 parcel substeps of 1e-4 s.  The default
run times 5 steps = 50 substeps with one kernel pass per substep.

Roofline of the dominant kernel (one pass = one sweep over every parcel):
  bytes/pass = N_p * B_state + N_c * (B_U + 48)
  B_state = 52 (FP32 state: read d,x,y,z,u,v,w, write x,y,z,u,v,w) or 104 (FP64)
  B_U = 12 (FP32) or 24 (FP64) cell-velocity bytes; 48 = FP64 source read + write.
Kernel time comes from HIP events recorded on the library's stream around the pass
launches of each step.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "scale-track_amd"), ROOT]

PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--parcels", type=int, default=1 << 24)
    ap.add_argument("--cells", type=int, default=128)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--method", choices=["sorted", "plain"], default="sorted")
    ap.add_argument("--substeps-per-pass", type=int, default=1)
    ap.add_argument("--brick", type=str, default="8,8,8")
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--flush", choices=["owner", "atomic"], default="owner")
    ap.add_argument("--only", action="store_true",
                    help="time only the headline configuration (for profiler runs)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-parcels", type=int, default=1 << 22)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run the streaming copy kernel over BYTES (traffic-counter calibration)")
    return ap.parse_args()


def bytes_per_pass(np_, nc, fp32):
    return np_ * (52 if fp32 else 104) + nc * ((12 if fp32 else 24) + 48)


def run_variant(stk, ctx, case, precision, p, steps, warmup, keep_state=False):
    """Time `steps` fluid steps after `warmup` untimed ones; returns a dict."""
    dt = np.float32 if precision == stk.FP32 else np.float64
    mesh = stk.Mesh(case.n, case.origin, case.ending)
    field = stk.TwoWayEulerian(ctx, mesh, precision)
    field.upload_U(case.U.astype(dt))
    chunk = stk.Chunk(ctx, case.np, case.mu_c, case.rho_d, case.rho_d_by_rho_c, precision)
    chunk.set_state(*[a.astype(dt) for a in case.state()])
    chunk.upload()
    srcs = []
    for _ in range(warmup):
        field.zero_source()
        stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p, want_bbox=False)
        if keep_state:
            srcs.append(field.download_source().copy())
    ctx.synchronize()
    kern_ms, launches = [], 0
    t0 = time.perf_counter()
    for _ in range(steps):
        field.zero_source()
        stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p, want_bbox=False)
        ms, n = ctx.last_kernel_ms()
        kern_ms.append(ms)
        launches = n
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    out = {
        "elapsed_s": elapsed,
        "ms_per_step": 1e3 * elapsed / steps,
        "updates_per_s": case.np * case.nsub * steps / elapsed,
        "kernel_ms_per_pass": float(np.mean(kern_ms)),
        "passes_per_step": launches,
        "stray_fraction": chunk.stray_fraction() if p.method == stk.SORTED else None,
    }
    if keep_state:
        out["srcs"] = srcs
        out["state"] = [a.copy() for a in chunk.download()]
    field.close()
    chunk.close()
    return out


def cpu_baseline(case, nparcels):
    """The oracle (C++ FP64 restatement, test infrastructure) on the first `nparcels`
    parcels of the same workload for one fluid step (nsub = 10 substeps), serial and
    OpenMP over the host's cores."""
    from oracle import st_oracle

    st = [a[:nparcels].copy() for a in case.state()]
    ncores = len(os.sched_getaffinity(0))
    res = {}
    for nt in (1, ncores):
        s = [a.copy() for a in st]
        t0 = time.perf_counter()
        st_oracle.evolve(s, case.U, case.n, case.origin, case.ending, case.dt_fluid, nsub=case.nsub,
                         boundary=case.boundary, mu_c=case.mu_c, rho_d=case.rho_d,
                         rho_d_by_rho_c=case.rho_d_by_rho_c, nthreads=nt)
        res[nt] = nparcels * case.nsub / (time.perf_counter() - t0)
    sample = (f"{nparcels} of the 2^24 parcels, 128^3 cells, 1 fluid step = {case.nsub} substeps, "
              "FP64 oracle")
    return ({"value": res[ncores], "unit": "parcel-updates/s", "cores": ncores, "kind": "port",
             "sample": sample},
            {"value": res[1], "unit": "parcel-updates/s", "cores": 1, "kind": "port", "sample": sample})


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the parcel step shards by spatial slabs, but this build measures one GPU)",
              file=sys.stderr)
        sys.exit(2)
    import scaletrack as stk
    from scaletrack import inputs

    if stk.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    ctx = stk.Context(0)
    if args.calibrate_copy:
        gbs = ctx.copy_bandwidth(args.calibrate_copy, reps=args.steps)
        print(json.dumps({"copy_bytes": args.calibrate_copy, "reps": args.steps, "GBs": gbs}))
        return
    copy_gbs = ctx.copy_bandwidth(1 << 30, reps=10)
    t0 = time.perf_counter()
    case = inputs.taylor_green_box(np_parcels=args.parcels, ncell=args.cells, boundary=stk.PERIODIC)
    print(f"# inputs generated in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)

    fp32 = args.precision == "fp32"
    precision = stk.FP32 if fp32 else stk.FP64
    brick = [int(v) for v in args.brick.split(",")]
    p = stk.params(boundary=stk.PERIODIC, method=stk.SORTED if args.method == "sorted" else stk.PLAIN,
                   substeps_per_pass=args.substeps_per_pass, brick=brick, block=args.block,
                   flush=stk.FLUSH_OWNER if args.flush == "owner" else stk.FLUSH_ATOMIC)
    head = run_variant(stk, ctx, case, precision, p, args.steps, args.warmup)
    print(f"# headline: {json.dumps(head)}", file=sys.stderr, flush=True)

    nc = case.ncell
    bpp = bytes_per_pass(case.np, nc, fp32)
    per = args.substeps_per_pass if args.substeps_per_pass > 0 else case.nsub
    achieved = bpp / (head["kernel_ms_per_pass"] * 1e-3) / 1e9
    line = {
        "metric": "parcel-updates/s at 16M parcels (fwd; fwd+adjoint); % HBM roofline",
        "value": head["updates_per_s"],
        "unit": "parcel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state, f64 face test and source" if fp32 else "f64",
        "data": "synthetic (seeded SplitMix64 inputs generated on the host)",
        "config": {
            "workload": "Taylor-Green periodic box, forward only",
            "parcels": case.np,
            "cells": list(case.n),
            "substeps_per_step": case.nsub,
            "substep_dt_s": case.dt_fluid / case.nsub,
            "substeps_per_pass": per,
            "method": args.method,
            "brick": brick,
            "block": args.block,
            "flush": args.flush,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": None,
            "kernel": "k_pass_sorted" if args.method == "sorted" else "k_pass_plain",
            "kernel_ms_per_pass": head["kernel_ms_per_pass"],
            "bytes_per_pass": bpp,
            "copy_kernel_GBs": copy_gbs,
        },
        "stray_fraction": head["stray_fraction"],
    }
    if not args.only:
        # the plain FP64 reference-shaped path and the other forms, same schedule
        pp = stk.params(boundary=stk.PERIODIC, method=stk.PLAIN, substeps_per_pass=per)
        plain64 = run_variant(stk, ctx, case, stk.FP64, pp, args.steps, args.warmup)
        ps = stk.params(boundary=stk.PERIODIC, method=stk.SORTED, substeps_per_pass=per, brick=brick,
                        block=args.block)
        sorted64 = run_variant(stk, ctx, case, stk.FP64, ps, args.steps, args.warmup)
        pf = stk.params(boundary=stk.PERIODIC, method=stk.SORTED, substeps_per_pass=0, brick=brick,
                        block=args.block)
        fused32 = run_variant(stk, ctx, case, stk.FP32, pf, args.steps, args.warmup)
        line["variants"] = {
            "fp64_plain": {k: plain64[k] for k in ("updates_per_s", "kernel_ms_per_pass")},
            "fp64_sorted": {k: sorted64[k] for k in ("updates_per_s", "kernel_ms_per_pass")},
            "fp32_sorted_fused10": {k: fused32[k] for k in ("updates_per_s", "kernel_ms_per_pass")},
        }
        line["variants"]["fp64_plain"]["roofline_frac"] = (
            bytes_per_pass(case.np, nc, False) / (plain64["kernel_ms_per_pass"] * 1e-3) / 1e9 / PEAK_HBM_GBS)
        line["variants"]["fp64_sorted"]["roofline_frac"] = (
            bytes_per_pass(case.np, nc, False) / (sorted64["kernel_ms_per_pass"] * 1e-3) / 1e9 / PEAK_HBM_GBS)
        line["variants"]["fp32_sorted_fused10"]["roofline_frac"] = (
            bytes_per_pass(case.np, nc, True) / (fused32["kernel_ms_per_pass"] * 1e-3) / 1e9 / PEAK_HBM_GBS)
        line["speedup_vs_fp64_plain"] = head["updates_per_s"] / plain64["updates_per_s"]
    if not args.no_cpu and not args.only:
        omp, serial = cpu_baseline(case, min(args.cpu_parcels, case.np))
        line["cpu_baseline"] = omp
        line["cpu_baseline_serial"] = serial
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
