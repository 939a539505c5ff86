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
# rocprofv3 summary of the headline kernel on the default forward workload: its
# FETCH_SIZE + WRITE_SIZE traffic per launch (calibrated against a copy of known size)
# fills roofline.traffic when this run's workload is the one profiled.
PROFILE = os.path.join(ROOT, "profiles", "k_pass_sorted_fp32_16M_128c.md")


def profiled_traffic(path, want):
    """HBM bytes per launch from a committed rocprofv3 summary, or None when the
    summary is missing or was measured on another workload."""
    from scaletrack.rocprof_summary import read_figures

    if not os.path.exists(path):
        return None
    fig = read_figures(path)
    if any(fig.get(k) != str(v) for k, v in want.items()) or "traffic_bytes" not in fig:
        return None
    return float(fig["traffic_bytes"])


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
    ap.add_argument("--brick", type=str, default="32,8,2")
    ap.add_argument("--block", type=int, default=512)
    ap.add_argument("--flush", choices=["owner", "atomic"], default="owner")
    ap.add_argument("--passes-per-launch", type=int, default=1,
                    help="passes of one persistent launch (each workgroup owning the same bricks)")
    ap.add_argument("--only", action="store_true",
                    help="time only the headline configuration (for profiler runs)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-parcels", type=int, default=1 << 23)
    ap.add_argument("--cpu-parcels-serial", type=int, default=1 << 20)
    ap.add_argument("--tune-cache", default=None, help="autotuner cache file (JSON lines)")
    ap.add_argument("--workload", choices=["forward", "batch", "adjoint", "channel"], default="forward",
                    help="with --only: which workload to time (profiler runs)")
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
    sorted_method = p.method == stk.SORTED
    resorts0 = chunk.resort_count() if sorted_method else 0
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        field.zero_source()
        stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p, want_bbox=False)
        field.merge_source()  # the step's source complete on the device, as the solver reads it
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    # pass durations of the timed steps, from the HIP events around their pass launches
    timed = ctx.kernel_times(steps)
    kern_ms = [ms for ms, _ in timed]
    launches = timed[-1][1]
    out = {
        "elapsed_s": elapsed,
        "ms_per_step": 1e3 * elapsed / steps,
        "updates_per_s": case.np * case.nsub * steps / elapsed,
        "kernel_ms_per_pass": float(np.mean(kern_ms)),
        "passes_per_step": launches,
        "stray_fraction": chunk.stray_fraction() if sorted_method else None,
        "resorts_in_timed_steps": chunk.resort_count() - resorts0 if sorted_method else None,
    }
    if keep_state:
        out["srcs"] = srcs
        out["state"] = [a.copy() for a in chunk.download()]
    field.close()
    chunk.close()
    return out


def graph_variant(stk, ctx, case, precision, p, steps, warmup):
    """The same steps replayed from one captured hipGraph of a fluid step (UTrans zeroed,
    nsub passes), each followed by the stray merge that completes the source as in the
    eager steps: wall-clock parcel-updates/s without per-launch host work."""
    dt = np.float32 if precision == stk.FP32 else np.float64
    field = stk.TwoWayEulerian(ctx, stk.Mesh(case.n, case.origin, case.ending), precision)
    field.upload_U(case.U.astype(dt))
    chunk = stk.Chunk(ctx, case.np, case.mu_c, case.rho_d, case.rho_d_by_rho_c, precision)
    chunk.set_state(*[a.astype(dt) for a in case.state()])
    chunk.upload()
    graph = stk.Graph(ctx, chunk, field, case.dt_fluid, case.nsub, p, zero_source=True)
    for _ in range(warmup):
        graph.launch()
        field.merge_source()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        graph.launch()
        field.merge_source()  # the step's source complete on the device, as the solver reads it
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    graph.close()
    field.close()
    chunk.close()
    return {"updates_per_s": case.np * case.nsub * steps / elapsed, "ms_per_step": 1e3 * elapsed / steps}


def precision_check(stk, ctx, case, p, calls=5, block=4):
    """FP32 form against the FP64 path on the same inputs and schedule (calls fluid
    steps of case.nsub substeps, source zeroed per call).  Both runs start from the
    FP32-representable inputs (state and cell velocities rounded to FP32), so every
    difference comes from the form's arithmetic.  Reported per call:
      * source relative L2 over cells, and max over cells against the largest |S|;
      * the same on block^3-cell blocks (a parcel that took the other side of a face
        moves its deposit to a neighbouring cell: blocks sum most such moves back);
      * total momentum source (sum over cells) relative to the sum of |S|;
    and at the end: position error in cell widths, fraction of parcels whose final
    cell differs, velocity relative L2."""
    runs = {}
    u32 = case.U.astype(np.float32)
    st32 = [a.astype(np.float32) for a in case.state()]
    for prec in (stk.FP64, stk.FP32):
        dt = np.float32 if prec == stk.FP32 else np.float64
        mesh = stk.Mesh(case.n, case.origin, case.ending)
        field = stk.TwoWayEulerian(ctx, mesh, prec)
        field.upload_U(u32.astype(dt))
        chunk = stk.Chunk(ctx, case.np, case.mu_c, case.rho_d, case.rho_d_by_rho_c, prec)
        chunk.set_state(*[a.astype(dt) for a in st32])
        chunk.upload()
        srcs = []
        for _ in range(calls):
            field.zero_source()
            stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p, want_bbox=False)
            srcs.append(field.download_source().astype(np.float64))
        runs[prec] = (srcs, [a.astype(np.float64) for a in chunk.download()])
        field.close()
        chunk.close()
    (s64, st64), (s32, stf) = runs[stk.FP64], runs[stk.FP32]
    nx, ny, nz = case.n

    def coarse(s):
        g = s.reshape(nz, ny, nx, 3)
        b = block
        return g.reshape(nz // b, b, ny // b, b, nx // b, b, 3).sum(axis=(1, 3, 5))

    out = {"substeps": calls * case.nsub, "block_cells": block, "per_call": []}
    for a, b in zip(s32, s64):
        ca, cb = coarse(a), coarse(b)
        out["per_call"].append({
            "source_rel_l2": float(np.linalg.norm(a - b) / np.linalg.norm(b)),
            "source_rel_max": float(np.abs(a - b).max() / np.abs(b).max()),
            "block_source_rel_l2": float(np.linalg.norm(ca - cb) / np.linalg.norm(cb)),
            "total_momentum_rel": float(np.abs(a.sum(0) - b.sum(0)).max() / np.abs(b).sum()),
        })
    delta = np.array([(e - o) / n for o, e, n in zip(case.origin, case.ending, case.n)])
    length = np.array([e - o for o, e in zip(case.origin, case.ending)])
    pos, pos64 = np.stack(stf[1:4], 1), np.stack(st64[1:4], 1)
    dpos = pos - pos64
    dpos -= length * np.round(dpos / length)  # a wrap taken in one run only is not an error
    cells = np.floor((pos - np.asarray(case.origin)) / delta)
    cells64 = np.floor((pos64 - np.asarray(case.origin)) / delta)
    vel, vel64 = np.stack(stf[4:7], 1), np.stack(st64[4:7], 1)
    out.update({
        "source_rel_l2_max": max(c["source_rel_l2"] for c in out["per_call"]),
        "block_source_rel_l2_max": max(c["block_source_rel_l2"] for c in out["per_call"]),
        "position_max_err_cells": float((np.abs(dpos) / delta).max()),
        "parcels_in_other_cell": float(np.any(cells != cells64, axis=1).mean()),
        "velocity_rel_l2": float(np.linalg.norm(vel - vel64) / np.linalg.norm(vel64)),
    })
    # the 1e-5 relative source target over the schedule the forward workload is quoted on
    # (5 fluid steps of 10 substeps), judged per cell, per block and for the total
    worst_total = max(c["total_momentum_rel"] for c in out["per_call"])
    out["gate"] = {
        "target": 1e-5,
        "per_cell": "met" if out["source_rel_l2_max"] <= 1e-5 else "unmet",
        "per_block": "met" if out["block_source_rel_l2_max"] <= 1e-5 else "unmet",
        "total_momentum": "met" if worst_total <= 1e-5 else "unmet",
        "total_momentum_rel_max": worst_total,
    }
    return out


def batch_workload(stk, inputs, ctx, nclouds=64, steps=5, warmup=1, seed=19891, tune_cache=None):
    """64 independent periodic clouds, parcel counts log-uniform in [1e3, 2e6], meshes
    n^3 with n uniform in {32..96}, Taylor-Green with seeded random phases, FP32 state;
    one persistent launch per substep, the fluid step captured as a hipGraph, brick and
    block sizes from the autotuner."""
    cases = inputs.batch_cases(nclouds, seed=seed, dtype=np.float32)
    counts = np.array([c.np for c in cases], dtype=np.int64)
    cells = np.array([c.n[0] for c in cases], dtype=np.int64)
    chunks, fields, bytes_pass = [], [], 0
    for c in cases:
        field = stk.TwoWayEulerian(ctx, stk.Mesh(c.n, c.origin, c.ending), stk.FP32)
        field.upload_U(c.U)
        chunk = stk.Chunk(ctx, c.np, c.mu_c, c.rho_d, c.rho_d_by_rho_c, stk.FP32)
        chunk.set_state(*c.state())
        chunk.upload()
        chunks.append(chunk)
        fields.append(field)
        bytes_pass += bytes_per_pass(c.np, c.ncell, True)
    p = stk.params(boundary=stk.PERIODIC, substeps_per_pass=1)
    t0 = time.perf_counter()
    best = stk.autotune_batch(ctx, chunks, fields, p, 1e-3, trials=5, cache_path=tune_cache)
    t_tune = time.perf_counter() - t0
    p.block = best["block"]
    batch = stk.Batch(ctx, chunks, fields, p, target_parcels_per_brick=best["target_parcels_per_brick"],
                      blocks_per_cu=best["blocks_per_cu"])
    graph = batch.graph(1e-3, 10, zero_source=True)
    for _ in range(warmup):
        graph.launch()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        graph.launch()
    ctx.synchronize()
    el = time.perf_counter() - t0
    # the stray merges that complete the 64 sources after a step (a caller that reads
    # them pays this once per step; not in ms_per_step)
    t0 = time.perf_counter()
    for f in fields:
        f.merge_source()
    ctx.synchronize()
    merge_ms = 1e3 * (time.perf_counter() - t0)
    # kernel time per pass from an eager step (events around the passes)
    for f in fields:
        f.zero_source()
    batch.evolve(1e-3, 10)
    kms, _ = ctx.last_kernel_ms()
    total = int(counts.sum())
    info = batch.info()
    graph.close()
    batch.close()
    for c in chunks:
        c.close()
    for f in fields:
        f.close()
    return {"clouds": nclouds, "parcels": total, "min_parcels": int(counts.min()), "max_parcels": int(counts.max()),
            "mesh_n_min": int(cells.min()), "mesh_n_max": int(cells.max()),
            "updates_per_s": total * 10 * steps / el, "ms_per_step": 1e3 * el / steps,
            "stray_merge_ms_per_step": merge_ms,
            "kernel_ms_per_pass": kms, "roofline_frac": bytes_pass / (kms * 1e-3) / 1e9 / PEAK_HBM_GBS,
            "tuned_over_kernel": best["ms_per_pass"] / kms,
            "bytes_per_pass": bytes_pass, "work_items": info["items"], "grid": info["grid"],
            "tuned": best, "tune_s": t_tune, "dtype": "f32 state, f64 face test and source"}


def channel_workload(stk, inputs, ctx, steps=5, warmup=1, np_parcels=1 << 25, inject_per_step=(1 << 25) // 100):
    """The channel at one GPU's shape: 2^25 parcels filling a 2 x 1 x 1 m duct of
    256 x 128 x 128 cells (walls on y and z, open inlet and outlet, frozen laminar
    carrier with 5% perturbations), FP32 state with FP64 face test and source.  One
    fluid step = the inflow's parcels injected at the inlet (2^25 / 100 per step, host
    arrays generated beforehand), UTrans zeroed, 10 one-substep passes."""
    case = inputs.channel_case(np_parcels=np_parcels, dtype=np.float32)
    field = stk.TwoWayEulerian(ctx, stk.Mesh(case.n, case.origin, case.ending), stk.FP32)
    field.upload_U(case.U)
    chunk = stk.Chunk(ctx, case.np, case.mu_c, case.rho_d, case.rho_d_by_rho_c, stk.FP32)
    chunk.set_state(*case.state())
    chunk.upload()
    chunk.reserve(case.np + (steps + warmup) * inject_per_step)
    news = [inputs.channel_injection(case, k + 1, inject_per_step) for k in range(steps + warmup)]
    p = stk.params(boundary=stk.CHANNEL, substeps_per_pass=1)
    for k in range(warmup):
        chunk.inject(*news[k])
        field.zero_source()
        stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p, want_bbox=False)
    resorts0 = chunk.resort_count()
    n0 = chunk.count()
    ctx.synchronize()
    t0 = time.perf_counter()
    for k in range(warmup, warmup + steps):
        chunk.inject(*news[k])
        field.zero_source()
        stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p, want_bbox=False)
    ctx.synchronize()
    el = time.perf_counter() - t0
    timed = ctx.kernel_times(steps)
    kms = float(np.mean([ms for ms, _ in timed]))
    n1 = chunk.count()
    removed = int((chunk.download()[0] < 0).sum())
    resorts = chunk.resort_count() - resorts0
    chunk.close()
    field.close()
    updates = sum(n0 + (k + 1) * inject_per_step for k in range(steps)) * case.nsub
    # parcels each timed step's passes sweep: n0 plus the step's own and earlier injections
    bpp = bytes_per_pass(n0 + inject_per_step * (steps + 1) // 2, case.ncell, True)
    return {"parcels_start": n0, "parcels_end": n1, "removed_total": removed, "inject_per_step": inject_per_step,
            "cells": list(case.n), "substeps_per_step": case.nsub, "updates_per_s": updates / el,
            "ms_per_step": 1e3 * el / steps, "kernel_ms_per_pass": kms,
            "roofline_frac": bpp / (kms * 1e-3) / 1e9 / PEAK_HBM_GBS, "resorts_in_timed_steps": resorts,
            "dtype": "f32 state, f64 face test and source",
            "timing": "wall clock of whole fluid steps: injection (host arrays over PCIe), zeroing, passes and "
                      "the sorts the injected parcels trigger; kernel_ms_per_pass from HIP events around the passes"}


def adjoint_workload(stk, inputs, ctx, np_parcels=1 << 20, ncell=64, nsub=200, K=20, seed=19891):
    """Forward + adjoint: frozen divergence-free field of 8 seeded Fourier modes, per-
    parcel d and rho, nsub substeps of 1e-4 s with a checkpoint every K; loss = L2 misfit
    of the deposited source against a seeded target; gradients for d, rho and cd."""
    from scaletrack.adjoint import Adjoint

    case = inputs.adjoint_case(np_parcels, ncell, nsub, seed=seed)
    n = case.n
    field = stk.TwoWayEulerian(ctx, stk.Mesh(case.n, case.origin, case.ending), stk.FP64)
    field.upload_U(case.U)
    chunk = stk.Chunk(ctx, np_parcels, case.mu_c, case.rho_d, case.rho_d_by_rho_c, stk.FP64, rho_p=case.rho,
                      rho_c=case.rho_c)
    chunk.set_state(*case.state())
    chunk.upload()
    # host arrays of the target and the gradients on page-locked memory, allocated once
    # (as a caller repeating the transfers would): they move by DMA at the bus's rate
    T = stk.pinned_copy(inputs.target_source(case, seed=seed))
    gout = {"d": stk.pinned_empty(np_parcels), "rho": stk.pinned_empty(np_parcels)}
    p = stk.params(boundary=stk.PERIODIC)
    adj = Adjoint(ctx, chunk, field, nsub, K, p)
    dt_fluid = case.dt_fluid

    def run():
        chunk.set_state(*case.state())
        chunk.upload()
        ctx.synchronize()
        ts = time.perf_counter()
        chunk.sort(ctx, field, p)  # once per upload of a fresh (unsorted) state
        ctx.synchronize()
        t0 = time.perf_counter()
        adj.forward(dt_fluid)
        ctx.synchronize()
        t1 = time.perf_counter()
        loss = adj.source_misfit(T)  # target upload (host array) and loss readback
        t2 = time.perf_counter()
        adj.backward()
        ctx.synchronize()
        t3 = time.perf_counter()
        bms, _ = ctx.last_kernel_ms()
        g = adj.gradients(out=gout)  # per-parcel gradients downloaded to host arrays
        t4 = time.perf_counter()
        return t1 - t0, t3 - t2, (t2 - t1) + (t4 - t3), t0 - ts, loss, g, bms

    run()  # warm-up (code load)
    tf, tb, th, tsort, loss, g, bms = run()
    adj.close()
    chunk.close()
    field.close()
    return {"parcels": np_parcels, "cells": list(n), "substeps": nsub, "checkpoint_every": K,
            "sort_s": tsort, "forward_s": tf, "backward_s": tb, "host_transfer_s": th,
            "fwd_adjoint_updates_per_s": np_parcels * nsub / (tf + tb),
            "fwd_adjoint_updates_per_s_with_host_arrays": np_parcels * nsub / (tf + tb + th),
            "forward_updates_per_s": np_parcels * nsub / tf, "reverse_kernel_ms_per_segment": bms,
            "loss": loss, "grad_cd": g["cd"], "dtype": "f64",
            "timing": "device compute: forward (with checkpoints) + backward (reverse sweep, parameter "
                      "gradients reduced on the device); host_transfer_s = target upload + gradient download "
                      "(page-locked host arrays allocated once); "
                      "sort_s = brick sort of the freshly uploaded state"}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, nparcels, nparcels_serial, calls=5, runs=3):
    """The oracle (C++ FP64 restatement of the reference's step, test infrastructure) on
    the first parcels of the same workload over the forward schedule: `calls` fluid steps
    of nsub = 10 substeps with the source zeroed per step, after one untimed warm-up
    step; the median of `runs` runs.  OpenMP over the host's cores on `nparcels`
    parcels, and the reference's serial loop shape on `nparcels_serial`."""
    from oracle import st_oracle

    ncores = len(os.sched_getaffinity(0))
    res = {}
    for nt, npar in ((ncores, nparcels), (1, nparcels_serial)):
        st0 = [a[:npar].copy() for a in case.state()]

        def run(s):
            for _ in range(calls):
                st_oracle.evolve(s, case.U, case.n, case.origin, case.ending, case.dt_fluid, nsub=case.nsub,
                                 boundary=case.boundary, mu_c=case.mu_c, rho_d=case.rho_d,
                                 rho_d_by_rho_c=case.rho_d_by_rho_c, nthreads=nt)

        st_oracle.evolve([a.copy() for a in st0], case.U, case.n, case.origin, case.ending, case.dt_fluid,
                         nsub=case.nsub, boundary=case.boundary, mu_c=case.mu_c, rho_d=case.rho_d,
                         rho_d_by_rho_c=case.rho_d_by_rho_c, nthreads=nt)  # warm-up
        rates = []
        for _ in range(runs):
            s = [a.copy() for a in st0]
            t0 = time.perf_counter()
            run(s)
            rates.append(npar * case.nsub * calls / (time.perf_counter() - t0))
        res[nt] = (float(np.median(rates)), npar)
    model = cpu_model()
    out = []
    for nt in (ncores, 1):
        rate, npar = res[nt]
        out.append({"value": rate, "unit": "parcel-updates/s", "cores": nt, "kind": "port",
                    "sample": f"{npar} of the 2^24 parcels, 128^3 cells, {calls} fluid steps x {case.nsub} "
                              f"substeps after one warm-up step, median of {runs} runs, FP64 oracle",
                    "cpu_model": model})
    return out[0], out[1]


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
    # the streaming copy runs in every process, so a counter run of any workload holds
    # the copy kernel the traffic counters are calibrated on
    copy_gbs = ctx.copy_bandwidth(1 << 30, reps=10)
    if args.only and args.workload == "batch":
        out = batch_workload(stk, inputs, ctx, steps=args.steps, warmup=args.warmup, tune_cache=args.tune_cache)
        out["calibration_copy_GBs"] = copy_gbs
        print(json.dumps(out))
        return
    if args.only and args.workload == "adjoint":
        print(json.dumps(adjoint_workload(stk, inputs, ctx)))
        return
    if args.only and args.workload == "channel":
        print(json.dumps(channel_workload(stk, inputs, ctx, steps=args.steps, warmup=args.warmup)))
        return
    t0 = time.perf_counter()
    case = inputs.taylor_green_box(np_parcels=args.parcels, ncell=args.cells, boundary=stk.PERIODIC)
    print(f"# inputs generated in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)

    fp32 = args.precision == "fp32"
    precision = stk.FP32 if fp32 else stk.FP64
    brick = [int(v) for v in args.brick.split(",")]
    p = stk.params(boundary=stk.PERIODIC, method=stk.SORTED if args.method == "sorted" else stk.PLAIN,
                   substeps_per_pass=args.substeps_per_pass, brick=brick, block=args.block,
                   passes_per_launch=args.passes_per_launch,
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
            "passes_per_launch": args.passes_per_launch,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": profiled_traffic(PROFILE, {
                "precision": args.precision, "method": args.method, "parcels": case.np, "cells": args.cells,
                "brick": ",".join(str(b) for b in brick), "block": args.block,
                "substeps_per_pass": args.substeps_per_pass, "flush": args.flush}),
            "kernel": "k_pass_sorted" if args.method == "sorted" else "k_pass_plain",
            "kernel_ms_per_pass": head["kernel_ms_per_pass"],
            "bytes_per_pass": bpp,
            # the streaming copy the traffic counters are calibrated on (a known byte count),
            # not a bandwidth ceiling
            "calibration_copy_GBs": copy_gbs,
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
        line["variants"]["fp32_sorted_graph"] = graph_variant(stk, ctx, case, precision, p, args.steps, args.warmup)
        line["fp32_vs_fp64"] = precision_check(stk, ctx, case, ps, calls=5)
        print(f"# fp32 vs fp64: {json.dumps(line['fp32_vs_fp64'])}", file=sys.stderr, flush=True)
    if not args.only:
        line["batch"] = batch_workload(stk, inputs, ctx, tune_cache=args.tune_cache)
        print(f"# batch: {json.dumps(line['batch'])}", file=sys.stderr, flush=True)
        line["channel"] = channel_workload(stk, inputs, ctx)
        print(f"# channel: {json.dumps(line['channel'])}", file=sys.stderr, flush=True)
        line["adjoint"] = adjoint_workload(stk, inputs, ctx)
        line["adjoint_16M"] = adjoint_workload(stk, inputs, ctx, np_parcels=1 << 24, ncell=128, nsub=50, K=10)
        print(f"# adjoint: {json.dumps(line['adjoint'])}", file=sys.stderr, flush=True)
    if not args.no_cpu and not args.only:
        omp, serial = cpu_baseline(case, min(args.cpu_parcels, case.np), min(args.cpu_parcels_serial, case.np))
        line["cpu_baseline"] = omp
        line["cpu_baseline_serial"] = serial
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
