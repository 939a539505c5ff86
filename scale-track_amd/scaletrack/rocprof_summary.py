"""Summaries of rocprofv3 runs of the parcel step, as committed next to the benchmark.

Reads the CSV files rocprofv3 writes (``--kernel-trace --stats`` gives
``*_kernel_stats.csv``; ``--pmc FETCH_SIZE`` and ``--pmc WRITE_SIZE`` give one
``*_counter_collection.csv`` each) and prints a Markdown summary:

* the kernel statistics table (calls, average and total duration per kernel);
* FETCH_SIZE and WRITE_SIZE per dispatch (KiB, summed over the counter's instances)
  for the measured kernel and for a streaming copy kernel of known size run in the same
  process, the calibration factor of each counter (copy bytes / counter bytes), and the
  measured kernel's HBM traffic per launch with both factors applied.

``--pmc`` adds a table of any other counters collected for the kernel (with VALU
instructions per wave, and with ``--simds`` the VALU issue rate per SIMD against one
wave64 instruction every 4 cycles at the peak clock).

The last block is ``key: value`` lines that bench.py reads back (``traffic_bytes``
and the workload keys it was measured on).

    python -m scaletrack.rocprof_summary --stats T/b_kernel_stats.csv \\
        --fetch F/b_counter_collection.csv --write W/b_counter_collection.csv \\
        --kernel k_pass_sorted --copy-kernel k_copy --copy-bytes 1073741824 \\
        --algorithmic-bytes 998244352 --set precision=fp32 ...
"""
import argparse
import collections
import csv
import sys


def kernel_stats(path):
    """Rows of a rocprofv3 kernel_stats.csv as dicts (name, calls, avg_ns, total_ns, pct)."""
    out = []
    with open(path) as f:
        for r in csv.DictReader(f):
            out.append({"name": r["Name"], "calls": int(r["Calls"]), "avg_ns": float(r["AverageNs"]),
                        "total_ns": float(r["TotalDurationNs"]), "pct": float(r["Percentage"])})
    return out


def counter_per_dispatch(path, counter):
    """{kernel name: [value per dispatch]} for one counter, summing its instances
    (rocprofv3 writes one row per dispatch and counter instance)."""
    acc = collections.defaultdict(float)
    names = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            acc[r["Dispatch_Id"]] += float(r["Counter_Value"])
            names[r["Dispatch_Id"]] = r["Kernel_Name"]
    out = collections.defaultdict(list)
    for did, v in acc.items():
        out[names[did]].append(v)
    return out


def pick(per_kernel, key):
    """Values of the one kernel whose name contains key."""
    hits = [k for k in per_kernel if key in k]
    if len(hits) != 1:
        raise SystemExit(f"expected one kernel matching {key!r}, found {hits}")
    return hits[0], per_kernel[hits[0]]


def mean(v):
    return sum(v) / len(v)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--stats", required=True, help="kernel_stats.csv of a --kernel-trace --stats run")
    ap.add_argument("--fetch", help="counter_collection.csv of a --pmc FETCH_SIZE run")
    ap.add_argument("--write", help="counter_collection.csv of a --pmc WRITE_SIZE run")
    ap.add_argument("--kernel", required=True, help="substring naming the measured kernel")
    ap.add_argument("--copy-kernel", default="k_copy", help="substring naming the calibration copy kernel")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30, help="bytes the copy kernel reads (and writes)")
    ap.add_argument("--algorithmic-bytes", type=float, default=None, help="algorithmic bytes per launch")
    ap.add_argument("--title", default="rocprofv3 summary")
    ap.add_argument("--command", action="append", default=[], help="command line recorded in the summary")
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE", help="workload key recorded")
    ap.add_argument("--simds", type=int, default=0, help="SIMDs of the device (CUs x 4) for the VALU utilisation")
    ap.add_argument("--clock-ghz", type=float, default=2.4, help="peak engine clock for the VALU issue bound")
    ap.add_argument("--pmc", action="append", default=[],
                    help="counter_collection.csv of a --pmc run: its counters for the kernel are tabulated")
    a = ap.parse_args(argv)

    w = sys.stdout.write
    w(f"# {a.title}\n\n")
    if a.command:
        w("Commands (run from the repository root):\n\n```\n" + "\n".join(a.command) + "\n```\n\n")
    stats = kernel_stats(a.stats)
    w("## Kernel statistics\n\n| kernel | calls | average (us) | total (ms) | share |\n|---|---|---|---|---|\n")
    for r in stats:
        w(f"| `{r['name'][:90]}` | {r['calls']} | {r['avg_ns'] / 1e3:.2f} | {r['total_ns'] / 1e6:.3f} | "
          f"{r['pct']:.1f}% |\n")
    kname, krow = None, None
    for r in stats:
        if a.kernel in r["name"]:
            kname, krow = r["name"], r
            break
    if krow is None:
        raise SystemExit(f"no kernel matching {a.kernel!r} in {a.stats}")
    keys = {"kernel": a.kernel, "kernel_avg_ns": f"{krow['avg_ns']:.0f}"}
    if a.fetch and a.write:
        fk, fv = pick(counter_per_dispatch(a.fetch, "FETCH_SIZE"), a.kernel)
        ck, cv = pick(counter_per_dispatch(a.fetch, "FETCH_SIZE"), a.copy_kernel)
        wk, wv = pick(counter_per_dispatch(a.write, "WRITE_SIZE"), a.kernel)
        cwk, cwv = pick(counter_per_dispatch(a.write, "WRITE_SIZE"), a.copy_kernel)
        f_fac = a.copy_bytes / (mean(cv) * 1024.0)
        w_fac = a.copy_bytes / (mean(cwv) * 1024.0)
        fetch_b = mean(fv) * 1024.0 * f_fac
        write_b = mean(wv) * 1024.0 * w_fac
        w("\n## HBM traffic counters (KiB per dispatch, mean over dispatches)\n\n")
        w("| kernel | dispatches | FETCH_SIZE | WRITE_SIZE |\n|---|---|---|---|\n")
        w(f"| `{kname[:90]}` | {len(fv)} | {mean(fv):.0f} | {mean(wv):.0f} |\n")
        w(f"| `{ck[:90]}` ({a.copy_bytes} B copy) | {len(cv)} | {mean(cv):.0f} | {mean(cwv):.0f} |\n")
        w(f"\nCalibration: the copy reads and writes {a.copy_bytes} B; FETCH_SIZE x {f_fac:.3f} and "
          f"WRITE_SIZE x {w_fac:.3f} give bytes.\n\n")
        w(f"Measured kernel per launch: {fetch_b / 1e6:.1f} MB read + {write_b / 1e6:.1f} MB written = "
          f"{(fetch_b + write_b) / 1e6:.1f} MB")
        keys.update({"fetch_bytes": f"{fetch_b:.0f}", "write_bytes": f"{write_b:.0f}",
                     "traffic_bytes": f"{fetch_b + write_b:.0f}", "fetch_factor": f"{f_fac:.4f}",
                     "write_factor": f"{w_fac:.4f}"})
        if a.algorithmic_bytes:
            w(f" ({(fetch_b + write_b) / a.algorithmic_bytes:.3f} x the {a.algorithmic_bytes / 1e6:.1f} MB "
              "algorithmic bytes)")
            keys["algorithmic_bytes"] = f"{a.algorithmic_bytes:.0f}"
        w(".\n")
    if a.pmc:
        per = {}
        for path in a.pmc:
            with open(path) as f:
                names = sorted({r["Counter_Name"] for r in csv.DictReader(f)})
            for c in names:
                _, v = pick(counter_per_dispatch(path, c), a.kernel)
                per[c] = mean(v)
        w("\n## Other counters (per dispatch, mean over dispatches, summed over instances)\n\n")
        w("| counter | value |\n|---|---|\n")
        for c, v in per.items():
            w(f"| {c} | {v:.4g} |\n")
        if "SQ_INSTS_VALU" in per and "SQ_WAVES" in per:
            w(f"\nVALU instructions per wave: {per['SQ_INSTS_VALU'] / per['SQ_WAVES']:.0f}.\n")
            keys["valu_per_wave"] = f"{per['SQ_INSTS_VALU'] / per['SQ_WAVES']:.0f}"
        if "SQ_INSTS_VALU" in per and a.simds:
            # a wave64 VALU instruction holds a 16-lane SIMD for 4 cycles: at most
            # clock / 4 instructions per SIMD per second over the traced duration
            rate = per["SQ_INSTS_VALU"] / a.simds / (krow["avg_ns"] * 1e-9)
            bound = a.clock_ghz * 1e9 / 4.0
            w(f"VALU instructions per SIMD per microsecond over the traced {krow['avg_ns'] / 1e3:.1f} us: "
              f"{rate / 1e6:.0f}, against {bound / 1e6:.0f} for one wave64 instruction every 4 cycles at "
              f"{a.clock_ghz} GHz ({100 * rate / bound:.0f}%).\n")
            keys["valu_issue_fraction"] = f"{rate / bound:.3f}"
    for kv in a.set:
        k, _, v = kv.partition("=")
        keys[k] = v
    w("\n## Figures read by bench.py\n\n```\n" + "".join(f"{k}: {v}\n" for k, v in keys.items()) + "```\n")


def read_figures(path):
    """The ``key: value`` block of a summary written by main()."""
    out = {}
    with open(path) as f:
        text = f.read()
    block = text.split("## Figures read by bench.py", 1)
    if len(block) < 2:
        return out
    body = block[1].split("```")
    if len(body) < 2:
        return out
    for line in body[1].splitlines():
        k, sep, v = line.partition(":")
        if sep:
            out[k.strip()] = v.strip()
    return out


if __name__ == "__main__":
    main()
