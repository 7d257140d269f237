"""Summaries of rocprofv3 CSV output for the solver's kernels.

  python -m agipc_learning_amd.prof_summary stats  <dir>/<name>_kernel_stats.csv  [--cmd "..."]
  python -m agipc_learning_amd.prof_summary busy   <dir>/<name>_kernel_trace.csv  [--kernels "a;b;..."]
  python -m agipc_learning_amd.prof_summary compare <bench.json> <dir>/<name>_kernel_trace.csv
  python -m agipc_learning_amd.prof_summary traffic <fetch.csv> <write.csv> <copy_fetch.csv> <copy_write.csv>
         [--cmd "..."] [--json out.json]

`stats` prints a markdown table of the kernel-trace statistics.  `busy`
reports each kernel's dispatch durations over the launches that did the work
(at least half the 95th-percentile duration; launches for systems already
converged exit at once and would pull a plain mean down), mean and median.
`compare` sets bench.py's per-kernel times (its JSON line's roofline
object) beside the busy medians of a trace of the same workload.  `traffic`
turns FETCH_SIZE / WRITE_SIZE (KiB per dispatch) into HBM bytes per launch,
each scaled by what the same counter reports for a plain device copy of known
size (the copy's true read and write bytes over its counter values), and
prints a markdown table; `--json` also writes the per-kernel bytes.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import sys

COPY_BYTES = 256 << 20  # bench.py --copy-calibration copies 256 MiB per launch


def _short(name):
    name = name.replace("(anonymous namespace)::", "")
    return name.split("(")[0].replace("void ", "").replace("masx::", "")


def stats(path, cmd=None, top=25):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = []
    if cmd:
        out.append(f"Command: `{cmd}`\n")
    out.append("| kernel | calls | total ms | avg us | min us | max us | % |")
    out.append("|---|---|---|---|---|---|---|")
    for r in rows[:top]:
        out.append(f"| {_short(r['Name'])} | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                   f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['MinNs']) / 1e3:.2f} | "
                   f"{float(r['MaxNs']) / 1e3:.2f} | {float(r['Percentage']):.2f} |")
    return "\n".join(out)


def busy(path, kernels=None):
    d = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        d[_short(r["Kernel_Name"])].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    names = kernels or sorted(d, key=lambda k: -sum(d[k]))
    out = ["| kernel | dispatches (busy/all) | busy avg us | busy median us | busy min us | busy max us |",
           "|---|---|---|---|---|---|"]
    for k in names:
        v = sorted(d.get(k, []))
        if not v:
            continue
        cut = 0.5 * v[min(len(v) - 1, int(0.95 * len(v)))]
        b = [x for x in v if x >= cut]
        out.append(f"| {k} | {len(b)}/{len(v)} | {sum(b) / len(b):.2f} | {b[len(b) // 2]:.2f} | {b[0]:.2f} | "
                   f"{b[-1]:.2f} |")
    return "\n".join(out)


def _busy_median(durations):
    v = sorted(durations)
    cut = 0.5 * v[min(len(v) - 1, int(0.95 * len(v)))]
    b = [x for x in v if x >= cut]
    return b[len(b) // 2]


def mas_local_variant(waves):
    """Template arguments after the element type of the FP32 MAS local kernel
    the variant code runs: 4 / 8 waves, 5 / 9 the same with the inverse
    staged through LDS."""
    return {4: "4, false", 8: "8, false", 5: "4, true", 9: "8, true"}[waves]


def compare(bench_json, trace_csv):
    """bench.py's per-kernel event times (the 64-system batch's FP32 kernels,
    or the 40^3 workload's for --fem40-only) beside the trace's busy medians."""
    line = [l for l in open(bench_json).read().splitlines() if l.strip().startswith("{")][-1]
    bench = json.loads(line)
    if bench.get("batch64"):
        obj = bench["batch64"]
        us = obj["fp32"]["kernels_us"]
    else:
        obj = bench["fem40c"]
        us = {k: v["us"] for k, v in obj["roofline"]["per_kernel"].items()}
    tiling = obj["spmv_tiling"]["entries_per_thread"]
    waves = obj["spmv_tiling"].get("mas_local", {}).get("fp32", {}).get("waves", 4)
    nt = "true" if obj["spmv_tiling"].get("nontemporal") else "false"
    names = {"spmv": f"k_spmv_fused<{tiling}, {nt},", "mas_local": f"k_mas_local<32, float, {mas_local_variant(waves)}>",
             "precond_finish": "k_precond_finish", "update": "k_update"}
    d = collections.defaultdict(list)
    for r in csv.DictReader(open(trace_csv)):
        d[_short(r["Kernel_Name"])].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    out = ["| kernel | bench.py events, unprofiled (us) | rocprofv3 busy median (us) | trace / events |",
           "|---|---|---|---|"]
    for k, n in names.items():
        e = us[k]
        full = [t for t in d if t == n or (n.endswith(",") and t.startswith(n))]
        if full:
            m = _busy_median(d[full[0]])
            out.append(f"| {full[0]} | {e:.2f} | {m:.2f} | {m / e:.3f} |")
    return "\n".join(out)


def _per_kernel(path, counter):
    d = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] == counter:
            d[r["Kernel_Name"]].append(float(r["Counter_Value"]) * 1024.0)
    return d


def _copy_factor(path, counter):
    d = _per_kernel(path, counter)
    vals = d.get("__amd_rocclr_copyBuffer", [])
    big = [v for v in vals if v > COPY_BYTES / 16]
    if not big:
        raise SystemExit(f"no copy kernel in {path}")
    return COPY_BYTES / (sum(big) / len(big))


def _busy(vals):
    """Median over the dispatches that did the work: those of at least half
    the 95th-percentile count (the early-exit launches of systems already
    converged read almost nothing; the median keeps the few heavier launches
    of a kernel, such as the SpMV that also forms the FP64 true residual once
    per solve, from shifting the figure of the regular launch)."""
    v = sorted(vals)
    cut = 0.5 * v[min(len(v) - 1, int(0.95 * len(v)))]
    busy = [x for x in v if x >= cut]
    return busy[len(busy) // 2], len(busy), len(vals)


def traffic(fetch, write, copy_fetch, copy_write, cmd=None, json_out=None):
    fr = _copy_factor(copy_fetch, "FETCH_SIZE")
    fw = _copy_factor(copy_write, "WRITE_SIZE")
    F = _per_kernel(fetch, "FETCH_SIZE")
    W = _per_kernel(write, "WRITE_SIZE")
    out = []
    if cmd:
        out.append(f"Commands: {cmd}\n")
    out.append(f"Copy calibration (256 MiB device copy): read factor {fr:.3f}, write factor {fw:.3f}.\n")
    out.append("| kernel | dispatches (busy/all) | read MB | write MB | HBM MB per launch |")
    out.append("|---|---|---|---|---|")
    res = {}
    for k in sorted(F, key=lambda k: -_busy(F[k])[0]):
        if k not in W:
            continue
        r, nb, na = _busy(F[k])
        w, _, _ = _busy(W[k])
        r, w = r * fr, w * fw
        res[_short(k)] = {"read_bytes": r, "write_bytes": w, "bytes": r + w}
        out.append(f"| {_short(k)} | {nb}/{na} | {r / 1e6:.2f} | {w / 1e6:.2f} | {(r + w) / 1e6:.2f} |")
    if json_out:
        json.dump({"read_factor": fr, "write_factor": fw, "kernels": res}, open(json_out, "w"), indent=1)
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    a = sub.add_parser("stats")
    a.add_argument("csv")
    a.add_argument("--cmd")
    cp = sub.add_parser("compare")
    cp.add_argument("bench_json")
    cp.add_argument("trace_csv")
    c = sub.add_parser("busy")
    c.add_argument("csv")
    c.add_argument("--kernels")
    b = sub.add_parser("traffic")
    b.add_argument("fetch")
    b.add_argument("write")
    b.add_argument("copy_fetch")
    b.add_argument("copy_write")
    b.add_argument("--cmd")
    b.add_argument("--json")
    args = ap.parse_args(argv)
    if args.what == "stats":
        print(stats(args.csv, args.cmd))
    elif args.what == "compare":
        print(compare(args.bench_json, args.trace_csv))
    elif args.what == "busy":
        print(busy(args.csv, args.kernels.split(";") if args.kernels else None))
    else:
        print(traffic(args.fetch, args.write, args.copy_fetch, args.copy_write, args.cmd, args.json))


if __name__ == "__main__":
    sys.exit(main())
