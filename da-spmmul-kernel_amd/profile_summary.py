#!/usr/bin/env python3
"""Turns rocprofv3 output of profile_workload.py into a short summary.

Inputs: a kernel-trace results database (or kernel_stats CSV) and the two
counter-collection CSVs (FETCH_SIZE, WRITE_SIZE); or, in the counter mode
(--pmc), any counter-collection CSVs, summarised per kernel with derived
busy fractions, hit rates and L1-miss latency.  Output: a JSON file with,
per kernel, the average duration and the average FETCH/WRITE KiB per
dispatch, plus the HBM traffic per dispatch scaled by the copy calibration
(the device-to-device copy of known size in the same run), and a Markdown
table beside it.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import sqlite3
from collections import defaultdict


def short(name: str) -> str:
    base = name.split("(")[0].replace("void ", "").strip()
    return base.replace("dspmm::", "")


def kernel_stats(path):
    out = {}
    if path.endswith(".db"):
        con = sqlite3.connect(path)
        for name, calls, avg in con.execute("select name, total_calls, average from top_kernels"):
            out[name] = {"calls": calls, "avg_ns": avg}
    else:
        with open(path) as f:
            for r in csv.DictReader(f):
                out[r["Name"]] = {"calls": int(r["Calls"]), "avg_ns": float(r["AverageNs"])}
    return out


def counters(path, counter):
    """Per kernel: average counter value per dispatch; for the runtime's copy
    kernel only the largest dispatch (the calibration copy) is kept."""
    vals = defaultdict(list)
    with open(path) as f:
        for r in csv.DictReader(f):
            if r.get("Counter_Name") != counter:
                continue
            vals[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    out = {}
    for k, v in vals.items():
        out[k] = max(v) if "copyBuffer" in k else sum(v) / len(v)
    return out


def last_dispatches(trace_csv, prefix, n):
    """Average duration (ns) of the last n dispatches whose kernel name
    (without 'void ' and namespaces) starts with prefix."""
    durs = []
    with open(trace_csv) as f:
        for r in csv.DictReader(f):
            if short(r["Kernel_Name"]).startswith(prefix):
                durs.append((int(r["Start_Timestamp"]),
                             int(r["End_Timestamp"]) - int(r["Start_Timestamp"]),
                             short(r["Kernel_Name"])))
    durs.sort()
    tail = durs[-n:]
    if not tail:
        return None
    return {"kernel": tail[-1][2], "dispatches": len(tail),
            "avg_ns": sum(d for _, d, _ in tail) / len(tail)}


def passes_after_copy(trace_csv, passes):
    """The dispatches after the last runtime copy (the calibration copy that
    precedes the timed passes), cut into `passes` equal runs: per run, the
    dispatch count, the wall time from first start to last end, and the sum
    of the kernels' own durations."""
    rows = []
    with open(trace_csv) as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    rows.sort()
    copies = [i for i, r in enumerate(rows) if "copyBuffer" in r[2]]
    after = rows[copies[-1] + 1:] if copies else rows
    n = len(after) // passes
    out = []
    for p in range(passes):
        seg = after[p * n:(p + 1) * n]
        if seg:
            out.append({"dispatches": len(seg), "wall_us": (seg[-1][1] - seg[0][0]) / 1e3,
                        "kernel_us": sum(e - b for b, e, _ in seg) / 1e3})
    return out


def trace_only(args):
    stats = kernel_stats(args.trace)
    rows = sorted(stats.items(), key=lambda kv: -kv[1]["avg_ns"] * kv[1]["calls"])
    with open(args.out, "w") as f:
        f.write(f"# rocprofv3 kernel trace: {args.title}\n\n")
        f.write(f"Command: `{args.command}`\n\n")
        if args.passes:
            trace_csv = args.trace.replace("kernel_stats.csv", "kernel_trace.csv")
            f.write("Timed passes (the dispatches after the calibration copy, in equal runs):\n\n")
            f.write("| pass | dispatches | wall us | summed kernel us |\n|---|---|---|---|\n")
            for i, p in enumerate(passes_after_copy(trace_csv, args.passes)):
                f.write(f"| {i} | {p['dispatches']} | {p['wall_us']:.1f} | {p['kernel_us']:.1f} |\n")
            f.write("\n")
        if args.last and args.kernel:
            trace_csv = args.trace.replace("kernel_stats.csv", "kernel_trace.csv")
            t = last_dispatches(trace_csv, args.kernel, args.last)
            if t:
                f.write(f"Timed region: the last {t['dispatches']} dispatches of `{t['kernel']}` "
                        f"average {t['avg_ns'] / 1e3:.1f} us\n\n")
        f.write("| kernel | calls | avg us | total us |\n|---|---|---|---|\n")
        for name, st in rows[: args.top]:
            f.write(f"| {short(name)} | {st['calls']} | {st['avg_ns'] / 1e3:.1f} | "
                    f"{st['avg_ns'] * st['calls'] / 1e3:.0f} |\n")
    print(json.dumps({"out": args.out}))


# Derived figures of the counter mode, from the raw counters they need.
def derived(v, cus, xcds):
    out = {}
    cyc = v.get("GRBM_GUI_ACTIVE")
    per_cu = cyc / xcds if cyc else None          # the counter sums the XCDs' clocks
    if per_cu:
        for name, key in (("TA busy", "TA_TA_BUSY_sum"), ("TD busy", "TD_TD_BUSY_sum"),
                          ("TD stalled on L1", "TD_TC_STALL_sum")):
            if key in v:
                out[name] = v[key] / cus / per_cu
    if "TCP_TCC_READ_REQ_sum" in v and v.get("TCP_TOTAL_CACHE_ACCESSES_sum"):
        out["L1 hit rate"] = 1.0 - v["TCP_TCC_READ_REQ_sum"] / v["TCP_TOTAL_CACHE_ACCESSES_sum"]
    if "TCC_HIT_sum" in v and "TCC_MISS_sum" in v:
        out["L2 hit rate"] = v["TCC_HIT_sum"] / max(1.0, v["TCC_HIT_sum"] + v["TCC_MISS_sum"])
    if v.get("TCP_TCC_READ_REQ_sum") and "TCP_TCC_READ_REQ_LATENCY_sum" in v:
        out["L1-miss latency (cycles)"] = (v["TCP_TCC_READ_REQ_LATENCY_sum"] /
                                           v["TCP_TCC_READ_REQ_sum"])
    if per_cu and "TCP_TCC_READ_REQ_sum" in v and "L1-miss latency (cycles)" in out:
        # Little's law: L1 misses in flight per CU
        out["L1 misses in flight per CU"] = (v["TCP_TCC_READ_REQ_sum"] / cus / per_cu *
                                             out["L1-miss latency (cycles)"])
    if v.get("SQ_INSTS_VMEM_RD") and v.get("SQ_WAVES"):
        out["vector-memory reads per wave"] = v["SQ_INSTS_VMEM_RD"] / v["SQ_WAVES"]
    if per_cu and "SQ_INSTS_VALU" in v:
        # a wave64 VALU instruction (no packed math) issues over 4 cycles on
        # one of a CU's 4 SIMDs: instructions / (CUs x cycles) is the share
        # of SIMD cycles spent issuing VALU work
        out["VALU issue share"] = v["SQ_INSTS_VALU"] / cus / per_cu
    if per_cu and "SQ_LDS_IDX_ACTIVE" in v:
        out["LDS active share"] = v["SQ_LDS_IDX_ACTIVE"] / cus / per_cu
    if v.get("SQ_LDS_IDX_ACTIVE") and "SQ_LDS_BANK_CONFLICT" in v:
        out["LDS bank-conflict share of LDS active"] = (v["SQ_LDS_BANK_CONFLICT"] /
                                                        v["SQ_LDS_IDX_ACTIVE"])
    if v.get("SQ_WAVE_CYCLES") and "SQ_WAIT_ANY" in v:
        out["wave cycles waiting"] = v["SQ_WAIT_ANY"] / v["SQ_WAVE_CYCLES"]
    return out


def pmc_mode(args):
    """Per kernel (fills and copies aside): the average of every counter per
    dispatch over all given counter-collection CSVs, and derived figures."""
    tables = []
    for spec in args.pmc:
        tag, pattern = spec.split("=", 1)
        vals = defaultdict(lambda: defaultdict(list))
        for path in sorted(glob.glob(pattern)):
            with open(path) as f:
                for r in csv.DictReader(f):
                    k = short(r["Kernel_Name"])
                    if any(x in k.lower() for x in ("fill", "copy", "to_bf16")):
                        continue
                    vals[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
        for k, cs in vals.items():
            avg = {c: sum(x) / len(x) for c, x in cs.items()}
            tables.append((tag, k, avg, derived(avg, args.cus, args.xcds)))
    with open(args.out, "w") as f:
        f.write(f"# rocprofv3 counters: {args.title}\n\n")
        f.write(f"Command: `{args.command}`\n\n")
        f.write(f"Per-dispatch averages; busy fractions are per CU over the kernel's clock "
                f"({args.cus} CUs, {args.xcds} XCDs).\n\n")
        for tag, k, avg, d in tables:
            f.write(f"## {tag}: `{k}`\n\n| figure | value |\n|---|---|\n")
            for name, x in d.items():
                f.write(f"| {name} | {x:.3g} |\n")
            for name, x in sorted(avg.items()):
                f.write(f"| {name} | {x:.4g} |\n")
            f.write("\n")
    print(json.dumps({"out": args.out, "kernels": [t[1] for t in tables]}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", help="kernel-trace .db or kernel_stats.csv")
    ap.add_argument("--pmc", nargs="*", default=[],
                    help="counter mode: TAG=GLOB of counter_collection CSVs, one per config")
    ap.add_argument("--cus", type=int, default=256)
    ap.add_argument("--xcds", type=int, default=8)
    ap.add_argument("--fetch", help="counter_collection.csv with FETCH_SIZE")
    ap.add_argument("--write", help="counter_collection.csv with WRITE_SIZE")
    ap.add_argument("--meta", help="JSON line printed by profile_workload.py")
    ap.add_argument("--out", required=True, help="output .json (a .md is written beside it)")
    ap.add_argument("--command", default="", help="command line that produced the inputs")
    ap.add_argument("--title", default="", help="trace-only mode: title of the table")
    ap.add_argument("--top", type=int, default=12)
    ap.add_argument("--kernel", default="", help="trace-only mode: kernel name prefix whose "
                    "last --last dispatches are averaged (the timed steps)")
    ap.add_argument("--last", type=int, default=0)
    ap.add_argument("--passes", type=int, default=0,
                    help="trace-only mode: cut the dispatches after the calibration copy into "
                         "this many timed passes")
    args = ap.parse_args()
    if args.pmc:
        return pmc_mode(args)
    if not args.trace:
        raise SystemExit("--trace is required outside the counter mode")
    if not args.fetch:
        return trace_only(args)

    def one(pattern):
        m = sorted(glob.glob(pattern, recursive=True))
        if not m:
            raise SystemExit(f"no file matches {pattern}")
        return m[0]

    meta = next(json.loads(l) for l in open(one(args.meta)) if l.startswith('{"workload"'))
    stats = kernel_stats(one(args.trace))
    fetch = counters(one(args.fetch), "FETCH_SIZE")
    write = counters(one(args.write), "WRITE_SIZE")
    copy = [k for k in fetch if "copyBuffer" in k or "copy" in k.lower()]
    if not copy:
        raise SystemExit("calibration copy kernel not found in counters")
    ck = copy[0]
    cal_f = meta["copy_bytes"] / (fetch[ck] * 1024.0)
    cal_w = meta["copy_bytes"] / (write[ck] * 1024.0)
    kernels = {}
    for name in sorted(set(fetch) | set(write)):
        s = short(name)
        st = next((v for k, v in stats.items() if short(k) == s), None)
        f, w = fetch.get(name, 0.0), write.get(name, 0.0)
        kernels[s] = {"full_name": name, "FETCH_SIZE_KiB": f, "WRITE_SIZE_KiB": w,
                      "traffic_bytes": f * 1024 * cal_f + w * 1024 * cal_w,
                      "avg_ns": st["avg_ns"] if st else None,
                      "calls": st["calls"] if st else None}
    cfg = meta["config"]
    main_kernel = None
    if cfg.get("variant") == "nnz_tiled":
        tb = "unsigned short" if meta.get("bf16") else "float"
        want = f"k_nnz_tiled<{tb}, {cfg['vec']}, {cfg['lanes']}, {4 if cfg['unroll'] <= 4 else 8}>"
        main_kernel = want if want in kernels else None
    if main_kernel is None:
        # otherwise the kernel with the most time per call, runtime copies and fills aside
        cand = {k: v for k, v in kernels.items()
                if v["avg_ns"] and "copyBuffer" not in k and "fill" not in k}
        main_kernel = max(cand, key=lambda k: cand[k]["avg_ns"]) if cand else None
    res = {"meta": meta, "main_kernel": main_kernel,
           "calibration": {"fetch": cal_f, "write": cal_w, "copy_kernel": ck},
           "command": args.command, "kernels": kernels}
    with open(args.out, "w") as f:
        json.dump(res, f, indent=1)
    md = os.path.splitext(args.out)[0] + ".md"
    with open(md, "w") as f:
        f.write(f"# rocprofv3 summary: {meta['workload']} N={meta['n']}"
                f"{' BF16-B' if meta.get('bf16') else ''}\n\n")
        f.write(f"Command: `{args.command}`\n\n")
        f.write(f"Config: `{json.dumps(meta['config'])}`; algorithmic bytes per SpMM "
                f"{meta['algorithmic_bytes']}\n\n")
        f.write(f"Calibration (copy of {meta['copy_bytes']} B): fetch x{cal_f:.3f}, "
                f"write x{cal_w:.3f}\n\n")
        if main_kernel:
            k = kernels[main_kernel]
            f.write(f"Main kernel `{main_kernel}`: HBM traffic {k['traffic_bytes']:.4e} B per "
                    f"launch = {k['traffic_bytes'] / meta['algorithmic_bytes']:.2f}x the "
                    f"algorithmic bytes\n\n")
        f.write("| kernel | calls | avg us | FETCH KiB | WRITE KiB | HBM bytes (calibrated) |\n")
        f.write("|---|---|---|---|---|---|\n")
        for s, k in kernels.items():
            avg = f"{k['avg_ns'] / 1e3:.1f}" if k["avg_ns"] else "-"
            f.write(f"| {s} | {k['calls']} | {avg} | {k['FETCH_SIZE_KiB']:.0f} | "
                    f"{k['WRITE_SIZE_KiB']:.0f} | {k['traffic_bytes']:.3e} |\n")
    print(json.dumps({"out": args.out, "kernels": list(kernels)}))


if __name__ == "__main__":
    main()
