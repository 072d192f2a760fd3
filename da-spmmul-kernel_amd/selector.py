#!/usr/bin/env python3
"""Data-aware kernel selection: a nearest-neighbour model fitted on sweep data.

The reference fixes the kernel by a static (alg, layout, nv) switch
(spmvspmm.h:72-197) and leaves the choice among its row-/element-balanced,
sequential/parallel variants to whoever reads the sweep CSV
(spmvspmm_benchmark.cu:277-326, run.sh).  This script closes that loop:

  collect  (on the GPU) times every candidate the autotuner knows
           (dspmm_autotune_table) for a corpus of synthetic matrices x dense
           widths, and writes one CSV row per (matrix, nv, candidate);
  fit      reads that CSV, evaluates the selector leave-one-matrix-out
           against the static heuristic and the per-matrix optimum, and
           writes csrc/selector_table.inc, which the library compiles in.

Model.  Features of a matrix (all available to the library without timing
anything): log2 rows, log2 mean row length, coefficient of variation of the
row length, log2(max / mean row length), and the column reuse of 16-row
windows (nonzeros per distinct column; 1 for random columns, ~2.5-3 for a
27-point stencil).  For a query (matrix, nv) the k = 3 nearest corpus
matrices at the nearest corpus nv (standardised Euclidean distance) vote:
the configuration with the smallest summed log(time / best time) over the
neighbours' top lists wins.  Row-major FP32 only; other operand types keep
the static heuristic.

  python3 da-spmmul-kernel_amd/selector.py collect --out sel.csv     (GPU)
  python3 da-spmmul-kernel_amd/selector.py fit --csv sel.csv \\
      --table da-spmmul-kernel_amd/csrc/selector_table.inc --report report.json
"""
from __future__ import annotations

import argparse
import csv
import json
import math
import os
import sys
from collections import defaultdict

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

NVS = (1, 8, 32, 64, 128)
TOP = 8            # configurations kept per corpus entry
K = 3              # neighbours that vote
FEATURES = ("log_rows", "log_mean", "cv", "log_skew", "reuse")
CFG_FIELDS = ("variant", "vec", "lanes", "tile_nnz", "unroll", "row_stride")


def corpus():
    """(name, Csr) pairs: power-law, uniform, fixed-degree and stencil
    families over several sizes (seeded)."""
    import workloads as W
    out = []
    seed = 900
    for s in (1.5, 1.8, 2.2):
        for m in (16384, 131072):
            seed += 1
            out.append((f"zipf{s}_{m}", W.gen_csr(W.GEN_ZIPF, m, m, seed, s, 1, 2000,
                                                  W.VAL_NORMAL)))
    for d in (2, 8, 32, 128):
        for m in (8192, 65536):
            seed += 1
            out.append((f"uniform{d}_{m}", W.gen_csr(W.GEN_BINOMIAL, m, m, seed, p0=float(d))))
    for d in (4, 16, 64):
        seed += 1
        out.append((f"fixed{d}_32768", W.gen_csr(W.GEN_FIXED, 32768, 32768, seed, p0=d)))
    for n in (20, 40, 64):
        out.append((f"stencil{n}", W.c3_matrix(n)))
    return out


def features(a):
    """The features of a Csr, computed as the library computes them."""
    import numpy as np
    lens = np.diff(a.rowPtr).astype(np.float64)
    mean = float(lens.mean()) if lens.size else 0.0
    cv = float(lens.std() / mean) if mean > 0 else 0.0
    mx = float(lens.max()) if lens.size else 0.0
    # column reuse over up to 256 windows of 16 rows spread over the matrix
    win = 16
    nwin = min(256, max(1, a.nr // win))
    nz = distinct = 0
    for w in range(nwin):
        r0 = (w * (a.nr - min(win, a.nr))) // max(1, nwin - 1)
        r1 = min(a.nr, r0 + win)
        p0, p1 = int(a.rowPtr[r0]), int(a.rowPtr[r1])
        if p1 <= p0:
            continue
        cols = a.colIdx[p0:p1]
        nz += cols.size
        distinct += np.unique(cols).size
    reuse = nz / distinct if distinct else 1.0
    return {"log_rows": math.log2(max(1, a.nr)), "log_mean": math.log2(max(mean, 1e-3)),
            "cv": cv, "log_skew": math.log2(max(mx, 1.0) / max(mean, 1e-3)), "reuse": reuse}


def collect(out_path, reps=5):
    import dspmm as D
    import workloads as W
    rows = []
    for name, a in corpus():
        f = features(a)
        m = D.Matrix(a.nr, a.nc, a.rowPtr, a.colIdx, a.values)
        nvmax = max(NVS)
        d_B = D.DeviceArray.from_numpy(W.gen_dense(7, a.nc * nvmax))
        d_C = D.DeviceArray.zeros(a.nr * nvmax)
        for nv in NVS:
            tab = m.autotune_table(d_B, d_C, nv, reps=reps)
            ok = [(c, ms) for c, ms in tab if ms is not None]
            for c, ms in ok:
                cfg = {k: getattr(c, k) for k in CFG_FIELDS}
                # the only row stride above 1 the autotuner tries for row blocks
                # is the matrix's detected one; stored as -1 ("detected") so it
                # transfers between matrices (a ring's row_stride is its walk
                # length and stays as it is)
                if cfg["row_stride"] > 1 and cfg["variant"] == 8:
                    cfg["row_stride"] = -1
                rows.append({"matrix": name, "nv": nv, **f, "nnz": a.nnz, **cfg, "ms": ms})
            best = min(ok, key=lambda x: x[1])
            print(f"{name} nv={nv}: {len(ok)} candidates, best {best[1] * 1e3:.1f} us "
                  f"{best[0].as_dict()}", flush=True)
        m.close()
    with open(out_path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=list(rows[0].keys()))
        w.writeheader()
        w.writerows(rows)


# ---- fitting ---------------------------------------------------------------

def load(path):
    """{(matrix, nv): {"f": features, "t": {cfg tuple: ms}}}"""
    data = defaultdict(lambda: {"f": None, "t": {}})
    with open(path) as fh:
        for r in csv.DictReader(fh):
            key = (r["matrix"], int(r["nv"]))
            data[key]["f"] = {k: float(r[k]) for k in FEATURES}
            cfg = tuple(int(r.get(k) or 0) for k in CFG_FIELDS)
            data[key]["t"][cfg] = float(r["ms"])
    return dict(data)


def heuristic_cfg(nv, table):
    """The library's static default for row-major FP32 (the configuration
    AUTO resolves to without the selector): SpMV tiles of 1024 for nv == 1,
    else the nnz-tiled kernel with its default vector width, lanes, tile 16 *
    lanes clamped to [16, 256] and unroll 8 -- looked up in the table by
    its defaults."""
    if nv == 1:
        return (6, 0, 0, 1024, 0, 0)
    cands = [c for c in table if c[0] == 1 and c[4] == 8]
    if not cands:
        return None
    vec = max(c[1] for c in cands)
    lanes = min(64, 1 << max(0, math.ceil(math.log2(max(1, -(-nv // vec))))))
    return (1, vec, lanes, max(16, min(256, 16 * lanes)), 8, 0)


def scales(data):
    import numpy as np
    X = np.array([[v["f"][k] for k in FEATURES] for v in data.values()])
    sd = X.std(axis=0)
    sd[sd == 0] = 1.0
    return dict(zip(FEATURES, sd.tolist()))


def top_list(table):
    best = min(table.values())
    ranked = sorted(table.items(), key=lambda kv: kv[1])[:TOP]
    return [(c, math.log(ms / best)) for c, ms in ranked]


def predict(fq, nv, entries, sd, k=K):
    """entries: [(matrix, nv, features, top list)]; returns the voted config."""
    nvs = sorted({e[1] for e in entries}, key=lambda x: abs(math.log2(x) - math.log2(nv)))
    pool = [e for e in entries if e[1] == nvs[0]]
    dist = sorted(pool, key=lambda e: sum(((fq[f] - e[2][f]) / sd[f]) ** 2 for f in FEATURES))
    nb = dist[:k]
    score = defaultdict(float)
    cands = {c for e in nb for c, _ in e[3]}
    for c in cands:
        for e in nb:
            d = dict(e[3])
            score[c] += d.get(c, 2.0)          # a config missing from a top list counts as e^2 slower
    return min(score, key=score.get)


def fit(csv_path, table_path, report_path):
    data = load(csv_path)
    sd = scales(data)
    entries = [(m, nv, v["f"], top_list(v["t"])) for (m, nv), v in data.items()]
    # leave-one-matrix-out
    res = []
    for (m, nv), v in sorted(data.items()):
        others = [e for e in entries if e[0] != m]
        pick = predict(v["f"], nv, others, sd)
        best = min(v["t"].values())
        t_pick = v["t"].get(pick)
        h = heuristic_cfg(nv, v["t"])
        t_h = v["t"].get(h)
        res.append({"matrix": m, "nv": nv, "best_us": best * 1e3,
                    "selector_us": None if t_pick is None else t_pick * 1e3,
                    "heuristic_us": None if t_h is None else t_h * 1e3,
                    "selector_cfg": list(pick), "heuristic_cfg": list(h) if h else None})

    def geo(key):
        xs = [r[key] / r["best_us"] for r in res if r[key] is not None]
        return math.exp(sum(math.log(x) for x in xs) / len(xs)) if xs else None, len(xs)

    summary = {"entries": len(res),
               "selector_geomean_slowdown_vs_best": geo("selector_us")[0],
               "selector_covered": geo("selector_us")[1],
               "heuristic_geomean_slowdown_vs_best": geo("heuristic_us")[0],
               "heuristic_covered": geo("heuristic_us")[1],
               "selector_worst": max((r["selector_us"] / r["best_us"] for r in res
                                      if r["selector_us"]), default=None),
               "heuristic_worst": max((r["heuristic_us"] / r["best_us"] for r in res
                                       if r["heuristic_us"]), default=None)}
    with open(report_path, "w") as fh:
        json.dump({"summary": summary, "leave_one_matrix_out": res}, fh, indent=1)
    print(json.dumps(summary, indent=1))
    write_table(entries, sd, table_path, summary)


def write_table(entries, sd, path, summary):
    lines = ["// selector_table.inc -- generated by da-spmmul-kernel_amd/selector.py fit;",
             "// do not edit.  Corpus entries of the nearest-neighbour kernel selector:",
             "// features, dense width and the entry's fastest configurations with",
             "// log(time / best time), measured on one MI355X.",
             f"// Leave-one-matrix-out: selector geomean slowdown vs the per-matrix best "
             f"{summary['selector_geomean_slowdown_vs_best']:.3f}, static heuristic "
             f"{summary['heuristic_geomean_slowdown_vs_best']:.3f}.",
             f"static const int kSelTop = {TOP};",
             f"static const int kSelK = {K};",
             "static const double kSelScale[5] = {" + ", ".join(f"{sd[f]:.6g}" for f in FEATURES)
             + "};",
             f"static const int kSelEntries = {len(entries)};",
             "struct SelEntry {",
             "    int nv;",
             "    double f[5];                    // " + ", ".join(FEATURES),
             "    int cfg[kSelTop][6];            // variant, vec, lanes, tile_nnz, unroll, row_stride",
             "    float logratio[kSelTop];        // log(time / best time); 99 = no entry",
             "};",
             "static const SelEntry kSelTable[] = {"]
    for m, nv, f, top in sorted(entries, key=lambda e: (e[1], e[0])):
        cfgs = [list(c) for c, _ in top] + [[0] * len(CFG_FIELDS)] * (TOP - len(top))
        lrs = [lr for _, lr in top] + [99.0] * (TOP - len(top))
        lines.append(f"    {{{nv}, {{{', '.join(f'{f[k]:.6g}' for k in FEATURES)}}},  // {m}")
        lines.append("     {" + ", ".join("{" + ", ".join(str(x) for x in c) + "}" for c in cfgs)
                     + "},")
        lines.append("     {" + ", ".join(f"{x:.4f}f" for x in lrs) + "}},")
    lines.append("};")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--out", required=True)
    c.add_argument("--reps", type=int, default=5)
    f = sub.add_parser("fit")
    f.add_argument("--csv", required=True)
    f.add_argument("--table", required=True)
    f.add_argument("--report", required=True)
    args = ap.parse_args()
    if args.cmd == "collect":
        collect(args.out, args.reps)
    else:
        fit(args.csv, args.table, args.report)


if __name__ == "__main__":
    main()
