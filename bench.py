#!/usr/bin/env python3
"""Headline benchmark: SpMM GFLOP/s and HBM-roofline fraction on one MI355X.

Headline workload (one "step" = one C = A*B over the whole matrix, inputs
resident in HBM): the 27-point 3-D stencil on a 100^3 grid (M = K = 10^6,
nnz = 26,463,592, values U[0.5,1.5]), dense B 10^6 x 64 FP32 row-major in
U[-1,1), C overwritten -- the largest single-SpMM configuration.  The
native path runs with the variant and tile the per-matrix autotuner picks.

Also reported, each with every output's error against an FP64 product
(relative to sum_p |a_p b_p|, the output's condition scale) and the count
above 1e-5 beside the reference host loop's own count:
  * the same stencil with B stored as BF16 (FP32 accumulation);
  * the power-law matrix (M = K = 262144, Zipf(1.8) rows on [1, 5000]) at
    N = 1, 32 and 128, with BF16 B at N = 128, column-major operands, the
    reference's eight algorithms, rocSPARSE and the untuned selector;
  * 512 uneven matrices in one batched launch against plain launches;
  * a Poisson request stream against the serving loop (FP32 and BF16 B);
  * a device-to-device stream copy, the measured bandwidth ceiling.

Prints ONE JSON line on stdout (everything else goes to stderr).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "da-spmmul-kernel_amd")]

PEAK_HBM_GBS = 8000.0
HEADLINE_TAG = "c3_n64_fp32"

# Outputs whose host-loop (reference cpuSpmvReference) error is reported
# beside the GPU's: filled by the legs, evaluated by the cpu_baseline leg.
HOST_CHECKS = []


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


# ---- accuracy ---------------------------------------------------------------

def fp64_product(np, a, B2d):
    """FP64 product and sum_p |a_p b_p| per output (scipy, row-major)."""
    import scipy.sparse as sp
    A64 = sp.csr_matrix((a.values.astype(np.float64), a.colIdx, a.rowPtr), shape=(a.nr, a.nc))
    B64 = np.asarray(B2d, np.float64)
    return np.asarray(A64 @ B64), np.asarray(abs(A64) @ np.abs(B64))


def rel_err(np, out, c64, absum):
    out = np.asarray(out, np.float64).reshape(c64.shape)
    scale = np.maximum(np.abs(c64), absum)
    err = np.abs(out - c64)
    return np.where(scale > 0, err / np.where(scale > 0, scale, 1.0), np.where(err > 0, np.inf, 0.0))


def err_fields(np, out, c64, absum):
    rel = rel_err(np, out, c64, absum)
    return {"max_rel_err_vs_fp64": float(rel.max()) if rel.size else 0.0,
            "outputs_over_1e-5": int((rel > 1e-5).sum()),
            "outputs_over_1e-2": int((rel > 1e-2).sum()), "outputs": int(rel.size)}


def host_check(dst, a, B2d, nv, c64, absum):
    """Registers an output set whose host-loop error the cpu_baseline leg adds to `dst`."""
    HOST_CHECKS.append((dst, a, B2d, nv, c64, absum))


# ---- profiles -----------------------------------------------------------------

def profile_meta(tag):
    try:
        with open(os.path.join(ROOT, "profiles", f"pmc_{tag}.json")) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def profiled_config(tag):
    """The kernel configuration of the committed rocprofv3 PMC summary, with
    every field of the current configuration struct."""
    d = profile_meta(tag)
    try:
        cfg = d["meta"]["config"]
    except (TypeError, KeyError):
        return None
    import dspmm as D
    return D.config_from_dict(cfg).as_dict()


def traffic_from_profile(tag, config):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the
    copy-kernel calibration in the same file), when that summary profiled
    exactly `config`; else None."""
    d = profile_meta(tag)
    if not d or profiled_config(tag) != config:
        return None
    try:
        return int(round(d["kernels"][d["main_kernel"]]["traffic_bytes"]))
    except (KeyError, TypeError, ValueError):
        return None


def prefer_profiled(D, np, m, d_B, d_C, nv, best, profiled, btype, rounds=4, reps=50):
    """Between the autotuner's pick and the configuration whose HBM traffic
    was profiled, keep the profiled one unless the pick is >4% faster
    (the better of interleaved region-timed rounds, timed the way `value`
    is, the order alternating between rounds): autotuning choices among
    configurations within a few percent of each other flip with timing
    noise."""
    if not profiled or profiled == best.as_dict():
        return best
    cand = [D.config_from_dict(profiled), best]
    t = [[], []]
    for r in range(rounds):
        for i in ((0, 1) if r % 2 == 0 else (1, 0)):
            c = cand[i]
            try:
                mean, _, _ = m.benchmark(d_B, d_C, nv, btype=btype, config=c, warmup=5, reps=reps,
                                         per_call=False)
            except D.DspmmError:
                return best
            t[i].append(float(mean))
    log(f"profiled config {profiled}: {min(t[0]) * 1e3:.1f} us; autotuned "
        f"{best.as_dict()}: {min(t[1]) * 1e3:.1f} us")
    return best if min(t[1]) < 0.96 * min(t[0]) else cand[0]


# ---- single-matrix legs ---------------------------------------------------------

def run_native(D, np, m, d_B, d_C, nv, warmup, steps, btype=None, tune_reps=5, profiled=None):
    """Autotune, then time `steps` calls bracketed by two events (ms per
    step), then the same number of calls with an event between calls and
    after each main kernel (per-call spread, main-kernel duration)."""
    btype = D.DSPMM_R_32F if btype is None else btype
    best, _ = m.autotune(d_B, d_C, nv, btype=btype, reps=tune_reps)
    D.synchronize()
    best = prefer_profiled(D, np, m, d_B, d_C, nv, best, profiled, btype)
    t0 = time.perf_counter()
    mean, _, _ = m.benchmark(d_B, d_C, nv, btype=btype, config=best, warmup=warmup, reps=steps,
                             per_call=False)
    D.synchronize()
    wall = time.perf_counter() - t0
    _, per, main = m.benchmark(d_B, d_C, nv, btype=btype, config=best, warmup=0, reps=steps)
    return best, mean, per, main, wall


def native_leg(D, W, np, a, m, d_B, d_C, B2d, nv, warmup, steps, btype=None, sb=4, ref=None,
               check_host=True, profiled=None):
    """One autotuned native SpMM: timing, roofline and error fields."""
    btype = D.DSPMM_R_32F if btype is None else btype
    best, mean, per, main, wall = run_native(D, np, m, d_B, d_C, nv, warmup, steps, btype=btype,
                                             profiled=profiled)
    ms_main = float(np.mean(main))
    by = W.algorithmic_bytes(a.nr, a.nc, a.nnz, nv, sb)
    m.spmm(d_B, d_C, nv, btype=btype, config=best)
    D.synchronize()
    out = d_C.numpy().reshape(-1)[:a.nr * nv].reshape(a.nr, nv)
    c64, absum = ref if ref is not None else fp64_product(np, a, B2d)
    res = {"config": best.as_dict(), "ms": mean, "ms_median": float(np.median(per)),
           "ms_main_kernel": ms_main, "GFLOP/s": W.flops(a.nnz, nv) / (mean * 1e-3) / 1e9,
           "GB/s_algorithmic": by / (ms_main * 1e-3) / 1e9,
           "roofline_frac": by / (ms_main * 1e-3) / 1e9 / PEAK_HBM_GBS,
           "roofline_frac_step": by / (mean * 1e-3) / 1e9 / PEAK_HBM_GBS,
           "algorithmic_bytes": by, "step_wall_s": wall}
    res.update(err_fields(np, out, c64, absum))
    if check_host:
        host_check(res, a, B2d, nv, c64, absum)
    return res, best, (c64, absum)


def stream_copy(D, np, nbytes=1 << 30, reps=20):
    """Device-to-device copy of `nbytes`, event-timed: read + write GB/s."""
    lib = D.lib()
    src = D.DeviceArray(nbytes, np.uint8)
    dst = D.DeviceArray(nbytes, np.uint8)
    D.check(lib.dspmm_fill_zero(src.ptr, nbytes, None), "fill")
    for _ in range(2):
        D.check(lib.dspmm_memcpy_d2d(dst.ptr, src.ptr, nbytes, None), "d2d")
    D.synchronize()
    t = ctypes.c_void_p()
    D.check(lib.dspmm_timer_create(ctypes.byref(t), 2), "timer")
    D.check(lib.dspmm_timer_mark(t, None), "mark")
    for _ in range(reps):
        D.check(lib.dspmm_memcpy_d2d(dst.ptr, src.ptr, nbytes, None), "d2d")
    D.check(lib.dspmm_timer_mark(t, None), "mark")
    ms = ctypes.c_float()
    D.check(lib.dspmm_timer_intervals(t, ctypes.byref(ms), 1), "intervals")
    lib.dspmm_timer_destroy(t)
    per = ms.value / reps
    src.free()
    dst.free()
    return {"bytes": nbytes, "ms": per, "GB/s_read_plus_write": 2 * nbytes / (per * 1e-3) / 1e9}


def bench_rocsparse(D, a, d_rp, d_ri, d_ci, d_v, d_B, d_C, widths, reps=20):
    """rocSPARSE generic SpMV/SpMM on the same matrix and dense operands (the
    comparator columns of the reference's sweep), GFLOP/s per column."""
    rs = D.RocSparse(a.nr, a.nc, a.nnz, d_rp, d_ri, d_ci, d_v, d_B, d_C)
    out = {}
    for n in widths:
        cols = {}
        for c in range(rs.columns(n)):
            ms = rs.benchmark(n, c, warmup=3, reps=reps)
            cols[rs.name(n, c)] = None if ms is None else 2.0 * a.nnz * n / (ms * 1e-3) / 1e9
        ok = {k: v for k, v in cols.items() if v}
        best = max(ok, key=ok.get) if ok else None
        out[f"N{n}"] = {"GFLOP/s": cols, "best": best, "best_GFLOP/s": ok.get(best)}
    rs.close()
    return out


def bench_selector(D, np, a, B, d_C, widths, legs, refs, warmup, steps):
    """The learned default (AUTO on a fresh, untuned handle) against the
    autotuned configuration measured above, region-timed the same way on the
    same packed operands (B's first n columns)."""
    m = D.Matrix(a.nr, a.nc, a.rowPtr, a.colIdx, a.values)
    out = {}
    for n in widths:
        cfg = m.resolve(n)
        d_B = D.DeviceArray.from_numpy(np.ascontiguousarray(B[:, :n]))
        ms, _, _ = m.benchmark(d_B, d_C, n, warmup=warmup, reps=steps, per_call=False)
        m.spmm(d_B, d_C, n)
        d_B.free()
        D.synchronize()
        got = d_C.numpy().reshape(-1)[:a.nr * n]
        tuned = legs.get(f"N{n}", {}).get("ms")
        out[f"N{n}"] = {"selected": cfg.as_dict(), "ms": ms, "autotuned_ms": tuned,
                        "ratio_to_autotuned": ms / tuned if tuned else None}
        out[f"N{n}"].update(err_fields(np, got, *refs[n]))
    m.close()
    return out


def bench_power_law(D, W, np, args, extra):
    """The power-law matrix: native N = 1 / 32 / 128 (FP32), BF16 B at 128,
    column-major at 128, the reference's eight algorithms, rocSPARSE and the
    selector."""
    t = time.perf_counter()
    a = W.c2_matrix()
    B = W.dense_b(2, a.nc, 128)
    log(f"power-law: M=K={a.nr} nnz={a.nnz} max row {int(np.diff(a.rowPtr).max())} "
        f"(generated in {time.perf_counter() - t:.1f}s)")
    m = D.Matrix(a.nr, a.nc, a.rowPtr, a.colIdx, a.values)
    d_C = D.DeviceArray.zeros((a.nr, 128))
    legs, refs = {}, {}
    for nv in (1, 32, 128):
        Bn = np.ascontiguousarray(B[:, :nv])
        d_B = D.DeviceArray.from_numpy(Bn)
        legs[f"N{nv}"], _, refs[nv] = native_leg(D, W, np, a, m, d_B, d_C, Bn, nv, args.warmup,
                                                 args.steps)
        log(f"power-law N={nv}: {legs[f'N{nv}']}")
        d_B.free()
    d_B = D.DeviceArray.from_numpy(B)
    nv = 128
    # BF16-stored B: error against FP64 on the original FP32 B
    d16 = D.to_bf16_device(d_B)
    legs["bf16_B_N128"], _, _ = native_leg(D, W, np, a, m, d16, d_C, B, nv, args.warmup,
                                           min(args.steps, 100), btype=D.DSPMM_R_16BF, sb=2,
                                           ref=refs[nv], check_host=False)
    log(f"power-law BF16 B N=128: {legs['bf16_B_N128']}")
    d16.free()
    # the reference's eight algorithms through the compat dispatcher (packed
    # operands, ld = nv), named as the reference's CSV (run.sh:8)
    fl = W.flops(a.nnz, nv)
    d_rp = D.DeviceArray.from_numpy(a.rowPtr)
    d_ri = D.DeviceArray.from_numpy(a.rowIdx())
    d_ci = D.DeviceArray.from_numpy(a.colIdx)
    d_v = D.DeviceArray.from_numpy(a.values)
    d_Bcm = D.DeviceArray.from_numpy(np.ascontiguousarray(B.T))      # nv columns of length nc
    prob = {D.DENSE_ROW_MAJOR: D.SpmvSpmmProblem(a.nr, a.nc, a.nnz, 128, d_rp, d_ri, d_ci, d_v,
                                                 d_B, d_C),
            D.DENSE_COL_MAJOR: D.SpmvSpmmProblem(a.nr, a.nc, a.nnz, 128, d_rp, d_ri, d_ci, d_v,
                                                 d_Bcm, d_C)}
    sweep = {}
    for layout in (D.DENSE_ROW_MAJOR, D.DENSE_COL_MAJOR):
        for alg in (D.ALG_CSR_SCALAR, D.ALG_CSR_VECTOR, D.ALG_COO_SCALAR, D.ALG_COO_VECTOR):
            t_ms = prob[layout].benchmark(layout, nv, alg, 3, 20)
            d_C.zero()                     # element-balanced classes accumulate into C
            prob[layout].run(layout, nv, alg)
            D.synchronize()
            got = d_C.numpy().reshape(-1)[:a.nr * nv]
            if layout == D.DENSE_COL_MAJOR:
                got = got.reshape(nv, a.nr).T
            r = {"GFLOP/s": fl / (t_ms * 1e-3) / 1e9}
            r.update(err_fields(np, got, *refs[nv]))
            sweep[D.CSV_NAMES[(alg, layout)]] = r
    legs["reference_algorithms_N128"] = sweep
    log(f"reference algorithms N=128: {sweep}")
    # column-major operands: the autotuned native kernel (row-major kernels on
    # transposed copies, or a reference class)
    cm = D.DENSE_COL_MAJOR
    best_cm, _ = m.autotune(d_Bcm, d_C, nv, layout=cm, reps=3)
    t_cm, _, _ = m.benchmark(d_Bcm, d_C, nv, layout=cm, config=best_cm, warmup=args.warmup,
                             reps=min(args.steps, 50), per_call=False)
    m.spmm(d_Bcm, d_C, nv, layout=cm, config=best_cm)
    D.synchronize()
    got = d_C.numpy().reshape(-1)[:a.nr * nv].reshape(nv, a.nr).T
    legs["column_major_N128"] = {"config": best_cm.as_dict(), "ms": t_cm,
                                 "GFLOP/s": fl / (t_cm * 1e-3) / 1e9,
                                 "best_reference_class_GFLOP/s":
                                     max(v["GFLOP/s"] for k, v in sweep.items() if k.endswith("_CM"))}
    legs["column_major_N128"].update(err_fields(np, got, *refs[nv]))
    host_check(legs["column_major_N128"], a, B, nv, *refs[nv])
    log(f"column-major N=128: {legs['column_major_N128']}")
    legs["rocsparse"] = bench_rocsparse(D, a, d_rp, d_ri, d_ci, d_v, d_B, d_C, (1, 32, 128))
    log(f"rocsparse: {legs['rocsparse']}")
    legs["selector"] = bench_selector(D, np, a, B, d_C, (1, 32, 128), legs,
                                      {n: refs[n] for n in (1, 32, 128)}, args.warmup, args.steps)
    log(f"selector: {legs['selector']}")
    m.close()
    extra["power_law"] = legs


# ---- batched --------------------------------------------------------------------

def bench_batched(D, W, np, count, reps=20, sample_every=8):
    """The batch workload: `count` independent matrices (M=K ~ U{256..16384},
    LogNormal mean row length, N ~ U{16..256}), each autotuned, then run (i)
    as one launch, (ii) as a replayed hipGraph of it, (iii) as one plain
    launch per matrix with the batch's configuration and (iv) with each
    matrix's own autotuned kernel.  Requests per second = count / time of one
    pass; speedups are against the faster plain mode.  Errors: every output
    of every `sample_every`-th matrix, batched and plain, against FP64."""
    t = time.perf_counter()
    m, d, n = W.c4_params(count)
    items, mats = [], []
    nnz_total = bytes_total = flops_total = 0
    for i in range(count):
        a = W.c4_matrix(i, m[i], d[i])
        nv = int(n[i])
        mat = D.Matrix(a.nr, a.nc, a.rowPtr, a.colIdx, a.values)
        Bi = W.dense_b(4000 + i, a.nc, nv)
        d_B = D.DeviceArray.from_numpy(Bi)
        d_C = D.DeviceArray.zeros((a.nr, nv))
        mat.autotune(d_B, d_C, nv, reps=2)
        items.append((mat, nv, d_B, d_C))
        mats.append((a, Bi if i % sample_every == 0 else None))
        nnz_total += a.nnz
        bytes_total += W.algorithmic_bytes(a.nr, a.nc, a.nnz, nv)
        flops_total += W.flops(a.nnz, nv)
    log(f"batched: {count} matrices, nnz {nnz_total}, built+tuned in {time.perf_counter() - t:.1f}s")
    b = D.Batch(items)
    info = b.describe()
    res = {"matrices": count, "nnz_total": nnz_total, "algorithmic_bytes": bytes_total,
           "schedule": {k: info[k] for k in info}}
    refs = {i: fp64_product(np, a, Bi) for i, (a, Bi) in enumerate(mats) if Bi is not None}
    sample = f"{len(refs)} of {count} matrices (every {sample_every}th), all outputs"

    def sampled_errors():
        worst = {"max_rel_err_vs_fp64": 0.0, "outputs_over_1e-5": 0, "outputs_over_1e-2": 0,
                 "outputs": 0, "sample": sample}
        for i, (c64, absum) in refs.items():
            f = err_fields(np, items[i][3].numpy(), c64, absum)
            worst["max_rel_err_vs_fp64"] = max(worst["max_rel_err_vs_fp64"],
                                               f["max_rel_err_vs_fp64"])
            for k in ("outputs_over_1e-5", "outputs_over_1e-2", "outputs"):
                worst[k] += f[k]
        return worst

    for name, mode in (("batched", D.BATCH_ONE_LAUNCH), ("graph", D.BATCH_GRAPH),
                       ("plain", D.BATCH_PLAIN), ("plain_own_best", D.BATCH_PLAIN_OWN)):
        for _, _, _, d_C in items:
            d_C.zero()
        ms, per = b.benchmark(mode, warmup=3, reps=reps)
        D.synchronize()
        per = np.asarray(per)
        res[name] = {"ms_mean": ms, "ms_p50": float(np.percentile(per, 50)),
                     "ms_p99": float(np.percentile(per, 99)), "req_per_s": count / (ms * 1e-3),
                     "GFLOP/s": flops_total / (ms * 1e-3) / 1e9,
                     "GB/s_algorithmic": bytes_total / (ms * 1e-3) / 1e9}
        res[name].update(sampled_errors())
    plain = min(res["plain"]["ms_mean"], res["plain_own_best"]["ms_mean"])
    res["speedup_graph_vs_plain"] = plain / res["graph"]["ms_mean"]
    res["speedup_batched_vs_plain"] = plain / res["batched"]["ms_mean"]
    res["speedup_batched_vs_plain_own_best"] = (res["plain_own_best"]["ms_mean"]
                                                / res["batched"]["ms_mean"])
    host = {"sample": sample}
    res["host_loop"] = host
    for i, (c64, absum) in refs.items():
        host_check(host, mats[i][0], mats[i][1], items[i][1], c64, absum)
    b.close()
    # the same pass with every B stored as BF16 (FP32 accumulation)
    b16 = [D.to_bf16_device(d_B) for _, _, d_B, _ in items]
    bb = D.Batch([(mat, nv, d16, d_C) for (mat, nv, _, d_C), d16 in zip(items, b16)],
                 btype=D.DSPMM_R_16BF)
    for _, _, _, d_C in items:
        d_C.zero()
    ms16, _ = bb.benchmark(D.BATCH_ONE_LAUNCH, warmup=3, reps=reps)
    D.synchronize()
    res["batched_bf16_B"] = {"ms_mean": ms16, "req_per_s": count / (ms16 * 1e-3),
                             "GFLOP/s": flops_total / (ms16 * 1e-3) / 1e9,
                             "speedup_vs_fp32_batched": res["batched"]["ms_mean"] / ms16}
    res["batched_bf16_B"].update(sampled_errors())
    bb.close()
    for d16 in b16:
        d16.free()
    for mat, _, d_B, d_C in items:
        mat.close()
        d_B.free()
        d_C.free()
    return res


# ---- serving --------------------------------------------------------------------

def serve_accuracy(D, np, srv, mids, allm, count, refs, n_small=16):
    """A burst of requests (every pool matrix once, in arrival order) in which
    n_small small ones and the first large one keep their C; their errors
    against FP64 on the device-generated FP32 B (before any BF16 storage).
    `refs` caches the FP64 products by pool index across servers."""
    rng = np.random.default_rng(77)
    keep = set(int(k) for k in rng.choice(count, min(n_small, count), replace=False))
    if len(allm) > count:
        keep.add(count)
    reqs = [D.Server.request(20_000_000 + i, mids[i], nv, 900 + i, discard=i not in keep)
            for i, (_, nv) in enumerate(allm)]
    srv.replay(reqs, [0.0] * len(reqs))
    comps = srv.poll()
    zero = {"max_rel_err_vs_fp64": 0.0, "outputs_over_1e-5": 0, "outputs_over_1e-2": 0,
            "outputs": 0, "requests": 0}
    out = {"small": dict(zero), "large": dict(zero)}
    checks = []
    ok = True
    for c in comps:
        i = c["id"] - 20_000_000
        ok &= c["status"] == 0
        if i not in keep or c["status"] != 0:
            continue
        a, nv = allm[i]
        got = D.device_read(c["C"], (a.nr, nv))
        srv.release(c["id"])
        B2d = D.seeded_uniform(900 + i, a.nc * nv).reshape(a.nc, nv)
        if i not in refs:
            refs[i] = fp64_product(np, a, B2d)
        c64, absum = refs[i]
        f = err_fields(np, got, c64, absum)
        g = out["large" if i >= count else "small"]
        g["max_rel_err_vs_fp64"] = max(g["max_rel_err_vs_fp64"], f["max_rel_err_vs_fp64"])
        for k in ("outputs_over_1e-5", "outputs_over_1e-2", "outputs"):
            g[k] += f[k]
        g["requests"] += 1
        checks.append((g, a, B2d, nv, c64, absum))
    out["all_ok"] = bool(ok)
    return out, checks


def bench_serving(D, W, np, count, n_req, n_large, rate=2000.0, p_large=0.05):
    """The serving workload: a seeded Poisson stream (`rate` req/s) drawing 95%
    of the requests from the batch workload's pool (`count` matrices, each
    with its N) and 5% from `n_large` large power-law matrices (M=K=2e6, ~40M
    nnz, N=64).  B is generated on the device per request (seeded by the
    request id).  The same stream is replayed against (i) the
    continuous-batching server, (ii) a plain server that launches each
    request on its own with the batch's configuration, and (iii) one that
    uses each matrix's own AUTO configuration; then every request is
    submitted at once for saturation throughput, and once more without the
    large requests.  Both FP32 and BF16-stored B; a burst with outputs
    kept gives the served results' errors."""
    t = time.perf_counter()
    m, d, n = W.c4_params(count)
    pool = [(W.c4_matrix(i, m[i], d[i]), int(n[i])) for i in range(count)]
    large = [(W.c5_large_matrix(seed=500 + k), 64) for k in range(n_large)]
    log(f"serving pool: {count} + {n_large} large (nnz {large[0][0].nnz if large else 0}) "
        f"built in {time.perf_counter() - t:.1f}s")
    at, cls = W.arrivals(n_req, rate, p_large)
    rng = np.random.default_rng(55)
    pick = [(count + int(rng.integers(max(1, n_large)))) if (c and n_large) else
            int(rng.integers(count)) for c in cls]
    allm = pool + large
    res = {"requests": n_req, "rate_offered": rate, "large_fraction": float(np.mean(cls)),
           "pool": count, "large_pool": n_large}
    refs = {}
    for btype, bname in ((D.DSPMM_R_32F, "fp32"), (D.DSPMM_R_16BF, "bf16_B")):
        leg = {}
        modes = (0, 1, 2) if btype == D.DSPMM_R_32F else (0, 2)
        for plain in modes:
            srv = D.Server(cache_bytes=200 << 30, operand_bytes=64 << 30, max_inflight=3,
                           max_batch=256, plain=plain, btype=btype)
            mids = [srv.add_matrix(a.nr, a.nc, a.rowPtr, a.colIdx, a.values, preload=True)
                    for a, _ in allm]
            warm = [D.Server.request(10_000_000 + i, mids[i], nv, i, discard=True)
                    for i, (_, nv) in enumerate(allm)]
            srv.replay(warm, [0.0] * len(warm))
            srv.poll()
            key = ("batched", "plain", "plain_own_best")[plain]
            small = [k for k in range(n_req) if pick[k] < count]
            for mode, ids, times in (("poisson", range(n_req), at),
                                     ("saturation", range(n_req), np.zeros(n_req)),
                                     ("saturation_small", small, np.zeros(len(small)))):
                reqs = [D.Server.request(k, mids[pick[k]], allm[pick[k]][1], k, discard=True)
                        for k in ids]
                srv.replay(reqs, times)
                comps = srv.poll()
                lat = np.array([c["t_done"] - c["t_submit"] for c in comps])
                big = np.array([pick[c["id"]] >= count for c in comps])
                span = max(c["t_done"] for c in comps) - min(c["t_submit"] for c in comps)
                r = {"completed": len(comps), "all_ok": all(c["status"] == 0 for c in comps),
                     "req_per_s": len(comps) / span,
                     "p50_ms": float(np.percentile(lat, 50) * 1e3),
                     "p99_ms": float(np.percentile(lat, 99) * 1e3),
                     "mean_batch": float(np.mean([c["batch_size"] for c in comps]))}
                if big.any() and (~big).any():
                    r["p50_ms_small"] = float(np.percentile(lat[~big], 50) * 1e3)
                    r["p99_ms_small"] = float(np.percentile(lat[~big], 99) * 1e3)
                    r["p50_ms_large"] = float(np.percentile(lat[big], 50) * 1e3)
                    r["p99_ms_large"] = float(np.percentile(lat[big], 99) * 1e3)
                leg[f"{key}_{mode}"] = r
                log(f"serving {bname} {key} {mode}: {r}")
            if key != "plain":
                acc, checks = serve_accuracy(D, np, srv, mids, allm, count, refs)
                leg[f"{key}_accuracy"] = acc
                if key == "batched" and btype == D.DSPMM_R_32F:
                    for g, a, B2d, nv, c64, absum in checks:
                        host_check(g.setdefault("host_loop", {}), a, B2d, nv, c64, absum)
            leg[f"{key}_stats"] = dict(srv.stats())
            srv.close()
        for mode in ("saturation", "saturation_small"):
            plains = [leg[f"{p}_{mode}"]["req_per_s"] for p in ("plain", "plain_own_best")
                      if f"{p}_{mode}" in leg]
            leg[f"throughput_ratio_{mode}"] = leg[f"batched_{mode}"]["req_per_s"] / max(plains)
        leg["p50_ratio_poisson_vs_plain_own_best"] = (leg["batched_poisson"]["p50_ms"]
                                                      / leg["plain_own_best_poisson"]["p50_ms"])
        res[bname] = leg
    return res


# ---- CPU baseline (the only leg that runs the oracle) ------------------------------

def cpu_baseline_leg(np, a, B2d, nv, seconds):
    """The restated reference host loop (cpuSpmvReference,
    spmvspmm_benchmark.cu:11-39; FP32, built -O3 -ffp-contract=off) timed on
    one thread over full SpMMs of the headline workload for about `seconds`;
    then, as the checker, the host loop's own error on every registered
    output set, reported beside the GPU's."""
    import importlib
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    O = importlib.import_module("oracle")
    dense = np.ascontiguousarray(B2d, np.float32).reshape(-1)
    t_one = O.time_host32(O.ROW_MAJOR, a.nr, a.nc, nv, a.rowPtr, a.colIdx, a.values, dense, reps=1)
    reps = max(1, int(math.ceil(seconds / max(t_one, 1e-3))) - 1)
    t_cpu = O.time_host32(O.ROW_MAJOR, a.nr, a.nc, nv, a.rowPtr, a.colIdx, a.values, dense,
                          reps=reps)
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    base = {
        "value": W_FLOPS(a.nnz, nv) / t_cpu / 1e9, "unit": "GFLOP/s", "cores": 1, "kind": "port",
        "sample": (f"headline workload (27-point stencil 100^3, N={nv}): {reps + 1} full SpMMs "
                   f"({t_one + reps * t_cpu:.1f} s) of oracle_spmm_host32 (restated "
                   f"cpuSpmvReference, FP32, gcc -O3 -ffp-contract=off) on 1 thread; host "
                   f"{cpu_model()}, {os.cpu_count()} logical CPUs, affinity {affinity}"),
    }
    log(f"cpu baseline: {base['value']:.2f} GFLOP/s ({t_cpu * 1e3:.0f} ms per SpMM)")
    for dst, m, B, n, c64, absum in HOST_CHECKS:
        h32 = O.spmm_host32(O.ROW_MAJOR, m.nr, m.nc, n, m.rowPtr, m.colIdx, m.values,
                            np.ascontiguousarray(B, np.float32).reshape(-1))
        rel = rel_err(np, h32, c64, absum)
        dst["host_loop_max_rel_err_vs_fp64"] = max(dst.get("host_loop_max_rel_err_vs_fp64", 0.0),
                                                   float(rel.max()) if rel.size else 0.0)
        dst["host_loop_outputs_over_1e-5"] = (dst.get("host_loop_outputs_over_1e-5", 0)
                                              + int((rel > 1e-5).sum()))
    return base


def W_FLOPS(nnz, n):
    return 2 * int(nnz) * int(n)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target CPU work of the cpu_baseline sample")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the power-law legs and the BF16 stencil")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched leg")
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-requests", type=int, default=4000)
    ap.add_argument("--serve-large", type=int, default=2, help="large matrices in the pool")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X: the SpMM path here runs on a single GPU and "
              f"--gpus {args.gpus} is not supported (across GPUs the path would shard A by "
              f"nonzero-balanced row ranges with B replicated).", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import dspmm as D
    import workloads as W

    info = D.device_info()
    log(f"device: {info['name']} {info['arch']} CUs={info['compute_units']}")
    extra = {"stream_copy": stream_copy(D, np)}
    log(f"stream copy: {extra['stream_copy']}")

    # ---- headline: 27-point stencil, N = 64, FP32 ----
    t = time.perf_counter()
    nv = 64
    a = W.c3_matrix(100)
    B = W.dense_b(3, a.nc, nv)
    log(f"stencil: M={a.nr} nnz={a.nnz} (generated in {time.perf_counter() - t:.1f}s)")
    m = D.Matrix(a.nr, a.nc, a.rowPtr, a.colIdx, a.values)
    d_B = D.DeviceArray.from_numpy(B)
    d_C = D.DeviceArray.zeros((a.nr, nv))
    t = time.perf_counter()
    ref = fp64_product(np, a, B)
    log(f"stencil FP64 product in {time.perf_counter() - t:.1f}s")
    head, best, _ = native_leg(D, W, np, a, m, d_B, d_C, B, nv, args.warmup, args.steps,
                               ref=ref, profiled=profiled_config(HEADLINE_TAG))
    log(f"headline: {head}")
    extra["headline"] = head
    if not args.no_extra:
        d16 = D.to_bf16_device(d_B)
        extra["stencil_bf16_B"], _, _ = native_leg(
            D, W, np, a, m, d16, d_C, B, nv, args.warmup, min(args.steps, 100),
            btype=D.DSPMM_R_16BF, sb=2, ref=ref, check_host=False,
            profiled=profiled_config("c3_n64_bf16"))
        log(f"stencil BF16 B: {extra['stencil_bf16_B']}")
        d16.free()
    m.close()
    d_B.free()
    d_C.free()

    if not args.no_extra:
        bench_power_law(D, W, np, args, extra)
    if not args.no_batch:
        extra["batched"] = bench_batched(D, W, np, args.batch_count)
        log(f"batched: {extra['batched']}")
    if not args.no_serve:
        extra["serving"] = bench_serving(D, W, np, args.batch_count, args.serve_requests,
                                         args.serve_large)

    cpu_baseline = cpu_baseline_leg(np, a, B, nv, args.cpu_seconds)

    ms = head["ms"]
    gflops = W.flops(a.nnz, nv) / (ms * 1e-3) / 1e9
    achieved = head["GB/s_algorithmic"]
    line = {
        "metric": "SpMM GFLOP/s + % HBM roofline; batched req/s & p99 latency vs plain launches",
        "value": gflops,
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator; the reference ships no matrices)",
        "config": {"workload": (f"27-point stencil SpMM: 100^3 grid, M=K={a.nr}, nnz={a.nnz}, "
                                f"N={nv}, FP32 A and B, row-major, C overwritten"),
                   "M": a.nr, "K": a.nc, "nnz": a.nnz, "N": nv, "layout": "row-major",
                   "variant": best.as_dict()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS,
                     "traffic": traffic_from_profile(HEADLINE_TAG, best.as_dict())},
        "cpu_baseline": cpu_baseline,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
