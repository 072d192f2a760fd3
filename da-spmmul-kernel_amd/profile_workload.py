#!/usr/bin/env python3
"""Runs one SpMM workload a few times, for use under rocprofv3.

Before the SpMM steps it performs one device-to-device copy of a known size
(256 MiB): its FETCH_SIZE / WRITE_SIZE readings calibrate the HBM byte
counters (bytes moved / counter reading) for the kernels measured in the
same run.

  rocprofv3 --kernel-trace --stats -d OUT -o run -- \
      python3 da-spmmul-kernel_amd/profile_workload.py --workload c2 --n 128 \
      --variant 1 --vec 1 --lanes 64 --tile 128 --unroll 8
  rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT -o fetch -- python3 ...
  rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT -o write -- python3 ...
"""
from __future__ import annotations

import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

COPY_BYTES = 256 << 20


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4"],
                    help="c4: the 512-matrix batch in one launch (--n is ignored)")
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--bf16", action="store_true")
    ap.add_argument("--variant", type=int, default=0, help="0 = autotune")
    ap.add_argument("--vec", type=int, default=0)
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--unroll", type=int, default=0)
    ap.add_argument("--row-stride", type=int, default=0)
    ap.add_argument("--batch-mode", default="one", choices=["one", "plain_own"],
                    help="c4: the batch's one launch, or one launch per matrix with each "
                         "matrix's own autotuned kernel (the matrices are autotuned first)")
    args = ap.parse_args()

    import numpy as np

    import dspmm as D
    import workloads as W

    if args.workload == "c4":
        return batch_c4(args, np, D, W)
    a = {"c1": W.c1_matrix, "c2": W.c2_matrix, "c3": W.c3_matrix}[args.workload]()
    nv = args.n
    B = W.dense_b(2, a.nc, nv)
    m = D.Matrix(a.nr, a.nc, a.rowPtr, a.colIdx, a.values)
    d_B = D.DeviceArray.from_numpy(B)
    btype = D.DSPMM_R_32F
    if args.bf16:
        d_B = D.to_bf16_device(d_B)
        btype = D.DSPMM_R_16BF
    d_C = D.DeviceArray.zeros((a.nr, nv))

    src = D.DeviceArray(COPY_BYTES, np.uint8)
    dst = D.DeviceArray(COPY_BYTES, np.uint8)
    D.check(D.lib().dspmm_memcpy_d2d(dst.ptr, src.ptr, COPY_BYTES, None), "d2d")
    D.synchronize()

    if args.variant:
        cfg = D.make_config(args.variant, args.vec, args.lanes, args.tile, args.unroll,
                            args.row_stride)
    else:
        cfg, _ = m.autotune(d_B, d_C, nv, btype=btype, reps=3)
    cfg = m.resolve(nv, btype=btype, config=cfg)
    D.synchronize()
    for _ in range(args.steps):
        m.spmm(d_B, d_C, nv, btype=btype, config=cfg)
    D.synchronize()
    print(json.dumps({"workload": args.workload, "n": nv, "bf16": args.bf16, "nnz": a.nnz,
                      "config": cfg.as_dict(), "copy_bytes": COPY_BYTES,
                      "algorithmic_bytes": W.algorithmic_bytes(a.nr, a.nc, a.nnz, nv,
                                                               2 if args.bf16 else 4)}))


def batch_c4(args, np, D, W, count=512):
    """The c4 batch (512 matrices, each with its N, static configuration) in
    one launch per step, after the same calibration copy."""
    m, d, n = W.c4_params(count)
    items = []
    nnz = alg = 0
    for i in range(count):
        a = W.c4_matrix(i, m[i], d[i])
        nv = int(n[i])
        mat = D.Matrix(a.nr, a.nc, a.rowPtr, a.colIdx, a.values)
        items.append((mat, nv, D.DeviceArray.from_numpy(W.dense_b(4000 + i, a.nc, nv)),
                      D.DeviceArray.zeros((a.nr, nv))))
        nnz += a.nnz
        alg += W.algorithmic_bytes(a.nr, a.nc, a.nnz, nv)
    if args.batch_mode == "plain_own":
        for mat, nv, d_B, d_C in items:
            mat.autotune(d_B, d_C, nv, reps=2)
    b = D.Batch(items)
    src = D.DeviceArray(COPY_BYTES, np.uint8)
    dst = D.DeviceArray(COPY_BYTES, np.uint8)
    D.check(D.lib().dspmm_memcpy_d2d(dst.ptr, src.ptr, COPY_BYTES, None), "d2d")
    D.synchronize()
    for _ in range(args.steps):
        if args.batch_mode == "plain_own":
            b.benchmark(D.BATCH_PLAIN_OWN, warmup=0, reps=1)
        else:
            b.run()
    D.synchronize()
    print(json.dumps({"workload": "c4", "n": 0, "bf16": False, "nnz": nnz,
                      "config": {"variant": "batched" if args.batch_mode == "one" else "plain_own_best",
                                 "matrices": count},
                      "copy_bytes": COPY_BYTES, "algorithmic_bytes": alg}))


if __name__ == "__main__":
    main()
