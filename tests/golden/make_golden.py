"""Writes the golden vectors in tests/golden/*.npz.

The reference ships no fixtures (its only check is the run-time comparison
against its host loop, spmvspmm_benchmark.cu:118-233), so the expected
outputs here come from independent libraries, not from this repository's
oracle:
  * C32  -- scipy.sparse CSR (float32) @ dense float32 B: scipy accumulates
            each output in float32, in nonzero order, with a separate
            multiply and add, i.e. the same arithmetic as the reference's
            host loop (spmvspmm_benchmark.cu:26-38);
  * C64  -- dense numpy float64 GEMM of the same operands;
  * ABS  -- dense float64 |A| @ |B| (the per-output condition scale).
Inputs are drawn with numpy's PCG64 so they do not depend on this
repository's generators.  Run:  python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))


def csr_from_lengths(rng, nr, nc, lens, int_values=False):
    rows, cols = [], []
    for r, n in enumerate(lens):
        c = np.sort(rng.choice(nc, size=int(n), replace=False))
        rows.append(np.full(int(n), r))
        cols.append(c)
    rows = np.concatenate(rows) if rows else np.zeros(0, int)
    cols = np.concatenate(cols) if cols else np.zeros(0, int)
    if int_values:
        vals = rng.integers(-8, 9, size=cols.size).astype(np.float32)
    else:
        vals = rng.uniform(-1, 1, size=cols.size).astype(np.float32)
    rowptr = np.zeros(nr + 1, np.int64)
    np.add.at(rowptr, rows + 1, 1)
    return np.cumsum(rowptr).astype(np.int32), cols.astype(np.int32), vals


def stencil27(n):
    rp, ci = [0], []
    for z in range(n):
        for y in range(n):
            for x in range(n):
                for dz in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            xx, yy, zz = x + dx, y + dy, z + dz
                            if 0 <= xx < n and 0 <= yy < n and 0 <= zz < n:
                                ci.append(xx + n * (yy + n * zz))
                rp.append(len(ci))
    return np.asarray(rp, np.int32), np.asarray(ci, np.int32)


def save(name, nr, nc, rowptr, colidx, values, b):
    A = sp.csr_matrix((values, colidx, rowptr), shape=(nr, nc))
    c32 = np.asarray(A @ b, dtype=np.float32)
    assert (A @ b).dtype == np.float32
    dense = A.toarray().astype(np.float64)
    c64 = dense @ b.astype(np.float64)
    absum = np.abs(dense) @ np.abs(b.astype(np.float64))
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), nr=nr, nc=nc, rowPtr=rowptr,
                        colIdx=colidx, values=values, B=b, C32=c32, C64=c64, ABS=absum)
    print(name, nr, nc, colidx.size, b.shape)


def main():
    rng = np.random.default_rng(20221)
    # ER pattern with exactly 16 per row (a small c1).
    nr = nc = 256
    rp, ci, v = csr_from_lengths(rng, nr, nc, [16] * nr)
    save("fixed16_256", nr, nc, rp, ci, v, rng.uniform(-1, 1, (nc, 8)).astype(np.float32))
    # Power-law rows (a small c2) with a few long rows.
    nr = nc = 2048
    k = np.arange(1, 501)
    p = k ** -1.8
    lens = rng.choice(k, size=nr, p=p / p.sum())
    lens[[3, 1000]] = [500, 400]
    rp, ci, v = csr_from_lengths(rng, nr, nc, lens)
    v = rng.standard_normal(ci.size).astype(np.float32)
    save("zipf_2048", nr, nc, rp, ci, v, rng.uniform(-1, 1, (nc, 4)).astype(np.float32))
    # 27-point stencil on 6^3 (a small c3).
    rp, ci = stencil27(6)
    v = rng.uniform(0.5, 1.5, ci.size).astype(np.float32)
    save("stencil_6", 216, 216, rp, ci, v, rng.uniform(-1, 1, (216, 3)).astype(np.float32))
    # Empty leading / middle / trailing rows and one long row.
    lens = [0, 0, 3, 0, 1, 0, 0, 200, 2, 0, 0, 5, 0, 0]
    rp, ci, v = csr_from_lengths(rng, len(lens), 300, lens)
    save("edge_rows", len(lens), 300, rp, ci, v, rng.uniform(-1, 1, (300, 5)).astype(np.float32))
    # Integer-valued operands: every sum is exact in FP32 in any order.
    nr, nc = 300, 200
    lens = rng.integers(0, 40, size=nr)
    rp, ci, v = csr_from_lengths(rng, nr, nc, lens, int_values=True)
    save("int_valued", nr, nc, rp, ci, v, rng.integers(-8, 9, (nc, 4)).astype(np.float32))


if __name__ == "__main__":
    main()
