"""Shared helpers of the GPU parity tests: cases uploaded once per process,
oracle results cached per (layout, nv), and the accuracy gates.

Accuracy gates (the tolerance the tests apply):
  * bit-exact: row-balanced sequential kernels (SR_RB, spmv_csr_scalar,
    spmm_csr_scalar_row/col) and the native nnz-tiled kernel on rows that
    fit one tile equal the reference host loop bit for bit;
  * FP32: |c - c64| <= max(1e-5, gamma_{n+1}) * sum_p |a_p b_p| per output,
    where n is the row length and gamma_k = k u / (1 - k u), u = 2^-24.
    This is the required 1e-5 relative error measured against the
    output's condition scale (|c64| alone is meaningless under
    cancellation); for rows longer than ~160 nonzeros 1e-5 is below the
    worst-case error of ANY FP32 summation order, including the
    reference's own host loop, so the rigorous bound applies there;
  * BF16-stored B: |c - c64| <= 1e-2 * sum_p |a_p b_p| against the FP64
    result on the original FP32 B;
  * integer-valued operands: exact, for every algorithm (atomics included).
"""
from __future__ import annotations

import sys

import numpy as np

import dspmm as D
import oracle as O
import workloads as W

U32 = 2.0 ** -24


def note(msg):
    print(f"[gpu] {msg}", flush=True)


class Case:
    def __init__(self, name, nr, nc, rowPtr, colIdx, values, maxnv, seed=123, B=None):
        self.name, self.nr, self.nc = name, int(nr), int(nc)
        self.rowPtr = np.ascontiguousarray(rowPtr, np.int32)
        self.colIdx = np.ascontiguousarray(colIdx, np.int32)
        self.values = np.ascontiguousarray(values, np.float32)
        self.nnz = int(self.colIdx.size)
        self.rowIdx = np.repeat(np.arange(self.nr, dtype=np.int32), np.diff(self.rowPtr))
        self.rowlen = np.diff(self.rowPtr).astype(np.float64)
        self.maxnv = maxnv
        if B is None:
            B = W.gen_dense(seed, self.nc * maxnv)
        self.B = np.ascontiguousarray(B, np.float32).reshape(-1)
        assert self.B.size >= self.nc * maxnv
        self.d_rp = D.DeviceArray.from_numpy(self.rowPtr)
        self.d_ri = D.DeviceArray.from_numpy(self.rowIdx)
        self.d_ci = D.DeviceArray.from_numpy(self.colIdx)
        self.d_v = D.DeviceArray.from_numpy(self.values)
        self.d_B = D.DeviceArray.from_numpy(self.B)
        self.d_C = D.DeviceArray.zeros(max(1, self.nr * maxnv))
        self._ref = {}

    def dense(self, nv):
        return self.B[: self.nc * nv]

    def ref(self, layout, nv):
        key = (layout, nv)
        if key not in self._ref:
            h32 = O.spmm_host32(layout, self.nr, self.nc, nv, self.rowPtr, self.colIdx,
                                self.values, self.dense(nv))
            c64, ab = O.spmm_fp64(layout, self.nr, self.nc, nv, self.rowPtr, self.colIdx,
                                  self.values, self.dense(nv))
            self._ref[key] = (h32.reshape(-1), c64.reshape(-1), ab.reshape(-1))
        return self._ref[key]

    def rowlen_flat(self, layout, nv):
        if layout == D.DENSE_ROW_MAJOR:
            return np.repeat(self.rowlen, nv)
        return np.tile(self.rowlen, nv)

    def output(self, nv):
        return self.d_C.numpy()[: self.nr * nv]

    def problem(self):
        return D.SpmvSpmmProblem(self.nr, self.nc, self.nnz, self.maxnv, self.d_rp, self.d_ri,
                                 self.d_ci, self.d_v, self.d_B, self.d_C)

    def fill_output(self, value):
        host = np.full(max(1, self.nr * self.maxnv), value, np.float32)
        D.check(D.lib().dspmm_memcpy_h2d(self.d_C.ptr, host.ctypes.data, host.nbytes), "h2d")


def bits(x):
    return np.ascontiguousarray(x, np.float32).view(np.uint32)


def assert_bitwise(out, h32, what):
    if not np.array_equal(bits(out), bits(h32)):
        bad = np.flatnonzero(bits(out) != bits(h32))
        raise AssertionError(f"{what}: {bad.size} outputs differ from the host loop; first at "
                             f"{bad[0]}: {out[bad[0]]!r} vs {h32[bad[0]]!r}")


def fp32_gate(out, c64, absum, rowlen):
    n = rowlen + 1.0
    gamma = n * U32 / (1.0 - n * U32)
    tol = np.maximum(1e-5, gamma) * absum
    err = np.abs(out.astype(np.float64) - c64)
    return err, tol


def assert_fp32(out, c64, absum, rowlen, what, h32=None):
    assert np.all(np.isfinite(out)), f"{what}: non-finite outputs"
    err, tol = fp32_gate(out, c64, absum, rowlen)
    bad = err > tol
    if bad.any():
        i = int(np.flatnonzero(bad)[0])
        raise AssertionError(f"{what}: {int(bad.sum())} outputs outside tolerance; first {i}: "
                             f"err {err[i]:.3e} tol {tol[i]:.3e} (row len {rowlen[i]:.0f})")
    # Report the literal 1e-5 relative error count beside the host loop's own.
    rel = O.rel_error(out, c64, absum)
    msg = f"{what}: max rel err {rel.max():.2e}, > 1e-5: {int((rel > 1e-5).sum())}"
    if h32 is not None:
        rel_h = O.rel_error(h32, c64, absum)
        msg += f" (host loop: max {rel_h.max():.2e}, > 1e-5: {int((rel_h > 1e-5).sum())})"
    note(msg)


def assert_bf16(out, c64, absum, what):
    assert np.all(np.isfinite(out)), f"{what}: non-finite outputs"
    err = np.abs(out.astype(np.float64) - c64)
    tol = 1e-2 * absum
    bad = err > tol
    assert not bad.any(), f"{what}: {int(bad.sum())} outputs beyond 1e-2 relative"
    rel = O.rel_error(out, c64, absum)
    note(f"{what}: BF16-B max rel err {rel.max():.2e} (vs FP64 on FP32 B)")


_CASES = {}


def golden_case(path, maxnv=None):
    key = ("golden", path)
    if key not in _CASES:
        g = np.load(path)
        b = g["B"]
        nv = b.shape[1]
        # dense input as the reference's flat buffer: row-major K x nv
        _CASES[key] = Case(path.split("/")[-1], int(g["nr"]), int(g["nc"]), g["rowPtr"],
                           g["colIdx"], g["values"], nv, B=b.reshape(-1))
    return _CASES[key]


def case(name):
    if name in _CASES:
        return _CASES[name]
    if name == "c1":
        a = W.c1_matrix()
        c = Case("c1", a.nr, a.nc, a.rowPtr, a.colIdx, a.values, 128, seed=1)
    elif name == "zipf_mid":
        a = W.gen_csr(W.GEN_ZIPF, 32768, 32768, 21, 1.8, 1, 5000, W.VAL_NORMAL)
        c = Case(name, a.nr, a.nc, a.rowPtr, a.colIdx, a.values, 128, seed=21)
    elif name == "c2":
        a = W.c2_matrix()
        c = Case(name, a.nr, a.nc, a.rowPtr, a.colIdx, a.values, 128, seed=2)
    elif name == "stencil_20":
        a = W.c3_matrix(20)
        c = Case(name, a.nr, a.nc, a.rowPtr, a.colIdx, a.values, 64, seed=3)
    elif name == "edge":
        # empty leading / middle / trailing rows, a 9000-long row, a 70-long row
        rng = np.random.default_rng(5)
        lens = np.zeros(3000, np.int64)
        lens[5:2000] = rng.integers(0, 12, 1995)
        lens[10] = 9000
        lens[1500] = 70
        lens[2500:2990] = rng.integers(0, 5, 490)
        rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        ci = np.concatenate([np.sort(rng.choice(12000, int(n), replace=False)) for n in lens
                             if n]).astype(np.int32)
        v = rng.uniform(-1, 1, ci.size).astype(np.float32)
        c = Case(name, 3000, 12000, rp, ci, v, 64, seed=5)
    elif name == "int_valued":
        rng = np.random.default_rng(8)
        lens = rng.integers(0, 300, 5000)
        lens[7] = 4000
        rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        ci = np.concatenate([np.sort(rng.choice(6000, int(n), replace=False)) for n in lens
                             if n]).astype(np.int32)
        v = rng.integers(-8, 9, ci.size).astype(np.float32)
        B = rng.integers(-8, 9, 6000 * 64).astype(np.float32)
        c = Case(name, 5000, 6000, rp, ci, v, 64, B=B)
    else:
        raise KeyError(name)
    _CASES[name] = c
    return c
