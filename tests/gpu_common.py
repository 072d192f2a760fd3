"""Shared helpers of the GPU parity tests: cases uploaded once per process,
oracle results cached per (layout, nv), and the accuracy gates.

Accuracy gates (the tolerance the tests apply):
  * bit-exact: row-balanced sequential kernels (SR_RB, spmv_csr_scalar,
    spmm_csr_scalar_row/col) and the native nnz-tiled kernel on rows that
    fit one tile equal the reference host loop bit for bit;
  * FP32: every output within 1e-5 * max(|c64|, sum_p |a_p b_p|) of the
    FP64 result (the required 1e-5 relative error, measured against the
    output's condition scale: |c64| alone is meaningless under
    cancellation).  Where the reference host loop itself misses that bar
    (long rows: past ~160 nonzeros 1e-5 is below the worst case of any
    FP32 summation order), the kernel may miss it on no more outputs than
    the host loop does, and every output stays within the rigorous bound
    gamma_{n+1} * sum_p |a_p b_p| (n the row length, gamma_k = k u / (1 - k u),
    u = 2^-24).  Paths whose summation order differs from the host loop's
    (split rows, tree reductions, atomics) are also pinned exactly by
    integer-valued operands, where every order gives the same bits;
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
    tol = gamma * absum
    err = np.abs(out.astype(np.float64) - c64)
    return err, tol


def assert_fp32(out, c64, absum, rowlen, what, h32):
    """The FP32 gate: the count of outputs beyond 1e-5 relative error is at
    most the host loop's own count on the same operands (h32), and every
    output is within its row's rigorous summation bound."""
    assert np.all(np.isfinite(out)), f"{what}: non-finite outputs"
    err, tol = fp32_gate(out, c64, absum, rowlen)
    bad = err > tol
    if bad.any():
        i = int(np.flatnonzero(bad)[0])
        raise AssertionError(f"{what}: {int(bad.sum())} outputs beyond the summation bound; "
                             f"first {i}: err {err[i]:.3e} bound {tol[i]:.3e} "
                             f"(row len {rowlen[i]:.0f})")
    rel = O.rel_error(out, c64, absum)
    rel_h = O.rel_error(h32, c64, absum)
    n_gpu, n_host = int((rel > 1e-5).sum()), int((rel_h > 1e-5).sum())
    note(f"{what}: max rel err {rel.max():.2e}, > 1e-5: {n_gpu} "
         f"(host loop: max {rel_h.max():.2e}, > 1e-5: {n_host})")
    assert n_gpu <= n_host, (f"{what}: {n_gpu} outputs beyond 1e-5 relative error, the host "
                             f"loop has {n_host}")


def assert_exact(out, c64, what):
    """Integer-valued operands with every partial sum below 2^24: every
    summation order gives the exact integer result."""
    want = c64.astype(np.float32)
    assert np.array_equal(want.astype(np.float64), c64), f"{what}: operands are not exact"
    if not np.array_equal(bits(out), bits(want)):
        bad = np.flatnonzero(bits(out) != bits(want))
        raise AssertionError(f"{what}: {bad.size} outputs differ from the exact result; first at "
                             f"{bad[0]}: {out[bad[0]]!r} vs {want[bad[0]]!r}")


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
    elif name == "int_zipf":
        # power-law rows up to 3000 long (many split at every tile size),
        # integer operands: |partial sums| <= 3000 * 64 < 2^24, so every
        # summation order is exact
        a = W.gen_csr(W.GEN_ZIPF, 20000, 20000, 31, 1.6, 1, 3000)
        rng = np.random.default_rng(31)
        v = rng.integers(-8, 9, a.nnz).astype(np.float32)
        B = rng.integers(-8, 9, 20000 * 128).astype(np.float32)
        c = Case(name, a.nr, a.nc, a.rowPtr, a.colIdx, v, 128, B=B)
    elif name == "int_spmv":
        # the SpMV plan's row classes at every tile size T = 512 .. 4096:
        # rows of <= 16, 17..256 (shared tiles), 257..16 T (a tile of their
        # own) and > 16 T nonzeros (carry chunks), integer operands
        rng = np.random.default_rng(33)
        lens = rng.integers(0, 17, 3000)
        lens[100:400] = rng.integers(17, 257, 300)
        lens[500:520] = rng.integers(257, 8193, 20)
        lens[600] = 8192 + 5        # > 16 T for T = 512
        lens[700] = 16384 + 77      # > 16 T for T <= 1024
        lens[800] = 32768 + 300     # > 16 T for T <= 2048
        lens[900] = 70000           # > 16 T for every T
        rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        ci = np.concatenate([np.sort(rng.choice(100000, int(n), replace=False)) for n in lens
                             if n]).astype(np.int32)
        v = rng.integers(-8, 9, ci.size).astype(np.float32)
        B = rng.integers(-8, 9, 100000 * 8).astype(np.float32)
        c = Case(name, 3000, 100000, rp, ci, v, 8, B=B)
    else:
        raise KeyError(name)
    _CASES[name] = c
    return c
