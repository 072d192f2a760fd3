"""Seeded synthetic inputs for the SpMM benchmark configurations.

Thin ctypes wrapper over libdspmm_gen.so (host/workloads.cpp).  The concrete
choices that the configurations leave open (seeds, dense-input range, how
"16 nnz/row" and the batch's row-length law are read) are fixed here:

  c1  M=K=4096, exactly 16 distinct uniform columns per row, A and B in
      U[-1,1), B 4096x32 row-major, seed 1.
  c2  M=K=262144, row length ~ Zipf(1.8) on [1,5000], uniform distinct
      columns, values N(0,1), B in U[-1,1), N in {1,32,128}, seed 2.
  c3  27-point stencil on a 100^3 grid (x fastest), values U[0.5,1.5),
      B 1e6 x 64 in U[-1,1), seed 3.
  c4  512 matrices: M=K ~ U{256..16384}, mean row length ~
      LogNormal(ln 16, 1) clipped to [1, 512], row length ~
      Binomial(K, d/K) clipped to >= 1, N ~ U{16,32,64,128,256},
      values U[-1,1), seed 4.
  c5  Poisson arrivals at 2000 req/s; 95% drawn from the c4 pool, 5% large
      (M=K=2e6, Zipf rows with exponent 1.7 on [1,5500] -> ~20 nnz/row,
      N=64), seed 5.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
GEN_PATH = os.path.join(_HERE, "lib", "libdspmm_gen.so")

GEN_FIXED, GEN_ZIPF, GEN_STENCIL27, GEN_BINOMIAL = 0, 1, 2, 3
VAL_UNIFORM, VAL_NORMAL = 0, 1


class GenParams(C.Structure):
    _fields_ = [("kind", C.c_int), ("nr", C.c_int), ("nc", C.c_int), ("value_dist", C.c_int),
                ("seed", C.c_uint64), ("p0", C.c_double), ("p1", C.c_double),
                ("p2", C.c_double), ("lo", C.c_float), ("hi", C.c_float)]


_gen = None


def gen_lib():
    global _gen
    if _gen is None:
        if not os.path.exists(GEN_PATH):
            raise RuntimeError(f"{GEN_PATH} is missing: build the package first")
        _gen = C.CDLL(GEN_PATH)
        P = C.POINTER(GenParams)
        _gen.dspmm_gen_rowptr.argtypes = [P, C.c_void_p]
        _gen.dspmm_gen_rowptr.restype = C.c_int64
        _gen.dspmm_gen_fill.argtypes = [P, C.c_void_p, C.c_void_p, C.c_void_p]
        _gen.dspmm_gen_fill.restype = C.c_int
        _gen.dspmm_gen_rowidx.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
        _gen.dspmm_gen_dense_uniform.argtypes = [C.c_uint64, C.c_int64, C.c_float, C.c_float,
                                                 C.c_void_p]
        _gen.dspmm_gen_batch_params.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int,
                                                C.c_double, C.c_double, C.c_double, C.c_double,
                                                C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                                C.c_void_p]
        _gen.dspmm_gen_arrivals.argtypes = [C.c_uint64, C.c_int, C.c_double, C.c_double,
                                            C.c_void_p, C.c_void_p]
        _gen.dspmm_gen_zipf_mean.argtypes = [C.c_double, C.c_int, C.c_int]
        _gen.dspmm_gen_zipf_mean.restype = C.c_double
    return _gen


@dataclass
class Csr:
    nr: int
    nc: int
    rowPtr: np.ndarray
    colIdx: np.ndarray
    values: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.colIdx.size)

    def rowIdx(self) -> np.ndarray:
        out = np.empty(self.nnz, np.int32)
        if self.nnz:
            gen_lib().dspmm_gen_rowidx(self.nr, self.rowPtr.ctypes.data, out.ctypes.data)
        return out


def gen_csr(kind, nr, nc, seed, p0=0.0, p1=0.0, p2=0.0, value_dist=VAL_UNIFORM, lo=-1.0,
            hi=1.0) -> Csr:
    g = gen_lib()
    prm = GenParams(kind, nr, nc, value_dist, seed, p0, p1, p2, lo, hi)
    rowPtr = np.empty(nr + 1, np.int32)
    nnz = g.dspmm_gen_rowptr(C.byref(prm), rowPtr.ctypes.data)
    if nnz < 0:
        raise ValueError("generator rejected parameters (or nnz exceeds int32)")
    colIdx = np.empty(nnz, np.int32)
    values = np.empty(nnz, np.float32)
    if g.dspmm_gen_fill(C.byref(prm), rowPtr.ctypes.data, colIdx.ctypes.data,
                        values.ctypes.data) != 0:
        raise RuntimeError("dspmm_gen_fill failed")
    return Csr(nr, nc, rowPtr, colIdx, values)


def gen_dense(seed, n, lo=-1.0, hi=1.0) -> np.ndarray:
    out = np.empty(int(n), np.float32)
    if n:
        gen_lib().dspmm_gen_dense_uniform(seed, int(n), lo, hi, out.ctypes.data)
    return out


def zipf_mean(s, kmin, kmax) -> float:
    return gen_lib().dspmm_gen_zipf_mean(s, kmin, kmax)


# ---- the benchmark configurations ---------------------------------------

def c1_matrix() -> Csr:
    return gen_csr(GEN_FIXED, 4096, 4096, seed=1, p0=16)


def c2_matrix(m=262144) -> Csr:
    return gen_csr(GEN_ZIPF, m, m, seed=2, p0=1.8, p1=1, p2=5000, value_dist=VAL_NORMAL)


def c3_matrix(n=100) -> Csr:
    return gen_csr(GEN_STENCIL27, n ** 3, n ** 3, seed=3, p0=n, lo=0.5, hi=1.5)


def c5_large_matrix(m=2_000_000, seed=5) -> Csr:
    return gen_csr(GEN_ZIPF, m, m, seed=seed, p0=1.7, p1=1, p2=5500, value_dist=VAL_NORMAL)


def dense_b(seed, k, n) -> np.ndarray:
    """Row-major K x N dense input in U[-1,1)."""
    return gen_dense(seed * 1000003 + 17, k * n).reshape(k, n)


N_CHOICES_C4 = (16, 32, 64, 128, 256)


def c4_params(count=512, seed=4):
    g = gen_lib()
    m = np.empty(count, np.int32)
    d = np.empty(count, np.float64)
    n = np.empty(count, np.int32)
    cand = np.asarray(N_CHOICES_C4, np.int32)
    g.dspmm_gen_batch_params(seed, count, 256, 16384, float(np.log(16.0)), 1.0, 1.0, 512.0,
                             cand.ctypes.data, cand.size, m.ctypes.data, d.ctypes.data,
                             n.ctypes.data)
    return m, d, n


def c4_matrix(i, m, d, seed=4) -> Csr:
    return gen_csr(GEN_BINOMIAL, int(m), int(m), seed=seed * 100003 + i, p0=float(d))


def arrivals(count, rate=2000.0, p_large=0.05, seed=5):
    t = np.empty(count, np.float64)
    cls = np.empty(count, np.int32)
    gen_lib().dspmm_gen_arrivals(seed, count, rate, p_large, t.ctypes.data, cls.ctypes.data)
    return t, cls


def algorithmic_bytes(m, k, nnz, n, b_bytes=4) -> int:
    """CSR compulsory traffic with B and C counted once:
    4(M+1) + 8 nnz + s_B K N + 4 M N."""
    return 4 * (m + 1) + 8 * nnz + b_bytes * k * n + 4 * m * n


def flops(nnz, n) -> int:
    """2 nnz N, the reference's flop count (spmvspmm_benchmark.cu:285)."""
    return 2 * nnz * n
