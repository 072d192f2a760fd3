"""ctypes binding of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker -- never by the library.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
ROW_MAJOR, COL_MAJOR = 0, 1
_lib = None


def build():
    subprocess.run(["make", "-B", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        P = C.c_void_p
        _lib.oracle_spmm_host32.argtypes = [C.c_int] * 5 + [P] * 5
        _lib.oracle_spmm_fp64.argtypes = [C.c_int] * 5 + [P] * 6
        _lib.oracle_csr_rowptr.argtypes = [C.c_int, C.c_int, P, P]
        _lib.oracle_csr_rowptr.restype = C.c_int
        _lib.oracle_check_result.argtypes = [C.c_int] * 4 + [P, P, C.c_float, P, P]
        _lib.oracle_check_result.restype = C.c_int
        _lib.oracle_round_bf16.argtypes = [C.c_int64, P, P]
        _lib.oracle_time_host32.argtypes = [C.c_int] * 5 + [P] * 5 + [C.c_int]
        _lib.oracle_time_host32.restype = C.c_double
    return _lib


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def spmm_host32(layout, nr, nc, nv, rowPtr, colIdx, values, dense):
    """cpuSpmvReference (spmvspmm_benchmark.cu:11-39).  `dense` is the packed
    input (K x nv row-major, or nv columns of length nc for col-major);
    returns the packed output (nr x nv, or nv x nr for col-major)."""
    rowPtr, colIdx, values, dense = _i32(rowPtr), _i32(colIdx), _f32(values), _f32(dense)
    out = np.empty(nr * nv, np.float32)
    lib().oracle_spmm_host32(layout, nr, nc, colIdx.size, nv, rowPtr.ctypes.data,
                             colIdx.ctypes.data, values.ctypes.data, dense.ctypes.data,
                             out.ctypes.data)
    return out.reshape((nr, nv) if layout == ROW_MAJOR else (nv, nr))


def spmm_fp64(layout, nr, nc, nv, rowPtr, colIdx, values, dense):
    """FP64 result and sum |a*b| per output, same packing as spmm_host32."""
    rowPtr, colIdx, values, dense = _i32(rowPtr), _i32(colIdx), _f32(values), _f32(dense)
    out = np.empty(nr * nv, np.float64)
    absum = np.empty(nr * nv, np.float64)
    lib().oracle_spmm_fp64(layout, nr, nc, colIdx.size, nv, rowPtr.ctypes.data,
                           colIdx.ctypes.data, values.ctypes.data, dense.ctypes.data,
                           out.ctypes.data, absum.ctypes.data)
    shape = (nr, nv) if layout == ROW_MAJOR else (nv, nr)
    return out.reshape(shape), absum.reshape(shape)


def csr_rowptr(nr, rowIdx):
    """generateCsrRowPtrArray (spmvspmm_benchmark.cu:41-76)."""
    rowIdx = _i32(rowIdx)
    out = np.empty(nr + 1, np.int32)
    lib().oracle_csr_rowptr(nr, rowIdx.size, rowIdx.ctypes.data, out.ctypes.data)
    return out


def check_result(host_layout, dev_layout, nrows, ncols, ref, out, eps=1e-3):
    """checkResult (spmvspmm_benchmark.cu:93-116): (passed, first_i, first_j)."""
    ref, out = _f32(ref), _f32(out)
    fi, fj = C.c_int(-1), C.c_int(-1)
    ok = lib().oracle_check_result(host_layout, dev_layout, nrows, ncols, ref.ctypes.data,
                                   out.ctypes.data, eps, C.byref(fi), C.byref(fj))
    return bool(ok), fi.value, fj.value


def round_bf16(a):
    a = _f32(a)
    out = np.empty_like(a)
    lib().oracle_round_bf16(a.size, a.ctypes.data, out.ctypes.data)
    return out


def time_host32(layout, nr, nc, nv, rowPtr, colIdx, values, dense, reps=1):
    """Seconds per oracle_spmm_host32 call on the calling thread."""
    rowPtr, colIdx, values, dense = _i32(rowPtr), _i32(colIdx), _f32(values), _f32(dense)
    out = np.empty(nr * nv, np.float32)
    return lib().oracle_time_host32(layout, nr, nc, colIdx.size, nv, rowPtr.ctypes.data,
                                    colIdx.ctypes.data, values.ctypes.data, dense.ctypes.data,
                                    out.ctypes.data, reps)


def rel_error(c, c64, absum):
    """Per-output error relative to max(|c64|, sum|a*b|) -- the relative error
    that stays meaningful under cancellation."""
    c = np.asarray(c, np.float64)
    scale = np.maximum(np.abs(c64), absum)
    err = np.abs(c - c64)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(scale > 0, err / np.where(scale > 0, scale, 1.0), np.where(err > 0, np.inf, 0.0))
    return r
