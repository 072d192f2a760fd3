"""Python host side of the MI355X SpMV/SpMM library (ctypes over libdspmm.so).

Mirrors the reference's host interface for this path:

* ``SpmvSpmmProblem`` -- same constructor arguments and the same ``run`` /
  ``benchmark`` / ``dry_run`` methods as the reference's struct
  (spmvspmm.h:13-50, :202-273).  As in the reference, a non power-of-two
  ``nv`` prints a message to stderr and returns (``benchmark`` returns -1),
  and ``dry_run`` also refuses nv > 32 (spmvspmm.h:209-213, :223-227,
  :252-256).  Where the reference exits the process on a CUDA error
  (helper.h:25-34) this raises ``DspmmError``.
* ``DeviceArray`` -- the owning upload of hgpu::CUArray (helper.h:78-154).
* ``Matrix`` -- the native device-resident CSR handle (any nv, BF16 B,
  autotuned variant).

Nothing here computes on the CPU: if the HIP library is missing, importing
works but every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libdspmm.so")
GEN_PATH = os.path.join(_HERE, "lib", "libdspmm_gen.so")
ROCSPARSE_LIB_PATH = os.path.join(_HERE, "lib", "libdspmm_rocsparse.so")
HARNESS_PATH = os.path.join(_HERE, "bin", "spmvspmm")
INCLUDE_DIR = os.path.join(os.path.dirname(_HERE), "include")

# Reference enums (spmvspmm.h:5-10, helper.h:256-264).
ALG_CSR_SCALAR, ALG_CSR_VECTOR, ALG_COO_SCALAR, ALG_COO_VECTOR = 0, 1, 2, 3
SPARSE_FORMAT_CSR, SPARSE_FORMAT_COO = 0, 1
DENSE_ROW_MAJOR, DENSE_COL_MAJOR = 0, 1
ALG_NAMES = {ALG_CSR_SCALAR: "CSR_SCALAR", ALG_CSR_VECTOR: "CSR_VECTOR",
             ALG_COO_SCALAR: "COO_SCALAR", ALG_COO_VECTOR: "COO_VECTOR"}
# CSV column names of the reference sweep (run.sh:8): {SR,PR}_{RB,EB}_{RM,CM}.
CSV_NAMES = {(ALG_CSR_SCALAR, 0): "SR_RB_RM", (ALG_CSR_VECTOR, 0): "PR_RB_RM",
             (ALG_COO_SCALAR, 0): "SR_EB_RM", (ALG_COO_VECTOR, 0): "PR_EB_RM",
             (ALG_CSR_SCALAR, 1): "SR_RB_CM", (ALG_CSR_VECTOR, 1): "PR_RB_CM",
             (ALG_COO_SCALAR, 1): "SR_EB_CM", (ALG_COO_VECTOR, 1): "PR_EB_CM"}

DSPMM_R_32F, DSPMM_R_16BF = 0, 1
VARIANT_AUTO, VARIANT_NNZ_TILED = 0, 1
VARIANT_CSR_SCALAR, VARIANT_CSR_VECTOR, VARIANT_COO_SCALAR, VARIANT_COO_VECTOR = 2, 3, 4, 5
VARIANT_SPMV_TILED = 6
VARIANT_NNZ_CACHED = 7
VARIANT_ROW_BLOCKED = 8
VARIANT_RING = 9
VARIANT_NAMES = {0: "auto", 1: "nnz_tiled", 2: "csr_scalar", 3: "csr_vector",
                 4: "coo_scalar", 5: "coo_vector", 6: "spmv_tiled", 7: "nnz_cached",
                 8: "row_blocked", 9: "ring"}
MATRIX_WITH_COO = 0x1
PREPARE_HANDLE_BUFFERS = 0x1
NOALLOC_SCRATCH_BYTES = 64 << 20

STATUS_SUCCESS = 0
STATUS_INVALID_VALUE = 1
STATUS_NOT_SUPPORTED = 2
STATUS_WORKSPACE_TOO_SMALL = 5
STATUS_NOT_PREPARED = 7

# The 12 reference-named launchers, in spmv.h / spmm.h order.
LAUNCHERS = ["spmv_csr_scalar", "spmv_coo_scalar", "spmv_csr_vector", "spmv_coo_vector",
             "spmm_csr_scalar_row", "spmm_csr_vector_row", "spmm_coo_scalar_row",
             "spmm_coo_vector_row", "spmm_csr_scalar_col", "spmm_csr_vector_col",
             "spmm_coo_scalar_col", "spmm_coo_vector_col"]


class DspmmError(RuntimeError):
    def __init__(self, status: int, what: str):
        msg = f"{what}: {status_string(status)} (status {status}"
        if status == 4 and _lib is not None:
            msg += f", hipError {_lib.dspmm_last_hip_error()}"
        super().__init__(msg + ")")
        self.status = status


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 32), ("compute_units", C.c_int),
                ("wave_size", C.c_int), ("clock_khz", C.c_int), ("l2_bytes", C.c_int),
                ("lds_per_block", C.c_size_t), ("total_mem", C.c_size_t),
                ("free_mem", C.c_size_t)]


class MatrixDesc(C.Structure):
    _fields_ = [("nr", C.c_int), ("nc", C.c_int), ("nnz", C.c_int), ("rowPtr", C.c_void_p),
                ("rowIdx", C.c_void_p), ("colIdx", C.c_void_p), ("values", C.c_void_p),
                ("hbm_bytes", C.c_size_t), ("max_row_len", C.c_int),
                ("mean_row_len", C.c_double), ("cv_row_len", C.c_double)]


class Config(C.Structure):
    _fields_ = [("variant", C.c_int), ("vec", C.c_int), ("lanes", C.c_int),
                ("tile_nnz", C.c_int), ("unroll", C.c_int), ("row_stride", C.c_int)]

    def as_dict(self):
        return {"variant": VARIANT_NAMES.get(self.variant, self.variant), "vec": self.vec,
                "lanes": self.lanes, "tile_nnz": self.tile_nnz, "unroll": self.unroll,
                "row_stride": self.row_stride}


class BatchItem(C.Structure):
    _fields_ = [("A", C.c_void_p), ("nv", C.c_int), ("B", C.c_void_p), ("ldb", C.c_int64),
                ("C", C.c_void_p), ("ldc", C.c_int64)]


class BatchInfo(C.Structure):
    _fields_ = [("count", C.c_int), ("units", C.c_int), ("segments", C.c_int),
                ("splits", C.c_int), ("blocks", C.c_int), ("units_per_xcd", C.c_int * 8),
                ("table_bytes", C.c_size_t), ("carry_bytes", C.c_size_t)]


class ServerConfig(C.Structure):
    _fields_ = [("cache_bytes", C.c_size_t), ("operand_bytes", C.c_size_t),
                ("max_inflight", C.c_int), ("max_batch", C.c_int), ("plain", C.c_int),
                ("btype", C.c_int), ("alone_nnz", C.c_int64)]


class Request(C.Structure):
    _fields_ = [("id", C.c_uint64), ("matrix", C.c_int), ("nv", C.c_int), ("seed", C.c_uint64),
                ("B", C.c_void_p), ("C", C.c_void_p), ("flags", C.c_uint)]


REQUEST_DISCARD_RESULT = 0x1


class Completion(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int), ("batch_size", C.c_int),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("C", C.c_void_p)]


class Coo(C.Structure):
    _fields_ = [("nr", C.c_int), ("nc", C.c_int), ("nnz", C.c_int),
                ("rowIdx", C.POINTER(C.c_int)), ("colIdx", C.POINTER(C.c_int)),
                ("values", C.POINTER(C.c_float))]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("rejected", C.c_int64),
                ("batches", C.c_int64), ("cache_hits", C.c_int64), ("cache_misses", C.c_int64),
                ("evictions", C.c_int64), ("cache_used", C.c_size_t), ("cache_peak", C.c_size_t),
                ("operand_peak", C.c_size_t), ("mean_batch", C.c_double)]


_lib = None
_P = C.c_void_p
_I = C.c_int
_I64 = C.c_int64
_SZ = C.c_size_t
_ST = C.c_int  # dspmm_status_t


def _declare(lib):
    launcher = [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P]
    for name in LAUNCHERS:
        f = getattr(lib, name)
        f.argtypes, f.restype = launcher, _ST
    sig = {
        "dspmm_get_error_string": ([_ST], C.c_char_p),
        "dspmm_last_hip_error": ([], _I),
        "dspmm_version": ([], C.c_char_p),
        "dspmm_spmm": ([_I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P], _ST),
        "dspmm_device_alloc": ([C.POINTER(C.c_void_p), _SZ], _ST),
        "dspmm_device_free": ([_P], _ST),
        "dspmm_memcpy_h2d": ([_P, _P, _SZ], _ST),
        "dspmm_memcpy_d2h": ([_P, _P, _SZ], _ST),
        "dspmm_memcpy_d2d": ([_P, _P, _SZ, _P], _ST),
        "dspmm_fill_zero": ([_P, _SZ, _P], _ST),
        "dspmm_stream_create": ([C.POINTER(C.c_void_p)], _ST),
        "dspmm_stream_destroy": ([_P], _ST),
        "dspmm_stream_synchronize": ([_P], _ST),
        "dspmm_device_synchronize": ([], _ST),
        "dspmm_get_device_info": ([C.POINTER(DeviceInfo)], _ST),
        "dspmm_timer_create": ([C.POINTER(C.c_void_p), _I], _ST),
        "dspmm_timer_destroy": ([_P], _ST),
        "dspmm_timer_reset": ([_P], _ST),
        "dspmm_timer_mark": ([_P, _P], _ST),
        "dspmm_timer_intervals": ([_P, C.POINTER(C.c_float), _I], _ST),
        "dspmm_benchmark_compat": ([_I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _I, _I,
                                    C.POINTER(C.c_float), C.POINTER(C.c_float), _P], _ST),
        "dspmm_dry_run_compat": ([_I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _I, _P], _ST),
        "dspmm_matrix_create": ([C.POINTER(C.c_void_p), _I, _I, _I, _P, _P, _P, _P, C.c_uint],
                                _ST),
        "dspmm_matrix_create_device": ([C.POINTER(C.c_void_p), _I, _I, _I, _P, _P, _P, _P,
                                        C.c_uint, _P], _ST),
        "dspmm_matrix_destroy": ([_P], _ST),
        "dspmm_matrix_describe": ([_P, C.POINTER(MatrixDesc)], _ST),
        "dspmm_spmm_ex": ([_P, _I, _I, _P, _I64, _I, _P, _I64, C.POINTER(Config), _P, _SZ, _P],
                          _ST),
        "dspmm_spmm_ex_noalloc": ([_P, _I, _I, _P, _I64, _I, _P, _I64, C.POINTER(Config), _P],
                                  _ST),
        "dspmm_matrix_prepare": ([_P, _I, _I, _P, _I64, _I, _P, _I64, C.POINTER(Config), C.c_uint],
                                 _ST),
        "dspmm_spmm_ex_workspace_size": ([_P, _I, _I, _I, C.POINTER(Config), C.POINTER(_SZ)],
                                         _ST),
        "dspmm_spmm_ex_resolve": ([_P, _I, _I, _I, C.POINTER(Config), C.POINTER(Config)], _ST),
        "dspmm_autotune": ([_P, _I, _I, _P, _I64, _I, _P, _I64, _I, C.POINTER(Config),
                            C.POINTER(C.c_float), _P], _ST),
        "dspmm_matrix_set_config": ([_P, _I, _I, _I, C.POINTER(Config)], _ST),
        "dspmm_autotune_table": ([_P, _I, _I, _P, _I64, _I, _P, _I64, _I, C.POINTER(Config),
                                  C.POINTER(C.c_float), _I, C.POINTER(C.c_int), _P], _ST),
        "dspmm_benchmark_ex": ([_P, _I, _I, _P, _I64, _I, _P, _I64, C.POINTER(Config), _I, _I,
                                C.POINTER(C.c_float), C.POINTER(C.c_float),
                                C.POINTER(C.c_float), _P], _ST),
        "dspmm_convert_f32_to_bf16": ([_P, _P, _I64, _P], _ST),
        "dspmm_batch_create": ([C.POINTER(C.c_void_p), C.POINTER(BatchItem), _I, _I, _I], _ST),
        "dspmm_batch_destroy": ([_P], _ST),
        "dspmm_batch_describe": ([_P, C.POINTER(BatchInfo)], _ST),
        "dspmm_batch_run": ([_P, _P], _ST),
        "dspmm_batch_item_config": ([_P, _I, C.POINTER(Config)], _ST),
        "dspmm_batch_graph_launch": ([_P, _P], _ST),
        "dspmm_batch_run_plain": ([_P, _P], _ST),
        "dspmm_batch_set_operands": ([_P, _I, _P, _P, _P], _ST),
        "dspmm_benchmark_batch": ([_P, _I, _I, _I, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                   _P], _ST),
        "dspmm_server_now": ([], C.c_double),
        "dspmm_server_create": ([C.POINTER(C.c_void_p), C.POINTER(ServerConfig)], _ST),
        "dspmm_server_destroy": ([_P], _ST),
        "dspmm_server_add_matrix": ([_P, _I, _I, _I, _P, _P, _P, _I, C.POINTER(C.c_int)], _ST),
        "dspmm_server_submit": ([_P, C.POINTER(Request)], _ST),
        "dspmm_server_replay": ([_P, C.POINTER(Request), C.POINTER(C.c_double), _I], _ST),
        "dspmm_server_poll": ([_P, C.POINTER(Completion), _I], _I),
        "dspmm_server_release": ([_P, C.c_uint64], _ST),
        "dspmm_server_drain": ([_P], _ST),
        "dspmm_server_get_stats": ([_P, C.POINTER(ServerStats)], _ST),
        "dspmm_mtx_read": ([C.c_char_p, C.c_uint, C.POINTER(Coo)], _ST),
        "dspmm_coo_free": ([C.POINTER(Coo)], None),
        "dspmm_csr_rowptr": ([_I, _I, _P, _P], _ST),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes, f.restype = args, res


def lib():
    """The loaded libdspmm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first "
                               "(python -c 'import __graft_entry__ as g; g.build()')")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def status_string(status: int) -> str:
    if _lib is None:
        return str(status)
    return _lib.dspmm_get_error_string(status).decode()


def check(status: int, what: str):
    if status != STATUS_SUCCESS:
        raise DspmmError(status, what)


def declared_symbols(header: str | None = None):
    """Function names declared with DSPMM_API in include/*.h."""
    import glob
    import re
    files = [header] if header else sorted(glob.glob(os.path.join(INCLUDE_DIR, "*.h")))
    names = []
    for path in files:
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = "\n".join(l for l in text.splitlines() if not l.lstrip().startswith("#"))
        for m in re.finditer(r"DSPMM_API\s+[^;(]*?\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
            names.append(m.group(1))
    return names


def ptr(x) -> int:
    """Device address of a DeviceArray, or an int passed through."""
    if x is None:
        return 0
    if isinstance(x, DeviceArray):
        return x.ptr
    return int(x)


class DeviceArray:
    """Owning device buffer (hgpu::CUArray): upload from / download to numpy."""

    def __init__(self, nbytes: int, dtype=np.float32, shape=None):
        self.dtype = np.dtype(dtype)
        self.nbytes = int(nbytes)
        self.shape = shape if shape is not None else (self.nbytes // self.dtype.itemsize,)
        p = C.c_void_p()
        check(lib().dspmm_device_alloc(C.byref(p), self.nbytes), "device_alloc")
        self.ptr = p.value or 0

    @classmethod
    def from_numpy(cls, a: np.ndarray):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes, a.dtype, a.shape)
        if a.nbytes:
            check(lib().dspmm_memcpy_h2d(d.ptr, a.ctypes.data, a.nbytes), "memcpy_h2d")
        return d

    @classmethod
    def zeros(cls, shape, dtype=np.float32):
        shape = tuple(np.atleast_1d(shape))
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        d = cls(n, dtype, shape)
        d.zero()
        return d

    def zero(self, stream=None):
        if self.nbytes:
            check(lib().dspmm_fill_zero(self.ptr, self.nbytes, stream), "fill_zero")
            check(lib().dspmm_stream_synchronize(stream), "sync")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().dspmm_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes), "memcpy_d2h")
        return out

    def free(self):
        if self.ptr and _lib is not None:
            _lib.dspmm_device_free(self.ptr)
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_info() -> dict:
    info = DeviceInfo()
    check(lib().dspmm_get_device_info(C.byref(info)), "get_device_info")
    return {"name": info.name.decode(), "arch": info.arch.decode(),
            "compute_units": info.compute_units, "wave_size": info.wave_size,
            "clock_khz": info.clock_khz, "l2_bytes": info.l2_bytes,
            "total_mem": info.total_mem, "free_mem": info.free_mem}


def synchronize():
    check(lib().dspmm_device_synchronize(), "device_synchronize")


class SpmvSpmmProblem:
    """The reference's problem descriptor (spmvspmm.h:13-50) over the C-ABI.

    Pointers are device addresses (ints) or DeviceArrays, non-owning."""

    def __init__(self, nr, nc, nnz, maxNv, rowPtr, rowIdx, colIdx, values, dnInput, dnOutput,
                 stream=None):
        self.nr, self.nc, self.nnz, self.maxNv = int(nr), int(nc), int(nnz), int(maxNv)
        self.rowPtr, self.rowIdx, self.colIdx = ptr(rowPtr), ptr(rowIdx), ptr(colIdx)
        self.values, self.dnInput, self.dnOutput = ptr(values), ptr(dnInput), ptr(dnOutput)
        self.stream = stream

    def _args(self, nv):
        return (self.nr, self.nc, self.nnz, nv, self.rowPtr, self.rowIdx, self.colIdx,
                self.values, self.dnInput, self.dnOutput)

    def run(self, layout, nv, alg):
        """spmvspmm.h:202-218: dispatch one launch (asynchronous)."""
        if nv & (nv - 1):
            print(f"nv must be power of 2, less than 32. Got nv = {nv}", file=sys.stderr)
            return
        check(lib().dspmm_spmm(alg, layout, *self._args(nv), self.stream), "SpmvSpmmProblem.run")

    def benchmark(self, layout, nv, alg, warmupIters=5, repeatIters=50, per_launch=None):
        """spmvspmm.h:220-243: mean milliseconds per launch over repeatIters."""
        if nv & (nv - 1):
            print(f"nv must be power of 2, less than 32. Got nv = {nv}", file=sys.stderr)
            return -1
        mean = C.c_float()
        arr = None
        if per_launch is not None:
            arr = (C.c_float * repeatIters)()
        check(lib().dspmm_benchmark_compat(alg, layout, *self._args(nv), warmupIters, repeatIters,
                                           C.byref(mean), arr, self.stream),
              "SpmvSpmmProblem.benchmark")
        if per_launch is not None:
            per_launch.extend(list(arr))
        return mean.value

    def dry_run(self, layout, nv, alg, warmupIters=0):
        """spmvspmm.h:245-273: warm-ups, then one profiled launch."""
        if (nv & (nv - 1)) or nv > 32:
            print(f"nv must be power of 2, less than 32. Got nv = {nv}", file=sys.stderr)
            return
        check(lib().dspmm_dry_run_compat(alg, layout, *self._args(nv), warmupIters, self.stream),
              "SpmvSpmmProblem.dry_run")


def launch(name, nr, nc, nnz, nv, rowPtr, rowIdx, colIdx, values, dnInput, dnOutput,
           stream=None):
    """Call one of the 12 reference-named launchers; returns the status."""
    f = getattr(lib(), name)
    return f(nr, nc, nnz, nv, ptr(rowPtr), ptr(rowIdx), ptr(colIdx), ptr(values), ptr(dnInput),
             ptr(dnOutput), stream)


def make_config(variant=VARIANT_AUTO, vec=0, lanes=0, tile_nnz=0, unroll=0, row_stride=0) -> Config:
    return Config(variant, vec, lanes, tile_nnz, unroll, row_stride)


def config_from_dict(d) -> Config:
    """Inverse of Config.as_dict."""
    names = {v: k for k, v in VARIANT_NAMES.items()}
    v = d["variant"]
    return make_config(names.get(v, v) if isinstance(v, str) else int(v), int(d.get("vec", 0)),
                       int(d.get("lanes", 0)), int(d.get("tile_nnz", 0)), int(d.get("unroll", 0)),
                       int(d.get("row_stride", 0)))


class Matrix:
    """Device-resident CSR matrix (native API)."""

    def __init__(self, nr, nc, rowPtr=None, colIdx=None, values=None, rowIdx=None, coo=False):
        rowPtr = None if rowPtr is None else np.ascontiguousarray(rowPtr, np.int32)
        rowIdx = None if rowIdx is None else np.ascontiguousarray(rowIdx, np.int32)
        colIdx = np.ascontiguousarray(colIdx, np.int32)
        values = np.ascontiguousarray(values, np.float32)
        self.nr, self.nc, self.nnz = int(nr), int(nc), int(colIdx.size)
        h = C.c_void_p()
        flags = MATRIX_WITH_COO if coo else 0
        check(lib().dspmm_matrix_create(
            C.byref(h), self.nr, self.nc, self.nnz,
            None if rowPtr is None else rowPtr.ctypes.data,
            None if rowIdx is None else rowIdx.ctypes.data,
            colIdx.ctypes.data if self.nnz else None,
            values.ctypes.data if self.nnz else None, flags), "matrix_create")
        self.handle = h.value

    def describe(self) -> dict:
        d = MatrixDesc()
        check(lib().dspmm_matrix_describe(self.handle, C.byref(d)), "matrix_describe")
        return {"nr": d.nr, "nc": d.nc, "nnz": d.nnz, "hbm_bytes": d.hbm_bytes,
                "max_row_len": d.max_row_len, "mean_row_len": d.mean_row_len,
                "cv_row_len": d.cv_row_len, "rowPtr": d.rowPtr, "rowIdx": d.rowIdx,
                "colIdx": d.colIdx, "values": d.values}

    def spmm(self, B, C_out, nv, layout=DENSE_ROW_MAJOR, ldb=None, ldc=None, btype=DSPMM_R_32F,
             config: Config | None = None, workspace=None, workspace_bytes=0, stream=None):
        ldb = ldb if ldb is not None else (nv if layout == DENSE_ROW_MAJOR else self.nc)
        ldc = ldc if ldc is not None else (nv if layout == DENSE_ROW_MAJOR else self.nr)
        st = lib().dspmm_spmm_ex(self.handle, layout, nv, ptr(B), ldb, btype, ptr(C_out), ldc,
                                 C.byref(config) if config is not None else None,
                                 ptr(workspace), workspace_bytes, stream)
        check(st, "spmm_ex")

    def prepare(self, B, C_out, nv, layout=DENSE_ROW_MAJOR, ldb=None, ldc=None,
                btype=DSPMM_R_32F, config: Config | None = None, handle_buffers=False):
        """dspmm_matrix_prepare: builds the plans (and optionally sizes the
        handle's buffers) a later spmm with these arguments needs."""
        ldb = ldb if ldb is not None else (nv if layout == DENSE_ROW_MAJOR else self.nc)
        ldc = ldc if ldc is not None else (nv if layout == DENSE_ROW_MAJOR else self.nr)
        check(lib().dspmm_matrix_prepare(self.handle, layout, nv, ptr(B), ldb, btype, ptr(C_out),
                                         ldc, C.byref(config) if config is not None else None,
                                         PREPARE_HANDLE_BUFFERS if handle_buffers else 0),
              "matrix_prepare")

    def resolve(self, nv, layout=DENSE_ROW_MAJOR, btype=DSPMM_R_32F, config=None) -> Config:
        out = Config()
        check(lib().dspmm_spmm_ex_resolve(self.handle, layout, nv, btype,
                                          C.byref(config) if config is not None else None,
                                          C.byref(out)), "spmm_ex_resolve")
        return out

    def workspace_size(self, nv, layout=DENSE_ROW_MAJOR, btype=DSPMM_R_32F, config=None) -> int:
        n = C.c_size_t()
        check(lib().dspmm_spmm_ex_workspace_size(self.handle, layout, nv, btype,
                                                 C.byref(config) if config is not None else None,
                                                 C.byref(n)), "spmm_ex_workspace_size")
        return n.value

    def autotune(self, B, C_out, nv, layout=DENSE_ROW_MAJOR, btype=DSPMM_R_32F, reps=5,
                 ldb=None, ldc=None, stream=None):
        ldb = ldb if ldb is not None else (nv if layout == DENSE_ROW_MAJOR else self.nc)
        ldc = ldc if ldc is not None else (nv if layout == DENSE_ROW_MAJOR else self.nr)
        best, ms = Config(), C.c_float()
        check(lib().dspmm_autotune(self.handle, layout, nv, ptr(B), ldb, btype, ptr(C_out), ldc,
                                   reps, C.byref(best), C.byref(ms), stream), "autotune")
        return best, ms.value

    def set_config(self, nv, config, layout=DENSE_ROW_MAJOR, btype=DSPMM_R_32F):
        """Pins `config` as the tuned configuration for (layout, nv, btype)."""
        check(lib().dspmm_matrix_set_config(self.handle, layout, nv, btype, C.byref(config)),
              "matrix_set_config")

    def autotune_table(self, B, C_out, nv, layout=DENSE_ROW_MAJOR, btype=DSPMM_R_32F, reps=5,
                       ldb=None, ldc=None, stream=None):
        """Every autotuner candidate with its median ms (None: not supported)."""
        ldb = ldb if ldb is not None else (nv if layout == DENSE_ROW_MAJOR else self.nc)
        ldc = ldc if ldc is not None else (nv if layout == DENSE_ROW_MAJOR else self.nr)
        n = C.c_int()
        args = (self.handle, layout, nv, ptr(B), ldb, btype, ptr(C_out), ldc, reps)
        check(lib().dspmm_autotune_table(*args, None, None, 0, C.byref(n), stream),
              "autotune_table")
        cfgs = (Config * n.value)()
        ms = (C.c_float * n.value)()
        check(lib().dspmm_autotune_table(*args, cfgs, ms, n.value, C.byref(n), stream),
              "autotune_table")
        return [(cfgs[i], ms[i] if ms[i] >= 0 else None) for i in range(n.value)]

    def benchmark(self, B, C_out, nv, layout=DENSE_ROW_MAJOR, btype=DSPMM_R_32F, config=None,
                  warmup=5, reps=50, ldb=None, ldc=None, stream=None, per_call=True):
        """Returns (mean ms per call, per-call ms list, per-call main-kernel ms
        list).  per_call=False times the region only (no events between
        calls); the two lists are then empty."""
        ldb = ldb if ldb is not None else (nv if layout == DENSE_ROW_MAJOR else self.nc)
        ldc = ldc if ldc is not None else (nv if layout == DENSE_ROW_MAJOR else self.nr)
        mean = C.c_float()
        arr = (C.c_float * reps)() if per_call else None
        main = (C.c_float * reps)() if per_call else None
        check(lib().dspmm_benchmark_ex(self.handle, layout, nv, ptr(B), ldb, btype, ptr(C_out),
                                       ldc, C.byref(config) if config is not None else None,
                                       warmup, reps, C.byref(mean), arr, main, stream),
              "benchmark_ex")
        return mean.value, (list(arr) if per_call else []), (list(main) if per_call else [])

    def close(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.dspmm_matrix_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def to_bf16_device(src: DeviceArray, stream=None) -> DeviceArray:
    """Device-side FP32 -> BF16 (RNE) copy of a dense operand."""
    n = src.nbytes // 4
    dst = DeviceArray(2 * n, np.uint16, src.shape)
    check(lib().dspmm_convert_f32_to_bf16(src.ptr, dst.ptr, n, stream), "convert_f32_to_bf16")
    check(lib().dspmm_stream_synchronize(stream), "sync")
    return dst


BATCH_ONE_LAUNCH, BATCH_GRAPH, BATCH_PLAIN, BATCH_PLAIN_OWN = 0, 1, 2, 3


class Batch:
    """Many row-major SpMMs in one launch (dspmm_batch_*).

    items: sequence of (Matrix, nv, B, C) or (Matrix, nv, B, C, ldb, ldc)."""

    def __init__(self, items, btype=DSPMM_R_32F, blocks_per_cu=0):
        arr = (BatchItem * max(1, len(items)))()
        self._keep = []
        for i, it in enumerate(items):
            m, nv, B, Cm = it[:4]
            ldb = it[4] if len(it) > 4 else nv
            ldc = it[5] if len(it) > 5 else nv
            arr[i] = BatchItem(m.handle, nv, ptr(B), ldb, ptr(Cm), ldc)
            self._keep.append(m)
        h = C.c_void_p()
        check(lib().dspmm_batch_create(C.byref(h), arr, len(items), btype, blocks_per_cu),
              "batch_create")
        self.handle = h.value
        self.count = len(items)

    def describe(self) -> dict:
        info = BatchInfo()
        check(lib().dspmm_batch_describe(self.handle, C.byref(info)), "batch_describe")
        return {"count": info.count, "units": info.units, "segments": info.segments,
                "splits": info.splits, "blocks": info.blocks,
                "units_per_xcd": list(info.units_per_xcd), "table_bytes": info.table_bytes,
                "carry_bytes": info.carry_bytes}

    def run(self, stream=None):
        check(lib().dspmm_batch_run(self.handle, stream), "batch_run")

    def item_config(self, index) -> "Config":
        """The configuration item `index` runs with in the batch."""
        out = Config()
        check(lib().dspmm_batch_item_config(self.handle, index, C.byref(out)), "batch_item_config")
        return out

    def graph_launch(self, stream=None):
        check(lib().dspmm_batch_graph_launch(self.handle, stream), "batch_graph_launch")

    def run_plain(self, stream=None):
        check(lib().dspmm_batch_run_plain(self.handle, stream), "batch_run_plain")

    def set_operands(self, index, B, C_out, stream=None):
        check(lib().dspmm_batch_set_operands(self.handle, index, ptr(B), ptr(C_out), stream),
              "batch_set_operands")

    def benchmark(self, mode=BATCH_ONE_LAUNCH, warmup=3, reps=20, stream=None):
        mean = C.c_float()
        arr = (C.c_float * reps)()
        check(lib().dspmm_benchmark_batch(self.handle, mode, warmup, reps, C.byref(mean), arr,
                                          stream), "benchmark_batch")
        return mean.value, list(arr)

    def close(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.dspmm_batch_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


STATUS_ADMISSION_REJECTED = 6


class Server:
    """The serving loop (dspmm_server_*): resident CSR cache, HBM-budgeted
    admission and continuous batching, run by a library thread."""

    def __init__(self, cache_bytes=200 << 30, operand_bytes=16 << 30, max_inflight=3,
                 max_batch=128, plain=False, btype=DSPMM_R_32F, alone_nnz=0):
        cfg = ServerConfig(cache_bytes, operand_bytes, max_inflight, max_batch, int(plain), btype,
                           alone_nnz)
        h = C.c_void_p()
        check(lib().dspmm_server_create(C.byref(h), C.byref(cfg)), "server_create")
        self.handle = h.value

    def add_matrix(self, nr, nc, rowPtr, colIdx, values, preload=False) -> int:
        rowPtr = np.ascontiguousarray(rowPtr, np.int32)
        colIdx = np.ascontiguousarray(colIdx, np.int32)
        values = np.ascontiguousarray(values, np.float32)
        mid = C.c_int()
        check(lib().dspmm_server_add_matrix(self.handle, nr, nc, int(colIdx.size),
                                            rowPtr.ctypes.data, colIdx.ctypes.data,
                                            values.ctypes.data, int(preload), C.byref(mid)),
              "server_add_matrix")
        return mid.value

    @staticmethod
    def request(rid, matrix, nv, seed, B=None, C_out=None, discard=False) -> Request:
        return Request(rid, matrix, nv, seed, ptr(B), ptr(C_out),
                       REQUEST_DISCARD_RESULT if discard else 0)

    def submit(self, req: Request) -> int:
        return lib().dspmm_server_submit(self.handle, C.byref(req))

    def replay(self, reqs, at):
        n = len(reqs)
        arr = (Request * max(1, n))(*reqs)
        t = np.ascontiguousarray(at, np.float64)
        check(lib().dspmm_server_replay(self.handle, arr, t.ctypes.data_as(C.POINTER(C.c_double)),
                                        n), "server_replay")

    def drain(self):
        check(lib().dspmm_server_drain(self.handle), "server_drain")

    def poll(self, max_n=4096):
        out = (Completion * max_n)()
        got = []
        while True:
            n = lib().dspmm_server_poll(self.handle, out, max_n)
            got.extend({"id": c.id, "status": c.status, "batch_size": c.batch_size,
                        "t_submit": c.t_submit, "t_admit": c.t_admit, "t_done": c.t_done,
                        "C": c.C} for c in out[:n])
            if n < max_n:
                return got

    def release(self, rid):
        check(lib().dspmm_server_release(self.handle, rid), "server_release")

    def stats(self) -> dict:
        st = ServerStats()
        check(lib().dspmm_server_get_stats(self.handle, C.byref(st)), "server_stats")
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.dspmm_server_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


MTX_KEEP_DIAGONAL = 0x1
MTX_ALIGN_VALUES = 0x2


def read_mtx(path, flags=0):
    """Reads a coordinate Matrix-Market file with the library's loader
    (readMtx<float>, util/spmatrix.hpp:286-333).  Returns
    (nr, nc, rowIdx, colIdx, values), sorted by (row, column)."""
    coo = Coo()
    check(lib().dspmm_mtx_read(os.fsencode(path), int(flags), C.byref(coo)), f"mtx_read {path}")
    try:
        n = coo.nnz
        r = np.ctypeslib.as_array(coo.rowIdx, (max(n, 1),))[:n].copy()
        c = np.ctypeslib.as_array(coo.colIdx, (max(n, 1),))[:n].copy()
        v = np.ctypeslib.as_array(coo.values, (max(n, 1),))[:n].copy()
        return coo.nr, coo.nc, r, c, v
    finally:
        lib().dspmm_coo_free(C.byref(coo))


def csr_rowptr(nr, rowIdx):
    """rowPtr[nr + 1] from sorted row indices (generateCsrRowPtrArray,
    spmvspmm_benchmark.cu:41-76)."""
    rows = np.ascontiguousarray(rowIdx, dtype=np.int32)
    out = np.empty(int(nr) + 1, np.int32)
    check(lib().dspmm_csr_rowptr(int(nr), int(rows.size), rows.ctypes.data if rows.size else None,
                                 out.ctypes.data), "csr_rowptr")
    return out


_rs_lib = None


def rocsparse_lib():
    """The loaded rocSPARSE comparator library (lib/libdspmm_rocsparse.so)."""
    global _rs_lib
    if _rs_lib is None:
        lib()  # status strings and the device helpers
        if not os.path.exists(ROCSPARSE_LIB_PATH):
            raise RuntimeError(f"{ROCSPARSE_LIB_PATH} is missing: build it first")
        r = C.CDLL(ROCSPARSE_LIB_PATH)
        sig = {
            "dspmm_rocsparse_create": ([C.POINTER(C.c_void_p), _I, _I, _I, _P, _P, _P, _P, _P, _P,
                                        _P], _ST),
            "dspmm_rocsparse_destroy": ([_P], _ST),
            "dspmm_rocsparse_columns": ([_I], _I),
            "dspmm_rocsparse_column_name": ([_I, _I], C.c_char_p),
            "dspmm_rocsparse_column_layout": ([_I, _I], _I),
            "dspmm_rocsparse_run": ([_P, _I, _I], _ST),
            "dspmm_rocsparse_benchmark": ([_P, _I, _I, _I, _I, C.POINTER(C.c_float)], _ST),
        }
        for name, (args, res) in sig.items():
            f = getattr(r, name)
            f.argtypes, f.restype = args, res
        _rs_lib = r
    return _rs_lib


class RocSparse:
    """rocSPARSE generic SpMV/SpMM over the caller's device arrays: the
    comparator columns of the sweep (CusparseSpmmProblem,
    cusparse_helper.h:120-437).  Dense operands as the reference lays them
    out: row-major ld = nv, column-major ld = nc (B) / nr (C)."""

    def __init__(self, nr, nc, nnz, rowPtr, rowIdx, colIdx, values, dnInput, dnOutput,
                 stream=None):
        self.rs = rocsparse_lib()
        self.h = C.c_void_p()
        check(self.rs.dspmm_rocsparse_create(C.byref(self.h), nr, nc, nnz, ptr(rowPtr),
                                             ptr(rowIdx), ptr(colIdx), ptr(values), ptr(dnInput),
                                             ptr(dnOutput), stream), "rocsparse_create")

    def columns(self, nv):
        return self.rs.dspmm_rocsparse_columns(nv)

    def name(self, nv, c):
        return self.rs.dspmm_rocsparse_column_name(nv, c).decode()

    def layout(self, nv, c):
        return self.rs.dspmm_rocsparse_column_layout(nv, c)

    def run(self, nv, c):
        """Runs column c; False when rocSPARSE does not support it."""
        st = self.rs.dspmm_rocsparse_run(self.h, nv, c)
        if st == STATUS_NOT_SUPPORTED:
            return False
        check(st, f"rocsparse run {self.name(nv, c)}")
        return True

    def benchmark(self, nv, c, warmup=20, reps=200):
        ms = C.c_float()
        st = self.rs.dspmm_rocsparse_benchmark(self.h, nv, c, warmup, reps, C.byref(ms))
        if st == STATUS_NOT_SUPPORTED:
            return None
        check(st, f"rocsparse benchmark {self.name(nv, c)}")
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            self.rs.dspmm_rocsparse_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def seeded_uniform(seed, n):
    """Host copy of the server's device generator of B (SplitMix64 of
    seed ^ i * golden, 24-bit uniform, mapped to [-1, 1))."""
    with np.errstate(over="ignore"):
        i = np.arange(int(n), dtype=np.uint64)
        z = np.uint64(seed) ^ (i * np.uint64(0x9E3779B97F4A7C15))
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -24)
    return (np.float32(-1.0) + np.float32(2.0) * u).astype(np.float32)


def device_read(ptr_value, shape, dtype=np.float32):
    """Copies a device buffer (raw address) to a new numpy array."""
    out = np.empty(shape, dtype)
    if out.nbytes:
        check(lib().dspmm_memcpy_d2h(out.ctypes.data, ptr_value, out.nbytes), "memcpy_d2h")
    return out
