"""C-ABI checks that need no GPU: the library loads, exports every function
include/*.h declares, keeps the reference's enum values, and the host-side
argument checks reject bad input without touching the device."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import dspmm as D

HEADER = os.path.join(D.INCLUDE_DIR, "dspmm.h")


RS_HEADER = os.path.join(D.INCLUDE_DIR, "dspmm_rocsparse.h")
LIBS = [(D.LIB_PATH, HEADER), (D.ROCSPARSE_LIB_PATH, RS_HEADER)]


def test_library_exports_every_declared_symbol(dlib):
    names = D.declared_symbols(HEADER)
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(dlib, n)]
    assert not missing, missing


def test_every_header_belongs_to_a_library():
    import glob
    assert sorted(glob.glob(os.path.join(D.INCLUDE_DIR, "*.h"))) == sorted(h for _, h in LIBS)


@pytest.mark.parametrize("lib,header", LIBS, ids=["dspmm", "rocsparse_comparator"])
def test_exported_symbols_are_exactly_the_declared_ones(lib, header):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    declared = set(D.declared_symbols(header))
    assert declared and declared <= exported
    extra = {s for s in exported - declared if not s.startswith("_")}
    assert not extra, f"undeclared exports: {sorted(extra)}"


def test_comparator_loads_and_names_the_reference_columns():
    rs = D.rocsparse_lib()
    assert rs.dspmm_rocsparse_columns(1) == 4 and rs.dspmm_rocsparse_columns(8) == 10
    assert rs.dspmm_rocsparse_columns(0) == 0
    names = [rs.dspmm_rocsparse_column_name(8, c).decode() for c in range(10)]
    assert len(set(names)) == 10 and all(names)
    lay = [rs.dspmm_rocsparse_column_layout(8, c) for c in range(10)]
    assert lay == [D.DENSE_ROW_MAJOR] * 4 + [D.DENSE_COL_MAJOR] * 6
    assert rs.dspmm_rocsparse_column_layout(8, 10) == -1
    h = C.c_void_p()
    assert rs.dspmm_rocsparse_create(C.byref(h), 4, 4, 1, None, None, None, None, None, None,
                                     None) == D.STATUS_INVALID_VALUE


def test_harness_usage_and_bad_file(tmp_path):
    exe = D.HARNESS_PATH
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 2 and "usage: spmvspmm" in out.stderr
    bad = os.path.join(os.path.dirname(__file__), "golden", "mtx", "bad_banner.mtx")
    out = subprocess.run([exe, bad], capture_output=True, text=True, timeout=60)
    assert out.returncode == 2 and "Could not read" in out.stderr


def test_reference_launcher_names_present():
    names = set(D.declared_symbols())
    for n in D.LAUNCHERS + ["dspmm_spmm"]:
        assert n in names


def test_enum_values_match_reference():
    text = open(HEADER).read()
    def val(name):
        return int(re.search(rf"\b{name}\s*=\s*(\d+)", text).group(1))
    # spmvspmm.h:5-10, helper.h:256-264
    assert [val(n) for n in ("ALG_CSR_SCALAR", "ALG_CSR_VECTOR", "ALG_COO_SCALAR",
                             "ALG_COO_VECTOR")] == [0, 1, 2, 3]
    assert (val("DENSE_ROW_MAJOR"), val("DENSE_COL_MAJOR")) == (0, 1)
    assert (val("SPARSE_FORMAT_CSR"), val("SPARSE_FORMAT_COO")) == (0, 1)
    assert (D.ALG_CSR_SCALAR, D.ALG_COO_VECTOR, D.DENSE_COL_MAJOR) == (0, 3, 1)


def test_error_strings(dlib):
    for s in range(7):
        assert dlib.dspmm_get_error_string(s)
    assert dlib.dspmm_get_error_string(0) == b"success"


# Host-side argument validation returns before any HIP call, so these run
# without a GPU.
@pytest.mark.parametrize("nv", [0, 3, 6, 12])
def test_dispatcher_rejects_non_power_of_two(dlib, nv):
    st = dlib.dspmm_spmm(0, 0, 4, 4, 4, nv, None, None, None, None, None, None, None)
    assert st == (D.STATUS_NOT_SUPPORTED if nv > 0 else 1)


def test_launchers_reject_bad_sizes(dlib):
    assert dlib.spmm_csr_scalar_row(-1, 4, 4, 1, None, None, None, None, None, None, None) == 1
    assert dlib.spmv_csr_scalar(4, 4, 0, 2, 1, None, None, None, None, 1, None) == 2   # nv != 1
    assert dlib.spmv_csr_vector(4, 4, 0, 3, 1, None, None, None, None, 1, None) == 2   # nv not 1/2/4
    # COO algorithms need rowIdx when nnz > 0
    assert dlib.spmm_coo_scalar_row(4, 4, 4, 2, 1, None, 1, 1, 1, 1, None) == 1
    # nothing to do is success, without touching the device
    assert dlib.spmm_coo_vector_row(4, 4, 0, 2, 1, None, None, None, None, 1, None) == 0
    assert dlib.spmm_csr_scalar_row(0, 4, 0, 2, None, None, None, None, None, None, None) == 0


def test_matrix_create_validates_on_host(dlib):
    h = C.c_void_p()
    rp = np.array([0, 2, 1], np.int32)            # decreasing rowPtr
    ci = np.array([0, 1], np.int32)
    v = np.ones(2, np.float32)
    assert dlib.dspmm_matrix_create(C.byref(h), 2, 2, 2, rp.ctypes.data, None, ci.ctypes.data,
                                    v.ctypes.data, 0) == 1
    rows = np.array([1, 0], np.int32)             # unsorted COO rows
    assert dlib.dspmm_matrix_create(C.byref(h), 2, 2, 2, None, rows.ctypes.data, ci.ctypes.data,
                                    v.ctypes.data, 0) == 1
    ci_bad = np.array([0, 5], np.int32)           # column out of range
    rp_ok = np.array([0, 1, 2], np.int32)
    assert dlib.dspmm_matrix_create(C.byref(h), 2, 2, 2, rp_ok.ctypes.data, None,
                                    ci_bad.ctypes.data, v.ctypes.data, 0) == 1


def test_python_problem_mirrors_reference_nv_rule(capsys):
    prob = D.SpmvSpmmProblem(4, 4, 4, 8, 0, 0, 0, 0, 0, 0)
    assert prob.run(D.DENSE_ROW_MAJOR, 3, D.ALG_CSR_SCALAR) is None
    assert prob.benchmark(D.DENSE_ROW_MAJOR, 6, D.ALG_CSR_SCALAR) == -1
    assert prob.dry_run(D.DENSE_ROW_MAJOR, 64, D.ALG_CSR_SCALAR) is None
    err = capsys.readouterr().err
    assert err.count("nv must be power of 2") == 3


def test_generators_are_deterministic():
    import workloads as W
    a = W.gen_csr(W.GEN_ZIPF, 3000, 3000, 9, 1.8, 1, 500, W.VAL_NORMAL)
    b = W.gen_csr(W.GEN_ZIPF, 3000, 3000, 9, 1.8, 1, 500, W.VAL_NORMAL)
    assert np.array_equal(a.rowPtr, b.rowPtr) and np.array_equal(a.colIdx, b.colIdx)
    assert np.array_equal(a.values, b.values)
    rp = a.rowPtr
    for r in range(0, 3000, 37):
        seg = a.colIdx[rp[r]:rp[r + 1]]
        assert np.all(np.diff(seg) > 0) and (seg.size == 0 or (seg[0] >= 0 and seg[-1] < 3000))
    s = W.c3_matrix(8)
    assert s.nnz == (3 * 8 - 2) ** 3


def test_cxx_caller_compiles_links_and_runs(tmp_path):
    """A C++ program using include/dspmm.h + include/spmvspmm.hpp builds with
    g++ against libdspmm.so and runs its GPU-free calls."""
    import shutil
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    src = os.path.join(os.path.dirname(__file__), "cxx", "abi_link.cpp")
    exe = str(tmp_path / "abi_link")
    libdir = os.path.dirname(D.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{D.INCLUDE_DIR}", src, "-o", exe,
                    f"-L{libdir}", "-ldspmm", f"-Wl,-rpath,{libdir}"], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert out.stdout.startswith("ok dspmm")
    assert out.stderr.count("nv must be power of 2") == 2
    # the ctypes mirror lays out every public struct as the C header does
    line = next(l for l in out.stdout.splitlines() if l.startswith("sizes "))
    sizes = dict(kv.split("=") for kv in line.split()[1:])
    import ctypes
    mirror = {"device_info": D.DeviceInfo, "matrix_desc": D.MatrixDesc, "config": D.Config,
              "batch_item": D.BatchItem, "batch_info": D.BatchInfo,
              "server_config": D.ServerConfig, "request": D.Request, "completion": D.Completion,
              "coo": D.Coo, "server_stats": D.ServerStats}
    for name, cls in mirror.items():
        assert int(sizes[name]) == ctypes.sizeof(cls), (name, sizes[name], ctypes.sizeof(cls))
