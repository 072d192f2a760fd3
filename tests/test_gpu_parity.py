"""GPU parity of the reference-compatible path: the 12 launchers and the
cudaSpmm-shaped dispatcher against the CPU oracle, through the C-ABI.

Mirrors the reference's own check (checkAlgorithms,
spmvspmm_benchmark.cu:118-233): every algorithm x layout x power-of-two nv,
output zeroed before each run, compared with the host loop -- here with the
stricter gates of gpu_common (bit-exact for the sequential row-balanced
kernels, gamma-bounded FP32 for the rest, exact for integer operands).
"""
import glob
import os

import numpy as np
import pytest

import dspmm as D
import oracle as O
from gpu_common import (assert_bitwise, assert_fp32, case, golden_case, note)

pytestmark = pytest.mark.gpu

ALGS = [D.ALG_CSR_SCALAR, D.ALG_CSR_VECTOR, D.ALG_COO_SCALAR, D.ALG_COO_VECTOR]
LAYOUTS = [D.DENSE_ROW_MAJOR, D.DENSE_COL_MAJOR]
GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def run_and_check(c, layout, nv, alg, zero_first=True):
    """One SPMV_SPMM_CHECK_SWITCH (spmvspmm_benchmark.cu:153-157)."""
    name = f"{c.name} {D.CSV_NAMES[(alg, layout)]} nv={nv}"
    note(name)
    if zero_first:
        c.fill_output(0.0)
    c.problem().run(layout, nv, alg)
    D.synchronize()
    out = c.output(nv)
    h32, c64, ab = c.ref(layout, nv)
    ok, fi, fj = O.check_result(layout, layout, c.nr, nv, h32, out)   # the reference's 1e-3 check
    assert ok, f"{name}: reference check failed at ({fi},{fj})"
    if alg == D.ALG_CSR_SCALAR:
        assert_bitwise(out, h32, name)
    else:
        assert_fp32(out, c64, ab, c.rowlen_flat(layout, nv), name, h32)
    return out


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
@pytest.mark.parametrize("layout", LAYOUTS, ids=["RM", "CM"])
@pytest.mark.parametrize("alg", ALGS, ids=["CSR_SCALAR", "CSR_VECTOR", "COO_SCALAR", "COO_VECTOR"])
def test_golden_all_algorithms(path, layout, alg):
    c = golden_case(path)
    g = np.load(path)
    for nv in (1, 2, 4, 8):
        if nv > c.maxnv:
            break
        out = run_and_check(c, layout, nv, alg)
        if nv == c.maxnv and layout == D.DENSE_ROW_MAJOR and alg == D.ALG_CSR_SCALAR:
            assert_bitwise(out, g["C32"].reshape(-1), "golden C32")
        if "int_valued" in path and nv == c.maxnv and layout == D.DENSE_ROW_MAJOR:
            np.testing.assert_array_equal(out, g["C32"].reshape(-1))


@pytest.mark.parametrize("layout", LAYOUTS, ids=["RM", "CM"])
@pytest.mark.parametrize("alg", ALGS, ids=["CSR_SCALAR", "CSR_VECTOR", "COO_SCALAR", "COO_VECTOR"])
def test_c1_sweep_all_nv(layout, alg):
    """The uniform matrix (M=K=4096, 16 nnz/row): nv = 1..128 as the reference sweeps it."""
    c = case("c1")
    for nv in (1, 2, 4, 8, 16, 32, 64, 128):
        run_and_check(c, layout, nv, alg)


@pytest.mark.parametrize("layout", LAYOUTS, ids=["RM", "CM"])
@pytest.mark.parametrize("alg", ALGS, ids=["CSR_SCALAR", "CSR_VECTOR", "COO_SCALAR", "COO_VECTOR"])
def test_power_law_and_edges(layout, alg):
    for name, nvs in (("zipf_mid", (1, 4, 32, 128)), ("edge", (1, 2, 16, 64))):
        c = case(name)
        for nv in nvs:
            run_and_check(c, layout, nv, alg)


@pytest.mark.parametrize("layout", LAYOUTS, ids=["RM", "CM"])
@pytest.mark.parametrize("alg", ALGS, ids=["CSR_SCALAR", "CSR_VECTOR", "COO_SCALAR", "COO_VECTOR"])
def test_integer_operands_exact(layout, alg):
    """Integer-valued A and B: every partial sum is exact, so even the atomic
    (nnz-balanced) algorithms must equal the host loop bit for bit."""
    c = case("int_valued")
    for nv in (1, 4, 64):
        out = run_and_check(c, layout, nv, alg)
        assert_bitwise(out, c.ref(layout, nv)[0], f"int_valued alg={alg} nv={nv}")


@pytest.mark.parametrize("nv", [1, 32, 128])
@pytest.mark.parametrize("alg", ALGS, ids=["CSR_SCALAR", "CSR_VECTOR", "COO_SCALAR", "COO_VECTOR"])
def test_c2_row_major(nv, alg):
    """The power-law matrix (M=K=262144, Zipf rows up to 5000): SpMV and SpMM."""
    run_and_check(case("c2"), D.DENSE_ROW_MAJOR, nv, alg)


def test_row_balanced_overwrites_and_nnz_balanced_accumulates():
    c = case("zipf_mid")
    nv = 8
    h32, c64, ab = c.ref(D.DENSE_ROW_MAJOR, nv)
    c.fill_output(np.nan)                       # CSR algorithms overwrite everything
    for alg in (D.ALG_CSR_SCALAR, D.ALG_CSR_VECTOR):
        c.problem().run(D.DENSE_ROW_MAJOR, nv, alg)
        D.synchronize()
        assert np.all(np.isfinite(c.output(nv)))
    c.fill_output(0.0)                          # COO algorithms accumulate
    for _ in range(2):
        c.problem().run(D.DENSE_ROW_MAJOR, nv, D.ALG_COO_SCALAR)
    D.synchronize()
    assert_fp32(c.output(nv) / 2, c64, ab, c.rowlen_flat(D.DENSE_ROW_MAJOR, nv), "2x accumulate",
                h32)


@pytest.mark.parametrize("name", D.LAUNCHERS)
def test_named_launchers_direct(name):
    """Each reference-named launcher on its own, packed operands."""
    c = case("c1")
    col = name.endswith("_col")
    spmv = name.startswith("spmv")
    nvs = (1,) if "scalar" in name and spmv else ((1, 2, 4) if spmv else (3, 8, 40))
    layout = D.DENSE_COL_MAJOR if col else D.DENSE_ROW_MAJOR
    for nv in nvs:
        note(f"{name} nv={nv}")
        c.fill_output(0.0)
        st = D.launch(name, c.nr, c.nc, c.nnz, nv, c.d_rp, c.d_ri, c.d_ci, c.d_v, c.d_B, c.d_C)
        assert st == 0, D.status_string(st)
        D.synchronize()
        h32, c64, ab = c.ref(layout, nv)
        out = c.output(nv)
        if "csr_scalar" in name:
            assert_bitwise(out, h32, name)
        else:
            assert_fp32(out, c64, ab, c.rowlen_flat(layout, nv), name, h32)


def test_empty_and_degenerate_shapes():
    # all rows empty: RB writes zeros, EB leaves the (zeroed) output alone
    nr, nc = 300, 50
    c0 = golden_case(GOLDEN[0])   # any device buffers big enough for the pointers
    rp = D.DeviceArray.from_numpy(np.zeros(nr + 1, np.int32))
    out = D.DeviceArray.from_numpy(np.full(nr * 4, 7.0, np.float32))
    for alg in ALGS:
        for layout in LAYOUTS:
            st = D.lib().dspmm_spmm(alg, layout, nr, nc, 0, 4, rp.ptr, 0, 0, 0, c0.d_B.ptr,
                                    out.ptr, None)
            assert st == 0
            D.synchronize()
            res = out.numpy()
            if alg in (D.ALG_CSR_SCALAR, D.ALG_CSR_VECTOR):
                assert np.all(res == 0)
            out = D.DeviceArray.from_numpy(np.full(nr * 4, 7.0, np.float32))
    # one row, one nonzero
    rp1 = D.DeviceArray.from_numpy(np.array([0, 1], np.int32))
    ri1 = D.DeviceArray.from_numpy(np.array([0], np.int32))
    ci1 = D.DeviceArray.from_numpy(np.array([2], np.int32))
    v1 = D.DeviceArray.from_numpy(np.array([3.0], np.float32))
    b1 = D.DeviceArray.from_numpy(np.arange(12, dtype=np.float32))
    o1 = D.DeviceArray.zeros(4)
    for alg in ALGS:
        o1.zero()
        assert D.lib().dspmm_spmm(alg, 0, 1, 3, 1, 4, rp1.ptr, ri1.ptr, ci1.ptr, v1.ptr, b1.ptr,
                                  o1.ptr, None) == 0
        D.synchronize()
        np.testing.assert_array_equal(o1.numpy(), 3.0 * np.arange(8, 12, dtype=np.float32))


def test_benchmark_and_dry_run_paths():
    c = case("c1")
    per = []
    ms = c.problem().benchmark(D.DENSE_ROW_MAJOR, 32, D.ALG_CSR_SCALAR, 2, 10, per_launch=per)
    assert ms > 0 and len(per) == 10 and all(x > 0 for x in per)
    c.problem().dry_run(D.DENSE_ROW_MAJOR, 32, D.ALG_COO_VECTOR, 1)
    D.synchronize()
