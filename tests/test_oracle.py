"""Pins the CPU oracle (oracle/oracle.c) -- CPU only.

How each oracle function is pinned:
  oracle_spmm_host32 (spmvspmm_benchmark.cu:11-39)
    - golden vectors from scipy.sparse float32 (bit-exact, independent lib);
    - closed forms: identity, diagonal, permutation, empty rows, integer
      operands (exact in any order), RM/CM agreement;
    - order and rounding: a cancellation case fixes left-to-right p order,
      an FMA-sensitive case fixes "multiply, round, add, round";
    - FP32 error within the recursive-summation bound of the FP64 result.
  oracle_spmm_fp64: dense float64 GEMM (golden) and brute force.
  oracle_csr_rowptr (:41-76): textbook count definition + edge cases.
  oracle_check_result (:93-116): eps boundary, layout mapping, first index.
  oracle_round_bf16: torch's bfloat16 conversion (RNE) + hand ties.
A planted-mistake test shows these checks fail for plausible oracle bugs.
"""
import glob
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))
RM, CM = O.ROW_MAJOR, O.COL_MAJOR


def rand_csr(rng, nr, nc, maxlen, int_values=False, empty_frac=0.0):
    lens = rng.integers(0, maxlen + 1, size=nr)
    lens[rng.random(nr) < empty_frac] = 0
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(nc, size=min(int(n), nc), replace=False))
                         for n in lens] or [np.zeros(0)]).astype(np.int32)
    if int_values:
        v = rng.integers(-8, 9, size=ci.size).astype(np.float32)
    else:
        v = rng.uniform(-1, 1, size=ci.size).astype(np.float32)
    return rp, ci, v


def dense_of(nr, nc, rp, ci, v):
    d = np.zeros((nr, nc), np.float64)
    for r in range(nr):
        for p in range(rp[r], rp[r + 1]):
            d[r, ci[p]] += v[p]
    return d


def bits(x):
    return np.ascontiguousarray(x, np.float32).view(np.uint32)


# ---------------------------------------------------------------- golden ----
@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_host32_matches_scipy_golden_bitwise(path):
    g = np.load(path)
    nr, nc, b = int(g["nr"]), int(g["nc"]), g["B"]
    nv = b.shape[1]
    out = O.spmm_host32(RM, nr, nc, nv, g["rowPtr"], g["colIdx"], g["values"], b)
    assert np.array_equal(bits(out), bits(g["C32"]))
    outc = O.spmm_host32(CM, nr, nc, nv, g["rowPtr"], g["colIdx"], g["values"], b.T.copy())
    assert np.array_equal(bits(outc.T), bits(g["C32"]))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_fp64_matches_dense_golden(path):
    g = np.load(path)
    nr, nc, b = int(g["nr"]), int(g["nc"]), g["B"]
    c64, absum = O.spmm_fp64(RM, nr, nc, b.shape[1], g["rowPtr"], g["colIdx"], g["values"], b)
    np.testing.assert_allclose(c64, g["C64"], rtol=0, atol=1e-12 * max(1.0, g["ABS"].max()))
    np.testing.assert_allclose(absum, g["ABS"], rtol=1e-12, atol=0)


# ----------------------------------------------------------- closed forms ----
@pytest.mark.parametrize("layout", [RM, CM])
def test_identity_returns_b_exactly(layout):
    rng = np.random.default_rng(0)
    n, nv = 97, 6
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    v = np.ones(n, np.float32)
    b = rng.standard_normal((n, nv)).astype(np.float32)
    dense = b if layout == RM else b.T.copy()
    out = O.spmm_host32(layout, n, n, nv, rp, ci, v, dense)
    assert np.array_equal(bits(out), bits(dense))


def test_diagonal_and_permutation_exact():
    rng = np.random.default_rng(1)
    n, nv = 64, 5
    perm = rng.permutation(n).astype(np.int32)
    d = rng.uniform(-2, 2, n).astype(np.float32)
    rp = np.arange(n + 1, dtype=np.int32)
    b = rng.standard_normal((n, nv)).astype(np.float32)
    out = O.spmm_host32(RM, n, n, nv, rp, perm, d, b)
    assert np.array_equal(bits(out), bits(d[:, None] * b[perm]))


def test_empty_rows_are_zero():
    rng = np.random.default_rng(2)
    rp, ci, v = rand_csr(rng, 50, 40, 6, empty_frac=0.5)
    b = rng.standard_normal((40, 3)).astype(np.float32)
    out = O.spmm_host32(RM, 50, 40, 3, rp, ci, v, b)
    empty = np.diff(rp) == 0
    assert empty.any() and np.all(out[empty] == 0)


def test_integer_operands_exact():
    rng = np.random.default_rng(3)
    nr, nc, nv = 120, 90, 7
    rp, ci, v = rand_csr(rng, nr, nc, 30, int_values=True)
    b = rng.integers(-8, 9, size=(nc, nv)).astype(np.float32)
    exact = (dense_of(nr, nc, rp, ci, v).astype(np.int64) @ b.astype(np.int64)).astype(np.float32)
    out = O.spmm_host32(RM, nr, nc, nv, rp, ci, v, b)
    assert np.array_equal(out, exact)


def test_row_and_column_major_agree_bitwise():
    rng = np.random.default_rng(4)
    nr, nc, nv = 80, 70, 9
    rp, ci, v = rand_csr(rng, nr, nc, 25)
    b = rng.standard_normal((nc, nv)).astype(np.float32)
    rm = O.spmm_host32(RM, nr, nc, nv, rp, ci, v, b)
    cm = O.spmm_host32(CM, nr, nc, nv, rp, ci, v, b.T.copy())
    assert np.array_equal(bits(rm), bits(cm.T))


def test_sum_is_left_to_right_in_nonzero_order():
    # (1e8 + 1) rounds back to 1e8 in FP32; then -1e8 gives 0.  Any other
    # order (or pairwise sums) gives 1.
    rp = np.array([0, 3], np.int32)
    ci = np.array([0, 1, 2], np.int32)
    v = np.array([1e8, 1.0, -1e8], np.float32)
    b = np.ones((3, 1), np.float32)
    assert O.spmm_host32(RM, 1, 3, 1, rp, ci, v, b)[0, 0] == 0.0
    v2 = np.array([1e8, -1e8, 1.0], np.float32)
    assert O.spmm_host32(RM, 1, 3, 1, rp, ci, v2, b)[0, 0] == 1.0


def test_multiply_and_add_are_rounded_separately():
    # acc = -1; acc += a*b with a = b = 1 + 2^-12: the exact product
    # 1 + 2^-11 + 2^-24 rounds (tie to even) to 1 + 2^-11, so the reference's
    # non-fused loop gives 2^-11; a fused multiply-add would give 2^-11 + 2^-24.
    x = np.float32(1 + 2 ** -12)
    rp = np.array([0, 2], np.int32)
    ci = np.array([0, 1], np.int32)
    v = np.array([-1.0, x], np.float32)
    b = np.array([[1.0], [x]], np.float32)
    assert O.spmm_host32(RM, 1, 2, 1, rp, ci, v, b)[0, 0] == np.float32(2 ** -11)


def test_host32_error_within_summation_bound():
    rng = np.random.default_rng(5)
    nr, nc, nv = 200, 300, 4
    rp, ci, v = rand_csr(rng, nr, nc, 200)
    b = rng.standard_normal((nc, nv)).astype(np.float32)
    c32 = O.spmm_host32(RM, nr, nc, nv, rp, ci, v, b).astype(np.float64)
    c64, absum = O.spmm_fp64(RM, nr, nc, nv, rp, ci, v, b)
    n = np.diff(rp).astype(np.float64)[:, None] + 1
    u = 2.0 ** -24
    gamma = n * u / (1 - n * u)
    assert np.all(np.abs(c32 - c64) <= gamma * absum + 1e-300)


def test_fp64_matches_bruteforce_dense():
    rng = np.random.default_rng(6)
    nr, nc, nv = 40, 33, 6
    rp, ci, v = rand_csr(rng, nr, nc, 12)
    b = rng.standard_normal((nc, nv)).astype(np.float32)
    d = dense_of(nr, nc, rp, ci, v)
    c64, absum = O.spmm_fp64(RM, nr, nc, nv, rp, ci, v, b)
    np.testing.assert_allclose(c64, d @ b.astype(np.float64), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(absum, np.abs(d) @ np.abs(b.astype(np.float64)), rtol=1e-13)
    c64c, _ = O.spmm_fp64(CM, nr, nc, nv, rp, ci, v, b.T.copy())
    np.testing.assert_array_equal(c64c.T, c64)


# -------------------------------------------------------------- rowPtr ------
def textbook_rowptr(nr, rows):
    return np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=nr))]).astype(np.int32)


@pytest.mark.parametrize("seed", range(6))
def test_rowptr_matches_count_definition(seed):
    rng = np.random.default_rng(seed)
    nr = int(rng.integers(1, 60))
    nnz = int(rng.integers(1, 200))
    rows = np.sort(rng.integers(0, nr, size=nnz)).astype(np.int32)
    assert np.array_equal(O.csr_rowptr(nr, rows), textbook_rowptr(nr, rows))


@pytest.mark.parametrize("nr,rows", [
    (1, [0]),                      # single row, nnz = 1
    (5, [0]),                      # nnz = 1 in the first row, trailing empties
    (5, [4]),                      # nnz = 1 in the last row, leading empties
    (6, [2, 2, 2]),                # everything in one middle row
    (7, [0, 3, 3, 6]),             # empty rows between runs
    (4, [0, 1, 2, 3]),             # one per row
])
def test_rowptr_edge_cases(nr, rows):
    rows = np.asarray(rows, np.int32)
    assert np.array_equal(O.csr_rowptr(nr, rows), textbook_rowptr(nr, rows))


# -------------------------------------------------------------- checkResult -
def test_check_result_semantics():
    ref = np.arange(12, dtype=np.float32).reshape(3, 4)
    assert O.check_result(RM, RM, 3, 4, ref, ref.copy())[0]
    out = ref.copy()
    out[1, 2] += np.float32(1e-3)          # |diff| <= eps passes (strict >)
    assert O.check_result(RM, RM, 3, 4, ref, out, 1e-3)[0]
    out[1, 2] += np.float32(1e-2)
    out[2, 0] += 1
    assert O.check_result(RM, RM, 3, 4, ref, out) == (False, 1, 2)   # first in (i, j) order
    # host row-major vs device column-major mapping
    assert O.check_result(RM, CM, 3, 4, ref, ref.T.copy())[0]
    assert not O.check_result(RM, RM, 3, 4, ref, ref.T.copy().reshape(3, 4))[0]
    nan = ref.copy()
    nan[0, 0] = np.nan                     # the reference's compare lets NaN through
    assert O.check_result(RM, RM, 3, 4, ref, nan)[0]


# -------------------------------------------------------------- bf16 --------
def test_round_bf16_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.standard_normal(5000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 5000),
                        np.array([0.0, -0.0, np.inf, -np.inf, 1.0, 65504.0, 3.4e38, 1e-40], np.float32)])
    ours = O.round_bf16(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(bits(ours), bits(ref))


def test_round_bf16_ties_to_even():
    one = np.float32(1.0)
    ulp = 2.0 ** -7                        # bf16 ulp at 1.0
    x = np.array([1 + ulp / 2, 1 + 1.5 * ulp, 1 + ulp / 2 + 2 ** -20], np.float32)
    np.testing.assert_array_equal(O.round_bf16(x), np.array([one, 1 + 2 * ulp, 1 + ulp], np.float32))
    assert np.isnan(O.round_bf16(np.array([np.nan], np.float32))[0])


# ------------------------------------------------------- planted mistakes --
def _mutant(kind):
    def run(nr, nc, nv, rp, ci, v, b):
        out = np.zeros((nr, nv), np.float32)
        for r in range(nr):
            acc = np.zeros(nv, np.float32)
            hi = rp[r + 1] - (1 if kind == "drop_last" else 0)
            for p in range(rp[r], hi):
                col = ci[p] if kind != "transpose" else r
                row_b = b[col] if kind != "wrong_stride" else b.reshape(-1)[ci[p]:ci[p] + nv]
                sign = -1 if (kind == "sign" and p == rp[r]) else 1
                acc = (acc + np.float32(sign) * v[p] * row_b).astype(np.float32)
            out[r] = acc
        return out
    return run


@pytest.mark.parametrize("kind", ["drop_last", "transpose", "wrong_stride", "sign"])
def test_planted_mistakes_are_caught(kind):
    """A restatement with a plausible bug fails the golden / closed-form
    checks above (the mistake is planted in a CPU copy, never in the oracle)."""
    g = np.load(GOLDEN[0])
    nr, nc, b = int(g["nr"]), int(g["nc"]), g["B"]
    bad = _mutant(kind)(nr, nc, b.shape[1], g["rowPtr"], g["colIdx"], g["values"], b)
    assert not np.array_equal(bits(bad), bits(g["C32"]))
    good = _mutant("none")(nr, nc, b.shape[1], g["rowPtr"], g["colIdx"], g["values"], b)
    assert np.array_equal(bits(good), bits(g["C32"]))
