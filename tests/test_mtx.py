"""CPU tests of the Matrix-Market loader (dspmm_mtx_read) and the sorted-COO
-> CSR row-offset build (dspmm_csr_rowptr).

Pinned three ways:
* default flags: equal, bit for bit, to the reference's readMtx<float>
  (util/spmatrix.hpp:286-333) on every fixture -- against the committed
  expected outputs, and against the reference loader itself when
  oracle/_ref/mtx_ref has been built;
* DSPMM_MTX_KEEP_DIAGONAL | DSPMM_MTX_ALIGN_VALUES: equal to scipy.io.mmread
  (full symmetric expansion, values with their entries);
* a hand-derived case of makeSymmetric's compaction (util/spmatrix.hpp:
  264-283), whose values stay where they were while the indices move.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest
import scipy.io
import scipy.sparse

import dspmm as D
import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
MTX = os.path.join(HERE, "golden", "mtx")
REF = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "mtx_ref")
GOOD = sorted(f[:-4] for f in os.listdir(MTX) if f.endswith(".mtx") and not f.startswith("bad_"))
CORRECTED = D.MTX_KEEP_DIAGONAL | D.MTX_ALIGN_VALUES


def path(name):
    return os.path.join(MTX, name + ".mtx")


def bits(x):
    return np.asarray(x, np.float32).view(np.uint32)


def same(got, exp):
    nr, nc, r, c, v = got
    assert (nr, nc) == (int(exp[0]), int(exp[1]))
    assert np.array_equal(r, exp[2]) and np.array_equal(c, exp[3])
    assert np.array_equal(bits(v), bits(exp[4]))


def test_fixture_set_is_complete():
    assert len(GOOD) >= 10
    for name in GOOD:
        assert os.path.exists(os.path.join(MTX, name + ".npz")), name


@pytest.mark.parametrize("name", GOOD)
def test_default_equals_reference_golden(name):
    z = np.load(os.path.join(MTX, name + ".npz"))
    same(D.read_mtx(path(name)), (z["nr"], z["nc"], z["rowIdx"], z["colIdx"], z["values"]))


@pytest.mark.skipif(not os.path.exists(REF), reason="reference loader not built "
                    "(make -C oracle -f Makefile.ref needs the reference sources)")
@pytest.mark.parametrize("name", GOOD)
def test_default_equals_reference_live(name):
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "o.bin")
        subprocess.run([REF, path(name), out], check=True, stdout=subprocess.DEVNULL)
        raw = np.fromfile(out, np.int32)
    n = int(raw[2])
    same(D.read_mtx(path(name)),
         (raw[0], raw[1], raw[3:3 + n], raw[3 + n:3 + 2 * n], raw[3 + 2 * n:].view(np.float32)))


def scipy_coo(name, banner):
    m = scipy.io.mmread(path(name))
    m = scipy.sparse.coo_matrix(m)
    r, c, v = m.row.astype(np.int32), m.col.astype(np.int32), m.data.astype(np.float32)
    if "pattern" in banner:
        v = np.ones_like(v)
    o = np.lexsort((c, r))
    return m.shape[0], m.shape[1], r[o], c[o], v[o]


@pytest.mark.parametrize("name", [n for n in GOOD if "skew" not in n and "both_halves" not in n])
def test_corrected_equals_scipy(name):
    with open(path(name)) as f:
        banner = f.readline().lower()
    same(D.read_mtx(path(name), CORRECTED), scipy_coo(name, banner))


def test_misaligned_values_by_hand():
    # sorted after mirroring: (0,0)1 (0,1)2 (1,0)2 (1,2)4 (2,1)4 (2,2)3; the
    # diagonal goes, indices shift left by one, values stay in place
    nr, nc, r, c, v = D.read_mtx(path("symmetric_tiny"))
    assert (nr, nc) == (3, 3)
    assert r.tolist() == [0, 1, 1, 2] and c.tolist() == [1, 0, 2, 1]
    assert v.tolist() == [1.0, 2.0, 2.0, 4.0]
    _, _, r2, c2, v2 = D.read_mtx(path("symmetric_tiny"), D.MTX_ALIGN_VALUES)
    assert r2.tolist() == r.tolist() and c2.tolist() == c.tolist()
    assert v2.tolist() == [2.0, 2.0, 4.0, 4.0]
    _, _, r3, c3, v3 = D.read_mtx(path("symmetric_tiny"), CORRECTED)
    assert list(zip(r3.tolist(), c3.tolist(), v3.tolist())) == [
        (0, 0, 1.0), (0, 1, 2.0), (1, 0, 2.0), (1, 2, 4.0), (2, 1, 4.0), (2, 2, 3.0)]


def test_symmetric_duplicates_and_self_loops():
    # both halves given: one copy of each mirrored pair, the diagonal dropped
    nr, nc, r, c, v = D.read_mtx(path("symmetric_both_halves"), D.MTX_ALIGN_VALUES)
    assert list(zip(r.tolist(), c.tolist(), v.tolist())) == [
        (0, 1, 0.5), (0, 3, 2.0), (1, 0, 0.5), (1, 2, -1.25), (2, 1, -1.25), (3, 0, 2.0)]
    _, _, r2, c2, v2 = D.read_mtx(path("sym_random_200"), D.MTX_ALIGN_VALUES)
    assert np.all(r2 != c2)
    key = r2.astype(np.int64) * 200 + c2
    assert np.all(np.diff(key) > 0)
    # mirrored pairs carry one value
    m = dict(zip(key.tolist(), v2.tolist()))
    assert all(m[cc * 200 + rr] == vv for rr, cc, vv in zip(r2.tolist(), c2.tolist(), v2.tolist()))


def test_values_are_float32_of_the_text():
    _, _, r, c, v = D.read_mtx(path("general_integer"))
    assert v[(r == 1) & (c == 0)][0] == np.float32(16777217)   # int -> float rounds
    _, _, r, c, v = D.read_mtx(path("general_real"))
    assert bits(v[(r == 1) & (c == 3)][0]) == bits(np.float32("0.333333343"))
    assert bits(v[(r == 1) & (c == 1)][0]) == bits(np.float32(-0.0))
    _, _, _, _, v = D.read_mtx(path("general_pattern"))
    assert np.all(v == 1.0)


@pytest.mark.parametrize("name,status", [
    ("bad_complex", D.STATUS_NOT_SUPPORTED), ("bad_array", D.STATUS_INVALID_VALUE),
    ("bad_truncated", D.STATUS_INVALID_VALUE), ("bad_banner", D.STATUS_INVALID_VALUE),
    ("bad_index", D.STATUS_INVALID_VALUE)])
def test_rejected_files(name, status):
    with pytest.raises(D.DspmmError) as e:
        D.read_mtx(path(name))
    assert e.value.status == status


def test_missing_file():
    with pytest.raises(D.DspmmError):
        D.read_mtx(os.path.join(MTX, "does_not_exist.mtx"))


@pytest.mark.parametrize("name", GOOD)
def test_rowptr_equals_oracle(name):
    nr, nc, r, c, v = D.read_mtx(path(name))
    rp = D.csr_rowptr(nr, r)
    assert np.array_equal(rp, np.concatenate([[0], np.cumsum(np.bincount(r, minlength=nr))]))
    if r.size:
        # the oracle's restatement of generateCsrRowPtrArray (defined for nnz > 0)
        assert np.array_equal(rp, O.csr_rowptr(nr, r))


def test_rowptr_edges():
    assert D.csr_rowptr(3, np.zeros(0, np.int32)).tolist() == [0, 0, 0, 0]
    assert D.csr_rowptr(0, np.zeros(0, np.int32)).tolist() == [0]
    assert D.csr_rowptr(4, np.array([3, 3, 3])).tolist() == [0, 0, 0, 0, 3]
    for bad in ([1, 0], [0, 4], [-1, 0]):
        with pytest.raises(D.DspmmError):
            D.csr_rowptr(4, np.array(bad))
