"""dspmm_spmm_ex_noalloc, through the C-ABI: the entry point that allocates
nothing (its scratch is a static device array) gives the same bits as
dspmm_spmm_ex with the same configuration, on two calls in a row -- the
second with a smaller input -- for split rows (carry scratch), column-major
operands (transposed copies in the scratch), AUTO and the B-row ring; it refuses an
unprepared handle and a call whose scratch would exceed the maximum."""
import ctypes as C

import numpy as np
import pytest

import dspmm as D
import workloads as W
from gpu_common import bits, note

pytestmark = pytest.mark.gpu
RM, CM = D.DENSE_ROW_MAJOR, D.DENSE_COL_MAJOR


class Op:
    """A matrix with its own dense operands (B: K x nv, C: M x nv).  Values
    and B are small integers, so every summation order (atomics included,
    which AUTO may pick) gives the same bits."""

    def __init__(self, a, nv, seed):
        rng = np.random.default_rng(seed)
        vals = rng.integers(-8, 9, a.nnz).astype(np.float32)
        self.a, self.nv = W.Csr(a.nr, a.nc, a.rowPtr, a.colIdx, vals), nv
        self.m = D.Matrix(a.nr, a.nc, a.rowPtr, a.colIdx, vals)
        self.d_B = D.DeviceArray.from_numpy(rng.integers(-8, 9, a.nc * nv).astype(np.float32))
        self.d_C = D.DeviceArray.zeros(a.nr * nv)

    def ld(self, layout):
        return (self.nv, self.nv) if layout == RM else (self.a.nc, self.a.nr)

    def noalloc(self, layout, cfg):
        ldb, ldc = self.ld(layout)
        return D.lib().dspmm_spmm_ex_noalloc(self.m.handle, layout, self.nv, self.d_B.ptr, ldb,
                                             D.DSPMM_R_32F, self.d_C.ptr, ldc,
                                             C.byref(cfg) if cfg is not None else None, None)

    def reference(self, layout, cfg):
        ldb, ldc = self.ld(layout)
        self.d_C.zero()
        self.m.spmm(self.d_B, self.d_C, self.nv, layout=layout, ldb=ldb, ldc=ldc, config=cfg)
        D.synchronize()
        return self.d_C.numpy().copy()

    def run_noalloc(self, layout, cfg):
        self.d_C.zero()
        st = self.noalloc(layout, cfg)
        assert st == D.STATUS_SUCCESS, D.status_string(st)
        D.synchronize()
        return self.d_C.numpy().copy()

    def close(self):
        self.m.close()


def zipf(n, seed, kmax):
    return W.gen_csr(W.GEN_ZIPF, n, n, seed, 1.7, 1, kmax, W.VAL_NORMAL)


@pytest.mark.parametrize("layout,cfg", [
    (RM, D.make_config(D.VARIANT_NNZ_TILED, tile_nnz=32)),     # split rows: carry scratch
    (CM, D.make_config(D.VARIANT_NNZ_TILED, tile_nnz=64)),     # transposes + carry in scratch
    (RM, D.make_config(D.VARIANT_SPMV_TILED, tile_nnz=512)),   # long SpMV rows: carry chunks
    (RM, None),                                                # AUTO, picked by prepare
], ids=["rm-nnz-tiled", "cm-transposed", "spmv", "auto"])
def test_noalloc_equals_spmm_ex_twice_second_smaller(layout, cfg):
    spmv = cfg is not None and cfg.variant == D.VARIANT_SPMV_TILED
    big = Op(zipf(30000, 41, 12000), 1 if spmv else 48, 41)
    small = Op(zipf(7000, 42, 9000), 1 if spmv else 24, 42)
    for op in (big, small):
        ldb, ldc = op.ld(layout)
        op.m.prepare(op.d_B, op.d_C, op.nv, layout=layout, ldb=ldb, ldc=ldc, config=cfg)
    D.synchronize()
    want = {id(op): op.reference(layout, cfg) for op in (big, small)}
    for op in (big, small):                  # two calls in a row, the second smaller
        got = op.run_noalloc(layout, cfg)
        assert np.array_equal(bits(got), bits(want[id(op)])), \
            f"{op.a.nr} rows nv={op.nv}: no-alloc result differs from dspmm_spmm_ex"
    # and once more in the same order: the scratch keeps the last call's
    # contents, which must not leak into the next result
    got = big.run_noalloc(layout, cfg)
    assert np.array_equal(bits(got), bits(want[id(big)]))
    note(f"no-alloc {layout} {cfg.as_dict() if cfg else 'auto'}: equal to spmm_ex, 3 calls")
    big.close()
    small.close()


def test_noalloc_ring_on_stencils():
    """The B-row ring needs no scratch: prepared, it runs through the
    no-alloc entry point with dspmm_spmm_ex's bits (27-point stencils, the
    second call smaller)."""
    cfg = D.make_config(D.VARIANT_RING, tile_nnz=4, unroll=8)
    big = Op(W.c3_matrix(20), 64, 45)
    small = Op(W.c3_matrix(12), 32, 46)
    for op in (big, small):
        op.m.prepare(op.d_B, op.d_C, op.nv, config=cfg)
    D.synchronize()
    want = {id(op): op.reference(RM, cfg) for op in (big, small)}
    for op in (big, small, big):
        got = op.run_noalloc(RM, cfg)
        assert np.array_equal(bits(got), bits(want[id(op)])), f"ring {op.a.nr} rows nv={op.nv}"
    note("no-alloc ring: equal to spmm_ex, 3 calls")
    big.close()
    small.close()


def test_noalloc_refuses_unprepared_and_oversized():
    op = Op(zipf(5000, 43, 3000), 16, 43)
    cfg = D.make_config(D.VARIANT_NNZ_TILED, tile_nnz=32)
    assert op.noalloc(RM, cfg) == D.STATUS_NOT_PREPARED            # no plan yet
    assert op.noalloc(RM, None) == D.STATUS_NOT_PREPARED           # AUTO not picked yet
    op.close()
    # column-major copies of a 300000 x 64 operand exceed the scratch
    n = 300000
    rp = np.arange(n + 1, dtype=np.int32)
    big = Op(W.Csr(n, n, rp, (np.arange(n, dtype=np.int32) * 7) % n,
                   np.ones(n, np.float32)), 64, 44)
    cfg = D.make_config(D.VARIANT_NNZ_TILED)
    big.m.prepare(big.d_B, big.d_C, 64, layout=CM, ldb=n, ldc=n, config=cfg)
    assert big.noalloc(CM, cfg) == D.STATUS_WORKSPACE_TOO_SMALL
    # the row-major call of the same size needs no scratch and runs
    big.m.prepare(big.d_B, big.d_C, 64, config=cfg)
    want = big.reference(RM, cfg)
    got = big.run_noalloc(RM, cfg)
    assert np.array_equal(bits(got), bits(want))
    big.close()
