"""GPU tests of the batched path: one launch over many matrices
equals the single-matrix kernel bit for bit (same per-matrix config), is
within the FP32 gate of the FP64 oracle, and its hipGraph replay follows
operand rebinding (set_operands) across replays with eager work between
them."""
import numpy as np
import pytest

import dspmm as D
import oracle as O
import workloads as W
from gpu_common import assert_bf16, assert_bitwise, assert_exact, assert_fp32, bits, case, note

pytestmark = pytest.mark.gpu
RM = D.DENSE_ROW_MAJOR


class Item:
    def __init__(self, i, a, nv, seed, bf16=False):
        self.a, self.nv = a, nv
        self.B = W.dense_b(seed, a.nc, nv)
        self.m = D.Matrix(a.nr, a.nc, a.rowPtr, a.colIdx, a.values)
        self.d_B = D.DeviceArray.from_numpy(self.B)
        if bf16:
            self.d_B = D.to_bf16_device(self.d_B)
        self.d_C = D.DeviceArray.from_numpy(np.full((a.nr, nv), np.nan, np.float32))
        self.d_C_ref = D.DeviceArray.from_numpy(np.full((a.nr, nv), np.nan, np.float32))

    def oracle(self, B=None):
        B = self.B if B is None else B
        c64, ab = O.spmm_fp64(RM, self.a.nr, self.a.nc, self.nv, self.a.rowPtr, self.a.colIdx,
                              self.a.values, B)
        rl = np.repeat(np.diff(self.a.rowPtr).astype(np.float64), self.nv)
        return c64.reshape(-1), ab.reshape(-1), rl

    def host32(self, B=None):
        B = self.B if B is None else B
        return O.spmm_host32(RM, self.a.nr, self.a.nc, self.nv, self.a.rowPtr, self.a.colIdx,
                             self.a.values, B).reshape(-1)


def make_items(count, seed=40, bf16=False, mmax=3000):
    m, d, n = W.c4_params(count, seed)
    items = []
    for i in range(count):
        mm = int(min(m[i], mmax))
        a = W.gen_csr(W.GEN_BINOMIAL, mm, mm, seed * 1000 + i, p0=float(min(d[i], mm)))
        items.append(Item(i, a, int(n[i]) if not bf16 else max(8, int(n[i]) // 2), seed + i, bf16))
    # plus shapes the c4 draw rarely makes: a long split row, empty rows, nv=1/3
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 6, 900)
    lens[[0, 1, 2, 899]] = 0
    lens[450] = 2500
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(3000, int(k), replace=False)) for k in lens if k]).astype(np.int32)
    v = rng.uniform(-1, 1, ci.size).astype(np.float32)
    for nv in ((1, 3, 64) if not bf16 else (8, 64)):
        items.append(Item(len(items), W.Csr(900, 3000, rp, ci, v), nv, seed + 77 + nv, bf16))
    # a row of 20000 nonzeros at nv = 100: 79 chunks of 256 in two column
    # blocks, which the single-matrix combine sums with a workgroup per block
    # and the batched combine with one wavefront
    lens = np.array([3, 0, 20000, 5, 70, 1])
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(30000, int(k), replace=False)) for k in lens if k]).astype(np.int32)
    v = rng.uniform(-1, 1, ci.size).astype(np.float32)
    items.append(Item(len(items), W.Csr(6, 30000, rp, ci, v), 100, seed + 99, bf16))
    return items


def test_batched_equals_single_and_oracle():
    items = make_items(24)
    b = D.Batch([(it.m, it.nv, it.d_B, it.d_C) for it in items])
    info = b.describe()
    note(f"batch: {info}")
    assert info["units"] > 0 and info["blocks"] % 8 == 0
    b.run()
    for i, it in enumerate(items):
        # single-matrix path with the configuration the batch runs the item
        # with (AUTO on a single launch may pick another kernel or tile)
        it.m.spmm(it.d_B, it.d_C_ref, it.nv, config=b.item_config(i))
    D.synchronize()
    for i, it in enumerate(items):
        out, ref = it.d_C.numpy().reshape(-1), it.d_C_ref.numpy().reshape(-1)
        assert_bitwise(out, ref, f"item {i} batched vs single")
        c64, ab, rl = it.oracle()
        assert_fp32(out, c64, ab, rl, f"item {i} (M={it.a.nr}, nv={it.nv})", it.host32())


def test_plain_launches_match_batched():
    items = make_items(12, seed=41)
    b = D.Batch([(it.m, it.nv, it.d_B, it.d_C) for it in items])
    b.run()
    D.synchronize()
    first = [it.d_C.numpy().copy() for it in items]
    for it in items:
        it.d_C.zero()
    b.run_plain()
    D.synchronize()
    for i, it in enumerate(items):
        assert np.array_equal(bits(first[i]), bits(it.d_C.numpy())), f"item {i}"


def test_graph_replays_follow_rebound_operands():
    items = make_items(10, seed=42)
    b = D.Batch([(it.m, it.nv, it.d_B, it.d_C) for it in items])
    for rep in range(4):
        # fresh B for every item each replay, bound without re-capturing
        news = []
        for i, it in enumerate(items):
            Bn = W.dense_b(1000 + 10 * rep + i, it.a.nc, it.nv)
            d_Bn = D.DeviceArray.from_numpy(Bn)
            news.append((Bn, d_Bn))
            b.set_operands(i, d_Bn, it.d_C)
        b.graph_launch()
        D.synchronize()
        # eager work between replays
        items[0].m.spmm(news[0][1], items[0].d_C_ref, items[0].nv,
                        config=D.make_config(D.VARIANT_NNZ_TILED))
        D.synchronize()
        assert np.array_equal(bits(items[0].d_C_ref.numpy()), bits(items[0].d_C.numpy()))
        for i, it in enumerate(items):
            c64, ab, rl = it.oracle(news[i][0])
            assert_fp32(it.d_C.numpy().reshape(-1), c64, ab, rl, f"replay {rep} item {i}",
                        it.host32(news[i][0]))


def test_batched_bf16():
    items = make_items(10, seed=43, bf16=True)
    b = D.Batch([(it.m, it.nv, it.d_B, it.d_C) for it in items], btype=D.DSPMM_R_16BF)
    b.run()
    D.synchronize()
    for i, it in enumerate(items):
        it.m.spmm(it.d_B, it.d_C_ref, it.nv, btype=D.DSPMM_R_16BF, config=b.item_config(i))
    D.synchronize()
    for i, it in enumerate(items):
        out = it.d_C.numpy().reshape(-1)
        assert_bitwise(out, it.d_C_ref.numpy().reshape(-1), f"bf16 item {i} batched vs single")
        c64, ab, _ = it.oracle()
        assert_bf16(out, c64, ab, f"bf16 item {i}")


def test_benchmark_modes_run():
    items = make_items(16, seed=44)
    b = D.Batch([(it.m, it.nv, it.d_B, it.d_C) for it in items])
    for mode in (D.BATCH_ONE_LAUNCH, D.BATCH_GRAPH, D.BATCH_PLAIN, D.BATCH_PLAIN_OWN):
        ms, per = b.benchmark(mode, warmup=1, reps=3)
        note(f"mode {mode}: {ms * 1e3:.1f} us")
        assert ms > 0 and len(per) == 3


def test_batch_with_empty_and_tiny_matrices():
    """A batch holding a matrix without nonzeros, a one-row one-nonzero
    matrix and a matrix without rows beside ordinary ones: the empty one's C
    is overwritten with zeros, the tiny one is exact, and every item equals
    its single-matrix launch bit for bit."""
    rng = np.random.default_rng(45)
    mats = [W.Csr(300, 50, np.zeros(301, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32)),
            W.Csr(1, 7, np.array([0, 1], np.int32), np.array([4], np.int32),
                  np.array([-1.5], np.float32)),
            W.Csr(0, 9, np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32))]
    mats += [W.gen_csr(W.GEN_BINOMIAL, 700, 700, 4500 + k, p0=12.0) for k in range(3)]
    items = []
    for k, a in enumerate(mats):
        nv = int(rng.choice([1, 5, 16]))
        it = Item(k, a, nv, 4600 + k)
        items.append(it)
    b = D.Batch([(it.m, it.nv, it.d_B, it.d_C) for it in items])
    b.run()
    D.synchronize()
    for i, it in enumerate(items):
        it.m.spmm(it.d_B, it.d_C_ref, it.nv, config=b.item_config(i))
    D.synchronize()
    for i, it in enumerate(items):
        if it.a.nr == 0:
            continue
        out = it.d_C.numpy().reshape(-1)
        assert_bitwise(out, it.d_C_ref.numpy().reshape(-1), f"item {i} batched vs single")
        h32 = O.spmm_host32(RM, it.a.nr, it.a.nc, it.nv, it.a.rowPtr, it.a.colIdx, it.a.values,
                            it.B).reshape(-1)
        if i < 2:
            assert np.array_equal(bits(out), bits(h32)), f"item {i} vs host loop"
        else:
            c64, ab, rl = it.oracle()
            assert_fp32(out, c64, ab, rl, f"item {i}", h32)
    b.close()


def test_batched_split_rows_exact():
    """The batched kernel and its split combine on integer operands (every
    summation order exact): power-law rows up to 3000 long, a 4000-long row,
    the SpMV ladder's rows up to 70000 long, and small matrices, in one
    launch and as plain launches -- every output equals the exact product."""
    items = []
    for name, nv in (("int_zipf", 128), ("int_zipf", 32), ("int_valued", 64), ("int_spmv", 8),
                     ("int_spmv", 1)):
        c = case(name)
        it = Item.__new__(Item)
        it.a = W.Csr(c.nr, c.nc, c.rowPtr, c.colIdx, c.values)
        it.nv = nv
        it.B = c.dense(nv).reshape(c.nc, nv)
        it.m = D.Matrix(c.nr, c.nc, c.rowPtr, c.colIdx, c.values)
        it.d_B = D.DeviceArray.from_numpy(it.B)
        it.d_C = D.DeviceArray.from_numpy(np.full((c.nr, nv), np.nan, np.float32))
        items.append(it)
    rng = np.random.default_rng(7)
    for k in range(6):
        mm = int(rng.integers(300, 3000))
        a = W.gen_csr(W.GEN_BINOMIAL, mm, mm, 700 + k, p0=float(rng.integers(2, 60)))
        vals = rng.integers(-8, 9, a.nnz).astype(np.float32)
        it = Item.__new__(Item)
        it.a = W.Csr(a.nr, a.nc, a.rowPtr, a.colIdx, vals)
        it.nv = int(rng.choice([16, 32, 64, 128]))
        it.B = rng.integers(-8, 9, (a.nc, it.nv)).astype(np.float32)
        it.m = D.Matrix(a.nr, a.nc, a.rowPtr, a.colIdx, vals)
        it.d_B = D.DeviceArray.from_numpy(it.B)
        it.d_C = D.DeviceArray.from_numpy(np.full((a.nr, it.nv), np.nan, np.float32))
        items.append(it)
    b = D.Batch([(it.m, it.nv, it.d_B, it.d_C) for it in items])
    info = b.describe()
    note(f"integer batch: {info}")
    assert info["splits"] > 0, "the batch must have split rows"
    for mode, run in (("batched", b.run), ("graph", b.graph_launch), ("plain", b.run_plain)):
        for it in items:
            it.d_C.zero()
        run()
        D.synchronize()
        for i, it in enumerate(items):
            c64, _, _ = it.oracle()
            assert_exact(it.d_C.numpy().reshape(-1), c64, f"{mode} item {i} (M={it.a.nr}, "
                                                           f"nv={it.nv})")
    b.close()
