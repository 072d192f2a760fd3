"""GPU parity at the full sizes of the measured workloads, through properties
that do not need the oracle to run the whole product:

* kernels that each sum every row in nonzero order (row-exact) must agree bit
  for bit on every output of the full matrix;
* a sample of rows equals the host loop (the oracle run on the sampled rows'
  sub-matrix) bit for bit, and the FP64 oracle within the FP32 gate;
* the batched launch over all 512 matrices of the batch workload equals one
  plain launch per matrix bit for bit;
* repeat calls give the same bits.
"""
import numpy as np
import pytest

import dspmm as D
import oracle as O
import workloads as W
from gpu_common import assert_bitwise, assert_fp32, bits, note

pytestmark = pytest.mark.gpu
RM = D.DENSE_ROW_MAJOR


def sample(a, count, seed):
    """Sub-matrix of `count` random rows: (rows, rowPtr, colIdx, values)."""
    rng = np.random.default_rng(seed)
    rows = np.sort(rng.choice(a.nr, size=min(count, a.nr), replace=False))
    lens = (a.rowPtr[rows + 1] - a.rowPtr[rows]).astype(np.int64)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    idx = np.repeat(a.rowPtr[rows].astype(np.int64) - rp[:-1], lens) + np.arange(rp[-1])
    return rows, rp, a.colIdx[idx], a.values[idx]


def check_sampled_rows(a, sub, B, nv, out, what, exact=True):
    rows, rp, ci, v = sub
    h32 = O.spmm_host32(RM, rows.size, a.nc, nv, rp, ci, v, B).reshape(-1)
    got = np.ascontiguousarray(out.reshape(a.nr, nv)[rows]).reshape(-1)
    if exact:
        assert_bitwise(got, h32, what + " (sampled rows vs host loop)")
    c64, ab = O.spmm_fp64(RM, rows.size, a.nc, nv, rp, ci, v, B)
    rl = np.repeat(np.diff(rp).astype(np.float64), nv)
    assert_fp32(got, c64.reshape(-1), ab.reshape(-1), rl, what + " (sampled rows vs FP64)", h32)


def fill_nan(d):
    host = np.full(d.nbytes // 4, np.nan, np.float32)
    D.check(D.lib().dspmm_memcpy_h2d(d.ptr, host.ctypes.data, host.nbytes), "h2d")


def run(m, d_B, d_C, nv, cfg, btype=D.DSPMM_R_32F):
    fill_nan(d_C)
    m.spmm(d_B, d_C, nv, btype=btype, config=cfg)
    D.synchronize()
    return d_C.numpy().reshape(-1).copy()


def test_stencil_100_cubed_row_exact_kernels_agree():
    """27-point stencil on a 100^3 grid (M = 1e6, 26.5 M nonzeros), N = 64:
    eight row-exact kernels -- the headline B-row ring among them, at the
    tile shapes and walk lengths the bench runs -- give the same 64 M
    outputs; sampled rows equal the host loop; the same with BF16-stored B
    (five kernels)."""
    a = W.c3_matrix(100)
    nv = 64
    B = W.dense_b(3, a.nc, nv)
    m = D.Matrix(a.nr, a.nc, a.rowPtr, a.colIdx, a.values)
    d_B = D.DeviceArray.from_numpy(B)
    d_C = D.DeviceArray.zeros((a.nr, nv))
    cfgs = {"reference-class SR_RB": D.make_config(D.VARIANT_CSR_SCALAR),
            "row-sequential": D.make_config(D.VARIANT_CSR_SCALAR, unroll=4),
            "nnz-tiled S=64": D.make_config(D.VARIANT_NNZ_TILED, tile_nnz=64),
            "row-blocked R=2": D.make_config(D.VARIANT_ROW_BLOCKED, tile_nnz=2, unroll=4),
            "row-blocked R=4, rows a plane apart":
                D.make_config(D.VARIANT_ROW_BLOCKED, tile_nnz=4, unroll=4, row_stride=10000),
            "ring 4x8, automatic walks": D.make_config(D.VARIANT_RING, tile_nnz=4, unroll=8),
            "ring 8x4, automatic walks": D.make_config(D.VARIANT_RING, tile_nnz=8, unroll=4),
            "ring 8x8, walks of 17":
                D.make_config(D.VARIANT_RING, tile_nnz=8, unroll=8, row_stride=17)}
    first = None
    for name, cfg in cfgs.items():
        note(f"c3 full {name}")
        out = run(m, d_B, d_C, nv, cfg)
        if first is None:
            first = out
            check_sampled_rows(a, sample(a, 3000, 31), B, nv, out, f"c3 {name}")
        else:
            assert np.array_equal(bits(out), bits(first)), f"c3 {name} differs from SR_RB"
    again = run(m, d_B, d_C, nv, cfgs["row-sequential"])
    assert np.array_equal(bits(again), bits(first)), "repeat call differs"
    d16 = D.to_bf16_device(d_B)
    outs = [run(m, d16, d_C, nv, cfg, D.DSPMM_R_16BF) for cfg in
            (D.make_config(D.VARIANT_CSR_SCALAR, unroll=4),
             D.make_config(D.VARIANT_NNZ_TILED, vec=8, lanes=8, tile_nnz=32, unroll=4),
             D.make_config(D.VARIANT_ROW_BLOCKED, tile_nnz=2, unroll=4, row_stride=10000),
             D.make_config(D.VARIANT_RING, tile_nnz=8, unroll=8, row_stride=17),
             D.make_config(D.VARIANT_RING, tile_nnz=4, unroll=8))]
    for o in outs[1:]:
        assert np.array_equal(bits(outs[0]), bits(o)), "c3 BF16: row kernels differ"
    rows, rp, ci, v = sample(a, 3000, 32)
    h32 = O.spmm_host32(RM, rows.size, a.nc, nv, rp, ci, v, O.round_bf16(B)).reshape(-1)
    got = np.ascontiguousarray(outs[0].reshape(a.nr, nv)[rows]).reshape(-1)
    assert_bitwise(got, h32, "c3 BF16 sampled rows vs host loop on the rounded B")
    m.close()


def test_large_power_law_row_exact_kernels_agree():
    """The served large matrix (M = K = 2e6, power-law rows, ~40 M nonzeros),
    N = 64: the reference-class and row-sequential kernels agree on every
    output and equal the host loop on sampled rows; the AUTO kernel (which
    splits long rows) is within the FP32 gate on the same rows."""
    a = W.c5_large_matrix()
    nv = 64
    note(f"c5 large: M={a.nr} nnz={a.nnz}")
    B = W.dense_b(5, a.nc, nv)
    m = D.Matrix(a.nr, a.nc, a.rowPtr, a.colIdx, a.values)
    d_B = D.DeviceArray.from_numpy(B)
    d_C = D.DeviceArray.zeros((a.nr, nv))
    ref = run(m, d_B, d_C, nv, D.make_config(D.VARIANT_CSR_SCALAR))
    seq = run(m, d_B, d_C, nv, D.make_config(D.VARIANT_CSR_SCALAR, unroll=8))
    assert np.array_equal(bits(ref), bits(seq)), "c5 large: row kernels differ"
    sub = sample(a, 2000, 51)
    check_sampled_rows(a, sub, B, nv, ref, "c5 large SR_RB")
    auto = run(m, d_B, d_C, nv, None)
    check_sampled_rows(a, sub, B, nv, auto, "c5 large AUTO", exact=False)
    m.close()


def test_batch_of_512_equals_plain_launches():
    """All 512 matrices of the batch workload (M = K in [256, 16384], N in
    {16 .. 256}): one batched launch equals one plain launch per matrix with
    the same configuration, bit for bit; a sample of matrices is checked
    against the FP64 oracle."""
    count = 512
    ms, ds, ns = W.c4_params(count)
    items, keep = [], []
    for i in range(count):
        a = W.c4_matrix(i, ms[i], ds[i])
        nv = int(ns[i])
        mat = D.Matrix(a.nr, a.nc, a.rowPtr, a.colIdx, a.values)
        B = W.dense_b(4000 + i, a.nc, nv)
        d_B = D.DeviceArray.from_numpy(B)
        d_C = D.DeviceArray.zeros((a.nr, nv))
        items.append((mat, nv, d_B, d_C))
        keep.append((a, B))
    b = D.Batch(items)
    note(f"c4 batch: {b.describe()}")
    b.run()
    D.synchronize()
    batched = [it[3].numpy().reshape(-1).copy() for it in items]
    for it in items:
        fill_nan(it[3])
    b.run_plain()
    D.synchronize()
    for i, it in enumerate(items):
        assert np.array_equal(bits(batched[i]), bits(it[3].numpy().reshape(-1))), f"matrix {i}"
    for i in np.random.default_rng(7).choice(count, 6, replace=False):
        a, B = keep[i]
        nv = items[i][1]
        c64, ab = O.spmm_fp64(RM, a.nr, a.nc, nv, a.rowPtr, a.colIdx, a.values, B)
        rl = np.repeat(np.diff(a.rowPtr).astype(np.float64), nv)
        h32 = O.spmm_host32(RM, a.nr, a.nc, nv, a.rowPtr, a.colIdx, a.values, B).reshape(-1)
        assert_fp32(batched[i], c64.reshape(-1), ab.reshape(-1), rl, f"c4 matrix {i}", h32)
    b.close()
    for it in items:
        it[0].close()


@pytest.mark.parametrize("btype", [D.DSPMM_R_32F, D.DSPMM_R_16BF])
def test_server_serves_the_large_matrix_alone_and_in_a_mix(btype):
    """The serving loop on the large power-law matrix (M = K = 2e6, ~40 M
    nonzeros, N = 64; it runs alone, with its own AUTO kernel and a
    workspace of its own): once with the server otherwise idle, once in a
    burst among small batched requests.  Both give the same bits (same
    seed), and sampled rows are within the FP32 gate of the FP64 product on
    the device-generated B (rounded to BF16 for BF16 storage); the small
    requests of the mix equal the same requests served alone."""
    a = W.c5_large_matrix(seed=500)
    nv = 64
    small = [(W.gen_csr(W.GEN_BINOMIAL, 3000, 3000, 7100 + k, p0=12.0), 32) for k in range(6)]
    pool = [(a, nv)] + small
    srv = D.Server(cache_bytes=8 << 30, operand_bytes=8 << 30, max_inflight=3, max_batch=64,
                   btype=btype)
    mids = [srv.add_matrix(x.nr, x.nc, x.rowPtr, x.colIdx, x.values, preload=True) for x, _ in pool]

    def served(comps, of):
        out = {}
        for c in comps:
            assert c["status"] == 0, c
            x, n = pool[of[c["id"]]]
            out[c["id"]] = D.device_read(c["C"], (x.nr * n,)).copy()
            srv.release(c["id"])
        return out

    note(f"large alone (btype {btype})")
    assert srv.submit(srv.request(1, mids[0], nv, 4242)) == 0
    srv.drain()
    comps = srv.poll()
    assert len(comps) == 1 and comps[0]["batch_size"] == 1
    alone = served(comps, {1: 0})[1]
    lone_small = {}
    for k in range(len(small)):
        assert srv.submit(srv.request(100 + k, mids[1 + k], 32, 600 + k)) == 0
        srv.drain()
        lone_small.update(served(srv.poll(), {100 + k: 1 + k}))
    note("large in a mix")
    reqs, of = [], {}
    for k in range(40):
        mi = 1 + k % len(small)
        reqs.append(srv.request(1000 + k, mids[mi], 32, 600 + (mi - 1)))
        of[1000 + k] = mi
        if k == 20:
            reqs.append(srv.request(2, mids[0], nv, 4242))
            of[2] = 0
    srv.replay(reqs, [0.0] * len(reqs))
    comps = srv.poll()
    assert len(comps) == len(reqs)
    assert max(c["batch_size"] for c in comps) > 1
    got = served(comps, of)
    assert np.array_equal(bits(got[2]), bits(alone)), "large request: mix vs alone"
    for rid, mi in of.items():
        if mi > 0:
            assert np.array_equal(bits(got[rid]), bits(lone_small[100 + mi - 1])), f"small {rid}"
    B = D.seeded_uniform(4242, a.nc * nv)
    if btype == D.DSPMM_R_16BF:
        B = O.round_bf16(B)
    check_sampled_rows(a, sample(a, 2000, 52), B, nv, alone, f"served large btype {btype}",
                       exact=False)
    srv.close()
