"""GPU tests of the serving loop: continuous batching returns, for every
request, exactly the result of the same request run alone; results match
the FP64 oracle on the device-generated B; a small cache evicts and stays
correct; oversized requests are rejected; plain mode gives the same bits."""
import numpy as np
import pytest

import dspmm as D
import oracle as O
import workloads as W
from gpu_common import assert_fp32, bits, note

pytestmark = pytest.mark.gpu
RM = D.DENSE_ROW_MAJOR


def make_pool(n=8, seed=60, mmax=2500):
    m, d, nvs = W.c4_params(n, seed)
    pool = []
    for i in range(n):
        mm = int(min(m[i], mmax))
        pool.append((W.gen_csr(W.GEN_BINOMIAL, mm, mm, seed * 1000 + i, p0=float(min(d[i], mm))),
                     int(nvs[i])))
    # one power-law matrix with split rows
    pool.append((W.gen_csr(W.GEN_ZIPF, 6000, 6000, seed + 7, 1.8, 1, 3000, W.VAL_NORMAL), 64))
    return pool


def register(srv, pool, preload=True):
    return [srv.add_matrix(a.nr, a.nc, a.rowPtr, a.colIdx, a.values, preload=preload)
            for a, _ in pool]


def results(srv, comps, pool, mids_of):
    out = {}
    for c in comps:
        assert c["status"] == 0, c
        a, nv = pool[mids_of[c["id"]]]
        out[c["id"]] = D.device_read(c["C"], (a.nr * nv,)).copy()
        srv.release(c["id"])
    return out


def test_server_alone_equals_in_the_mix():
    pool = make_pool()
    srv = D.Server(cache_bytes=1 << 30, operand_bytes=1 << 30, max_inflight=3, max_batch=16)
    mids = register(srv, pool)
    probe = [(1000 + k, k % len(pool), 777 + k) for k in range(len(pool))]
    alone = {}
    for rid, mi, seed in probe:                         # one at a time, server otherwise idle
        assert srv.submit(srv.request(rid, mids[mi], pool[mi][1], seed)) == 0
        srv.drain()
        comps = srv.poll()
        assert len(comps) == 1 and comps[0]["batch_size"] == 1
        alone.update(results(srv, comps, pool, {rid: mi}))
    # the same requests (same matrix, nv, seed) mixed into a burst of 120 others
    rng = np.random.default_rng(3)
    reqs, at, mi_of = [], [], {}
    for k in range(120):
        mi = int(rng.integers(len(pool)))
        reqs.append(srv.request(k, mids[mi], pool[mi][1], 5000 + k))
        at.append(k * 20e-6)
        mi_of[k] = mi
    for j, (rid, mi, seed) in enumerate(probe):
        reqs.insert(10 * j + 5, srv.request(2000 + rid, mids[mi], pool[mi][1], seed))
        at.insert(10 * j + 5, at[10 * j + 4])
        mi_of[2000 + rid] = mi
    srv.replay(reqs, at)
    comps = srv.poll()
    assert len(comps) == len(reqs)
    sizes = [c["batch_size"] for c in comps]
    note(f"batch sizes: max {max(sizes)}, mean {np.mean(sizes):.1f}; stats {srv.stats()}")
    assert max(sizes) > 1, "requests arriving together must share batches"
    got = results(srv, comps, pool, mi_of)
    for rid, mi, seed in probe:
        assert np.array_equal(bits(got[2000 + rid]), bits(alone[rid])), f"request {rid}"
    # accuracy of a few against the oracle on the same device-generated B
    for rid, mi, seed in probe[:4]:
        a, nv = pool[mi]
        B = D.seeded_uniform(seed, a.nc * nv)
        c64, ab = O.spmm_fp64(RM, a.nr, a.nc, nv, a.rowPtr, a.colIdx, a.values, B)
        rl = np.repeat(np.diff(a.rowPtr).astype(np.float64), nv)
        h32 = O.spmm_host32(RM, a.nr, a.nc, nv, a.rowPtr, a.colIdx, a.values, B).reshape(-1)
        assert_fp32(alone[rid], c64.reshape(-1), ab.reshape(-1), rl, f"served request {rid}", h32)
    srv.close()


def test_server_small_cache_evicts_and_stays_correct():
    pool = make_pool(6, seed=61)
    need = max(4 * (a.nr + 1) + 9 * a.nnz + a.nr for a, _ in pool) + (1 << 20)
    srv = D.Server(cache_bytes=2 * need + (1 << 20), operand_bytes=512 << 20, max_inflight=2,
                   max_batch=8)
    mids = register(srv, pool, preload=False)
    reqs = [srv.request(k, mids[k % len(pool)], pool[k % len(pool)][1], 90 + k)
            for k in range(3 * len(pool))]
    srv.replay(reqs, [0.0] * len(reqs))
    comps = srv.poll()
    st = srv.stats()
    note(f"stats {st}")
    assert st["evictions"] > 0 and st["completed"] == len(reqs)
    got = results(srv, comps, pool, {k: k % len(pool) for k in range(len(reqs))})
    for k in range(len(pool)):                          # the same seed always gives one answer
        a, nv = pool[k]
        B = D.seeded_uniform(90 + k, a.nc * nv)
        h32 = O.spmm_host32(RM, a.nr, a.nc, nv, a.rowPtr, a.colIdx, a.values, B).reshape(-1)
        c64, ab = O.spmm_fp64(RM, a.nr, a.nc, nv, a.rowPtr, a.colIdx, a.values, B)
        rl = np.repeat(np.diff(a.rowPtr).astype(np.float64), nv)
        assert_fp32(got[k], c64.reshape(-1), ab.reshape(-1), rl, f"evicting request {k}", h32)
    srv.close()


def test_server_rejects_what_never_fits():
    pool = make_pool(2, seed=62)
    srv = D.Server(cache_bytes=1 << 30, operand_bytes=4096, max_inflight=1, max_batch=4)
    mids = register(srv, pool)
    st = srv.submit(srv.request(1, mids[0], pool[0][1], 1))
    assert st == D.STATUS_ADMISSION_REJECTED
    assert srv.stats()["rejected"] == 1
    srv.close()


def test_plain_mode_same_bits_as_batched():
    pool = make_pool(5, seed=63)
    outs = []
    for plain in (False, True):
        srv = D.Server(cache_bytes=1 << 30, operand_bytes=1 << 30, max_inflight=3, max_batch=16,
                       plain=plain)
        mids = register(srv, pool)
        reqs = [srv.request(k, mids[k % len(pool)], pool[k % len(pool)][1], 300 + k)
                for k in range(2 * len(pool))]
        srv.replay(reqs, [0.0] * len(reqs))
        outs.append(results(srv, srv.poll(), pool, {k: k % len(pool) for k in range(len(reqs))}))
        srv.close()
    for k in outs[0]:
        assert np.array_equal(bits(outs[0][k]), bits(outs[1][k])), f"request {k}"


def test_plain_own_best_mode_is_correct():
    """plain = 2 launches each request with its matrix's own AUTO kernel:
    results within the FP32 gate of the oracle on the device-generated B."""
    pool = make_pool(4, seed=64)
    srv = D.Server(cache_bytes=1 << 30, operand_bytes=1 << 30, max_inflight=3, max_batch=16,
                   plain=2)
    mids = register(srv, pool)
    reqs = [srv.request(k, mids[k % len(pool)], pool[k % len(pool)][1], 700 + k)
            for k in range(2 * len(pool))]
    srv.replay(reqs, [0.0] * len(reqs))
    got = results(srv, srv.poll(), pool, {k: k % len(pool) for k in range(len(reqs))})
    for k in range(len(reqs)):
        a, nv = pool[k % len(pool)]
        B = D.seeded_uniform(700 + k, a.nc * nv)
        c64, ab = O.spmm_fp64(RM, a.nr, a.nc, nv, a.rowPtr, a.colIdx, a.values, B)
        rl = np.repeat(np.diff(a.rowPtr).astype(np.float64), nv)
        h32 = O.spmm_host32(RM, a.nr, a.nc, nv, a.rowPtr, a.colIdx, a.values, B).reshape(-1)
        assert_fp32(got[k], c64.reshape(-1), ab.reshape(-1), rl, f"plain-own request {k}", h32)
    srv.close()


def test_large_requests_run_alone_with_their_own_kernel():
    """Requests on matrices above alone_nnz launch one at a time with the
    matrix's AUTO kernel and a workspace of their own; mixed with batched
    small requests, every result is within the FP32 gate, and a request
    gives the same bits each time it runs alone."""
    pool = make_pool(5, seed=65)                       # the last entry is the split-row Zipf matrix
    threshold = pool[-1][0].nnz
    alone = {mi for mi, (a, _) in enumerate(pool) if a.nnz >= threshold}
    srv = D.Server(cache_bytes=1 << 30, operand_bytes=1 << 30, max_inflight=3, max_batch=16,
                   alone_nnz=threshold)
    mids = register(srv, pool)
    reqs, mi_of = [], {}
    for k in range(30):
        mi = k % len(pool)
        reqs.append(srv.request(k, mids[mi], pool[mi][1], 900 + (k if mi != len(pool) - 1 else 0)))
        mi_of[k] = mi
    srv.replay(reqs, [0.0] * len(reqs))
    comps = srv.poll()
    assert len(alone) < len(pool)
    assert all(c["batch_size"] == 1 for c in comps if mi_of[c["id"]] in alone)
    assert max(c["batch_size"] for c in comps) > 1
    got = results(srv, comps, pool, mi_of)
    big_ids = [k for k in mi_of if mi_of[k] == len(pool) - 1]
    for k in big_ids[1:]:                              # same seed 900: same bits every time
        assert np.array_equal(bits(got[k]), bits(got[big_ids[0]])), f"request {k}"
    for k in (0, 1, big_ids[0]):
        a, nv = pool[mi_of[k]]
        seed = 900 + (k if mi_of[k] != len(pool) - 1 else 0)
        B = D.seeded_uniform(seed, a.nc * nv)
        c64, ab = O.spmm_fp64(RM, a.nr, a.nc, nv, a.rowPtr, a.colIdx, a.values, B)
        rl = np.repeat(np.diff(a.rowPtr).astype(np.float64), nv)
        h32 = O.spmm_host32(RM, a.nr, a.nc, nv, a.rowPtr, a.colIdx, a.values, B).reshape(-1)
        assert_fp32(got[k], c64.reshape(-1), ab.reshape(-1), rl, f"mixed request {k}", h32)
    srv.close()


def test_server_bf16_operands():
    """BF16-stored B (generated on the device, rounded to nearest even): each
    result is the same alone and in a burst, within the FP32 gate of the FP64
    product on the rounded B, and within 1e-2 of the FP64 product on the
    FP32 B."""
    pool = make_pool(5, seed=64)
    srv = D.Server(cache_bytes=1 << 30, operand_bytes=1 << 30, max_inflight=3, max_batch=16,
                   btype=D.DSPMM_R_16BF)
    mids = register(srv, pool)
    alone = {}
    for k in range(len(pool)):
        assert srv.submit(srv.request(k, mids[k], pool[k][1], 500 + k)) == 0
        srv.drain()
        alone.update(results(srv, srv.poll(), pool, {k: k}))
    reqs = [srv.request(100 + k, mids[k % len(pool)], pool[k % len(pool)][1], 500 + k % len(pool))
            for k in range(4 * len(pool))]
    srv.replay(reqs, [0.0] * len(reqs))
    comps = srv.poll()
    assert max(c["batch_size"] for c in comps) > 1
    got = results(srv, comps, pool, {100 + k: k % len(pool) for k in range(len(reqs))})
    for k in range(len(reqs)):
        assert np.array_equal(bits(got[100 + k]), bits(alone[k % len(pool)])), f"request {k}"
    for k, (a, nv) in enumerate(pool):
        B = D.seeded_uniform(500 + k, a.nc * nv)
        rl = np.repeat(np.diff(a.rowPtr).astype(np.float64), nv)
        Br = O.round_bf16(B)
        c64r, abr = O.spmm_fp64(RM, a.nr, a.nc, nv, a.rowPtr, a.colIdx, a.values, Br)
        h32r = O.spmm_host32(RM, a.nr, a.nc, nv, a.rowPtr, a.colIdx, a.values, Br).reshape(-1)
        assert_fp32(alone[k], c64r.reshape(-1), abr.reshape(-1), rl, f"bf16 request {k}", h32r)
        c64, ab = O.spmm_fp64(RM, a.nr, a.nc, nv, a.rowPtr, a.colIdx, a.values, B)
        err = np.abs(alone[k].astype(np.float64) - c64.reshape(-1))
        assert np.all(err <= 1e-2 * ab.reshape(-1)), f"bf16 request {k}: beyond 1e-2"
    srv.close()


@pytest.mark.parametrize("bf16", [False, True], ids=["fp32", "bf16"])
def test_server_serves_a_stencil_on_its_ring(bf16):
    """A 36^3 27-point stencil beside random matrices: its own AUTO kernel is
    the B-row ring (plan built at admission), requests of it alone or mixed
    equal the host loop bit for bit on the device-generated B (with BF16-
    stored B: on B rounded to nearest even), and in a burst they share
    batches with the other matrices."""
    btype = D.DSPMM_R_16BF if bf16 else D.DSPMM_R_32F
    pool = make_pool(n=4, seed=61)
    pool.append((W.c3_matrix(36), 64))
    srv = D.Server(cache_bytes=1 << 30, operand_bytes=1 << 30, max_inflight=3, max_batch=16,
                   btype=btype)
    mids = register(srv, pool)
    si = len(pool) - 1
    a, nv = pool[si]
    m = D.Matrix(a.nr, a.nc, a.rowPtr, a.colIdx, a.values)
    assert m.resolve(nv, btype=btype).variant == D.VARIANT_RING
    m.close()
    reqs, at, mi_of = [], [], {}
    rng = np.random.default_rng(4)
    for k in range(40):
        mi = si if k % 4 == 0 else int(rng.integers(len(pool) - 1))
        reqs.append(srv.request(k, mids[mi], pool[mi][1], 900 + k))
        at.append(k * 10e-6)
        mi_of[k] = mi
    srv.replay(reqs, at)
    comps = srv.poll()
    assert len(comps) == len(reqs)
    got = results(srv, comps, pool, mi_of)
    for k in range(0, 40, 4):
        B = D.seeded_uniform(900 + k, a.nc * nv)
        if bf16:
            B = O.round_bf16(B)
        h32 = O.spmm_host32(RM, a.nr, a.nc, nv, a.rowPtr, a.colIdx, a.values, B).reshape(-1)
        assert np.array_equal(bits(got[k]), bits(h32)), f"stencil request {k}"
    note(f"stencil in the mix: stats {srv.stats()}")
    srv.close()
