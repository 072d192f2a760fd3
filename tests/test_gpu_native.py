"""GPU parity of the native path (dspmm_spmm_ex / Matrix): the nnz-tiled
kernel over tile shapes, the reference algorithm classes reached through
the native API, any nv, leading dimensions, BF16-stored B, the autotuner
and the caller-provided workspace."""
import numpy as np
import pytest

import dspmm as D
import oracle as O
import workloads as W
from gpu_common import assert_bf16, assert_bitwise, assert_exact, assert_fp32, bits, case, note

pytestmark = pytest.mark.gpu

RM, CM = D.DENSE_ROW_MAJOR, D.DENSE_COL_MAJOR
_MATS = {}


def matrix(name):
    if name not in _MATS:
        c = case(name)
        _MATS[name] = (c, D.Matrix(c.nr, c.nc, c.rowPtr, c.colIdx, c.values))
    return _MATS[name]


def split_rows(c, S):
    return c.rowlen > S


def run_native(c, m, nv, layout=RM, config=None, btype=D.DSPMM_R_32F, B=None):
    c.fill_output(np.nan)
    m.spmm(B if B is not None else c.d_B, c.d_C, nv, layout=layout, config=config, btype=btype)
    D.synchronize()
    return c.output(nv)


@pytest.mark.parametrize("name", ["c1", "zipf_mid", "edge", "stencil_20"])
@pytest.mark.parametrize("S,unroll", [(16, 4), (64, 8), (256, 8), (1024, 4)])
def test_nnz_tiled_tile_shapes(name, S, unroll):
    c, m = matrix(name)
    for nv in (1, 2, 3, 8, 32, 64):
        if nv > c.maxnv:
            continue
        cfg = D.make_config(D.VARIANT_NNZ_TILED, tile_nnz=S, unroll=unroll)
        what = f"{name} nnz_tiled S={S} U={unroll} nv={nv}"
        note(what)
        try:
            out = run_native(c, m, nv, config=cfg)
        except D.DspmmError as e:
            assert e.status == D.STATUS_NOT_SUPPORTED   # LDS budget for tiny lanes x big S
            continue
        h32, c64, ab = c.ref(RM, nv)
        split = np.repeat(split_rows(c, S), nv)
        # rows that fit a tile are summed exactly like the host loop
        assert_bitwise(out[~split], h32[~split], what + " (unsplit rows)")
        assert_fp32(out, c64, ab, c.rowlen_flat(RM, nv), what, h32)


@pytest.mark.parametrize("name", ["int_valued", "int_zipf"])
@pytest.mark.parametrize("S", [16, 64])
def test_split_rows_exact(name, S):
    """Rows longer than a tile are cut into chunks whose partial sums the
    split combine adds in its own fixed order.  Integer operands make every
    summation order exact, so every output -- split rows included -- must
    equal the exact product bit for bit (a dropped, doubled or misplaced
    chunk shows)."""
    c, m = matrix(name)
    assert (c.rowlen > S).sum() > 0, "the case must have split rows"
    for nv in (1, 32, 64, 128):
        if nv > c.maxnv:
            continue
        _, c64, _ = c.ref(RM, nv)
        for vec, lanes in ((0, 0), (1, 64), (4, 8), (2, 16)):
            cfg = D.make_config(D.VARIANT_NNZ_TILED, vec=vec, lanes=lanes, tile_nnz=S)
            what = f"{name} split S={S} vec={vec} lanes={lanes} nv={nv}"
            note(what)
            try:
                out = run_native(c, m, nv, config=cfg)
            except D.DspmmError as e:
                assert e.status == D.STATUS_NOT_SUPPORTED, what      # LDS budget
                continue
            assert_exact(out, c64, what)
    # BF16-stored B: small integers are exact in BF16, so the same holds
    nv = min(64, c.maxnv)
    _, c64, _ = c.ref(RM, nv)
    d16 = D.to_bf16_device(c.d_B)
    out = run_native(c, m, nv, config=D.make_config(D.VARIANT_NNZ_TILED, tile_nnz=S),
                     btype=D.DSPMM_R_16BF, B=d16)
    assert_exact(out, c64, f"{name} split S={S} bf16 nv={nv}")


@pytest.mark.parametrize("T", [512, 1024, 2048, 4096])
def test_spmv_tiled_exact(T):
    """Every row class of the SpMV plan on integer operands: rows of <= 16
    (one thread), 17..256 (the wavefront queue), 257..16 T (a tile of their
    own, tree-reduced) and > 16 T nonzeros (carry chunks and the combine) --
    all exact."""
    c, m = matrix("int_spmv")
    n = c.rowlen
    for lo, hi in ((0, 16), (17, 256), (257, 16 * T), (16 * T + 1, 1 << 30)):
        assert ((n >= lo) & (n <= hi)).any(), f"no row of {lo}..{hi} nonzeros"
    _, c64, _ = c.ref(RM, 1)
    for layout in (RM, CM):
        what = f"int_spmv spmv T={T} layout={layout}"
        note(what)
        out = run_native(c, m, 1, layout=layout, config=D.make_config(D.VARIANT_SPMV_TILED,
                                                                       tile_nnz=T))
        assert_exact(out, c64, what)
    for name in ("int_valued", "int_zipf"):
        ci, mi = matrix(name)
        _, c64, _ = ci.ref(RM, 1)
        out = run_native(ci, mi, 1, config=D.make_config(D.VARIANT_SPMV_TILED, tile_nnz=T))
        assert_exact(out, c64, f"{name} spmv T={T}")


@pytest.mark.parametrize("vec", [1, 2, 4])
def test_nnz_tiled_vector_widths_and_lanes(vec):
    c, m = matrix("zipf_mid")
    for nv in (4, 12, 128):
        for lanes in (0, 64):
            cfg = D.make_config(D.VARIANT_NNZ_TILED, vec=vec, lanes=lanes, tile_nnz=16)
            what = f"zipf_mid vec={vec} lanes={lanes} nv={nv}"
            note(what)
            out = run_native(c, m, nv, config=cfg)
            h32, c64, ab = c.ref(RM, nv)
            split = np.repeat(split_rows(c, 16), nv)
            assert_bitwise(out[~split], h32[~split], what)
            assert_fp32(out, c64, ab, c.rowlen_flat(RM, nv), what, h32)


@pytest.mark.parametrize("variant", [D.VARIANT_CSR_SCALAR, D.VARIANT_CSR_VECTOR,
                                     D.VARIANT_COO_SCALAR, D.VARIANT_COO_VECTOR])
@pytest.mark.parametrize("layout", [RM, CM], ids=["RM", "CM"])
def test_reference_classes_through_native_api(variant, layout):
    """Any nv (not only powers of two) and overwrite semantics for all."""
    c, m = matrix("zipf_mid")
    for nv in (1, 3, 5, 33, 100):
        what = f"zipf_mid {D.VARIANT_NAMES[variant]} layout={layout} nv={nv}"
        note(what)
        out = run_native(c, m, nv, layout=layout, config=D.make_config(variant))
        h32, c64, ab = c.ref(layout, nv)
        if variant == D.VARIANT_CSR_SCALAR:
            assert_bitwise(out, h32, what)
        else:
            assert_fp32(out, c64, ab, c.rowlen_flat(layout, nv), what, h32)


def test_leading_dimensions():
    c, m = matrix("zipf_mid")
    nv, ldb, ldc = 24, 40, 36
    B = W.gen_dense(77, c.nc * ldb).reshape(c.nc, ldb)
    d_B = D.DeviceArray.from_numpy(B)
    d_C = D.DeviceArray.from_numpy(np.full((c.nr, ldc), -3.0, np.float32))
    for variant in (D.VARIANT_NNZ_TILED, D.VARIANT_CSR_SCALAR, D.VARIANT_COO_VECTOR):
        note(f"ld variant={variant}")
        d_C = D.DeviceArray.from_numpy(np.full((c.nr, ldc), -3.0, np.float32))
        m.spmm(d_B, d_C, nv, ldb=ldb, ldc=ldc, config=D.make_config(variant))
        D.synchronize()
        out = d_C.numpy()
        assert np.all(out[:, nv:] == -3.0), "columns beyond nv must be left alone"
        dense = np.ascontiguousarray(B[:, :nv])
        c64, ab = O.spmm_fp64(RM, c.nr, c.nc, nv, c.rowPtr, c.colIdx, c.values, dense)
        h32 = O.spmm_host32(RM, c.nr, c.nc, nv, c.rowPtr, c.colIdx, c.values, dense)
        assert_fp32(np.ascontiguousarray(out[:, :nv]).reshape(-1), c64.reshape(-1),
                    ab.reshape(-1), np.repeat(c.rowlen, nv), "ld", h32.reshape(-1))


@pytest.mark.parametrize("name", ["zipf_mid", "stencil_20", "edge"])
def test_bf16_dense_input(name):
    c, m = matrix(name)
    nv = 64 if c.maxnv >= 64 else c.maxnv
    d_B16 = D.to_bf16_device(c.d_B)
    dense = c.dense(nv)
    rounded = O.round_bf16(dense)
    h32r = O.spmm_host32(RM, c.nr, c.nc, nv, c.rowPtr, c.colIdx, c.values, rounded).reshape(-1)
    c64, ab = O.spmm_fp64(RM, c.nr, c.nc, nv, c.rowPtr, c.colIdx, c.values, dense)
    for S in (32, 128):
        what = f"{name} bf16 S={S}"
        note(what)
        out = run_native(c, m, nv, config=D.make_config(D.VARIANT_NNZ_TILED, tile_nnz=S),
                         btype=D.DSPMM_R_16BF, B=d_B16)
        split = np.repeat(split_rows(c, S), nv)
        # = the host loop on the BF16-rounded B, bit for bit, on unsplit rows
        assert_bitwise(out[~split], h32r[~split], what)
        assert_bf16(out, c64.reshape(-1), ab.reshape(-1), what)


def test_device_conversion_matches_rne():
    c, _ = matrix("c1")
    d16 = D.to_bf16_device(c.d_B)
    got = d16.numpy().astype(np.uint32) << 16
    want = O.round_bf16(c.B).view(np.uint32)
    np.testing.assert_array_equal(got, want)


def test_autotune_then_auto_is_correct():
    c, m = matrix("zipf_mid")
    for nv, layout in ((32, RM), (1, RM), (8, CM)):
        note(f"autotune nv={nv} layout={layout}")
        best, ms = m.autotune(c.d_B, c.d_C, nv, layout=layout, reps=3)
        assert ms > 0 and best.variant != D.VARIANT_AUTO
        resolved = m.resolve(nv, layout=layout)
        assert resolved.as_dict() == best.as_dict()
        out = run_native(c, m, nv, layout=layout)
        h32, c64, ab = c.ref(layout, nv)
        assert_fp32(out, c64, ab, c.rowlen_flat(layout, nv), f"auto {best.as_dict()}", h32)


def test_caller_workspace_and_deterministic_repeat():
    c, m = matrix("edge")               # has split rows
    nv = 16
    cfg = D.make_config(D.VARIANT_NNZ_TILED, tile_nnz=64)
    need = m.workspace_size(nv, config=cfg)
    assert need > 0
    ws = D.DeviceArray(need, np.uint8)
    outs = []
    for _ in range(3):
        c.fill_output(np.nan)
        m.spmm(c.d_B, c.d_C, nv, config=cfg, workspace=ws, workspace_bytes=need)
        D.synchronize()
        outs.append(c.output(nv).copy())
    assert all(np.array_equal(bits(outs[0]), bits(o)) for o in outs[1:])
    with pytest.raises(D.DspmmError) as e:
        m.spmm(c.d_B, c.d_C, nv, config=cfg, workspace=ws, workspace_bytes=need - 4)
    assert e.value.status == 5


def test_c2_native_auto_all_n():
    c, m = matrix("c2")
    for nv in (1, 32, 128):
        note(f"c2 native nv={nv}")
        out = run_native(c, m, nv)
        h32, c64, ab = c.ref(RM, nv)
        assert_fp32(out, c64, ab, c.rowlen_flat(RM, nv), f"c2 native nv={nv}", h32)


def test_matrix_from_device_arrays_and_describe():
    c = case("zipf_mid")
    import ctypes as C
    h = C.c_void_p()
    D.check(D.lib().dspmm_matrix_create_device(C.byref(h), c.nr, c.nc, c.nnz, c.d_rp.ptr, 0,
                                               c.d_ci.ptr, c.d_v.ptr, 0, None), "create_device")
    d = D.MatrixDesc()
    D.check(D.lib().dspmm_matrix_describe(h.value, C.byref(d)), "describe")
    assert (d.nr, d.nc, d.nnz) == (c.nr, c.nc, c.nnz)
    assert d.max_row_len == int(c.rowlen.max())
    assert abs(d.mean_row_len - c.rowlen.mean()) < 1e-9
    assert d.hbm_bytes >= 4 * (c.nr + 1 + 2 * c.nnz)
    D.lib().dspmm_matrix_destroy(h.value)


@pytest.mark.parametrize("name", ["c1", "zipf_mid", "edge", "stencil_20", "c2"])
@pytest.mark.parametrize("T", [512, 1024, 2048, 4096])
def test_spmv_tiled(name, T):
    """SpMV kernel: rows of <= 16 nonzeros are summed like the host loop (bit
    for bit); longer rows are tree-reduced (and combined across tiles)."""
    c, m = matrix(name)
    h32, c64, ab = c.ref(RM, 1)
    cfg = D.make_config(D.VARIANT_SPMV_TILED, tile_nnz=T)
    for layout in (RM, CM):
        what = f"{name} spmv T={T} layout={layout}"
        note(what)
        out = run_native(c, m, 1, layout=layout, config=cfg)
        short = c.rowlen <= 16
        assert_bitwise(out[short], h32[short], what + " (rows <= 16)")
        assert_fp32(out, c64, ab, c.rowlen, what, h32)


def test_spmv_strided_and_bf16():
    c, m = matrix("zipf_mid")
    ld = 5
    X = W.gen_dense(91, c.nc * ld).reshape(c.nc, ld)
    d_X = D.DeviceArray.from_numpy(X)
    d_Y = D.DeviceArray.from_numpy(np.full((c.nr, 3), -1.0, np.float32))
    m.spmm(d_X, d_Y, 1, ldb=ld, ldc=3, config=D.make_config(D.VARIANT_SPMV_TILED))
    D.synchronize()
    y = d_Y.numpy()
    assert np.all(y[:, 1:] == -1.0)
    x0 = np.ascontiguousarray(X[:, 0])
    h32 = O.spmm_host32(RM, c.nr, c.nc, 1, c.rowPtr, c.colIdx, c.values, x0).reshape(-1)
    short = c.rowlen <= 16
    assert_bitwise(y[:, 0][short], h32[short], "strided spmv")
    d16 = D.to_bf16_device(c.d_B)
    out = run_native(c, m, 1, config=D.make_config(D.VARIANT_SPMV_TILED), btype=D.DSPMM_R_16BF,
                     B=d16)
    xr = O.round_bf16(c.dense(1))
    h32r = O.spmm_host32(RM, c.nr, c.nc, 1, c.rowPtr, c.colIdx, c.values, xr).reshape(-1)
    assert_bitwise(out[short], h32r[short], "bf16 spmv")


@pytest.mark.parametrize("name", ["stencil_20", "c1", "edge", "int_valued"])
@pytest.mark.parametrize("umax", [64, 128, 256])
def test_nnz_cached(name, umax):
    """Column-cached tiles: every row whose columns fit one tile is summed
    like the host loop (bit for bit); rows that become direct tiles (more
    distinct columns than the cache, or more nonzeros than a tile) are within
    the FP32 gate; integer operands are exact everywhere."""
    c, m = matrix(name)
    for nv in (2, 8, 16, 64):
        if nv > c.maxnv:
            continue
        h32, c64, ab = c.ref(RM, nv)
        what = f"{name} cached umax={umax} nv={nv}"
        note(what)
        try:
            out = run_native(c, m, nv, config=D.make_config(D.VARIANT_NNZ_CACHED, tile_nnz=umax))
        except D.DspmmError as e:
            assert e.status == D.STATUS_NOT_SUPPORTED, what
            note(f"{what}: not supported (LDS)")
            continue
        fits = np.repeat(c.rowlen <= umax, nv)
        assert_bitwise(out[fits], h32[fits], what + " (rows that fit a tile)")
        if name == "int_valued":
            assert np.array_equal(out, c64.astype(np.float32)), what
        assert_fp32(out, c64, ab, np.repeat(c.rowlen, nv), what, h32)


def test_nnz_cached_bf16_and_autotune():
    c, m = matrix("stencil_20")
    nv = 64
    d16 = D.to_bf16_device(c.d_B)
    out = run_native(c, m, nv, config=D.make_config(D.VARIANT_NNZ_CACHED, tile_nnz=128),
                     btype=D.DSPMM_R_16BF, B=d16)
    br = O.round_bf16(c.dense(nv))
    h32r = O.spmm_host32(RM, c.nr, c.nc, nv, c.rowPtr, c.colIdx, c.values, br).reshape(-1)
    assert_bitwise(out, h32r, "stencil_20 cached bf16")
    # the stencil's neighbouring rows share columns: the autotuner considers the cached kernel
    tab = m.autotune_table(c.d_B, c.d_C, nv, reps=2)
    names = {cfg.as_dict()["variant"] for cfg, ms in tab if ms is not None}
    assert "nnz_cached" in names and "nnz_tiled" in names
    note(f"stencil_20 nv=64 table: {len(tab)} candidates; best "
         f"{min((ms, cfg.as_dict()['variant'], cfg.tile_nnz) for cfg, ms in tab if ms is not None)}")


def test_learned_selector_resolves_and_runs():
    """Untuned AUTO on row-major FP32 resolves through the nearest-neighbour
    selector (not the static nnz-tiled default) and stays within the FP32
    gate; a fresh handle is used so no autotuned configuration is cached."""
    picks = {}
    for name in ("stencil_20", "c1", "zipf_mid", "edge"):
        c = case(name)
        m = D.Matrix(c.nr, c.nc, c.rowPtr, c.colIdx, c.values)
        for nv in (1, 8, 32, 64):
            if nv > c.maxnv:
                continue
            picks[(name, nv)] = m.resolve(nv).as_dict()
            out = run_native(c, m, nv)
            h32, c64, ab = c.ref(RM, nv)
            assert_fp32(out, c64, ab, c.rowlen_flat(RM, nv), f"selector {name} nv={nv}", h32)
        # BF16 and column-major keep the static default; column-major work of
        # at least 2^20 (nnz x nv) goes to the nnz-tiled kernel on transposed copies
        assert m.resolve(8, btype=D.DSPMM_R_16BF).as_dict()["variant"] == "nnz_tiled"
        want = "nnz_tiled" if c.nnz * 8 >= 1 << 20 else "csr_scalar"
        assert m.resolve(8, layout=CM).as_dict()["variant"] == want
        m.close()
    note(f"selector picks: {picks}")
    assert any(p["variant"] != "nnz_tiled" for p in picks.values())
    assert all(p["variant"] == "spmv_tiled" or k[1] != 1 or p["variant"] in
               ("csr_scalar", "csr_vector", "coo_scalar", "coo_vector") for k, p in picks.items())


@pytest.mark.parametrize("name", ["stencil_20", "c1", "edge", "int_valued", "zipf_mid"])
@pytest.mark.parametrize("R,unroll", [(2, 4), (4, 8), (8, 8)])
def test_row_blocked(name, R, unroll):
    """Row blocks: every row sums its nonzeros in nonzero order -- bit-exact
    with the host loop on all rows (no splits), FP32 and BF16-stored B."""
    c, m = matrix(name)
    cfg = D.make_config(D.VARIANT_ROW_BLOCKED, tile_nnz=R, unroll=unroll)
    for nv in (1, 3, 8, 64):
        if nv > c.maxnv:
            continue
        h32, c64, ab = c.ref(RM, nv)
        what = f"{name} row-blocked R={R} nv={nv}"
        note(what)
        out = run_native(c, m, nv, config=cfg)
        assert_bitwise(out, h32, what)
    nv = min(64, c.maxnv)
    d16 = D.to_bf16_device(c.d_B)
    out = run_native(c, m, nv, config=cfg, btype=D.DSPMM_R_16BF, B=d16)
    h32r = O.spmm_host32(RM, c.nr, c.nc, nv, c.rowPtr, c.colIdx, c.values,
                         O.round_bf16(c.dense(nv))).reshape(-1)
    assert_bitwise(out, h32r, f"{name} row-blocked R={R} bf16")


@pytest.mark.parametrize("name,strides", [("stencil_20", (20, 400, 7, 8000)),
                                          ("edge", (3, 1000, 3000)), ("zipf_mid", (5,)),
                                          ("int_valued", (2,))])
@pytest.mark.parametrize("R", [2, 4, 8])
def test_row_blocked_strided(name, strides, R):
    """Row blocks whose rows are row_stride apart, with strides that leave a
    partial last period, blocks past the last row and a stride equal to the
    row count: every output still equals the host loop bit for bit."""
    c, m = matrix(name)
    for rs in strides:
        cfg = D.make_config(D.VARIANT_ROW_BLOCKED, tile_nnz=R, unroll=4, row_stride=rs)
        for nv in (1, 8, 64):
            if nv > c.maxnv:
                continue
            h32, c64, ab = c.ref(RM, nv)
            what = f"{name} row-blocked R={R} stride={rs} nv={nv}"
            note(what)
            assert_bitwise(run_native(c, m, nv, config=cfg), h32, what)
    nv = min(64, c.maxnv)
    d16 = D.to_bf16_device(c.d_B)
    cfg = D.make_config(D.VARIANT_ROW_BLOCKED, tile_nnz=R, unroll=8, row_stride=strides[0])
    out = run_native(c, m, nv, config=cfg, btype=D.DSPMM_R_16BF, B=d16)
    h32r = O.spmm_host32(RM, c.nr, c.nc, nv, c.rowPtr, c.colIdx, c.values,
                         O.round_bf16(c.dense(nv))).reshape(-1)
    assert_bitwise(out, h32r, f"{name} row-blocked R={R} stride={strides[0]} bf16")


def test_row_blocked_stride_bounds_and_autotune():
    """row_stride outside [-1, nr] is rejected; on a 20^3 stencil the
    detected distance (row_stride -1) is one plane, 400 rows: sampled rows
    share 16.8 columns on average with the row a plane below, 16.0 with the
    row a line below, 11 or fewer at the other candidate distances.  The
    autotuner tries blocks of consecutive rows and blocks 400 rows apart."""
    c, m = matrix("stencil_20")
    nv = 8
    h32, _, _ = c.ref(RM, nv)
    detected = run_native(c, m, nv, config=D.make_config(D.VARIANT_ROW_BLOCKED, tile_nnz=4,
                                                         row_stride=-1))
    assert_bitwise(detected, h32, "stencil_20 row-blocked detected stride")
    p400 = run_native(c, m, nv, config=D.make_config(D.VARIANT_ROW_BLOCKED, tile_nnz=4,
                                                     row_stride=400))
    assert np.array_equal(bits(detected), bits(p400))
    for bad in (-2, c.nr + 1):
        with pytest.raises(D.DspmmError) as e:
            m.spmm(c.d_B, c.d_C, 8, config=D.make_config(D.VARIANT_ROW_BLOCKED, tile_nnz=4,
                                                         row_stride=bad))
        assert e.value.status == D.STATUS_INVALID_VALUE
    tab = m.autotune_table(c.d_B, c.d_C, 64, reps=2)
    strided = {(cfg.tile_nnz, cfg.row_stride) for cfg, ms in tab
               if cfg.as_dict()["variant"] == "row_blocked" and ms is not None}
    assert {(R, s) for R in (2, 4, 8) for s in (1, 400)} <= strided, strided


def test_row_blocked_rejects_repeated_columns():
    rp = np.array([0, 3, 4], np.int32)
    ci = np.array([1, 1, 2, 0], np.int32)
    v = np.ones(4, np.float32)
    m = D.Matrix(2, 3, rp, ci, v)
    d_B = D.DeviceArray.from_numpy(np.ones((3, 4), np.float32))
    d_C = D.DeviceArray.zeros((2, 4))
    with pytest.raises(D.DspmmError) as e:
        m.spmm(d_B, d_C, 4, config=D.make_config(D.VARIANT_ROW_BLOCKED, tile_nnz=4))
    assert e.value.status == D.STATUS_NOT_SUPPORTED
    m.spmm(d_B, d_C, 4)                               # AUTO still runs (another kernel)
    D.synchronize()
    assert np.array_equal(d_C.numpy(), np.array([[3] * 4, [1] * 4], np.float32))
    m.close()


@pytest.mark.parametrize("name", ["c1", "zipf_mid", "edge", "stencil_20", "int_valued"])
@pytest.mark.parametrize("unroll", [2, 4, 8, 16])
def test_row_sequential(name, unroll):
    """CSR_SCALAR with an unroll (direct index loads, padded last chunk):
    each row is summed in nonzero order, so every output equals the host
    loop bit for bit -- FP32 and BF16-stored B, any nv, and several column
    slabs when the lane count is capped."""
    c, m = matrix(name)
    cfg = D.make_config(D.VARIANT_CSR_SCALAR, unroll=unroll)
    for nv in (1, 3, 8, 64, 128):
        if nv > c.maxnv:
            continue
        h32, c64, ab = c.ref(RM, nv)
        what = f"{name} row-seq unroll={unroll} nv={nv}"
        note(what)
        assert_bitwise(run_native(c, m, nv, config=cfg), h32, what)
    # sliced staging windows (tile_nnz 1): wavefronts whose rows overflow one
    # window give each unfinished row a slice of every window
    sliced = D.make_config(D.VARIANT_CSR_SCALAR, unroll=unroll, tile_nnz=1)
    for nv in (1, 3, 8, 64):
        if nv > c.maxnv:
            continue
        h32, c64, ab = c.ref(RM, nv)
        what = f"{name} row-seq sliced unroll={unroll} nv={nv}"
        note(what)
        assert_bitwise(run_native(c, m, nv, config=sliced), h32, what)
    nv = min(64, c.maxnv)
    h32, c64, ab = c.ref(RM, nv)
    narrow = D.make_config(D.VARIANT_CSR_SCALAR, vec=2, lanes=4, unroll=unroll)   # 8 slabs
    assert_bitwise(run_native(c, m, nv, config=narrow), h32, f"{name} row-seq slabs")
    d16 = D.to_bf16_device(c.d_B)
    h32r = O.spmm_host32(RM, c.nr, c.nc, nv, c.rowPtr, c.colIdx, c.values,
                         O.round_bf16(c.dense(nv))).reshape(-1)
    for bf in (cfg, sliced, D.make_config(D.VARIANT_CSR_SCALAR)):
        out = run_native(c, m, nv, config=bf, btype=D.DSPMM_R_16BF, B=d16)
        assert_bitwise(out, h32r, f"{name} row-seq unroll={bf.unroll} bf16")


def test_row_sequential_leading_dimensions_and_bad_unroll():
    c, m = matrix("zipf_mid")
    nv, ldb, ldc = 24, 40, 36
    B = W.gen_dense(78, c.nc * ldb).reshape(c.nc, ldb)
    d_B = D.DeviceArray.from_numpy(B)
    d_C = D.DeviceArray.from_numpy(np.full((c.nr, ldc), -3.0, np.float32))
    m.spmm(d_B, d_C, nv, ldb=ldb, ldc=ldc, config=D.make_config(D.VARIANT_CSR_SCALAR, unroll=8))
    D.synchronize()
    out = d_C.numpy()
    assert np.all(out[:, nv:] == -3.0), "columns beyond nv must be left alone"
    h32 = O.spmm_host32(RM, c.nr, c.nc, nv, c.rowPtr, c.colIdx, c.values,
                        np.ascontiguousarray(B[:, :nv]))
    assert_bitwise(np.ascontiguousarray(out[:, :nv]).reshape(-1), h32.reshape(-1), "row-seq ld")
    for bad in (D.make_config(D.VARIANT_CSR_SCALAR, unroll=5),
                D.make_config(D.VARIANT_CSR_SCALAR, unroll=4, tile_nnz=2)):
        with pytest.raises(D.DspmmError) as e:
            m.spmm(d_B, d_C, nv, ldb=ldb, ldc=ldc, config=bad)
        assert e.value.status == D.STATUS_NOT_SUPPORTED


@pytest.mark.parametrize("name", ["c1", "zipf_mid", "edge", "stencil_20"])
def test_column_major_through_transposes(name):
    """Row-major kernels on column-major operands run on transposed copies of
    B and C: every output equals the row-major call's bit for bit (FP32 and
    BF16 B), with padded leading dimensions, through the handle's scratch or
    a caller workspace, and entries beyond the leading rows are untouched."""
    c, m = matrix(name)
    cfgs = [D.make_config(D.VARIANT_NNZ_TILED), D.make_config(D.VARIANT_CSR_SCALAR, unroll=4),
            D.make_config(D.VARIANT_CSR_SCALAR, unroll=8, tile_nnz=1),
            D.make_config(D.VARIANT_ROW_BLOCKED, tile_nnz=2, unroll=4),
            D.make_config(D.VARIANT_ROW_BLOCKED, tile_nnz=4, unroll=4, row_stride=-1)]
    for nv in (1, 3, 8, 64, 100):
        if nv > c.maxnv:
            continue
        Brm = c.dense(nv).reshape(c.nc, nv)
        ldb, ldc = c.nc + 5, c.nr + 3
        Bcm = np.zeros((nv, ldb), np.float32)
        Bcm[:, :c.nc] = Brm.T
        d_Bcm = D.DeviceArray.from_numpy(Bcm)
        for bt in (D.DSPMM_R_32F, D.DSPMM_R_16BF):
            bf = bt == D.DSPMM_R_16BF
            d_Bcm_t = D.to_bf16_device(d_Bcm) if bf else d_Bcm
            d_Brm_t = D.to_bf16_device(c.d_B) if bf else c.d_B
            for cfg in cfgs:
                what = f"{name} CM nv={nv} {'bf16' if bf else 'fp32'} {cfg.as_dict()['variant']}"
                note(what)
                ref = run_native(c, m, nv, config=cfg, btype=bt, B=d_Brm_t).reshape(c.nr, nv)
                for use_ws in (False, True):
                    d_C = D.DeviceArray.from_numpy(np.full((nv, ldc), -7.0, np.float32))
                    ws, wsb = None, 0
                    if use_ws:
                        wsb = m.workspace_size(nv, layout=CM, btype=bt, config=cfg)
                        if nv > 1:
                            assert wsb >= 4 * c.nr * nv, "the transposed copies need workspace"
                        ws = D.DeviceArray.zeros(wsb // 4 + 64)
                    m.spmm(d_Bcm_t, d_C, nv, layout=CM, ldb=ldb, ldc=ldc, btype=bt, config=cfg,
                           workspace=ws, workspace_bytes=wsb)
                    D.synchronize()
                    out = d_C.numpy()
                    assert np.all(out[:, c.nr:] == -7.0), what + ": padding written"
                    assert np.array_equal(bits(out[:, :c.nr].T), bits(ref)), what


def test_native_variants_on_empty_and_one_entry_matrices():
    """Every native kernel overwrites C on a matrix with no nonzeros (zeros
    everywhere, both layouts) and gets a one-nonzero matrix exactly right."""
    nr, nc = 300, 50
    empty = D.Matrix(nr, nc, np.zeros(nr + 1, np.int32), np.zeros(0, np.int32),
                     np.zeros(0, np.float32))
    rp = np.zeros(nr + 1, np.int32)
    rp[8:] = 1                                      # row 7 holds the one nonzero
    one = D.Matrix(nr, nc, rp, np.array([13], np.int32), np.array([2.5], np.float32))
    cfgs = [D.make_config(D.VARIANT_NNZ_TILED), D.make_config(D.VARIANT_CSR_SCALAR, unroll=4),
            D.make_config(D.VARIANT_ROW_BLOCKED, tile_nnz=2), D.make_config(D.VARIANT_NNZ_CACHED),
            D.make_config(D.VARIANT_SPMV_TILED), None]
    for nv in (1, 4):
        B = W.gen_dense(5, nc * nv).reshape(nc, nv)
        d_B = D.DeviceArray.from_numpy(B)
        for layout in (RM, CM):
            Bl = B if layout == RM else np.ascontiguousarray(B.T)
            d_Bl = D.DeviceArray.from_numpy(Bl)
            want = np.zeros((nr, nv), np.float32)
            want[7] = 2.5 * B[13]
            for cfg in cfgs:
                if cfg is not None and cfg.variant == D.VARIANT_SPMV_TILED and nv != 1:
                    continue
                what = f"nv={nv} layout={layout} {cfg.as_dict()['variant'] if cfg else 'auto'}"
                for m, ref in ((empty, np.zeros((nr, nv), np.float32)), (one, want)):
                    d_C = D.DeviceArray.from_numpy(np.full(nr * nv, 7.0, np.float32))
                    m.spmm(d_Bl, d_C, nv, layout=layout, config=cfg)
                    D.synchronize()
                    got = d_C.numpy().reshape((nr, nv) if layout == RM else (nv, nr))
                    got = got if layout == RM else got.T
                    assert np.array_equal(got, ref), what
    del d_B
    empty.close()
    one.close()


@pytest.mark.parametrize("name", ["zipf_mid", "edge"])
def test_wide_dense_operands(name):
    """nv = 300 (several column slabs, five combine column blocks): the
    nnz-tiled kernel within the FP32 gate, the row-exact kernels bit-exact
    with the host loop, FP32 and BF16-stored B."""
    c, m = matrix(name)
    nv = 300
    B = W.gen_dense(301, c.nc * nv).reshape(c.nc, nv)
    d_B = D.DeviceArray.from_numpy(B)
    d_C = D.DeviceArray.from_numpy(np.full((c.nr, nv), np.nan, np.float32))
    h32 = O.spmm_host32(RM, c.nr, c.nc, nv, c.rowPtr, c.colIdx, c.values, B).reshape(-1)
    c64, ab = O.spmm_fp64(RM, c.nr, c.nc, nv, c.rowPtr, c.colIdx, c.values, B)
    rl = np.repeat(c.rowlen, nv)
    for cfg in (D.make_config(D.VARIANT_NNZ_TILED), D.make_config(D.VARIANT_NNZ_TILED, tile_nnz=16),
                D.make_config(D.VARIANT_CSR_SCALAR, unroll=4),
                D.make_config(D.VARIANT_CSR_SCALAR, unroll=8, tile_nnz=1),
                D.make_config(D.VARIANT_ROW_BLOCKED, tile_nnz=2, unroll=4),
                D.make_config(D.VARIANT_CSR_SCALAR)):
        what = f"{name} nv={nv} {cfg.as_dict()}"
        note(what)
        m.spmm(d_B, d_C, nv, config=cfg)
        D.synchronize()
        out = d_C.numpy().reshape(-1)
        if cfg.variant == D.VARIANT_NNZ_TILED:
            assert_fp32(out, c64.reshape(-1), ab.reshape(-1), rl, what, h32)
        else:
            assert_bitwise(out, h32, what)
    d16 = D.to_bf16_device(d_B)
    h32r = O.spmm_host32(RM, c.nr, c.nc, nv, c.rowPtr, c.colIdx, c.values,
                         O.round_bf16(B)).reshape(-1)
    m.spmm(d16, d_C, nv, btype=D.DSPMM_R_16BF, config=D.make_config(D.VARIANT_CSR_SCALAR, unroll=4))
    D.synchronize()
    assert_bitwise(d_C.numpy().reshape(-1), h32r, f"{name} nv={nv} bf16 row-seq")


def test_split_combine_repeats_and_matches_batched():
    """Split rows on float data (summation order matters): repeated calls,
    calls covering nv in several column slabs and the SpMV tiles give the
    same bits each time, within the FP32 gate, and the single-matrix
    combine equals the batched launch's bit for bit (both sum a row's chunk
    carries in one order fixed by nv and the chunk count)."""
    c, m = matrix("zipf_mid")
    assert (c.rowlen > 64).sum() > 0
    for nv, cfg in ((128, D.make_config(D.VARIANT_NNZ_TILED, vec=4, lanes=16, tile_nnz=64)),
                    (64, D.make_config(D.VARIANT_NNZ_TILED, vec=1, lanes=64, tile_nnz=16)),
                    (1, D.make_config(D.VARIANT_SPMV_TILED, tile_nnz=512))):
        what = f"zipf_mid split combine nv={nv} {cfg.as_dict()}"
        note(what)
        first = run_native(c, m, nv, config=cfg).copy()
        for _ in range(3):
            assert np.array_equal(bits(first), bits(run_native(c, m, nv, config=cfg))), what
        h32, c64, absum = c.ref(RM, nv)
        assert_fp32(first, c64, absum, c.rowlen_flat(RM, nv), what, h32)
    # single launch against a one-item batch, the batch's configuration
    # for the item on both
    nv = 96
    d_C2 = D.DeviceArray.zeros((c.nr, nv))
    b = D.Batch([(m, nv, c.d_B, d_C2)])
    b.run()
    D.synchronize()
    m.spmm(c.d_B, c.d_C, nv, config=b.item_config(0))
    D.synchronize()
    single = c.output(nv).copy()
    assert np.array_equal(bits(single), bits(d_C2.numpy().reshape(-1))), "single vs batched"
