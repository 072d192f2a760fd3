"""GPU parity of the B-row ring kernel (DSPMM_VARIANT_RING): every row is
summed in nonzero order with a separately rounded multiply and add, so every
output equals the reference host loop bit for bit -- FP32 and BF16-stored B,
grid stencils with two row distances (line and plane), with one distance,
matrices without structure (consecutive tiles), several tile shapes and walk
lengths, column-major operands through transposed copies.  Operands the
kernel does not take (nv not a multiple of the vector width, a tile with
more nonzeros than one register load of metadata) are refused and AUTO falls
back."""
import numpy as np
import pytest

import dspmm as D
import oracle as O
import workloads as W
from gpu_common import Case, assert_bitwise, bits, case, note

pytestmark = pytest.mark.gpu
RM, CM = D.DENSE_ROW_MAJOR, D.DENSE_COL_MAJOR
_MATS = {}


def stencil2d(n, seed=7):
    """9-point stencil on an n x n grid (one shared row distance, n)."""
    rows, cols = [], []
    for y in range(n):
        for x in range(n):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    xx, yy = x + dx, y + dy
                    if 0 <= xx < n and 0 <= yy < n:
                        rows.append(y * n + x)
                        cols.append(yy * n + xx)
    rows = np.asarray(rows)
    rp = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n * n))]).astype(np.int32)
    v = np.random.default_rng(seed).uniform(0.5, 1.5, len(cols)).astype(np.float32)
    return Case(f"stencil2d_{n}", n * n, n * n, rp, np.asarray(cols, np.int32), v, 64, seed=seed)


def matrix(name):
    if name not in _MATS:
        c = stencil2d(120) if name == "stencil2d" else case(name)
        _MATS[name] = (c, D.Matrix(c.nr, c.nc, c.rowPtr, c.colIdx, c.values))
    return _MATS[name]


def run(c, m, nv, cfg, layout=RM, btype=D.DSPMM_R_32F, B=None):
    c.fill_output(np.nan)
    m.spmm(B if B is not None else c.d_B, c.d_C, nv, layout=layout, config=cfg, btype=btype)
    D.synchronize()
    return c.output(nv)


SHAPES = [(8, 4, 25), (8, 8, 10), (4, 4, 3), (16, 4, 25)]


@pytest.mark.parametrize("name", ["stencil_20", "stencil2d", "c1"])
@pytest.mark.parametrize("w,h,z", SHAPES)
def test_ring_bit_exact(name, w, h, z):
    c, m = matrix(name)
    cfg = D.make_config(D.VARIANT_RING, tile_nnz=w, unroll=h, row_stride=z)
    ran = 0
    for nv in (4, 8, 16, 32, 64):
        if nv > c.maxnv:
            continue
        what = f"{name} ring w={w} h={h} z={z} nv={nv}"
        note(what)
        try:
            out = run(c, m, nv, cfg)
        except D.DspmmError as e:
            # a tile of an unstructured matrix needs many B rows: wide rows
            # may not fit LDS; grid stencils always fit
            assert e.status == D.STATUS_NOT_SUPPORTED and name == "c1", what
            note(f"{what}: ring does not fit LDS")
            continue
        h32, _, _ = c.ref(RM, nv)
        assert_bitwise(out, h32, what)
        ran += 1
    # grid stencils run at every width; an unstructured matrix's tiles and
    # walks may need more LDS than a CU has (refused: AUTO falls back)
    assert ran >= 3 or name == "c1"
    # BF16-stored B: the host loop on the rounded B, bit for bit
    d16 = D.to_bf16_device(c.d_B)
    for nv in (8, 64):
        if nv > c.maxnv:
            continue
        try:
            out = run(c, m, nv, cfg, btype=D.DSPMM_R_16BF, B=d16)
        except D.DspmmError as e:
            assert e.status == D.STATUS_NOT_SUPPORTED and name == "c1"
            continue
        h32r = O.spmm_host32(RM, c.nr, c.nc, nv, c.rowPtr, c.colIdx, c.values,
                             O.round_bf16(c.dense(nv))).reshape(-1)
        assert_bitwise(out, h32r, f"{name} ring w={w} h={h} bf16 nv={nv}")


def banded_chunks(T, longest, seed=11):
    """Consecutive T-row tiles whose longest rows are the given lengths (the
    other rows of a tile shorter, some empty), columns inside a 64-wide band
    near the row."""
    rng = np.random.default_rng(seed)
    band = 64
    lens = []
    for lm in longest:
        tile = rng.integers(0, lm + 1, T)
        tile[rng.integers(0, T)] = lm
        lens += list(tile)
    lens = np.asarray(lens)
    nc = len(lens) + 2 * band
    cols = [np.sort(rng.choice(np.arange(r + band // 2, r + 3 * band // 2), int(k), replace=False))
            for r, k in enumerate(lens)]
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate(cols).astype(np.int32)
    v = rng.uniform(-1, 1, ci.size).astype(np.float32)
    return Case(f"banded_chunks_{T}", len(lens), nc, rp, ci, v, 128, seed=seed)


# (w, h, longest rows of the tiles, FP32 widths): 32-row tiles take one row
# per 16-lane group at nv = 64, 64-row tiles two (nv = 128 would need eight
# per 32-lane group there, which the kernel refuses)
CHUNK_CASES = [(8, 4, [1, 3, 4, 15, 16, 17, 31, 32, 33, 47, 48], (32, 64, 128)),
               (8, 8, [1, 4, 15, 16, 17, 31, 32], (32, 64))]


@pytest.mark.parametrize("w,h,longest,widths", CHUNK_CASES)
def test_ring_metadata_chunks(w, h, longest, widths):
    """Tiles whose longest rows end anywhere in the ring kernel's four-entry
    steps and in 16-entry runs of metadata, one and two rows per lane group,
    lane groups of 8, 16 and 32 lanes (FP32 nv = 32, 64, 128; BF16 nv = 128):
    bit for bit against the host loop."""
    c = banded_chunks(w * h, longest)
    m = D.Matrix(c.nr, c.nc, c.rowPtr, c.colIdx, c.values)
    cfg = D.make_config(D.VARIANT_RING, tile_nnz=w, unroll=h, row_stride=3)
    for nv in widths:
        out = run(c, m, nv, cfg)
        h32, _, _ = c.ref(RM, nv)
        assert_bitwise(out, h32, f"{c.name} ring {w}x{h} nv={nv}")
    d16 = D.to_bf16_device(c.d_B)
    out = run(c, m, 128, cfg, btype=D.DSPMM_R_16BF, B=d16)
    h32r = O.spmm_host32(RM, c.nr, c.nc, 128, c.rowPtr, c.colIdx, c.values,
                         O.round_bf16(c.dense(128))).reshape(-1)
    assert_bitwise(out, h32r, f"{c.name} ring {w}x{h} bf16 nv=128")


@pytest.mark.parametrize("name", ["stencil_20", "stencil2d", "c1"])
def test_ring_auto_stream_length(name):
    """row_stride 0: the library picks the tiles per walk for this GPU's
    CU count; the sums do not depend on it (bit for bit the host
    loop, and the same bits as walks of 10 tiles)."""
    c, m = matrix(name)
    auto = D.make_config(D.VARIANT_RING, tile_nnz=4, unroll=8, row_stride=0)
    fixed = D.make_config(D.VARIANT_RING, tile_nnz=4, unroll=8, row_stride=10)
    for nv in (16, 64):
        try:
            out = run(c, m, nv, auto).copy()
        except D.DspmmError as e:
            assert e.status == D.STATUS_NOT_SUPPORTED and name == "c1"
            continue
        h32, _, _ = c.ref(RM, nv)
        assert_bitwise(out, h32, f"{name} ring auto nv={nv}")
        try:
            ref = run(c, m, nv, fixed)
        except D.DspmmError as e:
            # walks of 10 random-row tiles may need more fills per tile
            # than the kernel's register passes carry
            assert e.status == D.STATUS_NOT_SUPPORTED and name == "c1"
            continue
        assert np.array_equal(bits(out), bits(ref)), f"{name} nv={nv}"


def test_untuned_auto_takes_the_ring_on_a_stencil():
    """Without autotuning, AUTO on a 27-point stencil at nv = 64 resolves (by
    the learned selector, whose corpus times the ring on stencils) to the
    B-row ring, with FP32 and with BF16-stored B, and the call is bit for bit
    the host loop."""
    a = W.c3_matrix(48)
    c = Case("stencil48", a.nr, a.nc, a.rowPtr, a.colIdx, a.values, 64, seed=5)
    m = D.Matrix(c.nr, c.nc, c.rowPtr, c.colIdx, c.values)
    cfg = m.resolve(64)
    note(f"stencil48 nv=64 AUTO resolves to {cfg.as_dict()}")
    assert cfg.variant == D.VARIANT_RING
    out = run(c, m, 64, None)
    h32, _, _ = c.ref(RM, 64)
    assert_bitwise(out, h32, "stencil48 AUTO nv=64")
    # BF16-stored B: the ring pick carries over, bit for bit the host loop on
    # the rounded B
    assert m.resolve(64, btype=D.DSPMM_R_16BF).variant == D.VARIANT_RING
    d16 = D.to_bf16_device(c.d_B)
    out = run(c, m, 64, None, btype=D.DSPMM_R_16BF, B=d16)
    h32r = O.spmm_host32(RM, c.nr, c.nc, 64, c.rowPtr, c.colIdx, c.values,
                         O.round_bf16(c.dense(64))).reshape(-1)
    assert_bitwise(out, h32r, "stencil48 AUTO bf16 nv=64")


def test_ring_integer_exact_and_column_major():
    c, m = matrix("int_valued")          # its 4000-long row does not fit a tile's metadata
    cfg = D.make_config(D.VARIANT_RING, tile_nnz=4, unroll=2, row_stride=8)
    with pytest.raises(D.DspmmError) as e:
        run(c, m, 32, cfg)
    assert e.value.status == D.STATUS_NOT_SUPPORTED
    # AUTO never lands on an unsupported configuration
    out = run(c, m, 32, None)
    _, c64, _ = c.ref(RM, 32)
    assert np.array_equal(out, c64.astype(np.float32))
    # column-major operands through transposed copies: the row-major bits
    s, ms = matrix("stencil_20")
    cfg = D.make_config(D.VARIANT_RING, tile_nnz=8, unroll=4, row_stride=10)
    rm = run(s, ms, 32, cfg).copy()
    Bcm = np.ascontiguousarray(s.dense(32).reshape(s.nc, 32).T)
    d_Bcm = D.DeviceArray.from_numpy(Bcm)
    s.fill_output(np.nan)
    ms.spmm(d_Bcm, s.d_C, 32, layout=CM, config=cfg)
    D.synchronize()
    cm = s.output(32).reshape(32, s.nr).T.reshape(-1)
    assert np.array_equal(bits(cm), bits(rm)), "column-major ring differs from row-major"
    # the same with BF16-stored B (its transposed copy is BF16 too)
    rm16 = run(s, ms, 32, cfg, btype=D.DSPMM_R_16BF, B=D.to_bf16_device(s.d_B)).copy()
    d_Bcm16 = D.to_bf16_device(d_Bcm)
    s.fill_output(np.nan)
    ms.spmm(d_Bcm16, s.d_C, 32, layout=CM, config=cfg, btype=D.DSPMM_R_16BF)
    D.synchronize()
    cm16 = s.output(32).reshape(32, s.nr).T.reshape(-1)
    assert np.array_equal(bits(cm16), bits(rm16)), "column-major BF16 ring differs from row-major"


def test_ring_runs_unstructured_small_tiles():
    c, m = matrix("c1")
    cfg = D.make_config(D.VARIANT_RING, tile_nnz=4, unroll=4, row_stride=3)
    for nv in (4, 16, 32):
        out = run(c, m, nv, cfg)
        h32, _, _ = c.ref(RM, nv)
        assert_bitwise(out, h32, f"c1 ring 4x4 z3 nv={nv}")


def grid3(nx, ny, nz, seed=9):
    """27-point stencil on an nx x ny x nz grid (x fastest)."""
    rows, cols = [], []
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                for dz in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            xx, yy, zz = x + dx, y + dy, z + dz
                            if 0 <= xx < nx and 0 <= yy < ny and 0 <= zz < nz:
                                rows.append(x + nx * (y + ny * z))
                                cols.append(xx + nx * (yy + ny * zz))
    n = nx * ny * nz
    rp = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n))]).astype(np.int32)
    v = np.random.default_rng(seed).uniform(-1, 1, len(cols)).astype(np.float32)
    return Case(f"grid{nx}x{ny}x{nz}", n, n, rp, np.asarray(cols, np.int32), v, 64, seed=seed)


def grid7(nx, ny, nz, seed=10):
    """7-point stencil on an nx x ny x nz grid (nz = 1: the 5-point one)."""
    rows, cols = [], []
    offs = ((0, 0, -1), (0, -1, 0), (-1, 0, 0), (0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1))
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                for dx, dy, dz in offs:
                    xx, yy, zz = x + dx, y + dy, z + dz
                    if 0 <= xx < nx and 0 <= yy < ny and 0 <= zz < nz:
                        rows.append(x + nx * (y + ny * z))
                        cols.append(xx + nx * (yy + ny * zz))
    n = nx * ny * nz
    rp = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n))]).astype(np.int32)
    v = np.random.default_rng(seed).uniform(-1, 1, len(cols)).astype(np.float32)
    return Case(f"grid7pt_{nx}x{ny}x{nz}", n, n, rp, np.asarray(cols, np.int32), v, 64, seed=seed)


def random_rows(name, nr, nc, maxlen, seed, keep_every=1):
    """Rows of 0..maxlen sorted distinct random columns; with keep_every k,
    all but one row in k are empty."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, maxlen + 1, nr)
    lens = np.minimum(lens, nc)
    if keep_every > 1:
        lens[np.arange(nr) % keep_every != keep_every // 2] = 0
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    parts = [np.sort(rng.choice(nc, int(k), replace=False)) for k in lens if k]
    ci = (np.concatenate(parts) if parts else np.zeros(0)).astype(np.int32)
    v = rng.uniform(-1, 1, ci.size).astype(np.float32)
    return Case(name, nr, nc, rp, ci, v, 64, seed=seed)


EDGE = {
    "grid2x2x2": lambda: grid3(2, 2, 2),          # smaller than one tile
    "grid3x3x3": lambda: grid3(3, 3, 3),
    "grid1x1x5": lambda: grid3(1, 1, 5),          # a line: one row distance
    "grid7x9x11": lambda: grid3(7, 9, 11),        # partial tiles on every edge
    "grid13x5x3": lambda: grid3(13, 5, 3),
    "grid37x2x2": lambda: grid3(37, 2, 2),
    "grid7pt_20x20x20": lambda: grid7(20, 20, 20),   # 7-point stencil
    "grid7pt_13x7x9": lambda: grid7(13, 7, 9),
    "grid7pt_120x120x1": lambda: grid7(120, 120, 1),  # 5-point stencil: one row distance
    "one_row": lambda: random_rows("one_row", 1, 50, 20, 3),
    "nine_rows": lambda: random_rows("nine_rows", 9, 9, 9, 5),
    "wide": lambda: random_rows("wide", 4000, 100000, 30, 7),     # nc = 25 nr
    "narrow": lambda: random_rows("narrow", 6000, 12, 12, 9),     # every row shares 12 B rows
    "mostly_empty": lambda: random_rows("mostly_empty", 3000, 3000, 6, 13, keep_every=100),
}


@pytest.mark.parametrize("name", sorted(EDGE))
def test_ring_edge_shapes(name):
    """Tiny, ragged and lopsided matrices, and 7-point and 5-point stencils,
    through the ring: bit for bit the host loop at every tile shape, with
    FP32 and BF16-stored B.  A shape
    whose tile needs more B rows than LDS holds is refused; grids never
    are."""
    c = EDGE[name]()
    m = D.Matrix(c.nr, c.nc, c.rowPtr, c.colIdx, c.values)
    d16 = D.to_bf16_device(c.d_B)
    ran = 0
    for w, h, z in SHAPES + [(4, 8, 0)]:
        cfg = D.make_config(D.VARIANT_RING, tile_nnz=w, unroll=h, row_stride=z)
        for nv in (4, 16, 64):
            what = f"{name} ({c.nr}x{c.nc}, nnz {c.nnz}) ring w={w} h={h} z={z} nv={nv}"
            note(what)
            try:
                out = run(c, m, nv, cfg)
            except D.DspmmError as e:
                assert e.status == D.STATUS_NOT_SUPPORTED and not name.startswith("grid"), what
                continue
            h32, _, _ = c.ref(RM, nv)
            assert_bitwise(out, h32, what)
            ran += 1
            if nv == 64:
                out = run(c, m, nv, cfg, btype=D.DSPMM_R_16BF, B=d16)
                h32r = O.spmm_host32(RM, c.nr, c.nc, nv, c.rowPtr, c.colIdx, c.values,
                                     O.round_bf16(c.dense(nv))).reshape(-1)
                assert_bitwise(out, h32r, what + " bf16")
    assert ran >= 5, name
    m.close()


def test_ring_refuses_an_empty_matrix():
    """No nonzeros: nothing to stage, the ring refuses and AUTO writes zeros."""
    c = Case("empty", 40, 30, np.zeros(41, np.int32), np.zeros(0, np.int32),
             np.zeros(0, np.float32), 64)
    m = D.Matrix(c.nr, c.nc, c.rowPtr, c.colIdx, c.values)
    with pytest.raises(D.DspmmError) as e:
        run(c, m, 16, D.make_config(D.VARIANT_RING))
    assert e.value.status == D.STATUS_NOT_SUPPORTED
    out = run(c, m, 16, None)
    assert np.array_equal(bits(out), np.zeros(c.nr * 16, np.uint32))
    m.close()


def test_ring_refuses_what_it_cannot_take():
    c, m = matrix("stencil_20")
    cfg = D.make_config(D.VARIANT_RING)
    for nv in (1, 2, 6):                 # not whole 16-byte pieces per B row
        with pytest.raises(D.DspmmError) as e:
            run(c, m, nv, cfg)
        assert e.value.status == D.STATUS_NOT_SUPPORTED, nv


def test_autotuner_times_ring_candidates():
    c, m = matrix("stencil_20")
    tab = m.autotune_table(c.d_B, c.d_C, 64, reps=2)
    ring = [(cfg.as_dict(), ms) for cfg, ms in tab if cfg.variant == D.VARIANT_RING]
    assert ring and all(ms is not None for _, ms in ring), ring
    best = min(ring, key=lambda x: x[1])
    others = min(ms for cfg, ms in tab if ms is not None and cfg.variant != D.VARIANT_RING)
    note(f"stencil_20 nv=64: best ring {best[1] * 1e3:.1f} us {best[0]}, best other "
         f"{others * 1e3:.1f} us")
