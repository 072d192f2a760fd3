"""dspmm_matrix_prepare: after it, the first dspmm_spmm_ex call with the
prepared arguments allocates nothing and copies nothing to the host, so it
can be captured into a hipGraph; every replay (with eager work and
synchronisations between replays) equals the eager result bit for bit.
Without prepare, a call under capture returns DSPMM_STATUS_NOT_PREPARED
instead of breaking the capture."""
import ctypes as C

import numpy as np
import pytest

import dspmm as D
from gpu_common import bits, case, note

pytestmark = pytest.mark.gpu
RM, CM = D.DENSE_ROW_MAJOR, D.DENSE_COL_MAJOR
CAPTURE_GLOBAL = 0     # hipStreamCaptureModeGlobal


def hip():
    h = C.CDLL("libamdhip64.so")
    h.hipStreamBeginCapture.argtypes = [C.c_void_p, C.c_int]
    h.hipStreamEndCapture.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
    h.hipGraphInstantiate.argtypes = [C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_size_t]
    h.hipGraphLaunch.argtypes = [C.c_void_p, C.c_void_p]
    h.hipGraphExecDestroy.argtypes = [C.c_void_p]
    h.hipGraphDestroy.argtypes = [C.c_void_p]
    return h


def stream():
    s = C.c_void_p()
    D.check(D.lib().dspmm_stream_create(C.byref(s)), "stream_create")
    return s.value


def capture(h, st, fn):
    assert h.hipStreamBeginCapture(st, CAPTURE_GLOBAL) == 0
    try:
        status = fn()
    finally:
        graph = C.c_void_p()
        end = h.hipStreamEndCapture(st, C.byref(graph))
    assert end == 0, f"hipStreamEndCapture: {end}"
    return status, graph


CASES = [
    ("int_zipf", 64, RM, D.make_config(D.VARIANT_NNZ_TILED, tile_nnz=64)),     # split rows
    ("int_zipf", 32, RM, None),                                                # AUTO
    ("stencil_20", 64, RM, D.make_config(D.VARIANT_ROW_BLOCKED, tile_nnz=4, row_stride=-1)),
    ("zipf_mid", 1, RM, D.make_config(D.VARIANT_SPMV_TILED, tile_nnz=512)),
    ("edge", 16, CM, D.make_config(D.VARIANT_NNZ_TILED, tile_nnz=32)),         # transposed copies
    ("int_zipf", 8, RM, D.make_config(D.VARIANT_COO_VECTOR)),                  # zeroes C, rowIdx
    # the B-row ring: its plan (and walk length) built by prepare
    ("stencil_20", 64, RM, D.make_config(D.VARIANT_RING, tile_nnz=4, unroll=8)),
    ("stencil_20", 32, CM, D.make_config(D.VARIANT_RING, tile_nnz=8, unroll=4, row_stride=10)),
]


@pytest.mark.parametrize("name,nv,layout,cfg", CASES,
                         ids=[f"{c[0]}-n{c[1]}-{'RM' if c[2] == RM else 'CM'}-"
                              f"{D.VARIANT_NAMES[c[3].variant] if c[3] else 'auto'}" for c in CASES])
def test_prepared_first_call_is_captured(name, nv, layout, cfg):
    h = hip()
    c = case(name)
    st = stream()
    # eager result on a handle of its own
    ref_m = D.Matrix(c.nr, c.nc, c.rowPtr, c.colIdx, c.values)
    c.fill_output(np.nan)
    ref_m.spmm(c.d_B, c.d_C, nv, layout=layout, config=cfg)
    D.synchronize()
    want = c.output(nv).copy()
    ref_m.close()
    # a fresh handle: prepared, then its very first call is captured
    m = D.Matrix(c.nr, c.nc, c.rowPtr, c.colIdx, c.values)
    m.prepare(c.d_B, c.d_C, nv, layout=layout, config=cfg, handle_buffers=True)
    D.synchronize()
    ldb = nv if layout == RM else c.nc
    ldc = nv if layout == RM else c.nr

    def call():
        return D.lib().dspmm_spmm_ex(m.handle, layout, nv, c.d_B.ptr, ldb, D.DSPMM_R_32F,
                                     c.d_C.ptr, ldc, C.byref(cfg) if cfg is not None else None,
                                     None, 0, st)

    status, graph = capture(h, st, call)
    assert status == D.STATUS_SUCCESS, f"captured call: {D.status_string(status)}"
    ex = C.c_void_p()
    assert h.hipGraphInstantiate(C.byref(ex), graph, None, None, 0) == 0
    for rep in range(3):
        c.fill_output(np.nan)                       # eager work between replays
        assert h.hipGraphLaunch(ex, st) == 0
        D.check(D.lib().dspmm_stream_synchronize(st), "sync")
        got = c.output(nv)
        assert np.array_equal(bits(got), bits(want)), f"{name} nv={nv} replay {rep}"
    note(f"{name} nv={nv}: 3 replays equal the eager call")
    h.hipGraphExecDestroy(ex)
    h.hipGraphDestroy(graph)
    m.close()
    D.lib().dspmm_stream_destroy(st)


def test_unprepared_call_under_capture_is_refused():
    h = hip()
    c = case("int_zipf")
    st = stream()
    m = D.Matrix(c.nr, c.nc, c.rowPtr, c.colIdx, c.values)
    cfg = D.make_config(D.VARIANT_NNZ_TILED, tile_nnz=64)
    status, graph = capture(h, st, lambda: D.lib().dspmm_spmm_ex(
        m.handle, RM, 32, c.d_B.ptr, 32, D.DSPMM_R_32F, c.d_C.ptr, 32, C.byref(cfg), None, 0, st))
    assert status == D.STATUS_NOT_PREPARED, D.status_string(status)
    h.hipGraphDestroy(graph)
    # outside capture the same call builds its plan and runs
    m.spmm(c.d_B, c.d_C, 32, config=cfg)
    D.synchronize()
    m.close()
    D.lib().dspmm_stream_destroy(st)


def test_prepare_fixes_auto_and_sizes_the_fallback_workspace():
    c = case("zipf_mid")
    m = D.Matrix(c.nr, c.nc, c.rowPtr, c.colIdx, c.values)
    m.prepare(c.d_B, c.d_C, 32)
    picked = m.resolve(32).as_dict()
    assert picked["variant"] != "auto"
    assert m.resolve(32).as_dict() == picked
    # AUTO's workspace covers the static default it may fall back to
    heur = m.workspace_size(32, config=D.make_config(D.VARIANT_NNZ_TILED))
    assert m.workspace_size(32) >= heur
    m.close()
