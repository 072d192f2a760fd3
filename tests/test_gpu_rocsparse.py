"""GPU tests of the rocSPARSE comparator columns and of the sweep harness
(bin/spmvspmm, run.sh): every comparator column is within the FP32 gate of
the FP64 oracle (exact on integer-valued operands), and the harness checks
all algorithms and prints the reference's CSV rows."""
import csv
import math
import os
import subprocess

import numpy as np
import pytest

import dspmm as D
from gpu_common import assert_fp32, case, note

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
MTX = os.path.join(HERE, "golden", "mtx")
RUN_SH = os.path.join(os.path.dirname(D.LIB_PATH), "..", "run.sh")
SENTINEL = 7.0   # C is overwritten (beta = 0): no sentinel may survive


@pytest.mark.parametrize("nv", [1, 8, 32])
@pytest.mark.parametrize("name", ["zipf_mid", "edge", "int_valued"])
def test_comparator_columns_within_fp32_gate(name, nv):
    c = case(name)
    rs = D.RocSparse(c.nr, c.nc, c.nnz, c.d_rp, c.d_ri, c.d_ci, c.d_v, c.d_B, c.d_C)
    ran = 0
    for col in range(rs.columns(nv)):
        layout = rs.layout(nv, col)
        c.fill_output(SENTINEL)
        what = f"rocsparse {rs.name(nv, col)} {name} nv={nv}"
        if not rs.run(nv, col):
            note(f"{what}: not supported by rocSPARSE")
            continue
        D.synchronize()
        h32, c64, ab = c.ref(layout, nv)
        out = c.output(nv)
        if name == "int_valued":
            assert np.array_equal(out, c64.astype(np.float32)), what
        else:
            assert_fp32(out, c64, ab, c.rowlen_flat(layout, nv), what, h32)
        ran += 1
    assert ran >= (3 if nv == 1 else 6), f"only {ran} comparator columns ran"
    rs.close()


def test_comparator_benchmark_reports_time():
    c = case("zipf_mid")
    rs = D.RocSparse(c.nr, c.nc, c.nnz, c.d_rp, c.d_ri, c.d_ci, c.d_v, c.d_B, c.d_C)
    ms = rs.benchmark(32, 0, warmup=2, reps=5)
    assert ms is not None and ms > 0
    note(f"rocsparse {rs.name(32, 0)} zipf_mid nv=32: {ms * 1e3:.1f} us")
    rs.close()


def _floats(fields):
    return [float(x) for x in fields]


def test_harness_sweep_csv(tmp_path):
    files = [os.path.join(MTX, "general_random_300x250.mtx"), os.path.join(MTX, "sym_random_200.mtx")]
    log = tmp_path / "log.csv"
    env = dict(os.environ, LOG_FILE=str(log))
    out = subprocess.run(["bash", RUN_SH, "--native", *files], capture_output=True, text=True,
                         timeout=300, env=env)
    note(out.stderr[-2000:])
    assert out.returncode == 0, out.stderr[-2000:]
    rows = list(csv.reader(open(log)))
    header = rows[0]
    assert header[:2] == ["filename", "KDense"] and len(header) == 21 and header[-1] == "native"
    body = rows[1:]
    assert len(body) == 2 * 8        # nv = 1, 2, ..., 128 for each file
    for r in body:
        nv = int(r[1])
        algs = _floats(r[2:10])
        assert all(v > 0 for v in algs), r
        comp = r[10:20]
        if nv == 1:
            assert all(x == "" for x in comp[4:]), r
            comp = comp[:4]
        vals = _floats(comp)
        assert sum(not math.isnan(v) for v in vals) >= (3 if nv == 1 else 6), r
        assert float(r[20]) > 0, r
    note(f"harness row: {body[-1]}")


def test_harness_profile_mode():
    f = os.path.join(MTX, "general_random_300x250.mtx")
    out = subprocess.run([D.HARNESS_PATH, f, "profile", "4"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    assert "maxNv = 4" in out.stderr
