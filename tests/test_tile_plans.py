"""The native kernels' host plans (csrc/tile_plans.cpp), checked on the CPU
by the C++ driver tests/cxx/tile_plans_check.cpp: nnz tiles with split rows,
SpMV tiles, column-cached tiles and strided row blocks cover every row and
nonzero once and in order and keep their kernels' limits, and row blocks
rebuild every row bit for bit -- on random rows, very long rows, empty rows,
one row and no rows.  Run plainly and with the host code under
AddressSanitizer."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "da-spmmul-kernel_amd", "csrc")
CHECKS = 7


def build(tmp_path, name, extra):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path / name)
    out = subprocess.run([hipcc, "-std=c++17", "-x", "hip", "--offload-arch=gfx950", *extra,
                          f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
                          os.path.join(ROOT, "tests", "cxx", "tile_plans_check.cpp"),
                          os.path.join(CSRC, "tile_plans.cpp"), "-o", exe],
                         capture_output=True, text=True)
    return exe, out


def test_tile_plans_cover_every_row_in_order(tmp_path):
    exe, b = build(tmp_path, "tile_plans_check", ["-O2"])
    assert b.returncode == 0, b.stderr[-2000:]
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert f"{CHECKS} of {CHECKS} checks ok" in out.stdout, out.stdout


def test_tile_plans_are_memory_clean(tmp_path):
    exe, b = build(tmp_path, "tile_plans_check_asan",
                   ["-O1", "-g", "-Xarch_host", "-fsanitize=address",
                    "-Xarch_host", "-fno-omit-frame-pointer"])
    if b.returncode != 0 and "sanitizer" in b.stderr.lower():
        pytest.skip("host AddressSanitizer runtime not available")
    assert b.returncode == 0, b.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert out.returncode == 0 and "Sanitizer" not in out.stderr, out.stderr[-4000:]
    assert f"{CHECKS} of {CHECKS} checks ok" in out.stdout
