"""The serving arena's first-fit bookkeeping (csrc/first_fit.hpp), checked
on the CPU by the C++ driver tests/cxx/first_fit_check.cpp: random
allocations and releases against a model of the live blocks -- aligned,
disjoint, free blocks coalesced, the region tiled exactly, the lowest
fitting block chosen, used and peak counted.  Run plainly and under
AddressSanitizer."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "da-spmmul-kernel_amd", "csrc")
CHECKS = 4


def build(tmp_path, name, extra):
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("g++ not available")
    exe = str(tmp_path / name)
    out = subprocess.run([cxx, "-std=c++17", *extra, f"-I{CSRC}",
                          os.path.join(ROOT, "tests", "cxx", "first_fit_check.cpp"), "-o", exe],
                         capture_output=True, text=True)
    return exe, out


@pytest.mark.parametrize("asan", [False, True], ids=["plain", "asan"])
def test_first_fit_keeps_the_region_tiled(tmp_path, asan):
    flags = ["-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer"] if asan else ["-O2"]
    exe, b = build(tmp_path, "first_fit_check", flags)
    if asan and b.returncode != 0 and "sanitizer" in b.stderr.lower():
        pytest.skip("AddressSanitizer runtime not available")
    assert b.returncode == 0, b.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert out.returncode == 0 and "Sanitizer" not in out.stderr, out.stdout + out.stderr[-2000:]
    assert f"{CHECKS} of {CHECKS} checks ok" in out.stdout, out.stdout
