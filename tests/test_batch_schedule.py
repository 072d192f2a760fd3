"""The batched launch's host schedule (assemble_schedule, in
csrc/batch_schedule.cpp), checked on the CPU by the C++ driver
tests/cxx/batch_schedule_check.cpp: every item's work units are covered once
across the eight XCD lists, each list's running counts and offsets agree,
only heavy items are cut (into contiguous ranges in XCD order), and carry
offsets and split counts are prefix sums -- on 512 random items, one
dominant item, empty items, equal items and one item.  Run plainly and with
the host code under AddressSanitizer."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "da-spmmul-kernel_amd", "csrc")
CHECKS = 5


def build(tmp_path, name, extra):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path / name)
    out = subprocess.run([hipcc, "-std=c++17", "-x", "hip", "--offload-arch=gfx950", *extra,
                          f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
                          os.path.join(ROOT, "tests", "cxx", "batch_schedule_check.cpp"),
                          os.path.join(CSRC, "batch_schedule.cpp"), "-o", exe],
                         capture_output=True, text=True)
    return exe, out


def test_batch_schedule_covers_every_unit_once(tmp_path):
    exe, b = build(tmp_path, "batch_schedule_check", ["-O2"])
    assert b.returncode == 0, b.stderr[-2000:]
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert f"{CHECKS} of {CHECKS} checks ok" in out.stdout, out.stdout


def test_batch_schedule_is_memory_clean(tmp_path):
    exe, b = build(tmp_path, "batch_schedule_check_asan",
                   ["-O1", "-g", "-Xarch_host", "-fsanitize=address",
                    "-Xarch_host", "-fno-omit-frame-pointer"])
    if b.returncode != 0 and "sanitizer" in b.stderr.lower():
        pytest.skip("host AddressSanitizer runtime not available")
    assert b.returncode == 0, b.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert out.returncode == 0 and "Sanitizer" not in out.stderr, out.stderr[-4000:]
    assert f"{CHECKS} of {CHECKS} checks ok" in out.stdout
