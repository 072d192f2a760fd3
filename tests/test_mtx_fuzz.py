"""The Matrix-Market reader (csrc/mtx.cpp) under AddressSanitizer on the
CPU, by the C++ driver tests/cxx/mtx_fuzz_check.cpp: every committed fixture
with every flag setting (malformed ones refused), then 150 mutations of each
readable fixture (truncated, bytes overwritten, lines repeated or dropped),
each either refused with an error status or read into entries inside the
declared shape with sorted rows -- and no out-of-bounds access or leak."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "da-spmmul-kernel_amd", "csrc")


def test_mtx_reader_survives_mutated_files(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "mtx_fuzz_check")
    b = subprocess.run([hipcc, "-O1", "-g", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
                        "-Xarch_host", "-fsanitize=address",
                        "-Xarch_host", "-fno-omit-frame-pointer",
                        f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
                        os.path.join(ROOT, "tests", "cxx", "mtx_fuzz_check.cpp"),
                        os.path.join(CSRC, "mtx.cpp"), "-o", exe], capture_output=True, text=True)
    if b.returncode != 0 and "sanitizer" in b.stderr.lower():
        pytest.skip("host AddressSanitizer runtime not available")
    assert b.returncode == 0, b.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1")
    out = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "mtx"), str(tmp_path)],
                         capture_output=True, text=True, timeout=600, env=env)
    assert out.returncode == 0 and "Sanitizer" not in out.stderr, out.stdout + out.stderr[-4000:]
    assert "none inconsistent" in out.stdout, out.stdout
