"""The B-row ring plan (csrc/ring_plan.cpp), checked on the CPU: the C++
driver tests/cxx/ring_plan_check.cpp replays every stream's fills into a
simulated ring, checks that each nonzero's slot holds its column when its
tile runs, that rows appear once and the kernel's limits hold, and that C
evaluated through the plan equals the host loop bit for bit -- on grid
stencils (ragged tile shapes, one and two row distances, a fixed larger
ring, grids smaller than one tile, non-cubic grids, 7-point and 5-point
stencils), random rows in
consecutive tiles (one row, a partial tile, far more columns than rows, a
dozen columns shared by every row, mostly empty rows), random rows under
arbitrary row-distance pairs (the library samples them, so an unstructured
matrix may bring any), and a ring too small to be accepted."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "da-spmmul-kernel_amd", "csrc")


def test_ring_plan_replays_exactly(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "ring_plan_check")
    subprocess.run([hipcc, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
                    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
                    os.path.join(ROOT, "tests", "cxx", "ring_plan_check.cpp"),
                    os.path.join(CSRC, "ring_plan.cpp"), "-o", exe], check=True,
                   capture_output=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = [l for l in out.stdout.splitlines() if ": ok," in l]
    assert len(lines) == 67, out.stdout


def test_ring_plan_builder_is_memory_clean(tmp_path):
    """The same replay with the host code built under AddressSanitizer (the
    planner's vectors, the replay's indexing): no out-of-bounds access or
    leak on any of the plans above."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "ring_plan_check_asan")
    build = subprocess.run([hipcc, "-O1", "-g", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
                            "-Xarch_host", "-fsanitize=address",
                            "-Xarch_host", "-fno-omit-frame-pointer",
                            f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
                            os.path.join(ROOT, "tests", "cxx", "ring_plan_check.cpp"),
                            os.path.join(CSRC, "ring_plan.cpp"), "-o", exe],
                           capture_output=True, text=True)
    if build.returncode != 0 and "sanitizer" in build.stderr.lower():
        pytest.skip("host AddressSanitizer runtime not available")
    assert build.returncode == 0, build.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert out.returncode == 0 and "Sanitizer" not in out.stderr, out.stderr[-4000:]
    assert sum(": ok," in l for l in out.stdout.splitlines()) == 67
