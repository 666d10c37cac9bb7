"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the two
planners (piquasso_amd/csrc/plan.cpp, poly_plan.cpp) driven by
tests/sanitize/host_driver.cpp, and the C oracle driven by
tests/sanitize/oracle_driver.c, each built by the host compiler alone (no
device code) with -fsanitize=address,undefined and no recovery, and the
oracle's OpenMP loops under ThreadSanitizer.  A report from any sanitizer
aborts the driver and fails the test."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]


def _run(cmd, cwd, env=None):
    res = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=600, env=env)
    return res.returncode, res.stdout + res.stderr


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_planners_under_sanitizers(tmp_path):
    csrc = os.path.join(ROOT, "piquasso_amd", "csrc")
    exe = str(tmp_path / "host_driver")
    rc, log = _run(["g++", "-std=c++17", *SAN, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-I" + csrc, "-I" + os.path.join(ROOT, "include"), "-include", "algorithm",
                    "-o", exe, os.path.join(ROOT, "tests", "sanitize", "host_driver.cpp"),
                    os.path.join(csrc, "plan.cpp"), os.path.join(csrc, "poly_plan.cpp")], ROOT)
    assert rc == 0, log
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    rc, log = _run([exe], ROOT, env)
    assert rc == 0 and "runtime error" not in log and "AddressSanitizer" not in log, log
    assert "0 failures" in log and "plans built 400/400" in log


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no host C compiler")
def test_oracle_under_sanitizers(tmp_path):
    odir = os.path.join(ROOT, "oracle")
    exe = str(tmp_path / "oracle_driver")
    srcs = [os.path.join(odir, f) for f in ("glynn_port.c", "glynn_dd.c", "tensor_dd.c", "perm_poly_dd.c")]
    rc, log = _run(["gcc", "-std=c11", *SAN, "-mfma", "-fopenmp", "-I" + odir, "-o", exe,
                    os.path.join(ROOT, "tests", "sanitize", "oracle_driver.c"), *srcs, "-lm"], ROOT)
    assert rc == 0, log
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", OMP_NUM_THREADS="2")
    rc, log = _run([exe], ROOT, env)
    assert rc == 0 and "runtime error" not in log and "AddressSanitizer" not in log, log
    assert "oracle: ok" in log


CLANG = "/opt/rocm/llvm/bin/clang"
ARCHER = "/opt/rocm/llvm/lib/libarcher.so"


@pytest.mark.skipif(not (os.path.exists(CLANG) and os.path.exists(ARCHER)),
                    reason="no clang with libomp and the Archer OpenMP tool")
def test_oracle_under_thread_sanitizer(tmp_path):
    """The oracle's OpenMP loops (per-job slots, ordered reductions) under
    ThreadSanitizer.  libomp's own synchronisation reaches TSan through the
    Archer tool, and the uninstrumented runtime is ignored; races inside the
    instrumented oracle are still reported (a shared accumulator in an
    ``omp parallel for`` is)."""
    odir = os.path.join(ROOT, "oracle")
    exe = str(tmp_path / "oracle_driver_tsan")
    srcs = [os.path.join(odir, f) for f in ("glynn_port.c", "glynn_dd.c", "tensor_dd.c", "perm_poly_dd.c")]
    rc, log = _run([CLANG, "-std=c11", "-O1", "-g", "-fsanitize=thread", "-fopenmp", "-mfma",
                    "-I" + odir, "-o", exe, os.path.join(ROOT, "tests", "sanitize", "oracle_driver.c"),
                    *srcs, "-lm", "-Wl,-rpath," + os.path.dirname(ARCHER)], ROOT)
    assert rc == 0, log
    env = dict(os.environ, OMP_NUM_THREADS="4", OMP_TOOL_LIBRARIES=ARCHER,
               TSAN_OPTIONS="halt_on_error=1 ignore_noninstrumented_modules=1")
    rc, log = _run([exe], ROOT, env)
    assert rc == 0 and "ThreadSanitizer" not in log, log
    assert "oracle: ok" in log
