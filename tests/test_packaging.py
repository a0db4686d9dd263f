"""Packaging parity (CPU): the CMake project builds the gfx950 library with
the reference's options, installs it with a pbe_cudaConfig.cmake, and a
separate project that knows only find_package(pbe_cuda) and the
pbe_cuda::pbe_cuda target (tests/cmake_consumer) builds the reference's unit
tests against it (reference CMakeLists.txt:42-46,136-176,
cmake/pbe_cudaConfig.cmake.in).  Compiling the library takes about two
minutes on 8 cores; the consumer binary runs on the GPU box
(ctest -L gpu)."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(cmd, cwd=None, timeout=900):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=timeout)
    assert r.returncode == 0, f"{' '.join(cmd)} failed:\n{r.stdout[-3000:]}"
    return r.stdout


@pytest.mark.skipif(shutil.which("cmake") is None, reason="cmake not installed")
def test_find_package_consumer_builds(tmp_path):
    build = tmp_path / "build"
    jobs = str(max(1, min(8, os.cpu_count() or 1)))
    _run(["cmake", "-S", REPO, "-B", str(build), "-DPBE_BUILD_STATIC=ON",
          "-DPBE_BUILD_SHARED=OFF"])
    _run(["cmake", "--build", str(build), "-j", jobs])
    out = _run(["ctest", "--test-dir", str(build), "-LE", "gpu", "--output-on-failure"])
    assert "100% tests passed" in out, out
    prefix = build / "test_install"
    cfg = prefix / "lib" / "cmake" / "pbe_cuda"
    for f in ("pbe_cudaConfig.cmake", "pbe_cudaConfigVersion.cmake", "pbe_cudaTargets.cmake"):
        assert (cfg / f).exists(), f
    for h in ("include/pbe_cuda/pbe_cuda.cuh", "include/pbe_cuda/aggregation.cuh",
              "include/pbe_cuda/breakage.cuh", "include/pbe_hip/pbe_hip.h",
              "lib/libpbe_hip.a"):
        assert (prefix / h).exists(), h
    exe = build / "consumer" / "test_pbe_api"
    assert exe.exists()
    syms = _run(["nm", str(exe)])
    assert "pbe_launch_aggregation_rhs" in syms and "pbe_launch_breakage_rhs" in syms
