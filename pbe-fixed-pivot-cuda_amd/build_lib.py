"""Build libpbe_hip.so (gfx950) and the C++ API test program in-tree.

    python pbe-fixed-pivot-cuda_amd/build_lib.py

Each HIP source is compiled to an object in parallel, then linked into
lib/libpbe_hip.so.  Flags: FP64 code for gfx950 only; -munsafe-fp-atomics so
FP64 atomic adds are the hardware instructions (device memory from hipMalloc
is coarse-grained, where they are exact); packed-FP32 generation disabled.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "_build")
INCLUDE = os.path.join(REPO, "include")

SOURCES = ["agg_kernels.hip", "brk_kernels.hip", "batch.hip", "integrate.hip", "rhs_fused.hip",
           "vjp.hip", "det_kernels.hip", "capi.hip"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", f"-I{INCLUDE}", f"-I{CSRC}", "-Wall",
          "-Wno-unused-function"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_library(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *COMMON, "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(len(cmds), 8)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out:
                print(out)
    lib = os.path.join(LIBDIR, "libpbe_hip.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, lib)
    return lib


def build_cpp_tests():
    """tests/cpp/test_pbe_api.cpp: the reference's unit tests restated
    against the C++ drop-in header (needs a GPU to run, not to build)."""
    src = os.path.join(REPO, "tests", "cpp", "test_pbe_api.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(REPO, "tests", "cpp", "test_pbe_api")
    _run([HIPCC, "-O2", "-std=c++17", f"-I{INCLUDE}", src, "-o", out + ".tmp",
          f"-L{LIBDIR}", "-lpbe_hip", f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath,$ORIGIN/../../"
          "pbe-fixed-pivot-cuda_amd/lib"])
    os.replace(out + ".tmp", out)
    return out


def build_host_checks():
    """tests/cpp/test_roots.cpp: host-side check of the device cube root;
    tests/cpp/test_strip_plan.cpp: host-side check of the strip launch's
    work split."""
    for name in ("test_roots", "test_strip_plan"):
        src = os.path.join(REPO, "tests", "cpp", name + ".cpp")
        out = os.path.join(REPO, "tests", "cpp", name)
        _run([HIPCC, "-O2", "-std=c++17", f"-I{INCLUDE}", f"-I{CSRC}", src, "-o", out + ".tmp"])
        os.replace(out + ".tmp", out)


def build_bench_helpers():
    """bench/batched_loop.cpp: batched launch vs a loop of plain calls (the
    bench's batched form)."""
    src = os.path.join(HERE, "bench", "batched_loop.cpp")
    out = os.path.join(HERE, "bench", "batched_loop")
    _run([HIPCC, "-O2", "-std=c++17", f"-I{INCLUDE}", src, "-o", out + ".tmp",
          f"-L{LIBDIR}", "-lpbe_hip", "-Wl,-rpath,$ORIGIN/../lib"])
    os.replace(out + ".tmp", out)
    return out


EXAMPLES = ["aggregation_constant", "aggregation_sum", "aggregation_product",
            "breakage_linear_symmetricbinary", "breakage_linear_uniform",
            "flocculation_two_stage"]


def build_examples():
    """examples/<name>/main.cpp -> examples/<name>/example_<name>: the
    reference's example programs over the drop-in headers (host code linked
    to lib/libpbe_hip.so)."""
    cmds, outs = [], []
    for name in EXAMPLES:
        src = os.path.join(HERE, "examples", name, "main.cpp")
        out = os.path.join(HERE, "examples", name, "example_" + name)
        cmds.append([HIPCC, "-O2", "-std=c++17", f"-I{INCLUDE}", src, "-o", out,
                     f"-L{LIBDIR}", "-lpbe_hip", "-Wl,-rpath,$ORIGIN/../../lib"])
        outs.append(out)
    with cf.ThreadPoolExecutor(max_workers=min(len(cmds), 8)) as ex:
        list(ex.map(_run, cmds))
    return outs


def build_tools():
    """Measurement tools, built on request (build_lib.py --tools), not by
    build(): bench/atomic_rates.hip (LDS and global FP64 atomic-add rates)
    and bench/fp64_issue.hip (FP64 FMA issue rates by waves per SIMD and
    independent chains per wave)."""
    outs = []
    for name in ("atomic_rates", "fp64_issue"):
        src = os.path.join(HERE, "bench", name + ".hip")
        exe = os.path.join(HERE, "bench", name)
        _run([HIPCC, *COMMON, src, "-o", exe + ".tmp"])
        os.replace(exe + ".tmp", exe)
        outs.append(exe)
    return outs


def build(verbose=False):
    lib = build_library(verbose)
    exe = build_cpp_tests()
    build_host_checks()
    build_bench_helpers()
    build_examples()
    return lib, exe


if __name__ == "__main__":
    print(build(verbose=True))
    if "--tools" in sys.argv[1:]:
        print(build_tools())
    sys.exit(0)
