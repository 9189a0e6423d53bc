"""Build the HIP library grasph2_amd/libgrasph2_amd.so for gfx950.

hipcc compiles every csrc/*.hip unit to an object in parallel and links them
into one shared library that exports the C ABI of include/grasph2.h.  Runs
here (hipcc cross-compiles without a GPU); the built .so travels with the tree.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgrasph2_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32; the compiler would otherwise
# form them from scalar float code.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wall", "-Wno-unused-function",
    "-I", os.path.join(ROOT, "include"),
]

SOURCES = ["capi.hip", "sweep.hip", "sym_sweep.hip", "nbr_passes.hip", "integrate.hip", "reorder.hip", "batch.hip"]


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    return obj


def build_library(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(6, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


CPP_EXAMPLES = ("api_roundtrip", "snapshot_roundtrip", "ic_generate", "ic_device_roundtrip")


def build_cpp_example(verbose=True):
    """Compile the C++ programs under tests/cpp: api_roundtrip (the C++
    template API driven like the reference's main loop) and
    snapshot_roundtrip (snapshot files, host only).  Host code: g++ and the
    HIP runtime."""
    for name in CPP_EXAMPLES:
        src = os.path.join(ROOT, "tests", "cpp", name + ".cpp")
        exe = os.path.join(ROOT, "tests", "cpp", name)
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
               "-I" + os.path.join(ROOT, "include"), src, "-L" + HERE, "-lgrasph2_amd", "-L/opt/rocm/lib",
               "-lamdhip64", "-pthread", "-Wl,-rpath,$ORIGIN/../../grasph2_amd", "-Wl,-rpath,/opt/rocm/lib", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("C++ example build failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
        if verbose:
            print("built", exe)


def build_host_checks(verbose=True):
    """Compile tests/cpp/sym_geometry_check (host code that includes the
    symmetric sweep's geometry header; hipcc for its __device__ __host__
    functions, no GPU needed to run it)."""
    src = os.path.join(ROOT, "tests", "cpp", "sym_geometry_check.cpp")
    exe = os.path.join(ROOT, "tests", "cpp", "sym_geometry_check")
    cmd = [HIPCC, "-O2", "-std=c++17", "-Wno-unused-command-line-argument", src, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host check build failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    if verbose:
        print("built", exe)


def build_oracle(verbose=True):
    """Build the CPU oracle (test infrastructure) with its Makefile."""
    r = subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n%s" % r.stderr)
    if verbose:
        print("built", os.path.join(ROOT, "oracle", "liboracle.so"))


if __name__ == "__main__":
    build_library()
    if "--with-oracle" in sys.argv:
        build_oracle()
