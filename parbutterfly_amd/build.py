"""Builds libparbutterfly_amd.so (HIP kernels for gfx950 + host C++) and the CLI in-tree.

The library is the product: a C-ABI shared object (include/parbutterfly.h).  HIP sources are
compiled with hipcc for gfx950 only; the host-side C++ (IO, generators) with g++ -fopenmp.
Packed-FP32 instructions are disabled for all device code.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libparbutterfly_amd.so")
CLI = os.path.join(HERE, "bin", "bipartite_butterflies")
KONECT = os.path.join(HERE, "bin", "konect_to_hyperadj")

HIP_SOURCES = ["scan.hip", "rank.hip", "count.hip", "peel.hip", "batch.hip", "mlp.hip", "sparse.hip", "textio.hip", "api.hip"]
HOST_SOURCES = ["host.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wno-unused-function",
              "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout)
    if verbose and r.stdout.strip():
        print(r.stdout)
    return r.stdout


def _compile(src, verbose):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    else:
        cmd = ["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    _run(cmd, verbose)
    return obj


def build(verbose=False, jobs=None):
    """Compiles every source (always; this is the build check) and links the library + CLI."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    srcs = [s for s in HIP_SOURCES + HOST_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(8, len(srcs))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lgomp"], verbose)
    os.replace(tmp, LIB)
    for src, exe in (("cli.cpp", CLI), ("konect.cpp", KONECT)):
        path = os.path.join(CSRC, src)
        if os.path.exists(path):
            _run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), path, "-o", exe + ".tmp",
                  "-L" + HERE, "-lparbutterfly_amd", "-Wl,-rpath,$ORIGIN/.."], verbose)
            os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
