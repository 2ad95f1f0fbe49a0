"""Build libpmcd.so (host C++ + HIP kernels for gfx950) in-tree.

Usage: python -m parmoon2022_amd.build  (or build() from __graft_entry__)
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpmcd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
OMP_LIB_DIR = "/opt/rocm/lib/llvm/lib"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-I" + CSRC, "-I" + os.path.join(HERE, "..", "include")]
# packed FP32 instructions are disabled for all device code
HIP_FLAGS = ["-x", "hip", "--offload-arch=gfx950", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOST_FLAGS = ["-ffp-contract=off", "-fopenmp"]


def _sources():
    srcs = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".cpp") or f.endswith(".hip"):
            srcs.append(os.path.join(CSRC, f))
    return srcs


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + COMMON + HIP_FLAGS + ["-c", src, "-o", obj]
    else:
        cmd = [HIPCC] + COMMON + HOST_FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = _sources()
    jobs = min(len(srcs), max(1, min(16, os.cpu_count() or 4)))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fopenmp", "--offload-arch=gfx950", "-o", tmp] + objs + [
        "-L" + OMP_LIB_DIR, "-Wl,-rpath," + OMP_LIB_DIR]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
