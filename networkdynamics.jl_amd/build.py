"""Builds libnd_hip.so in-tree for gfx950.

Host C++ (index layer, graph generators) is compiled with g++; the HIP sources
with hipcc --offload-arch=gfx950; the shared library is linked by hipcc.
-ffp-contract=off keeps every model's arithmetic in the reference's operation
order (no FMA contraction); packed FP32 instructions are disabled for all
device code.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnd_hip.so")
ARCH = os.environ.get("ND_OFFLOAD_ARCH", "gfx950")

HOST_SOURCES = ["nd_network.cpp", "nd_graphgen.cpp"]
HIP_SOURCES = ["nd_plan.hip", "nd_ensemble.hip", "nd_newton.hip", "nd_mpnn.hip"]

HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wextra"]
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]


def _hipcc():
    for c in (os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc"), "hipcc"):
        if os.path.sep not in c or os.path.exists(c):
            return c
    return "hipcc"


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    objs = []
    for s in HOST_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        jobs.append(["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        jobs.append([hipcc] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout + r.stderr, file=sys.stderr)
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
