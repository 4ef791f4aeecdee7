"""Build ``markovkit_amd/libmarkovkit_amd.so`` for gfx950 (in-tree, no JIT cache).

HIP sources are compiled with hipcc for ``--offload-arch=gfx950`` with packed-FP32 code
generation disabled; the host setup code is compiled as plain C++. Run as
``python -m markovkit_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmarkovkit_amd.so")
ARCH = os.environ.get("MK_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}", "-pthread",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-result", f"-I{ROOT}/include"]
# no contraction of products into sums on the host: setup sums round as the reference's default
# (FMA-less) build does, and as the device assembly (compiled with contraction off) does
CXX_FLAGS = ["-O3", "-std=c++20", "-fPIC", "-march=x86-64-v3", "-ffp-contract=off", "-pthread", "-Wall",
             "-Wextra", f"-I{ROOT}/include"]

HIP_SOURCES = ["kernels.hip", "batch_kernels.hip", "api.hip", "batch.hip", "expv.hip", "server.hip",
               "assemble.hip", "boxmg.hip"]
CXX_SOURCES = ["setup.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        o = os.path.join(OBJ, src + ".o")
        objs.append(o)
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
    for src in CXX_SOURCES:
        o = os.path.join(OBJ, src + ".o")
        objs.append(o)
        jobs.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
    with ThreadPoolExecutor(max_workers=min(4, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-pthread", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
