"""Build libkigpu.so (HIP, gfx950) in-tree with hipcc.

Usage: python kern-interp-old_amd/build.py [-j N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libkigpu.so")

SOURCES = ["alloc.cpp", "scan.hip", "tree.hip", "tree_import.cpp", "factor.hip", "border.hip", "dense_inv.hip", "solve.hip", "treecode.hip", "blocks.hip", "adjoint.hip", "domain.hip", "capi.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-std=c++17", "-O3", "-fPIC", "--offload-arch=gfx950",
    # keep the compiler from emitting packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


# per-file additions: the matrix-core loops of solve.hip (kernel apply) and
# factor.hip (elimination products) keep their loop-carried accumulators in
# the VGPR form of the MFMA instructions (the default heuristic put them in
# AGPRs and copied them back and forth every step)
FILE_FLAGS = {"solve.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
              "factor.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, *FILE_FLAGS.get(src, [])]
    if src.endswith(".cpp"):
        cmd += ["-x", "hip"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-8000:]}")
    return obj


API_SOURCES = ["runtime.cpp", "ki_mat.cpp", "boundary.cpp", "cubic_boundary.cpp", "quadtree.cpp", "kernel.cpp",
               "skel_factorization.cpp", "linear_solve.cpp", "output_dump.cpp"]
API_LIB = os.path.join(HERE, "libkern_interp.so")
DROPIN_TEST = os.path.join(BUILD, "dropin_test")
EXPERIMENTS = ["experiment_one", "experiment_two", "experiment_three"]
CXX = os.environ.get("CXX", "g++")
CXXFLAGS = ["-std=c++17", "-O2", "-fPIC", "-Wall", "-I" + os.path.join(ROOT, "include")]


def _compile_api(src: str) -> str:
    obj = os.path.join(BUILD, "api_" + os.path.splitext(src)[0] + ".o")
    cmd = [CXX, *CXXFLAGS, "-c", os.path.join(HERE, "src", src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed on {src}:\n{r.stderr[-8000:]}")
    return obj


def build_api(jobs: int = 4) -> str:
    """libkern_interp.so (the drop-in C++ API over libkigpu.so) and the
    C++ drop-in test program."""
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_api, API_SOURCES))
    tmp = API_LIB + ".tmp"
    cmd = [CXX, "-shared", "-o", tmp, *objs, "-L" + HERE, "-lkigpu", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link of libkern_interp.so failed:\n{r.stderr[-8000:]}")
    os.replace(tmp, API_LIB)
    cmd = [CXX, *CXXFLAGS, os.path.join(ROOT, "tests", "cpp", "dropin_test.cpp"), "-o",
           DROPIN_TEST, "-L" + HERE, "-lkern_interp", "-lkigpu",
           "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link," + HERE]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"dropin_test link failed:\n{r.stderr[-8000:]}")
    # the reference's experiment drivers over the drop-in API
    for ex in EXPERIMENTS:
        cmd = [CXX, *CXXFLAGS, "-fopenmp", os.path.join(HERE, "experiments", ex + ".cpp"), "-o",
               os.path.join(BUILD, ex), "-L" + HERE, "-lkern_interp", "-lkigpu",
               "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link," + HERE]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{ex} link failed:\n{r.stderr[-8000:]}")
    return API_LIB


def build(jobs: int = 4) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-8000:]}")
    os.replace(tmp, LIB)
    build_api(jobs)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=4)
    a = ap.parse_args()
    try:
        print(build(a.j))
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
