"""Builds the gfx950 backend library (hussar_amd/libhussar_hip.so) in-tree with hipcc.

The library is the C ABI of include/hussar_hip.h: host C++ (BVH build, guiding tree, runtime)
and the HIP kernels, compiled for gfx950 only, without floating-point contraction (the
kernels' triangle tests and geometry chain must match the host restatement bit for bit) and
with packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhussar_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["kernels.hip", "wavefront.hip", "primary.hip", "post.hip", "api.cpp", "serve.cpp", "bvh.cpp", "guide.cpp"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument"]


def build(verbose: bool = False) -> str:
    """Compile every source into the shared library; returns its path."""
    def compile_one(src: str) -> str:
        obj = os.path.join(CSRC, "_build", src + ".o")
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{err}")
        if verbose and err:
            print(err, file=sys.stderr)
        return obj

    # one hipcc process per source, at most four at a time
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as pool:
        objs = list(pool.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx",
           "-Wl,-rpath,/opt/rocm/lib", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    build_cpp_front_end()
    return LIB


CPP_DEMO_SRC = os.path.join(ROOT, "tests", "cpp", "drop_in_dihedral.cpp")
CPP_DEMO = os.path.join(ROOT, "tests", "cpp", "drop_in_dihedral")


def build_cpp_front_end() -> str:
    """Compiles the dihedral sweep written against include/hussar_hip.hpp (the header-only
    C++ front end) with g++ and links it to the library through an $ORIGIN-relative rpath,
    so that the binary finds the in-tree library wherever the tree is copied."""
    tmp = CPP_DEMO + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
           CPP_DEMO_SRC, "-L", HERE, "-lhussar_hip", "-Wl,-rpath,$ORIGIN/../../hussar_amd", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed on the C++ front end:\n" + r.stderr)
    os.replace(tmp, CPP_DEMO)
    return CPP_DEMO


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
