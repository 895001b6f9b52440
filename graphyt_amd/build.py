"""Build the gfx950 shared library libgraphyt_amd.so in-tree with hipcc, and the C++
host module graphyt_amd/graphyt*.so (the graphyt Python surface over that library) with
g++ and pybind11.

    python -m graphyt_amd.build            # compile (always)

The C ABI is declared in include/graphyt_amd.h. Compiler flags: -O3, gfx950 code object,
and packed FP32 instructions disabled for device code (-target-feature -packed-fp32-ops).
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libgraphyt_amd.so")
SOURCES = ["gy_kernels.hip", "gy_plane.hip", "gy_thermal.hip", "gy_ctx.hip", "gy_mlp.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GY_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(LIBDIR, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + FLAGS + ["-I", os.path.join(ROOT, "include"), "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_host(verbose: bool = False) -> str:
    """C++ host classes + pybind11 binding -> graphyt_amd/graphyt<EXT_SUFFIX>."""
    import pybind11
    src = os.path.join(PKG, "host", "graphyt_module.cpp")
    out = os.path.join(PKG, "graphyt" + sysconfig.get_config_var("EXT_SUFFIX"))
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
           "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"], "-I", os.path.join(ROOT, "include"),
           src, "-o", out + ".tmp", "-L", LIBDIR, "-lgraphyt_amd", "-Wl,-rpath,$ORIGIN/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host module build failed:\n{r.stderr}")
    os.replace(out + ".tmp", out)
    return out


def build_all(verbose: bool = False) -> list:
    return [build(verbose), build_host(verbose)]


if __name__ == "__main__":
    print("\n".join(build_all(verbose="-v" in sys.argv)))
