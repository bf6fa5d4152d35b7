"""Build librgf.so (HIP kernels + C ABI) in-tree for gfx950 with hipcc.

Usage: ``python -m rgf_amd.build`` or ``rgf_amd.build.build()``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librgf.so")
DRIVER = os.path.join(HERE, "rgf_inverse_driver")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

SOURCES = ["rgf_kernels.hip", "rgf_channels.cpp", "rgf_api.cpp", "rgf_batch.cpp"]
MEX_GATEWAY = os.path.join(HERE, "mex", "cal_smatrix_RGF_mex.c")
MEX_HARNESS = os.path.join(os.path.dirname(HERE), "tests", "mex_harness")

# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 in device code.
HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    """Compile every source and link librgf.so; returns the library path."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    cc = hipcc()
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [cc, *HIPCC_FLAGS, *lang, "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    # the vendor dense solver (pivoted slice inverses, the fallback for slices
    # the blocked kernels refuse)
    rocm_lib = os.path.join(os.path.dirname(os.path.dirname(cc)), "lib")
    cmd = [cc, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-L" + rocm_lib, "-Wl,-rpath," + rocm_lib,
           "-lrocsolver", "-lrocblas", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    # host program for hardware-counter runs of the slice inverse
    cmd = ["gcc", "-O2", "-I", INCLUDE, os.path.join(CSRC, "inverse_driver.c"), "-o", DRIVER + ".tmp", "-L", HERE,
           "-lrgf", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(DRIVER + ".tmp", DRIVER)
    build_mex_harness(verbose)
    return LIB


def build_mex_harness(verbose: bool = False) -> str | None:
    """The MEX gateway compiled against the tests' host implementation of the
    MEX API (tests/mex_harness), so that it runs without MATLAB."""
    if not os.path.isdir(MEX_HARNESS):
        return None
    out = os.path.join(MEX_HARNESS, "libmexgw.so")
    cmd = ["gcc", "-O2", "-std=gnu11", "-Wall", "-shared", "-fPIC", "-I", MEX_HARNESS, "-I", INCLUDE, MEX_GATEWAY,
           os.path.join(MEX_HARNESS, "mex_harness.c"), "-o", out + ".tmp", "-L", HERE, "-lrgf",
           "-Wl,-rpath,$ORIGIN/../../rgf_amd"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
