"""Build libszx_amd.so (gfx950) in-tree with hipcc.

Usage: python -m szx_amd.build   (or szx_amd.build.build() from Python)

Flags that matter for exactness:
  -ffp-contract=off                  no FMA contraction of the codec's
                                     sub/mul/add (the reference does none)
  -fno-gpu-flush-denormals-to-zero   fp32 denormals are data here
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                     keep the compiler from emitting packed
                                     fp32 instructions
  -mllvm -amdgpu-atomic-optimizer-strategy=None
                                     atomics stay as written (see DEVICE)
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libszx_amd.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HOSTCXX = os.environ.get("CXX", "g++")
ARCH = os.environ.get("SZX_AMD_ARCH", "gfx950")

SOURCES = ["szx_compress.hip", "szx_decompress.hip", "szx_index.hip", "szx_generic.hip", "szx_nonblocked.hip", "szx_cuszx.hip", "szx_fxr.hip", "szx_mlp.hip", "szx_noalloc.hip", "szx_probe.hip",
           "szx_api.cpp", "szx_fxr_host.cpp", "szx_rate.cpp", "szx_batch.cpp"]
HEADERS = ["szx_common.hpp", "szx_internal.hpp", "szx_wave.hpp", "szx_fxr.hpp",
           "szx_fxr_host.hpp", "szx_mlp.hpp", "szx_rate.hpp", "szx_index.hpp", "szx_batch.hpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
          "-Wno-unused-result"]
DEVICE = [f"--offload-arch={ARCH}", "-fno-gpu-flush-denormals-to-zero",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          # keep single-lane atomics single-lane: the compressor's look-back
          # issues one and reads its result much later, which the combining
          # rewrite (a wave-wide broadcast of the result) would force at once
          "-mllvm", "-amdgpu-atomic-optimizer-strategy=None"]


def build(verbose: bool = False) -> str:
    """Compile every source and link the shared library; returns its path."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, src + ".o")
        if src.endswith(".hip"):
            cmd = [HIPCC, "-x", "hip", *COMMON, *DEVICE, "-I" + INCLUDE, "-c", path, "-o", obj]
        else:   # host-only C++ against the HIP runtime headers
            cmd = [HOSTCXX, *COMMON, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                   "-I" + INCLUDE, "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        cmds.append(cmd)
        objs.append(obj)
    # the translation units are independent: compile them side by side
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
        for r in pool.map(lambda c: subprocess.run(c), cmds):
            if r.returncode != 0:
                raise subprocess.CalledProcessError(r.returncode, r.args)
    tmp = LIB + ".tmp"
    # export the C ABI of include/szx_amd.h only
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}",
           "-Wl,--version-script=" + os.path.join(CSRC, "exports.map"), "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
