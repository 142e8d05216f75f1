"""Build the product library ``crystal2_amd/_lib/libcrystal2.so`` in-tree.

Host C++ (scene/config loaders, BVH builder, C-ABI glue) is compiled with g++;
the HIP kernels with ``hipcc --offload-arch=gfx950``. Every object is compiled
with ``-ffp-contract=off`` (no fused multiply-add) and without fast-math: the
reference's fp32 operation order is part of the contract. HIP code is built
with ``-Xclang -target-feature -Xclang -packed-fp32-ops`` so the compiler
never emits packed-fp32 instructions.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libcrystal2.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("CRYSTAL2_ARCH", "gfx950")

HOST_SOURCES = ["c2_host.cpp", "c2_bvh.cpp", "c2_image.cpp", "c2_film.cpp", "c2_kdtree.cpp"]
HIP_SOURCES = ["c2_device.hip", "c2_nrc.hip"]
HEADERS = ["c2_host.h", "c2_vec.h", "c2_bvh.h", "c2_device.h", "c2_image.h", "c2_kdtree.h"]

CXXFLAGS = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wno-unused-function",
            "-pthread", f"-I{os.path.join(HERE, '..', 'include')}"]
HIPFLAGS = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-function", "-Wno-unused-command-line-argument",
            f"-I{os.path.join(HERE, '..', 'include')}"]


# The MLP kernels keep their MFMA accumulators in VGPRs (the compiler's default puts them in
# AGPRs, and every bias init and ReLU then copies them across: 2 instructions per element).
EXTRA_HIPFLAGS = {"c2_nrc.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    os.makedirs(os.path.join(OUT_DIR, "obj"), exist_ok=True)
    objs = []
    for src in HOST_SOURCES:
        p = os.path.join(CSRC, src)
        if not os.path.exists(p):
            continue
        o = os.path.join(OUT_DIR, "obj", src + ".o")
        _run(["g++", *CXXFLAGS, "-c", p, "-o", o])
        objs.append(o)
    for src in HIP_SOURCES:
        p = os.path.join(CSRC, src)
        if not os.path.exists(p):
            continue
        o = os.path.join(OUT_DIR, "obj", src + ".o")
        _run([HIPCC, *HIPFLAGS, *EXTRA_HIPFLAGS.get(src, []), "-c", p, "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-pthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
