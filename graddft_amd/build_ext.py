"""Build the native libraries in-tree.

* ``graddft_amd/libgdx.so``       HIP kernels + C-ABI (``include/gdx.h``), gfx950 only.
* ``graddft_amd/libgdx_ints.so``  host C++ integrals (S, T, V, packed ERI), OpenMP.

``python -m graddft_amd.build_ext`` or ``build()`` from Python.  Always recompiles.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)

HIP_SOURCES = ["gdx_ao.hip", "gdx_xc.hip", "gdx_scf.hip", "gdx_hf.hip", "gdx_eri.hip", "gdx_int1e.hip",
               "gdx_becke.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are disabled for every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip(out=None):
    out = out or os.path.join(HERE, "libgdx.so")
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = out + ".tmp"
    _run([HIPCC] + HIP_FLAGS + srcs + ["-o", tmp])
    os.replace(tmp, out)
    return out


def build_ints(out=None):
    out = out or os.path.join(HERE, "libgdx_ints.so")
    tmp = out + ".tmp"
    _run(["g++", "-O3", "-std=c++17", "-fopenmp", "-fPIC", "-shared",
          os.path.join(CSRC, "gdx_ints.cpp"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build():
    build_ints()
    build_hip()


if __name__ == "__main__":
    build()
    sys.exit(0)
