"""Build librcg.so (HIP, gfx950) in-tree.

    python -m rayleighchebyshevgpgpu_amd.build

hipcc cross-compiles for gfx950 without a GPU.  Packed FP32 instructions are
disabled for all device code (-packed-fp32-ops).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librcg.so")
ROOT = os.path.dirname(HERE)

SOURCES = ["rcg_kernels.hip", "rcg_batch.hip", "rcg_solver.cpp", "rcg_io.cpp"]
HEADERS = ["rcg_internal.hpp", "rcg_csr_rows.hpp", "rcg_cpoly.hpp", "rcg_dense.hpp"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-command-line-argument",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # host code: complex products without the C99 inf/nan recovery calls (the small
    # Hermitian eigensolver and Cholesky run twice as fast; finite results unchanged)
    "-Xarch_host", "-fcx-limited-range",
    # host code: AVX2 + FMA (the oracle is built for the same level) -- the
    # small eigensolvers run 25-30% faster at b = 64-96
    "-Xarch_host", "-march=x86-64-v3",
    "-I" + os.path.join(ROOT, "include"),
]


def build(verbose: bool = True, out: str = LIB, defines=()) -> str:
    objs = []
    tag = "".join("_" + d.lower() for d in defines)
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + tag + ".o")
        cmd = [HIPCC, *FLAGS, *("-D" + d for d in defines), "-x", "hip", "-c",
               os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,-soname,librcg.so", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    for o in objs:
        os.remove(o)
    return out


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
