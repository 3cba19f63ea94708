"""Build libtyphon_gpu.so (gfx950) in-tree with hipcc.

Usage: python -m typhon_amd.build   (or typhon_amd.build.build())
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtyphon_gpu.so")
SOURCES = ["ty_kernels_f64.hip", "ty_kernels_f32.hip", "ty_krylov.hip", "ty_adjoint.hip", "ty_setup.hip", "ty_probe.hip",
           "ty_host.cpp"]
ARCH = os.environ.get("TY_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: keep the compiler from emitting packed FP32 VALU ops
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-pass-failed", "-I", os.path.join(ROOT, "include")]

# The derivative kernels (tangent and adjoint) have no reference counterpart to match bit for bit:
# let them multiply by reciprocals and use the rcp + Newton FP64 division (about 1 ulp).  The
# primal kernels keep IEEE division and square root.
EXTRA = {"ty_adjoint.hip": ["-freciprocal-math", "-fapprox-func"]}


def build(verbose: bool = False) -> str:
    objs, procs = [], []
    for src in SOURCES:  # translation units compile in parallel
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmd = ["hipcc", *HIPFLAGS, *EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd.insert(1, "-x")
            cmd.insert(2, "hip")
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = [src for src, p in procs if p.wait() != 0]
    if failed:
        raise RuntimeError(f"hipcc failed on {failed}")
    tmp = LIB + ".tmp"
    cmd = ["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
