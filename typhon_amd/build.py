"""Build libtyphon_gpu.so (gfx950) in-tree with hipcc.

Usage: python -m typhon_amd.build   (or typhon_amd.build.build())
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtyphon_gpu.so")
SOURCES = ["ty_adjoint.hip", "ty_kernels_f64.hip", "ty_kernels_f32.hip", "ty_krylov.hip", "ty_setup.hip",
           "ty_probe.hip", "ty_meshbuild.hip", "ty_order.hip", "ty_host.cpp"]
GEN = os.path.join(HERE, "_build")  # generated instantiation units (not sources of their own)

# The residual kernels' variants (ty_kernels_impl.h: rec_variant<DIM, R, NSM, DENSE, GOUT>,
# gradient_variant<DIM, R, NSM, DENSE>, flux_variant<DIM, R, NSM, JAC, DENSE>) are instantiated one per
# translation unit so that they
# compile in parallel; one more unit per (dimension, precision) holds launch_gradient/launch_flux.
GRAD_VARIANTS = [(4, "true"), (6, "true"), (4, "false"), ("MAX_SLOTS", "false")]
FLUX_VARIANTS = [(4, "true", "false"), ("MAX_SLOTS", "true", "false"), (4, "false", "true"), (6, "false", "true"),
                 (4, "false", "false"), ("MAX_SLOTS", "false", "false")]


def _unit(decls):
    return '#include "ty_kernels_impl.h"\nnamespace ty {\n' + "".join(d + "\n" for d in decls) + "}\n"


def generated_units():
    """(file name, text) of the instantiation units"""
    units = []
    for r in ("double", "float"):
        for dim in (2, 3):
            tag = f"{dim}_{'f64' if r == 'double' else 'f32'}"
            gsig = [f"void gradient_variant<{dim}, {r}, {ns}, {de}>(const KArgs<{r}>&, hipStream_t);"
                    for ns, de in GRAD_VARIANTS]
            rsig = [f"void rec_variant<{dim}, {r}, {ns}, {de}, {go}>(const KArgs<{r}>&, hipStream_t);"
                    for ns, de in GRAD_VARIANTS for go in ("true", "false")]
            for i, sig in enumerate(rsig):
                units.append((f"inst_rec{tag}_{i}.hip", _unit(["template " + sig])))
            fsig = [f"void flux_variant<{dim}, {r}, {ns}, {jac}, {de}>(const KArgs<{r}>&, hipStream_t);"
                    for ns, jac, de in FLUX_VARIANTS]
            for i, sig in enumerate(gsig):
                units.append((f"inst_grad{tag}_{i}.hip", _unit(["extern template " + x for x in rsig]
                                                               + ["template " + sig])))
            for i, sig in enumerate(fsig):
                units.append((f"inst_flux{tag}_{i}.hip", _unit(["template " + sig])))
            units.append((f"inst_launch{tag}.hip", _unit(
                ["extern template " + sig for sig in gsig + fsig]
                + [f"template hipError_t launch_gradient<{dim}, {r}>(const KArgs<{r}>&, hipStream_t);",
                   f"template hipError_t launch_flux<{dim}, {r}>(const KArgs<{r}>&, hipStream_t);"])))
    return units
ARCH = os.environ.get("TY_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: keep the compiler from emitting packed FP32 VALU ops
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-pass-failed", "-I", os.path.join(ROOT, "include")]

# The derivative kernels (tangent and adjoint) have no reference counterpart to match bit for bit:
# let them multiply by reciprocals and use the rcp + Newton FP64 division (about 1 ulp).  The
# primal kernels keep IEEE division and square root.
EXTRA = {"ty_adjoint.hip": ["-freciprocal-math", "-fapprox-func"],
         # the device mesh build reproduces the host build's geometry bit for bit: no FMA contraction
         "ty_meshbuild.hip": ["-ffp-contract=off"]}


def build(verbose: bool = False) -> str:
    os.makedirs(GEN, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    for name, text in generated_units():
        path = os.path.join(GEN, name)
        with open(path, "w") as f:
            f.write(text)
        srcs.append(path)
    srcs.sort(key=lambda p: (0 if "inst_rec" in p else 1))  # the slowest units first

    def compile_one(src):
        obj = os.path.join(GEN, os.path.splitext(os.path.basename(src))[0] + ".o")
        cmd = ["hipcc", *HIPFLAGS, "-I", CSRC, *EXTRA.get(os.path.basename(src), []), "-c", src, "-o", obj]
        if src.endswith(".cpp"):
            cmd.insert(1, "-x")
            cmd.insert(2, "hip")
        if verbose:
            print(" ".join(cmd), flush=True)
        return obj, subprocess.run(cmd).returncode

    # translation units compile in parallel, one per core
    with ThreadPoolExecutor(max_workers=max(1, min(16, os.cpu_count() or 1))) as ex:
        results = list(ex.map(compile_one, srcs))
    failed = [src for src, (_, rc) in zip(srcs, results) if rc != 0]
    if failed:
        raise RuntimeError(f"hipcc failed on {failed}")
    objs = [o for o, _ in results]
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
