"""The C-ABI boundary (CPU only, no compute calls): the built library exports every
function include/typhon_gpu.h declares, and the C header, the ctypes mirror
(typhon_amd/_lib.py) and the Fortran iso_c_binding module (fortran/typhon_gpu.f90)
agree on every struct's size and field offsets."""
import ctypes as C
import json
import os
import re
import shutil
import subprocess

import pytest

from typhon_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "typhon_gpu.h")
FMOD = os.path.join(ROOT, "fortran", "typhon_gpu.f90")

STRUCTS = {
    "ty_boco_desc": (_lib.TyBocoDesc, ["type", "nface", "iface", "ipart", "param"]),
    "ty_mesh_desc": (_lib.TyMeshDesc, ["dim", "ncell_int", "ncell", "nface", "nface_int", "facecell", "face",
                                       "centre", "volume", "nboco", "boco", "cell_order"]),
    "ty_gas": (_lib.TyGas, ["gamma", "r_const", "prandtl", "typ_visc", "visc_dyn", "visc_kin", "tref", "tsuth",
                            "typ_fluid"]),
    "ty_scheme": (_lib.TyScheme, ["sch_hyp", "sch_dis", "method", "limiter", "gradmeth", "postlimiter", "venkat_k",
                                  "nstage", "rk_coef", "cfl", "precision"]),
    "ty_krylov": (_lib.TyKrylov, ["nkrylov", "max_restart", "max_outer", "inner_precision", "precond", "inner_tol",
                                  "outer_tol"]),
    "ty_krylov_stats": (_lib.TyKrylovStats, ["outer_iters", "inner_iters", "rel_res", "rhs_norm"]),
    "ty_unsteady_stats": (_lib.TyUnsteadyStats, ["forward_steps", "adjoint_steps", "max_snapshots",
                                                 "snapshots_allocated", "ms"]),
    "ty_steady": (_lib.TySteady, ["local_dt", "maxit", "check_every", "cfl", "cfl_max", "maxres", "residu_ref",
                                  "wlog", "refsca", "refvec"]),
    "ty_steady_stats": (_lib.TySteadyStats, ["iterations", "converged", "cur_res", "residu_ref", "cfl"]),
}


def header_functions():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ty_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_function():
    if not os.path.exists(_lib.LIBPATH):
        pytest.skip("libtyphon_gpu.so not built")
    lib = C.CDLL(_lib.LIBPATH)
    names = header_functions()
    assert len(names) >= 14
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.EXPORTS)


def _c_layout(tmp_path):
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){", 'printf("{");']
    first = True
    for sname, (_, fields) in STRUCTS.items():
        sep = "" if first else ","
        first = False
        lines.append(f'printf("{sep}\\"{sname}\\": {{\\"size\\": %zu", sizeof({sname}));')
        for f in fields:
            lines.append(f'printf(", \\"{f}\\": %zu", offsetof({sname}, {f}));')
        lines.append('printf("}");')
    lines += ['printf("}\\n");', "return 0;}"]
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", str(src), "-o", str(exe)], check=True)
    return json.loads(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)


def test_ctypes_mirror_matches_c_header(tmp_path):
    lay = _c_layout(tmp_path)
    for sname, (cls, fields) in STRUCTS.items():
        assert C.sizeof(cls) == lay[sname]["size"], sname
        for f in fields:
            assert getattr(cls, f).offset == lay[sname][f], (sname, f)


@pytest.mark.skipif(shutil.which("flang") is None, reason="no Fortran compiler")
def test_fortran_module_matches_c_header(tmp_path):
    lay = _c_layout(tmp_path)
    prog = tmp_path / "sizes.f90"
    body = ["program sizes", "use iso_c_binding", "use TYPHON_GPU", "implicit none"]
    for sname in STRUCTS:
        body.append(f"type({sname}), target :: v_{sname}")
    body.append("integer(c_intptr_t) :: base")
    for sname, (_, fields) in STRUCTS.items():
        body.append(f"print '(a,1x,i0)', '{sname}', c_sizeof(v_{sname})")
        body.append(f"base = transfer(c_loc(v_{sname}), base)")
        for f in fields:
            body.append(f"print '(a,1x,i0)', '{sname}.{f}', transfer(c_loc(v_{sname}%{f}), base) - base")
    body.append("end program sizes")
    prog.write_text("\n".join(body))
    exe = tmp_path / "sizes"
    subprocess.run(["flang", "-module-dir", str(tmp_path), FMOD, str(prog), "-o", str(exe)], check=True,
                   cwd=str(tmp_path))
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {ln.split()[0]: int(ln.split()[1]) for ln in out if ln.strip()}
    for sname, (_, fields) in STRUCTS.items():
        assert got[sname] == lay[sname]["size"], sname
        for f in fields:
            assert got[f"{sname}.{f}"] == lay[sname][f], (sname, f)


def test_fortran_module_binds_every_function():
    txt = open(FMOD).read()
    bound = set(re.findall(r'bind\(C, name="(ty_\w+)"\)', txt))
    assert bound == set(header_functions())


def test_package_refuses_to_run_without_library(monkeypatch, tmp_path):
    """The product path has no fallback: a missing library is a loud error."""
    monkeypatch.setattr(_lib, "LIBPATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError):
        _lib.load()


def test_bench_refuses_more_than_one_gpu():
    """bench.py --gpus N > 1 stops with a message before touching a GPU: nothing here shards"""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run(["python3", os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
