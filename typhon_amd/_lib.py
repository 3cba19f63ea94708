"""ctypes binding of libtyphon_gpu.so (include/typhon_gpu.h).

The library must be built in-tree (python -m typhon_amd.build); loading fails
loudly otherwise -- there is no CPU fallback."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libtyphon_gpu.so")

I32P = C.POINTER(C.c_int32)
F64P = C.POINTER(C.c_double)

EXPORTS = ["ty_create", "ty_destroy", "ty_last_error", "ty_mesh_upload", "ty_set_scheme", "ty_state_set",
           "ty_state_get", "ty_timestep", "ty_residual", "ty_rk_step", "ty_advance", "ty_sync", "ty_timing",
           "ty_mesh_info", "ty_jacobian_diag", "ty_jvp", "ty_linear_solve", "ty_implicit_step",
           "ty_implicit_bdlu_step", "ty_residual_tangent", "ty_residual_adjoint", "ty_unsteady_adjoint",
           "ty_refcons", "ty_local_timestep", "ty_steady_cycle", "ty_probe_define", "ty_probe_eval", "ty_probe_eval_noalloc",
           "ty_cell_order", "ty_meshbuild_create", "ty_meshbuild_destroy", "ty_meshbuild_error", "ty_meshbuild_faces",
           "ty_meshbuild_boundary", "ty_meshbuild_finish"]


class TyBocoDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("nface", C.c_int32), ("iface", I32P), ("ipart", I32P),
                ("param", C.c_double * 8)]


class TyMeshDesc(C.Structure):
    _fields_ = [("dim", C.c_int32), ("ncell_int", C.c_int32), ("ncell", C.c_int32), ("nface", C.c_int32),
                ("nface_int", C.c_int32), ("facecell", I32P), ("face", F64P), ("centre", F64P),
                ("volume", F64P), ("nboco", C.c_int32), ("boco", C.POINTER(TyBocoDesc)),
                ("cell_order", C.c_int32)]


class TyGas(C.Structure):
    _fields_ = [("gamma", C.c_double), ("r_const", C.c_double), ("prandtl", C.c_double),
                ("typ_visc", C.c_int32), ("visc_dyn", C.c_double), ("visc_kin", C.c_double),
                ("tref", C.c_double), ("tsuth", C.c_double), ("typ_fluid", C.c_int32)]


class TyScheme(C.Structure):
    _fields_ = [("sch_hyp", C.c_int32), ("sch_dis", C.c_int32), ("method", C.c_char), ("limiter", C.c_char),
                ("gradmeth", C.c_int32), ("postlimiter", C.c_int32), ("venkat_k", C.c_double),
                ("nstage", C.c_int32), ("rk_coef", F64P), ("cfl", C.c_double), ("precision", C.c_int32)]


class TyKrylov(C.Structure):
    _fields_ = [("nkrylov", C.c_int32), ("max_restart", C.c_int32), ("max_outer", C.c_int32),
                ("inner_precision", C.c_int32), ("precond", C.c_int32), ("inner_tol", C.c_double),
                ("outer_tol", C.c_double)]


class TyKrylovStats(C.Structure):
    _fields_ = [("outer_iters", C.c_int32), ("inner_iters", C.c_int32), ("rel_res", C.c_double),
                ("rhs_norm", C.c_double)]


class TyUnsteadyStats(C.Structure):
    _fields_ = [("forward_steps", C.c_int32), ("adjoint_steps", C.c_int32), ("max_snapshots", C.c_int32),
                ("snapshots_allocated", C.c_int32), ("ms", C.c_double)]


class TySteady(C.Structure):
    _fields_ = [("local_dt", C.c_int32), ("maxit", C.c_int32), ("check_every", C.c_int32), ("cfl", C.c_double),
                ("cfl_max", C.c_double), ("maxres", C.c_double), ("residu_ref", C.c_double), ("wlog", C.c_double),
                ("refsca", C.c_double * 2), ("refvec", C.c_double)]


class TySteadyStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("cur_res", C.c_double),
                ("residu_ref", C.c_double), ("cfl", C.c_double)]


class TyError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"typhon_gpu error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load the in-tree HIP library (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(f"{LIBPATH} is missing: build it with `python -m typhon_amd.build`")
    lib = C.CDLL(LIBPATH)
    vp = C.c_void_p
    lib.ty_create.argtypes = [C.POINTER(vp), C.c_int]
    lib.ty_destroy.argtypes = [vp]
    lib.ty_last_error.restype = C.c_char_p
    lib.ty_last_error.argtypes = [vp]
    lib.ty_mesh_upload.argtypes = [vp, C.POINTER(TyMeshDesc)]
    lib.ty_set_scheme.argtypes = [vp, C.POINTER(TyGas), C.POINTER(TyScheme)]
    lib.ty_state_set.argtypes = [vp, F64P, F64P, F64P]
    lib.ty_state_get.argtypes = [vp, F64P, F64P, F64P]
    lib.ty_timestep.argtypes = [vp, C.c_double, F64P]
    lib.ty_residual.argtypes = [vp, C.c_double, F64P, F64P, F64P]
    lib.ty_rk_step.argtypes = [vp, C.c_double]
    lib.ty_advance.argtypes = [vp, C.c_int32, C.c_double, F64P]
    lib.ty_sync.argtypes = [vp]
    lib.ty_timing.argtypes = [vp, C.c_int32, F64P, I32P]
    lib.ty_mesh_info.argtypes = [vp, I32P, I32P, I32P, I32P]
    lib.ty_cell_order.argtypes = [vp, I32P]
    lib.ty_jacobian_diag.argtypes = [vp, C.c_double, F64P]
    lib.ty_jvp.argtypes = [vp, C.c_double, F64P, F64P]
    lib.ty_linear_solve.argtypes = [vp, C.c_double, C.POINTER(TyKrylov), F64P, F64P, C.POINTER(TyKrylovStats)]
    lib.ty_implicit_step.argtypes = [vp, C.c_double, C.POINTER(TyKrylov), C.POINTER(TyKrylovStats)]
    lib.ty_implicit_bdlu_step.argtypes = [vp, C.c_double, C.c_int32, C.c_double, C.c_int32, C.POINTER(C.c_int32)]
    lib.ty_residual_tangent.argtypes = [vp, C.c_double, F64P, F64P, F64P, F64P, F64P]
    lib.ty_residual_adjoint.argtypes = [vp, C.c_double, F64P, F64P, F64P, F64P, F64P]
    lib.ty_unsteady_adjoint.argtypes = [vp, C.c_int32, F64P, C.c_int32, F64P, F64P, F64P,
                                        C.POINTER(TyUnsteadyStats)]
    lib.ty_refcons.argtypes = [vp, F64P]
    lib.ty_local_timestep.argtypes = [vp, F64P, F64P]
    lib.ty_steady_cycle.argtypes = [vp, C.POINTER(TySteady), F64P, C.POINTER(TySteadyStats)]
    lib.ty_probe_define.argtypes = [vp, C.c_char_p, C.c_int32, I32P]
    lib.ty_probe_eval.argtypes = [vp, C.c_double, C.c_int32, I32P, F64P]
    lib.ty_probe_eval_noalloc.argtypes = [vp, C.c_double, C.c_int32, I32P, F64P]
    lib.ty_meshbuild_create.argtypes = [C.POINTER(vp), C.c_int]
    lib.ty_meshbuild_destroy.argtypes = [vp]
    lib.ty_meshbuild_error.restype = C.c_char_p
    lib.ty_meshbuild_error.argtypes = [vp]
    lib.ty_meshbuild_faces.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, I32P, C.c_int32, F64P, I32P, I32P]
    lib.ty_meshbuild_boundary.argtypes = [vp, I32P, F64P]
    lib.ty_meshbuild_finish.argtypes = [vp, I32P, I32P, I32P, F64P, F64P, F64P]
    for name in EXPORTS:
        if name not in ("ty_last_error", "ty_meshbuild_error"):
            getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib
