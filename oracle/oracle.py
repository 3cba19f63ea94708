"""ctypes wrapper of the oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker of the HIP path.  The product
(typhon_amd/) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

I32P = C.POINTER(C.c_int32)
F64P = C.POINTER(C.c_double)


class OrBoco(C.Structure):
    _fields_ = [("type", C.c_int32), ("nface", C.c_int32), ("iface", I32P), ("ipart", I32P),
                ("param", C.c_double * 8)]


class OrConfig(C.Structure):
    _fields_ = [("dim", C.c_int32), ("ncell_int", C.c_int32), ("ncell", C.c_int32),
                ("nface", C.c_int32), ("nface_int", C.c_int32),
                ("ic1", I32P), ("ic2", I32P), ("fnorm", F64P), ("fcent", F64P), ("fsurf", F64P),
                ("ccent", F64P), ("cvol", F64P), ("nboco", C.c_int32), ("boco", C.POINTER(OrBoco)),
                ("gamma", C.c_double), ("rgas", C.c_double), ("prandtl", C.c_double),
                ("visc_law", C.c_int32), ("visc_dyn", C.c_double), ("visc_kin", C.c_double),
                ("tref", C.c_double), ("tsuth", C.c_double), ("viscous", C.c_int32), ("sch_dis", C.c_int32),
                ("sch_hyp", C.c_int32), ("method", C.c_char), ("limiter", C.c_char),
                ("postlimiter", C.c_int32), ("gradmeth", C.c_int32), ("venkat", C.c_int32),
                ("venkat_k", C.c_double), ("nstage", C.c_int32), ("coef", C.c_double * 64),
                ("cfl", C.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB):
        build()
    lib = C.CDLL(_LIB)
    lib.or_threads.restype = C.c_int
    lib.or_threads.argtypes = []
    lib.or_set_threads.argtypes = [C.c_int]
    lib.or_create.restype = C.c_void_p
    lib.or_create.argtypes = [C.POINTER(OrConfig)]
    lib.or_destroy.argtypes = [C.c_void_p]
    lib.or_set_cons.argtypes = [C.c_void_p, F64P]
    lib.or_get_cons.argtypes = [C.c_void_p, F64P]
    lib.or_timestep.restype = C.c_double
    lib.or_timestep.argtypes = [C.c_void_p, C.c_double, C.c_double]
    lib.or_calc_rhs.argtypes = [C.c_void_p, C.c_double, F64P]
    lib.or_rk_step.argtypes = [C.c_void_p, C.c_double]
    lib.or_rk_step_local.argtypes = [C.c_void_p]
    lib.or_residual_norm.restype = C.c_double
    lib.or_residual_norm.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
    lib.or_refcons.argtypes = [C.c_void_p, F64P]
    for name in ("or_prim", "or_grad", "or_face_left", "or_face_right", "or_flux"):
        getattr(lib, name).restype = F64P
        getattr(lib, name).argtypes = [C.c_void_p]
    lib.or_flux_face.argtypes = [C.POINTER(OrConfig), F64P, F64P, F64P, F64P]
    lib.or_limiter.restype = C.c_double
    lib.or_limiter.argtypes = [C.c_char, C.c_double, C.c_double]
    lib.or_venkat_phi.restype = C.c_double
    lib.or_venkat_phi.argtypes = [C.c_double] * 4
    lib.or_viscosity.restype = C.c_double
    lib.or_viscosity.argtypes = [C.POINTER(OrConfig), C.c_double, C.c_double]
    lib.or_jac_hll.argtypes = [C.POINTER(OrConfig), F64P, F64P, F64P, F64P, F64P]
    lib.or_jac_diag.argtypes = [C.c_void_p, C.c_double, F64P]
    lib.or_bdlu_assemble.argtypes = [C.c_void_p, C.c_double, F64P, F64P, F64P]
    lib.or_bdlu_gmres.restype = C.c_int
    lib.or_bdlu_gmres.argtypes = [C.c_void_p, F64P, F64P, F64P, F64P, C.c_int, C.c_double, C.c_int]
    lib.or_implicit_bdlu_step.restype = C.c_int
    lib.or_implicit_bdlu_step.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_int]
    lib.or_res_all.restype = F64P
    lib.or_res_all.argtypes = [C.c_void_p]
    lib.or_dtloc.restype = F64P
    lib.or_dtloc.argtypes = [C.c_void_p]
    lib.or_meshgeom.restype = C.c_int
    lib.or_meshgeom.argtypes = [C.c_int, C.c_int, F64P, C.c_int, I32P, I32P, I32P, C.c_int,
                                F64P, F64P, F64P, F64P, F64P]
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _p64(a):
    return a.ctypes.data_as(F64P)


def _p32(a):
    return a.ctypes.data_as(I32P)


def make_config(mesh=None, scheme=None, gas=None) -> tuple:
    """OrConfig from a typhon_amd.mesh.UstMesh-like object and dicts; returns
    (config, keepalive)."""
    keep = []
    cfg = OrConfig()
    g = dict(gamma=1.4, rgas=287.14, prandtl=0.72, visc_law=13, visc_dyn=0.0, visc_kin=0.0,
             tref=273.0, tsuth=110.5, viscous=0, sch_dis=1)
    if gas:
        g.update(gas)
    for k, v in g.items():
        setattr(cfg, k, v)
    s = dict(sch_hyp=25, method="N", limiter="X", postlimiter=0, gradmeth=5, venkat=0, venkat_k=5.0,
             nstage=2, coef=None, cfl=0.8)
    if scheme:
        s.update(scheme)
    cfg.sch_hyp = s["sch_hyp"]
    cfg.method = s["method"].encode()
    cfg.limiter = s["limiter"].encode()
    cfg.postlimiter = s["postlimiter"]
    cfg.gradmeth = s["gradmeth"]
    cfg.venkat = s["venkat"]
    cfg.venkat_k = s["venkat_k"]
    cfg.cfl = s["cfl"]
    if s["coef"] is not None:
        coef = np.asarray(s["coef"], dtype=np.float64)
        ns = coef.shape[0]
        cfg.nstage = ns
        for i in range(ns):
            for j in range(ns):
                cfg.coef[i * 8 + j] = coef[i, j]
    if mesh is not None:
        m = mesh
        arrs = dict(ic1=np.ascontiguousarray(m.ic1, dtype=np.int32), ic2=np.ascontiguousarray(m.ic2, dtype=np.int32),
                    fnorm=np.ascontiguousarray(m.fnorm), fcent=np.ascontiguousarray(m.fcent),
                    fsurf=np.ascontiguousarray(m.fsurf), ccent=np.ascontiguousarray(m.ccent),
                    cvol=np.ascontiguousarray(m.cvol))
        keep.append(arrs)
        cfg.dim, cfg.ncell_int, cfg.ncell = m.dim, m.ncell_int, m.ncell
        cfg.nface, cfg.nface_int = m.nface, m.nface_int
        cfg.ic1, cfg.ic2 = _p32(arrs["ic1"]), _p32(arrs["ic2"])
        cfg.fnorm, cfg.fcent, cfg.fsurf = _p64(arrs["fnorm"]), _p64(arrs["fcent"]), _p64(arrs["fsurf"])
        cfg.ccent, cfg.cvol = _p64(arrs["ccent"]), _p64(arrs["cvol"])
        bocos = (OrBoco * max(1, len(m.bocos)))()
        for k, b in enumerate(m.bocos):
            fi = np.ascontiguousarray(b.iface, dtype=np.int32)
            keep.append(fi)
            bocos[k].type = b.type
            bocos[k].nface = fi.shape[0]
            bocos[k].iface = _p32(fi)
            if b.ipart is not None:
                pi = np.ascontiguousarray(b.ipart, dtype=np.int32)
                keep.append(pi)
                bocos[k].ipart = _p32(pi)
            for j in range(8):
                bocos[k].param[j] = float(b.param[j])
        keep.append(bocos)
        cfg.nboco = len(m.bocos)
        cfg.boco = bocos
    return cfg, keep


class Oracle:
    """One oracle context over a mesh (reference form) and a scheme."""

    def __init__(self, mesh, scheme=None, gas=None):
        self.mesh = mesh
        self.cfg, self._keep = make_config(mesh, scheme, gas)
        self.lib = lib()
        self.h = self.lib.or_create(C.byref(self.cfg))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.or_destroy(self.h)
            self.h = None

    def set_cons(self, cons):
        c = np.ascontiguousarray(cons, dtype=np.float64).reshape(5, self.mesh.ncell_int)
        self.lib.or_set_cons(self.h, _p64(c))

    def get_cons(self):
        out = np.empty((5, self.mesh.ncell_int))
        self.lib.or_get_cons(self.h, _p64(out))
        return out

    def timestep(self, cfl, dtmax=np.inf):
        return self.lib.or_timestep(self.h, cfl, dtmax)

    def calc_rhs(self, dt):
        out = np.empty((5, self.mesh.ncell_int))
        self.lib.or_calc_rhs(self.h, dt, _p64(out))
        return out

    def rk_step(self, dt):
        self.lib.or_rk_step(self.h, dt)

    def rk_step_local(self):
        """RK step with the per-cell time steps of the last timestep() call (loc_stab_cond)"""
        self.lib.or_rk_step_local(self.h)

    def local_dt(self):
        """per-cell time steps of the last timestep() call [ncell_int]"""
        return np.ctypeslib.as_array(self.lib.or_dtloc(self.h), shape=(self.mesh.ncell_int,)).copy()

    def residual_norm(self, refsca0, refsca1, refvec):
        return self.lib.or_residual_norm(self.h, refsca0, refsca1, refvec)

    def refcons(self):
        out = np.empty(3)
        self.lib.or_refcons(self.h, _p64(out))
        return out

    def view(self, name, nvar, n):
        p = getattr(self.lib, name)(self.h)
        return np.ctypeslib.as_array(p, shape=(nvar * n,)).reshape(nvar, n).copy()

    def prim(self):
        return self.view("or_prim", 5, self.mesh.ncell)

    def grad(self):
        return self.view("or_grad", 15, self.mesh.ncell)

    def face_states(self):
        return self.view("or_face_left", 5, self.mesh.nface), self.view("or_face_right", 5, self.mesh.nface)

    def flux(self):
        return self.view("or_flux", 5, self.mesh.nface)

    def jac_diag(self, dt):
        """diagonal blocks [ncell_int, 5, 5] of I + dt/V dF.A/dU after calc_rhs(dt)"""
        out = np.empty((self.mesh.ncell_int, 5, 5))
        self.lib.or_jac_diag(self.h, dt, _p64(out))
        return out

    # A15: the reference's assembled implicit path
    def bdlu_assemble(self, dt):
        """diag [ncell, 25], upper/lower [nface, 25] (row-major 5x5 blocks) after calc_rhs(dt)"""
        m = self.mesh
        d = np.empty((m.ncell, 25))
        up = np.empty((m.nface, 25))
        lo = np.empty((m.nface, 25))
        self.lib.or_bdlu_assemble(self.h, dt, _p64(d), _p64(up), _p64(lo))
        return d, up, lo

    def bdlu_gmres(self, diag, upper, lower, rhs, nkrylov=50, maxres=1e-4, max_it=50):
        """bdlu_gmres on rhs [ncell, 5] (block-major); returns (solution [ncell, 5], info)"""
        sol = np.ascontiguousarray(rhs, dtype=np.float64).copy()
        info = self.lib.or_bdlu_gmres(self.h, _p64(np.ascontiguousarray(diag)), _p64(np.ascontiguousarray(upper)),
                                      _p64(np.ascontiguousarray(lower)), _p64(sol), nkrylov, maxres, max_it)
        return sol, info

    def implicit_bdlu_step(self, dt, nkrylov=50, maxres=1e-4, max_it=50):
        return self.lib.or_implicit_bdlu_step(self.h, dt, nkrylov, maxres, max_it)

    def res_all(self):
        """residual of every cell after calc_rhs, ghost rows included [5, ncell]"""
        return self.view("or_res_all", 5, self.mesh.ncell)


def flux_face(qL, qR, n, sch_hyp=25, gamma=1.4):
    cfg, _ = make_config(None, dict(sch_hyp=sch_hyp), dict(gamma=gamma))
    out = np.empty(5)
    a, b, c = (np.ascontiguousarray(x, dtype=np.float64) for x in (qL, qR, n))
    lib().or_flux_face(C.byref(cfg), _p64(a), _p64(b), _p64(c), _p64(out))
    return out


def jac_hll(qL, qR, n, sch_hyp=25, gamma=1.4):
    cfg, _ = make_config(None, dict(sch_hyp=sch_hyp), dict(gamma=gamma))
    jl, jr = np.empty(25), np.empty(25)
    a, b, c = (np.ascontiguousarray(x, dtype=np.float64) for x in (qL, qR, n))
    lib().or_jac_hll(C.byref(cfg), _p64(a), _p64(b), _p64(c), _p64(jl), _p64(jr))
    return jl.reshape(5, 5), jr.reshape(5, 5)


def limiter(lim, x1, x2):
    return lib().or_limiter(lim.encode(), x1, x2)


def venkat_phi(dmax, dmin, d2, eps2):
    return lib().or_venkat_phi(dmax, dmin, d2, eps2)


def viscosity(rho, T, **gas):
    cfg, _ = make_config(None, None, gas)
    return lib().or_viscosity(C.byref(cfg), rho, T)


def meshgeom(dim, vtx, fvtx, ic1, ic2, ncell_int):
    """Run the C restatement of MESHGEOM on face vertex lists (ic2 = -1 for boundary)."""
    vtx = np.ascontiguousarray(vtx, dtype=np.float64)
    fvtx = np.ascontiguousarray(fvtx, dtype=np.int32)
    a = np.ascontiguousarray(ic1, dtype=np.int32).copy()
    b = np.ascontiguousarray(ic2, dtype=np.int32).copy()
    nf = fvtx.shape[0]
    fn, fc, fs = np.empty((nf, 3)), np.empty((nf, 3)), np.empty(nf)
    cc, cv = np.empty((ncell_int, 3)), np.empty(ncell_int)
    lib().or_meshgeom(dim, vtx.shape[0], _p64(vtx), nf, _p32(fvtx), _p32(a), _p32(b), ncell_int,
                      _p64(fn), _p64(fc), _p64(fs), _p64(cc), _p64(cv))
    return dict(ic1=a, ic2=b, fnorm=fn, fcent=fc, fsurf=fs, ccent=cc, cvol=cv)
