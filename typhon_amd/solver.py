"""Host-side mirror of Typhon's EQNS residual / time-integration interface.

``Solver`` takes the mesh as Typhon holds it after init_world (a
``typhon_amd.mesh.UstMesh``) and the resolved solver menus, and drives the
MI355X residual path through the C-ABI (include/typhon_gpu.h):

    calc_rhs(dt)      -> SOURCE/MAIN/calc_rhs.f90:10 (residual -dt/V sum F.A)
    timestep(dtmax)   -> SOURCE/ZONE/calc_zonetimestep.f90:57 (stab_cond)
    rk_step(dt)       -> SOURCE/MAIN/treelevel_rungekutta.f90:8 + update_field
    advance(n, dtmax) -> n device-resident cycles of the two above

Conservative state arrays are [5, ncell_int] (rho, rhoE, rho u, rho v, rho w)
in the caller's cell order.  Errors from the library raise TyError, mirroring
the reference's error_stop.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .scheme import EQ_NSLAM, PREC_FP64, default_gas, default_scheme, rk_tableau


def _p64(a):
    return a.ctypes.data_as(_lib.F64P)


def _p32(a):
    return a.ctypes.data_as(_lib.I32P)


class Solver:
    CELL_ORDER = {"rcm": 0, "hilbert": 1, "none": 2}

    def __init__(self, mesh, scheme: dict | None = None, gas: dict | None = None, device: int = 0,
                 cell_order: str = "hilbert"):
        self.lib = _lib.load()
        self.mesh = mesh
        self.ni = mesh.ncell_int
        self.h = C.c_void_p()
        self._check(self.lib.ty_create(C.byref(self.h), device), "ty_create")
        self._upload_mesh(mesh, self.CELL_ORDER[cell_order])
        self.set_scheme(scheme or {}, gas or {})

    # -- plumbing ---------------------------------------------------------
    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.ty_last_error(self.h).decode() if self.h else what
            raise _lib.TyError(rc, f"{what}: {msg}")

    def close(self):
        if self.h:
            self.lib.ty_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _upload_mesh(self, m, cell_order):
        fc = np.ascontiguousarray(np.concatenate([m.ic1 + 1, m.ic2 + 1]).astype(np.int32))
        face = np.ascontiguousarray(np.column_stack([m.fnorm, m.fcent, m.fsurf]))
        centre = np.ascontiguousarray(m.ccent, dtype=np.float64)
        volume = np.ascontiguousarray(m.cvol, dtype=np.float64)
        keep = [fc, face, centre, volume]
        bocos = (_lib.TyBocoDesc * max(1, len(m.bocos)))()
        for k, b in enumerate(m.bocos):
            fi = np.ascontiguousarray(b.iface + 1, dtype=np.int32)
            keep.append(fi)
            bocos[k].type = b.type
            bocos[k].nface = fi.shape[0]
            bocos[k].iface = _p32(fi)
            if b.ipart is not None:
                pi = np.ascontiguousarray(b.ipart + 1, dtype=np.int32)
                keep.append(pi)
                bocos[k].ipart = _p32(pi)
            for j in range(8):
                bocos[k].param[j] = float(b.param[j])
        d = _lib.TyMeshDesc(m.dim, m.ncell_int, m.ncell, m.nface, m.nface_int, _p32(fc), _p64(face),
                            _p64(centre), _p64(volume), len(m.bocos), bocos, cell_order)
        self._check(self.lib.ty_mesh_upload(self.h, C.byref(d)), "ty_mesh_upload")
        del keep

    def upload(self, mesh, cell_order: str = "hilbert"):
        """replace the mesh (e.g. the same mesh with moved vertices or other boundary parameters)
        and set the current scheme again; the state must be set again afterwards"""
        self._upload_mesh(mesh, self.CELL_ORDER[cell_order])
        self.mesh = mesh
        self.ni = mesh.ncell_int
        self.set_scheme(self.scheme, self.gas)

    def set_scheme(self, scheme: dict, gas: dict):
        s = default_scheme(**scheme)
        g = default_gas(**gas)
        coef = s.get("rk_coef")
        coef = np.asarray(rk_tableau(s["tps_meth"]) if coef is None else coef, dtype=np.float64)
        self.nstage = coef.shape[0]
        cm = np.ascontiguousarray(coef.T)  # column-major rk%coef
        sd = _lib.TyScheme(s["sch_hyp"], s["sch_dis"], s["method"].encode(), s["limiter"].encode(),
                           s["gradmeth"], s["postlimiter"], s["venkat_k"], self.nstage, _p64(cm), s["cfl"],
                           s["precision"])
        gd = _lib.TyGas(g["gamma"], g["r_const"], g["prandtl"], g["typ_visc"], g["visc_dyn"], g["visc_kin"],
                        g["tref"], g["tsuth"], g["typ_fluid"])
        self._check(self.lib.ty_set_scheme(self.h, C.byref(gd), C.byref(sd)), "ty_set_scheme")
        self.scheme, self.gas = s, g

    @property
    def viscous(self) -> bool:
        """Navier-Stokes residual (the viscous flux reads cell gradients)"""
        return self.gas["typ_fluid"] == EQ_NSLAM

    # -- reference interface ----------------------------------------------
    def set_state(self, cons):
        cons = np.asarray(cons, dtype=np.float64).reshape(5, self.ni)
        rho = np.ascontiguousarray(cons[0])
        rhoe = np.ascontiguousarray(cons[1])
        mom = np.ascontiguousarray(cons[2:5].T)
        self._check(self.lib.ty_state_set(self.h, _p64(rho), _p64(rhoe), _p64(mom)), "ty_state_set")

    def get_state(self):
        rho, rhoe, mom = np.empty(self.ni), np.empty(self.ni), np.empty((self.ni, 3))
        self._check(self.lib.ty_state_get(self.h, _p64(rho), _p64(rhoe), _p64(mom)), "ty_state_get")
        return np.vstack([rho, rhoe, mom.T])

    def timestep(self, dtmax=np.inf):
        dt = C.c_double()
        self._check(self.lib.ty_timestep(self.h, dtmax, C.byref(dt)), "ty_timestep")
        return dt.value

    def calc_rhs(self, dt):
        rho, rhoe, mom = np.empty(self.ni), np.empty(self.ni), np.empty((self.ni, 3))
        self._check(self.lib.ty_residual(self.h, dt, _p64(rho), _p64(rhoe), _p64(mom)), "ty_residual")
        return np.vstack([rho, rhoe, mom.T])

    def rk_step(self, dt):
        self._check(self.lib.ty_rk_step(self.h, dt), "ty_rk_step")

    def advance_fp64(self, U0, dts):
        """state [5, ncell_int] after len(dts) FP64 RK steps from U0 with the given time steps; a
        mixed-precision solver takes its FP64 form for the sweep and returns to its own form.  The
        solver's state is U0 on return."""
        form = self.scheme["precision"]
        if form != PREC_FP64:
            self.set_scheme(dict(self.scheme, precision=PREC_FP64), self.gas)
        self.set_state(U0)
        for dt in dts:
            self.rk_step(float(dt))
        UN = self.get_state()
        if form != PREC_FP64:
            self.set_scheme(dict(self.scheme, precision=form), self.gas)
        self.set_state(U0)
        return UN

    def advance(self, nsteps, dtmax=np.inf, want_dt=False):
        if want_dt:
            dt = C.c_double()
            self._check(self.lib.ty_advance(self.h, nsteps, dtmax, C.byref(dt)), "ty_advance")
            return dt.value
        self._check(self.lib.ty_advance(self.h, nsteps, dtmax, None), "ty_advance")
        return None

    # -- steady mode (time_steady) ----------------------------------------
    def refcons(self):
        """calc_refcons + define_refcons at the current state: (rho_ref, rhoE_ref, momentum ref)"""
        out = np.empty(3)
        self._check(self.lib.ty_refcons(self.h, _p64(out)), "ty_refcons")
        return out

    def local_timestep(self):
        """per-cell time steps of loc_stab_cond [ncell_int] and their minimum"""
        dt = np.empty(self.ni)
        dmin = C.c_double()
        self._check(self.lib.ty_local_timestep(self.h, _p64(dt), C.byref(dmin)), "ty_local_timestep")
        return dt, dmin.value

    def steady_cycle(self, maxit, cfl=None, cfl_max=None, maxres=1e-6, local_dt=True, residu_ref=None,
                     wlog=0.0, ref=None, check_every=16):
        """one steady cycle (integration_cyclezone for time_steady) on the device: returns (history of
        cur_res per iteration, stats dict).  residu_ref defaults to the first cycle's value (tiny), ref
        to refcons() at the current state."""
        ref = self.refcons() if ref is None else np.asarray(ref, dtype=np.float64)
        cfl = self.scheme["cfl"] if cfl is None else cfl
        st = _lib.TySteady(1 if local_dt else 0, maxit, check_every, cfl, cfl if cfl_max is None else cfl_max, maxres,
                           np.finfo(float).tiny if residu_ref is None else residu_ref, wlog,
                           (C.c_double * 2)(ref[0], ref[1]), ref[2])
        hist = np.zeros(maxit)
        out = _lib.TySteadyStats()
        self._check(self.lib.ty_steady_cycle(self.h, C.byref(st), _p64(hist), C.byref(out)), "ty_steady_cycle")
        stats = dict(iterations=out.iterations, converged=bool(out.converged), cur_res=out.cur_res,
                     residu_ref=out.residu_ref, cfl=out.cfl)
        return hist[:out.iterations], stats

    # -- volume probes (BLOCK:PROBE) ------------------------------------------
    PROBE_TYPES = {"VOL_AVERAGE": 21, "VOL_MIN": 22, "VOL_MAX": 23}

    def define_probe(self, quantity, type="VOL_AVERAGE"):
        """compile a probe QUANTITY (FCT expression of RHO, PS, V, V_X, V_Y, V_Z, X, Y, Z, T) of the
        given TYPE (def_probe.f90:34-80); returns its id"""
        code = self.PROBE_TYPES[type.upper()] if isinstance(type, str) else int(type)
        pid = C.c_int32()
        self._check(self.lib.ty_probe_define(self.h, quantity.encode(), code, C.byref(pid)), "ty_probe_define")
        return pid.value

    def probe(self, ids, t=0.0, noalloc=False):
        """the probes ids (an int or a sequence) on the current state at time t (prb_zone_vol);
        noalloc: through ty_probe_eval_noalloc (no device allocation, fixed per-GPU workspace)"""
        scalar = np.isscalar(ids)
        idv = np.ascontiguousarray(np.atleast_1d(ids), dtype=np.int32)
        out = np.empty(idv.size)
        fn, name = ((self.lib.ty_probe_eval_noalloc, "ty_probe_eval_noalloc") if noalloc
                    else (self.lib.ty_probe_eval, "ty_probe_eval"))
        self._check(fn(self.h, float(t), idv.size, _p32(idv), _p64(out)), name)
        return float(out[0]) if scalar else out

    def sync(self):
        self._check(self.lib.ty_sync(self.h), "ty_sync")

    def timing(self, enable=True):
        ms = (C.c_double * 5)()
        cnt = (C.c_int32 * 5)()
        self._check(self.lib.ty_timing(self.h, 1 if enable else 0, ms, cnt), "ty_timing")
        return dict(grad_ms=ms[0], flux_ms=ms[1], dt_ms=ms[2], tangent_ms=ms[3], adjoint_ms=ms[4],
                    grad_n=cnt[0], flux_n=cnt[1], dt_n=cnt[2], tangent_n=cnt[3], adjoint_n=cnt[4])

    # -- implicit path -----------------------------------------------------
    def jacobian_diag(self, dt):
        """diagonal blocks [ncell_int, 5, 5] of I + dt/V d(sum F.A)/dU (HLL face Jacobians)"""
        out = np.empty((self.ni, 5, 5))
        self._check(self.lib.ty_jacobian_diag(self.h, dt, _p64(out)), "ty_jacobian_diag")
        return out

    def jvp(self, dt, v):
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(5, self.ni)
        out = np.empty((5, self.ni))
        self._check(self.lib.ty_jvp(self.h, dt, _p64(v), _p64(out)), "ty_jvp")
        return out

    def residual_tangent(self, dt, du=None, ds=None, dvol=None, dq=None):
        """exact directional derivative of the first-order residual at the current state and a
        fixed dt: du [5, ncell_int], ds [nface, 3] face area vectors S = A n, dvol [ncell_int],
        dq [nboco, 5] supersonic-inlet ghost state (rho, p, u, v, w); absent = zero.  Viscous
        residuals take no ds (their flux also reads cell and face centres)."""
        m = self.mesh

        def arr(x, shape):
            return None if x is None else np.ascontiguousarray(x, dtype=np.float64).reshape(shape)
        du = arr(du, (5, self.ni))
        ds = arr(ds, (m.nface, 3))
        dvol = arr(dvol, (self.ni,))
        dq = arr(dq, (max(1, len(m.bocos)), 5))
        out = np.empty((5, self.ni))
        p = [None if x is None else _p64(x) for x in (du, ds, dvol, dq)]
        self._check(self.lib.ty_residual_tangent(self.h, dt, *p, _p64(out)), "ty_residual_tangent")
        return out

    def residual_adjoint(self, dt, lam, area_vectors=True, volumes=True):
        """transpose of residual_tangent: dict u [5, ncell_int], s [nface, 3] (None unless
        area_vectors; not provided for viscous residuals), vol [ncell_int] (None unless volumes),
        q [nboco, 5] with the products (dR/d.)^T lam"""
        m = self.mesh
        lam = np.ascontiguousarray(lam, dtype=np.float64).reshape(5, self.ni)
        gu = np.empty((5, self.ni))
        gs = np.empty((m.nface, 3)) if area_vectors else None
        gv = np.empty(self.ni) if volumes else None
        gq = np.empty((max(1, len(m.bocos)), 5))
        self._check(self.lib.ty_residual_adjoint(self.h, dt, _p64(lam), _p64(gu), None if gs is None else _p64(gs),
                                                 None if gv is None else _p64(gv), _p64(gq)), "ty_residual_adjoint")
        return dict(u=gu, s=gs, vol=gv, q=gq[:len(m.bocos)])

    def unsteady_adjoint(self, dts, lamN, snaps=20):
        """dJ/dU_0 and dJ/dq (inlet ghost states) of an objective of the state after len(dts) RK
        steps from the current state, given lamN = dJ/dU_N [5, ncell_int]; checkpoints in device
        memory (at most ``snaps`` besides U_0).  Returns dict(u, q, stats); the state is U_0 again."""
        m = self.mesh
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        lamN = np.ascontiguousarray(lamN, dtype=np.float64).reshape(5, self.ni)
        lam0 = np.empty((5, self.ni))
        gq = np.empty((max(1, len(m.bocos)), 5))
        st = _lib.TyUnsteadyStats()
        self._check(self.lib.ty_unsteady_adjoint(self.h, dts.shape[0], _p64(dts), snaps, _p64(lamN), _p64(lam0),
                                                 _p64(gq), C.byref(st)), "ty_unsteady_adjoint")
        stats = dict(forward_steps=st.forward_steps, adjoint_steps=st.adjoint_steps,
                     max_snapshots=st.max_snapshots, snapshots_allocated=st.snapshots_allocated, ms=st.ms)
        return dict(u=lam0, q=gq[:len(m.bocos)], stats=stats)

    def linear_solve(self, dt, b, kry=None, allow_unconverged=False):
        """solve (I - J) x = b at the current state; returns (x, stats)"""
        kry = kry or krylov()
        b = np.ascontiguousarray(b, dtype=np.float64).reshape(5, self.ni)
        x = np.empty((5, self.ni))
        st = _lib.TyKrylovStats()
        rc = self.lib.ty_linear_solve(self.h, dt, C.byref(kry), _p64(b), _p64(x), C.byref(st))
        if not (allow_unconverged and rc == 5):
            self._check(rc, "ty_linear_solve")
        return x, _stats(st)

    def implicit_step(self, dt, kry=None):
        kry = kry or krylov()
        st = _lib.TyKrylovStats()
        self._check(self.lib.ty_implicit_step(self.h, dt, C.byref(kry), C.byref(st)), "ty_implicit_step")
        return _stats(st)

    def implicit_bdlu_step(self, dt, nkrylov=50, maxres=1e-4, max_it=50):
        """one implicit step of the reference's assembled path (build_implicit_bdlu + bdlu_gmres,
        INV_RES = maxres, NKRYLOV, IMP_MAX_IT = max_it); returns the GMRES info"""
        info = C.c_int32()
        self._check(self.lib.ty_implicit_bdlu_step(self.h, dt, nkrylov, maxres, max_it, C.byref(info)),
                    "ty_implicit_bdlu_step")
        return info.value

    def mesh_info(self):
        v = [C.c_int32() for _ in range(4)]
        self._check(self.lib.ty_mesh_info(self.h, *[C.byref(x) for x in v]), "ty_mesh_info")
        return dict(ncell=v[0].value, nface=v[1].value, nslot=v[2].value, nbface=v[3].value)

    def cell_order(self):
        """the device cell order: the caller's (0-based) cell index of each device cell"""
        out = np.empty(self.ni, dtype=np.int32)
        self._check(self.lib.ty_cell_order(self.h, _p32(out)), "ty_cell_order")
        return out - 1


def krylov(nkrylov=30, max_restart=4, max_outer=8, inner_precision=1, precond=1, inner_tol=1e-3, outer_tol=1e-8):
    """Settings of the matrix-free implicit solve (ty_krylov): GMRES(nkrylov) with block-Jacobi,
    FP32 inner iterations (inner_precision=1) inside FP64 refinement to outer_tol."""
    return _lib.TyKrylov(nkrylov, max_restart, max_outer, inner_precision, precond, inner_tol, outer_tol)


def _stats(st):
    return dict(outer_iters=st.outer_iters, inner_iters=st.inner_iters, rel_res=st.rel_res, rhs_norm=st.rhs_norm)


def prim_to_cons(rho, p, vel, gamma=1.4):
    """calc_varcons_ns (SOURCE/EQNS/calc_varcons_ns.f90:35-45) on arrays: vel [3, n]."""
    vel = np.asarray(vel, dtype=np.float64)
    mom = rho * vel
    rhoe = p / (gamma - 1.0) + .5 * rho * np.einsum("ij,ij->j", vel, vel)
    return np.vstack([rho, rhoe, mom])
