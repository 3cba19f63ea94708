"""Resolved scheme identifiers and Runge-Kutta tableaux.

Integer and character IDs are the ones of SOURCE/PARAM/MENU_NUM.f90:24-147 (the
values def_spat / def_time resolve from main.rpm); the tableaux restate
init_rungekutta (MENU_NUM.f90:368-449) as rk%coef(1:nstage, 1:nstage).
"""
from __future__ import annotations

import numpy as np

# time integration (MENU_NUM.f90:29-39)
TPS_EXPL, TPS_RK2, TPS_RK2SSP, TPS_RK3SSP, TPS_RK4 = 10, 30, 31, 33, 35
TPS_LSRK25BB, TPS_LSRK26BB, TPS_LSRK12BS, TPS_LSRK13BS = 41, 42, 51, 52
# numerical fluxes (MENU_NUM.f90:42-63)
SCH_RUSANOV, SCH_ROE, SCH_HLLE, SCH_HLLEB, SCH_HLLC, SCH_HLLCB = 5, 10, 21, 22, 25, 26
SCH_HLLEK, SCH_HLLEKB, SCH_HLLCK, SCH_HLLCKB = 23, 24, 27, 28  # kinetic wave speeds (MENU_NUM.f90:49-54)
# gradients (MENU_NUM.f90:128-132)
GRAD_NONE, GRAD_LSQ, GRAD_LSQW = 5, 10, 12
# post-limiters (MENU_NUM.f90:95-100)
POSTLIM_NONE, POSTLIM_BARTH, POSTLIM_SUPERBARTH = 0, 20, 21
POSTLIM_MONOTONIC0, POSTLIM_MONOTONIC1, POSTLIM_MONOTONIC2 = 10, 11, 12
# dissipative fluxes (MENU_NUM.f90:120-125)
DIS_NOFLUX, DIS_CELLAVG2, DIS_CELLFULL = 1, 5, 10
DIS_COMPACT, DIS_FACECENTERED, DIS_FACEPENALTY = 2, 20, 22  # refused for NS, as by the reference
# fluid models (EQNS.f90:27-32) and viscosity laws (EQNS.f90:39-43)
EQ_EULER, EQ_NSLAM = 10, 15
VISC_SUTH, VISC_DYNCST, VISC_KINCST = 11, 13, 14
# precision forms
PREC_FP64, PREC_MIXED, PREC_MIXED_F32SUM = 0, 1, 2


def rk_tableau(method: int) -> np.ndarray:
    """rk%coef for a time-integration method (init_rungekutta, MENU_NUM.f90:368-449)."""
    if method == TPS_EXPL:
        return np.array([[1.0]])
    if method == TPS_RK2:
        c = np.zeros((2, 2)); c[0, 0] = .5; c[1, :] = (0.0, 1.0); return c
    if method == TPS_RK2SSP:
        c = np.zeros((2, 2)); c[0, 0] = 1.0; c[1, :] = (0.5, 0.5); return c
    if method == TPS_RK3SSP:
        c = np.zeros((3, 3)); c[0, 0] = 1.0; c[1, :2] = (0.25, 0.25); c[2, :] = np.array((1.0, 1.0, 4.0)) / 6.0
        return c
    if method == TPS_RK4:
        c = np.zeros((4, 4)); c[0, 0] = .5; c[1, :2] = (0.0, .5); c[2, :3] = (0.0, 0.0, 1.0)
        c[3, :] = np.array((1.0, 2.0, 2.0, 1.0)) / 6.0
        return c
    if method == TPS_LSRK25BB:
        return np.diag([0.1815754863270908, 0.238260222208392, 0.330500707328, 0.5, 1.0])
    if method == TPS_LSRK26BB:
        return np.diag([0.11797990162882, 0.18464696649448, 0.24662360430959, 0.33183954253762, 0.5, 1.0])
    if method == TPS_LSRK12BS:
        return np.diag([1.0, 1.0])
    if method == TPS_LSRK13BS:
        return np.diag([.15, .4, 1.0])
    raise ValueError(f"unknown RUNGE KUTTA method {method}")


def default_gas(**kw) -> dict:
    """Perfect gas; GAS=AIR constants of def_model_ns.f90:148-176."""
    g = dict(gamma=1.4, r_const=287.14, prandtl=0.72, typ_visc=VISC_DYNCST, visc_dyn=0.0, visc_kin=0.0,
             tref=273.0, tsuth=110.5, typ_fluid=EQ_EULER)
    g.update(kw)
    return g


def default_scheme(**kw) -> dict:
    s = dict(sch_hyp=SCH_HLLC, sch_dis=DIS_NOFLUX, method="N", limiter="X", gradmeth=GRAD_NONE,
             postlimiter=POSTLIM_NONE, venkat_k=5.0, tps_meth=TPS_RK2, cfl=0.8, precision=PREC_FP64)
    s.update(kw)
    if s["method"] != "N" and s["gradmeth"] == GRAD_NONE:
        s["gradmeth"] = GRAD_LSQ
    return s
