"""Replay of the reference's NRG TIME/RUNGEKUTTA cases (driver logic of
SOURCE/MAIN/integration.f90:100-130 and integration_cyclezone.f90:80-167):
cycles of dtbase = DURATION/NCYCLE, inside each cycle global-dt RK steps with
dt = min(dt_CFL, remaining cycle time), then the VOL_AVERAGE probe
(CFDTOOLS/Mesh/PROBECALC.f90:66-90) on the updated state (write_monitors_cycle.f90:45-57)."""
from __future__ import annotations

import os

import numpy as np

from typhon_amd import mesh as M
from typhon_amd import scheme as S
from typhon_amd.solver import prim_to_cons

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = {"RK2": (S.TPS_RK2, 0.8), "RK2-SSP": (S.TPS_RK2SSP, 0.8), "RK3-SSP": (S.TPS_RK3SSP, 1.5),
         "RK4": (S.TPS_RK4, 2.0), "RK25BB": (S.TPS_LSRK25BB, 2.5), "RK26BB": (S.TPS_LSRK26BB, 2.5),
         "RK12BS": (S.TPS_LSRK25BB, 2.5), "RK13BS": (S.TPS_LSRK25BB, 2.5)}


def golden(case):
    return np.loadtxt(os.path.join(HERE, "golden", "nrg_rk", f"{case}.tmon"))


def quad20x20(from_file=True):
    """NRG/COMMON/quad20x20.tym (the unit square, vertices on the 0.05 grid) read from the copy in
    tests/golden/meshes, periodic in x and y as the cases' CONNECTION blocks set it; or its
    synthetic twin (test_tyfmt.py checks that the two agree)"""
    if not from_file:
        return M.periodic_quad(20, 20, 1.0, 1.0, 0.0, 1)
    from typhon_amd import tyfmt
    bocos = [dict(family="IMIN", type=M.BC_PERIOD, partner="IMAX", period=(1.0, 0.0, 0.0)),
             dict(family="IMAX", type=M.BC_PERIOD, partner="IMIN", period=(-1.0, 0.0, 0.0)),
             dict(family="JMIN", type=M.BC_PERIOD, partner="JMAX", period=(0.0, 1.0, 0.0)),
             dict(family="JMAX", type=M.BC_PERIOD, partner="JMIN", period=(0.0, -1.0, 0.0))]
    return tyfmt.tym_ustmesh(os.path.join(HERE, "golden", "meshes", "quad20x20.tym"), bocos)


def setup(case, from_file=True):
    """the case's mesh, scheme, gas and initial state (NRG/TIME/RUNGEKUTTA/<case>/main.rpm)"""
    tps, cfl = CASES[case]
    m = quad20x20(from_file)
    x = m.ccent[:m.ncell_int]
    rho = 1.0 + (np.sin(2 * np.pi * x[:, 0]) * np.sin(2 * np.pi * x[:, 1])) ** 2 / 2.0
    n = m.ncell_int
    vel = np.zeros((3, n))
    vel[0] = 50.0
    cons = prim_to_cons(rho, np.full(n, 100000 / 14.), vel)
    sch = S.default_scheme(sch_hyp=S.SCH_HLLC, method="F", limiter="A", gradmeth=S.GRAD_LSQ, tps_meth=tps, cfl=cfl)
    return m, sch, S.default_gas(), cons


# the cases' BLOCK:PROBE QUANTITY strings (main.rpm), evaluated on the device by ty_probe_eval
RK_PROBE = "sqr(RHO-(1+sqr(sin(2*_pi*(x-50*t))*sin(2*_pi*y))/2))"
VISC_PROBES = ("sqr(V_X-(erf(y/2/sqrt(t+1.))))", "2*50*(1-sqr(V_X))")
MESH_PROBES = {"TYM-QUAD": "PS", "PER2D-TRANS": "(RHO-1.25)^2"}


def rhoerr(m, rho, t):
    x = m.ccent[:m.ncell_int]
    q = (rho - (1 + (np.sin(2 * np.pi * (x[:, 0] - 50 * t)) * np.sin(2 * np.pi * x[:, 1])) ** 2 / 2)) ** 2
    v = m.cvol[:m.ncell_int]
    return (q * v).sum() / v.sum()


def replay(step_fn, dt_fn, get_rho, m, ncycle=100, duration=0.1):
    dtbase = duration / ncycle
    curtps = 0.0
    out = []
    for icycle in range(1, ncycle + 1):
        cycle_time = 0.0
        end = False
        while not end:
            remaining = dtbase - cycle_time
            dt = dt_fn(remaining)
            if dt >= dtbase - cycle_time:
                end = True
                dt = dtbase - cycle_time
            step_fn(dt)
            cycle_time = cycle_time + dt
        curtps = curtps + dtbase
        out.append((icycle, curtps, rhoerr(m, get_rho(), curtps)))
    return np.array(out)


# ---------------------------------------------------------------------------
# NRG/NS/NUM/FLUX/VISCOUS/{CELLFULL,CELLAVG}: the Stokes layer between two moving adiabatic
# walls (main.rpm of both cases): mxlay2x20quad-h50.tym (2 x 20 quads, y in [-50, 50] clustered
# at y = 0), periodic in x (translation 1), ADIABATIC_WALL at JMIN / JMAX moving at -1 / +1,
# laminar air with constant kinematic viscosity 1, HLLC + MUSCL-FAST + van Albada, LSQ
# gradients, RK25BB at CFL 0.5, rho = 1, p = 100/1.4, u = erf(y/2) at t = 0 (the layer's age
# is 1), 10 cycles over a duration of 10.  Probes: VXERR = VOL_AVERAGE (u - erf(y/2/sqrt(t+1)))^2
# and DELTA = VOL_AVERAGE 2*50*(1 - u^2).

VISC_CASES = {"CELLFULL": S.DIS_CELLFULL, "CELLAVG": S.DIS_CELLAVG2}


def golden_visc(case, probe):
    return np.loadtxt(os.path.join(HERE, "golden", "nrg_visc", case, f"{probe}.tmon"))


def mxlay():
    from typhon_amd import tyfmt
    bocos = [dict(family="IMIN", type=M.BC_PERIOD, partner="IMAX", period=(1.0, 0.0, 0.0)),
             dict(family="IMAX", type=M.BC_PERIOD, partner="IMIN", period=(-1.0, 0.0, 0.0)),
             dict(family="JMIN", type=M.BC_WALL_ADIAB, param=[0.0, 0.0, 0.0, 0.0, -1.0, 0.0, 0.0]),
             dict(family="JMAX", type=M.BC_WALL_ADIAB, param=[0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0])]
    return tyfmt.tym_ustmesh(os.path.join(HERE, "golden", "meshes", "mxlay2x20quad-h50.tym"), bocos)


def setup_visc(case):
    from math import erf
    m = mxlay()
    n = m.ncell_int
    y = m.ccent[:n, 1]
    vel = np.zeros((3, n))
    vel[0] = [erf(v / 2.0) for v in y]
    cons = prim_to_cons(np.ones(n), np.full(n, 100 / 1.4), vel)
    sch = S.default_scheme(sch_hyp=S.SCH_HLLC, method="F", limiter="A", gradmeth=S.GRAD_LSQ,
                           tps_meth=S.TPS_LSRK25BB, cfl=0.5, sch_dis=VISC_CASES[case])
    gas = S.default_gas(typ_fluid=S.EQ_NSLAM, typ_visc=S.VISC_KINCST, visc_kin=1.0)
    return m, sch, gas, cons


def visc_probes(m, cons, t):
    from math import erf, sqrt
    n = m.ncell_int
    y = m.ccent[:n, 1]
    v = m.cvol[:n]
    vx = cons[2] / cons[0]
    exact = np.array([erf(a / 2.0 / sqrt(t + 1.0)) for a in y])
    vxerr = ((vx - exact) ** 2 * v).sum() / v.sum()
    delta = (2 * 50 * (1 - vx ** 2) * v).sum() / v.sum()
    return vxerr, delta


def replay_probes(step_fn, dt_fn, probe_fn, ncycle, duration):
    """the cycle loop of replay() with a probe function of the cycle end time"""
    dtbase = duration / ncycle
    curtps = 0.0
    out = []
    for icycle in range(1, ncycle + 1):
        cycle_time = 0.0
        end = False
        while not end:
            dt = dt_fn(dtbase - cycle_time)
            if dt >= dtbase - cycle_time:
                end = True
                dt = dtbase - cycle_time
            step_fn(dt)
            cycle_time = cycle_time + dt
        curtps = curtps + dtbase
        out.append((icycle, curtps, *probe_fn(curtps)))
    return np.array(out)


# ---------------------------------------------------------------------------
# NRG/MESH/ELEM/TYM-QUAD (quad20x10.tym, every boundary SYMMETRY, Euler air, first-order HLLC,
# RK3-SSP at CFL 1; P = 1e5, T = 300, MACH = 0.2 along x; PAVG = VOL_AVERAGE of the static
# pressure, 10 cycles over 5e-3) and NRG/MESH/CON/PER2D-TRANS (quad20x20.tym periodic in x and y,
# Rusanov, RK3-SSP at CFL 1; rho = 1 + 9 x(x-1) y(y-1), p = 1e5/14, u = 10; RHODEV = VOL_AVERAGE
# (rho - 1.25)^2, 100 cycles over 0.1).

MESH_CASES = {"TYM-QUAD": ("PAVG", 10, 0.005), "PER2D-TRANS": ("RHODEV", 100, 0.1)}


def golden_mesh(case):
    return np.loadtxt(os.path.join(HERE, "golden", "nrg_mesh", case, f"{MESH_CASES[case][0]}.tmon"))


def setup_mesh_case(case):
    from typhon_amd import tyfmt
    if case == "TYM-QUAD":
        bocos = [dict(family=f, type=M.BC_SYM) for f in ("IMIN", "IMAX", "JMIN", "JMAX")]
        m = tyfmt.tym_ustmesh(os.path.join(HERE, "golden", "meshes", "quad20x10.tym"), bocos)
        n = m.ncell_int
        g = S.default_gas()
        T, p, mach = 300.0, 1e5, 0.2
        rho = p / (g["r_const"] * T)
        vel = np.zeros((3, n))
        vel[0] = mach * np.sqrt(g["gamma"] * g["r_const"] * T)
        cons = prim_to_cons(np.full(n, rho), np.full(n, p), vel)
        sch = S.default_scheme(sch_hyp=S.SCH_HLLC, tps_meth=S.TPS_RK3SSP, cfl=1.0)
        return m, sch, g, cons
    m = quad20x20()
    n = m.ncell_int
    x = m.ccent[:n]
    rho = 1.0 + 9 * x[:, 0] * (x[:, 0] - 1) * x[:, 1] * (x[:, 1] - 1)
    vel = np.zeros((3, n))
    vel[0] = 10.0
    cons = prim_to_cons(rho, np.full(n, 100000 / 14.), vel)
    sch = S.default_scheme(sch_hyp=S.SCH_RUSANOV, tps_meth=S.TPS_RK3SSP, cfl=1.0)
    return m, sch, S.default_gas(), cons


def mesh_case_probe(case, m, cons):
    n = m.ncell_int
    v = m.cvol[:n]
    if case == "TYM-QUAD":
        g = S.default_gas()["gamma"]
        q = (g - 1) * (cons[1] - 0.5 * (cons[2] ** 2 + cons[3] ** 2 + cons[4] ** 2) / cons[0])
    else:
        q = (cons[0] - 1.25) ** 2
    return ((q * v).sum() / v.sum(),)
