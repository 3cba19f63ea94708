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
