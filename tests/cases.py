"""Workloads shared by the tests, the smoke check and the benchmark.

Each builder returns (mesh, scheme, gas, cons0) with the mesh in Typhon's form
(typhon_amd.mesh.UstMesh) and the initial conservative state [5, ncell_int].
"""
from __future__ import annotations

import numpy as np

from typhon_amd import mesh as M
from typhon_amd import scheme as S
from typhon_amd.solver import prim_to_cons

GAMMA = 1.4
R_AIR = 287.14


def freestream_inlet(rho, p, mach, direction=(1.0, 0.0, 0.0), gamma=GAMMA, r=R_AIR):
    """Supersonic-inlet parameters (ptot, ttot, mach, dir) reproducing a given static state
    through pi_ti_mach_dir2nspri (SOURCE/EQNS/EQNS.f90:139-160)."""
    t = p / (rho * r)
    fac = 1.0 + 0.5 * (gamma - 1.0) * mach * mach
    return [p * fac ** (gamma / (gamma - 1.0)), t * fac, mach, *direction]


def channel_case(nx=256, ny=128, seed=1, jitter=0.1, **scheme):
    """2D Euler channel: 1 x 0.5 jittered quads, M=0.5 freestream, Gaussian density bump,
    slip walls (SYMMETRY), supersonic-inlet Dirichlet at x=0, static-pressure outlet at x=1."""
    p0 = 1.0 / GAMMA
    bocos = [dict(family="JMIN", type=M.BC_SYM), dict(family="JMAX", type=M.BC_SYM),
             dict(family="IMIN", type=M.BC_INLET_SUP, param=freestream_inlet(1.0, p0, 0.5)),
             dict(family="IMAX", type=M.BC_OUTLET_SUB, param=[p0])]
    m = M.channel_quad(nx, ny, 1.0, 0.5, jitter, seed, bocos)
    rng_c = M.uniform_pm1(seed, np.array([7, 8]), 2)
    xc = np.array([0.5 + 0.25 * rng_c[0], 0.25 + 0.125 * rng_c[1]])
    x = m.ccent[:m.ncell_int]
    r2 = (x[:, 0] - xc[0]) ** 2 + (x[:, 1] - xc[1]) ** 2
    rho = 1.0 + 0.1 * np.exp(-r2 / 0.05 ** 2)
    n = m.ncell_int
    vel = np.zeros((3, n))
    vel[0] = 0.5
    cons = prim_to_cons(rho, np.full(n, p0), vel)
    sch = dict(sch_hyp=S.SCH_HLLC, method="N", tps_meth=S.TPS_RK2, cfl=0.8)
    sch.update(scheme)
    return m, S.default_scheme(**sch), S.default_gas(), cons


def vortex_state(x, L, beta=5.0, gamma=GAMMA):
    """Shu isentropic vortex, axis parallel to z through (L/2, L/2), freestream rho=p=1, u=0.5 sqrt(gamma)."""
    dx = x[:, 0] - 0.5 * L
    dy = x[:, 1] - 0.5 * L
    r2 = dx * dx + dy * dy
    e = np.exp(0.5 * (1.0 - r2))
    u = 0.5 * np.sqrt(gamma) - beta / (2 * np.pi) * dy * e
    v = beta / (2 * np.pi) * dx * e
    T = 1.0 - (gamma - 1.0) * beta * beta / (8 * gamma * np.pi * np.pi) * np.exp(1.0 - r2)
    rho = T ** (1.0 / (gamma - 1.0))
    p = rho ** gamma
    vel = np.vstack([u, v, np.zeros_like(u)])
    return prim_to_cons(rho, p, vel, gamma)


def vortex_case(n=100, L=10.0, seed=2, jitter=0.05, device=None, **scheme):
    """3D Euler, n^3 hexes x 6 tets, periodic in x/y/z, isentropic vortex (strength 5),
    second-order MUSCL with LSQ gradients and the Venkatakrishnan limiter (K=5), RK4.
    device: GPU index to build the mesh's faces and geometry on (None: on the host)."""
    m = M.box_tets(n, L, jitter, seed, periodic=True, device=device)
    cons = vortex_state(m.ccent[:m.ncell_int], L)
    sch = dict(sch_hyp=S.SCH_HLLC, method="M", limiter="W", gradmeth=S.GRAD_LSQ, venkat_k=5.0,
               tps_meth=S.TPS_RK4, cfl=0.8)
    sch.update(scheme)
    return m, S.default_scheme(**sch), S.default_gas(), cons


def periodic_quad_case(nx=20, ny=20, jitter=0.0, seed=1, **scheme):
    """The NRG TIME/RUNGEKUTTA layout: unit square, periodic in x and y,
    rho = 1 + (sin 2pi x sin 2pi y)^2 / 2, p = 1e5/14, u = 50 (NRG/TIME/RUNGEKUTTA/RK2/main.rpm)."""
    m = M.periodic_quad(nx, ny, 1.0, 1.0, jitter, seed)
    x = m.ccent[:m.ncell_int]
    rho = 1.0 + (np.sin(2 * np.pi * x[:, 0]) * np.sin(2 * np.pi * x[:, 1])) ** 2 / 2.0
    n = m.ncell_int
    vel = np.zeros((3, n))
    vel[0] = 50.0
    cons = prim_to_cons(rho, np.full(n, 100000 / 14.), vel)
    sch = dict(sch_hyp=S.SCH_HLLC, method="F", limiter="A", gradmeth=S.GRAD_LSQ, tps_meth=S.TPS_RK2, cfl=0.8)
    sch.update(scheme)
    return m, S.default_scheme(**sch), S.default_gas(), cons


def tgv_case(n=16, seed=3, mach=0.3, re=1e4, noise=0.01, jitter=0.0, walls=False, device=None, **scheme):
    """3D compressible Navier-Stokes Taylor-Green field on a periodic (2 pi L)^3 hex box, air with
    Sutherland's law (T0 = 300 K, p0 = 1e5 Pa), Re = rho0 U0 L / mu(T0), Pr = 0.72, 1% seeded noise.
    walls: slip walls at y = 0 and y = 2 pi L instead of periodicity in y (the field's v vanishes
    there).  device: GPU index to build the mesh's faces and geometry on (None: on the host)."""
    gamma, r = GAMMA, R_AIR
    T0, p0 = 300.0, 1e5
    rho0 = p0 / (r * T0)
    a0 = np.sqrt(gamma * r * T0)
    U0 = mach * a0
    mu0, tref, ts = 1.68e-5, 273.0, 110.5
    mu = mu0 * (T0 / tref) ** 1.5 * (tref + ts) / (T0 + ts)
    Lr = re * mu / (rho0 * U0)
    m = M.box_hex(n, 2 * np.pi * Lr, jitter, seed, periodic=(True, not walls, True), device=device)
    x = m.ccent[:m.ncell_int] / Lr
    ni = m.ncell_int
    ids = np.arange(ni)
    u = U0 * np.sin(x[:, 0]) * np.cos(x[:, 1]) * np.cos(x[:, 2]) * (1 + noise * M.uniform_pm1(seed + 100, ids, 0))
    v = -U0 * np.cos(x[:, 0]) * np.sin(x[:, 1]) * np.cos(x[:, 2]) * (1 + noise * M.uniform_pm1(seed + 100, ids, 1))
    w = U0 * noise * M.uniform_pm1(seed + 100, ids, 2)
    p = p0 + rho0 * U0 * U0 / 16.0 * (np.cos(2 * x[:, 0]) + np.cos(2 * x[:, 1])) * (np.cos(2 * x[:, 2]) + 2.0)
    rho = p / (r * T0)
    cons = prim_to_cons(rho, p, np.vstack([u, v, w]))
    sch = dict(sch_hyp=S.SCH_HLLC, sch_dis=S.DIS_CELLFULL, method="N", gradmeth=S.GRAD_LSQ,
               tps_meth=S.TPS_RK4, cfl=0.5)
    sch.update(scheme)
    gas = S.default_gas(typ_fluid=S.EQ_NSLAM, typ_visc=S.VISC_SUTH, visc_dyn=mu0, tref=tref, tsuth=ts)
    return m, S.default_scheme(**sch), gas, cons


def bump_case(nx=63, ny=33, nz=20, inflow=(1.0, 0.5 * np.sqrt(GAMMA), 1.0), vtx=None, noise=0.01, seed=4,
              p_out=None, **scheme):
    """3D Euler tet channel [0,3]x[0,1]x[0,0.6] with the bottom-wall bump y_w = 0.05 sin^2(pi(x-1)),
    x in [1,2]: inflow (rho, u, p) imposed at x=0 (supersonic-inlet Dirichlet), static pressure p at
    x=3 (subsonic outlet; ``p_out`` if given), slip walls elsewhere; first-order HLLC, RK2
    (reference midpoint).
    Initial state: the inflow state with seeded relative noise (no exact ties in the flux's
    wave-speed selection).  ``vtx`` replaces the vertex coordinates (displaced wall vertices)."""
    rho, u, p = inflow
    mach = u / np.sqrt(GAMMA * p / rho)
    bocos = [dict(family="IMIN", type=M.BC_INLET_SUP, param=freestream_inlet(rho, p, mach)),
             dict(family="IMAX", type=M.BC_OUTLET_SUB, param=[p if p_out is None else p_out]),
             dict(family="JMIN", type=M.BC_SYM), dict(family="JMAX", type=M.BC_SYM),
             dict(family="KMIN", type=M.BC_SYM), dict(family="KMAX", type=M.BC_SYM)]
    m = M.bump_channel_tets(nx, ny, nz, bocos=bocos, vtx=vtx, seed=seed)
    n = m.ncell_int
    ids = np.arange(n)
    vel = np.zeros((3, n))
    vel[0] = u * (1.0 + noise * M.uniform_pm1(seed + 100, ids, 0))
    vel[1] = u * noise * M.uniform_pm1(seed + 100, ids, 1)
    vel[2] = u * noise * M.uniform_pm1(seed + 100, ids, 2)
    cons = prim_to_cons(rho * (1.0 + noise * M.uniform_pm1(seed + 100, ids, 3)),
                        p * (1.0 + noise * M.uniform_pm1(seed + 100, ids, 4)), vel)
    sch = dict(sch_hyp=S.SCH_HLLC, method="N", tps_meth=S.TPS_RK2, cfl=0.8)
    sch.update(scheme)
    return m, S.default_scheme(**sch), S.default_gas(), cons


def oracle_scheme(scheme: dict) -> dict:
    """Oracle configuration (oracle/typhon_oracle.h) for a resolved scheme dict."""
    coef = scheme.get("rk_coef")
    coef = S.rk_tableau(scheme["tps_meth"]) if coef is None else coef
    lim = scheme["limiter"]
    return dict(sch_hyp=scheme["sch_hyp"], method=scheme["method"], limiter="X" if lim == "W" else lim,
                postlimiter=scheme["postlimiter"], gradmeth=scheme["gradmeth"], venkat=1 if lim == "W" else 0,
                venkat_k=scheme["venkat_k"], coef=coef, cfl=scheme["cfl"])


def oracle_gas(gas: dict, scheme: dict) -> dict:
    return dict(gamma=gas["gamma"], rgas=gas["r_const"], prandtl=gas["prandtl"], visc_law=gas["typ_visc"],
                visc_dyn=gas["visc_dyn"], visc_kin=gas["visc_kin"], tref=gas["tref"], tsuth=gas["tsuth"],
                viscous=1 if gas["typ_fluid"] == S.EQ_NSLAM else 0, sch_dis=scheme["sch_dis"])


def make_oracle(mesh, scheme, gas):
    from oracle.oracle import Oracle
    return Oracle(mesh, oracle_scheme(scheme), oracle_gas(gas, scheme))


def rel_err(a, b):
    """max over the field groups rho, rhoE and momentum (as one vector field) of
    ||a - b||_inf / ||b||_inf; momentum components are normalised together so that a
    component that is pure round-off (e.g. v in a flow with v = 0) is not divided by itself."""
    a = np.asarray(a)
    b = np.asarray(b)
    groups = [[0], [1], list(range(2, a.shape[0]))]
    out = 0.0
    for g in groups:
        s = np.abs(b[g]).max()
        d = np.abs(a[g] - b[g]).max()
        out = max(out, d / s if s > 0 else d)
    return out
