"""Wall and subsonic-inlet boundary conditions (setboco_ns_isoth.f90, setboco_ns_flux.f90 with
ns_bocoflux.f90, setboco_ns_inlet_sub.f90).

CPU: the oracle against closed forms chosen so that a wrong sign, a dropped term or a wrong
ghost rule fails --
  * plane Couette flow between moving walls on a uniform quad mesh: every interior cell's
    momentum residual vanishes and its energy residual is the viscous dissipation
    dt mu (dU/dy)^2, for isothermal and adiabatic walls;
  * an imposed wall heat flux changes only the wall cells' energy residual, by dt q A / V;
  * isothermal ghost states average to the wall temperature and velocity on the face;
  * subsonic-inlet ghost states satisfy the isentropic total-pressure / total-temperature
    relations at the interior static pressure, along the given direction.
The reference's own goldens NRG/NS/NUM/FLUX/VISCOUS/{CELLFULL,CELLAVG} (Stokes layer between
moving adiabatic walls) are replayed in test_nrg_golden.py.
GPU: the HIP path against the oracle (1e-12 relative in FP64) on channels with every kind.
"""
import numpy as np
import pytest

import cases as K
from typhon_amd import mesh as M
from typhon_amd import scheme as S

GAMMA, R = 1.4, 287.14


def couette(kind, nx=6, ny=8, h=2.0, du=3.0, T0=300.0, p0=1e5, mu=0.05, q=None, jitter=0.0, dis=S.DIS_CELLFULL,
            method="N", visc=S.VISC_DYNCST, noise=0.0):
    """Plane Couette flow u = du (y/h - 1/2) between walls at y = 0 (moving at -du/2) and
    y = h (moving at +du/2), periodic in x, uniform p and T; ``noise``: seeded relative
    perturbation of rho, p and u (and v = noise * du), so that residuals are not pure round-off."""
    uw = 0.5 * du
    if kind == "isoth":
        bot = dict(family="JMIN", type=M.BC_WALL_ISOTH, param=[T0, 0, 0, 0, -uw, 0, 0])
        top = dict(family="JMAX", type=M.BC_WALL_ISOTH, param=[T0, 0, 0, 0, uw, 0, 0])
    elif kind == "adiab":
        bot = dict(family="JMIN", type=M.BC_WALL_ADIAB, param=[0, 0, 0, 0, -uw, 0, 0])
        top = dict(family="JMAX", type=M.BC_WALL_ADIAB, param=[0, 0, 0, 0, uw, 0, 0])
    else:  # heat-flux walls, flux in the algorithm's convention (outward)
        bot = dict(family="JMIN", type=M.BC_WALL_FLUX, param=[q, 0, 0, 0, -uw, 0, 0])
        top = dict(family="JMAX", type=M.BC_WALL_FLUX, param=[0.0, 0, 0, 0, uw, 0, 0])
    bocos = [dict(family="IMIN", type=M.BC_PERIOD, partner="IMAX"),
             dict(family="IMAX", type=M.BC_PERIOD, partner="IMIN"), bot, top]
    m = M.channel_quad(nx, ny, 1.0, h, jitter, 7, bocos)
    n = m.ncell_int
    y = m.ccent[:n, 1]
    rho0 = p0 / (R * T0)
    vel = np.zeros((3, n))
    ids = np.arange(n)
    vel[0] = du * (y / h - 0.5) * (1 + noise * M.uniform_pm1(31, ids, 0))
    vel[1] = du * noise * M.uniform_pm1(31, ids, 1)
    from typhon_amd.solver import prim_to_cons
    cons = prim_to_cons(rho0 * (1 + noise * M.uniform_pm1(31, ids, 2)), p0 * (1 + noise * M.uniform_pm1(31, ids, 3)),
                        vel)
    sch = S.default_scheme(sch_hyp=S.SCH_HLLC, method=method, limiter="A" if method != "N" else "X",
                           gradmeth=S.GRAD_LSQ, sch_dis=dis, tps_meth=S.TPS_RK2, cfl=0.5)
    gas = S.default_gas(typ_fluid=S.EQ_NSLAM, typ_visc=visc, visc_dyn=mu, visc_kin=mu / rho0)
    return m, sch, gas, cons


@pytest.mark.parametrize("kind", ["isoth", "adiab"])
def test_couette_dissipation_closed_form(kind):
    du, h, mu = 3.0, 2.0, 0.05
    m, s, g, u = couette(kind, du=du, h=h, mu=mu)
    o = K.make_oracle(m, s, g)
    o.set_cons(u)
    dt = 1e-3
    r = o.calc_rhs(dt)
    # momentum: zero up to the round-off of the cancelling pressure fluxes p A dt / V
    scale = dt * 1e5 / (h / 8)
    assert np.abs(r[2:]).max() <= 1e-13 * scale
    assert np.abs(r[0]).max() <= 1e-13 * scale
    # energy: the dissipation, up to the round-off of the cancelling convective energy fluxes
    # (|u| (rho E + p) ~ 5e5 per unit area, 1e-16 of it relative to the 1e-4 signal)
    expect = dt * mu * (du / h) ** 2
    np.testing.assert_allclose(r[1], expect, rtol=2e-8)


def test_heat_flux_wall_closed_form():
    """q (outward, the algorithm's convention) on the bottom wall: the wall cells' energy
    residual differs from the adiabatic one by -dt q A / V and nothing else changes."""
    q = 7.5
    ma, s, g, ua = couette("adiab")
    mq, _, _, uq = couette("flux", q=q)
    oa = K.make_oracle(ma, s, g)
    oq = K.make_oracle(mq, s, g)
    oa.set_cons(ua)
    oq.set_cons(uq)
    dt = 1e-3
    ra = oa.calc_rhs(dt)
    rq = oq.calc_rhs(dt)
    n = ma.ncell_int
    bot = [b for b in mq.bocos if b.family == "JMIN"][0]
    expect = np.zeros(n)
    for f in bot.iface:
        c = mq.ic1[f]
        expect[c] += -dt * q * mq.fsurf[f] / mq.cvol[c]
    d = rq - ra
    scale = np.abs(expect).max()
    assert np.abs(d[[0, 2, 3, 4]]).max() <= 1e-12 * scale
    np.testing.assert_allclose(d[1], expect, rtol=1e-10, atol=1e-12 * scale)


def test_isothermal_ghost_averages_to_wall_state():
    Tw = 350.0
    m, s, g, u = couette("isoth", T0=300.0)
    for b in m.bocos:
        if b.type == M.BC_WALL_ISOTH:
            b.param[0] = Tw
    o = K.make_oracle(m, s, g)
    o.set_cons(u)
    o.calc_rhs(1e-3)
    prim = o.prim()
    for b in m.bocos:
        if b.type != M.BC_WALL_ISOTH:
            continue
        ci, cg = m.ic1[b.iface], m.ic2[b.iface]
        T = prim[1] / (prim[0] * R)
        np.testing.assert_allclose(0.5 * (T[ci] + T[cg]), Tw, rtol=1e-13)
        np.testing.assert_allclose(prim[1][cg], prim[1][ci], rtol=0)
        np.testing.assert_allclose(0.5 * (prim[2][ci] + prim[2][cg]), b.param[4], rtol=1e-13)
        np.testing.assert_allclose(0.5 * (prim[3][ci] + prim[3][cg]), 0.0, atol=1e-14)


def subsonic_channel(second=False, nx=12, ny=6, jitter=0.1, viscous=True):
    """2D jittered channel: subsonic inlet (PI, TI, DIRECTION) at x = 0, static-pressure
    outlet at x = 1, isothermal bottom wall, moving adiabatic top wall (NS) or slip walls."""
    T0, p0 = 300.0, 1e5
    pi, ti = 1.06 * p0, 1.02 * T0
    d = np.array([1.0, 0.2, 0.0])
    if viscous:
        walls = [dict(family="JMIN", type=M.BC_WALL_ISOTH, param=[310.0, 0, 0, 0, 0.0, 0, 0]),
                 dict(family="JMAX", type=M.BC_WALL_ADIAB, param=[0, 0, 0, 0, 20.0, 0, 0])]
    else:
        walls = [dict(family="JMIN", type=M.BC_SYM), dict(family="JMAX", type=M.BC_SYM)]
    bocos = [dict(family="IMIN", type=M.BC_INLET_SUB, param=[pi, ti, 0.0, *d]),
             dict(family="IMAX", type=M.BC_OUTLET_SUB, param=[p0])] + walls
    m = M.channel_quad(nx, ny, 1.0, 0.5, jitter, 11, bocos)
    n = m.ncell_int
    ids = np.arange(n)
    rho0 = p0 / (R * T0)
    vel = np.zeros((3, n))
    vel[0] = 80.0 * (1 + 0.05 * M.uniform_pm1(21, ids, 0))
    vel[1] = 5.0 * M.uniform_pm1(21, ids, 1)
    from typhon_amd.solver import prim_to_cons
    cons = prim_to_cons(rho0 * (1 + 0.02 * M.uniform_pm1(21, ids, 2)), p0 * (1 + 0.01 * M.uniform_pm1(21, ids, 3)),
                        vel)
    sch = S.default_scheme(sch_hyp=S.SCH_HLLC, method="M" if second else "N", limiter="V" if second else "X",
                           gradmeth=S.GRAD_LSQ, sch_dis=S.DIS_CELLFULL if viscous else S.DIS_NOFLUX,
                           tps_meth=S.TPS_RK2, cfl=0.5)
    gas = S.default_gas(typ_fluid=S.EQ_NSLAM if viscous else S.EQ_EULER, typ_visc=S.VISC_SUTH, visc_dyn=1.68e-5)
    return m, sch, gas, cons


def test_subsonic_inlet_isentropic_ghost():
    m, s, g, u = subsonic_channel(viscous=False)
    o = K.make_oracle(m, s, g)
    o.set_cons(u)
    o.calc_rhs(1e-6)
    qL, qR = o.face_states()
    b = [b for b in m.bocos if b.type == M.BC_INLET_SUB][0]
    pi, ti = b.param[0], b.param[1]
    d = b.param[3:6] / np.linalg.norm(b.param[3:6])
    f = b.iface
    rho, p = qR[0][f], qR[1][f]
    v = qR[2:5][:, f]
    np.testing.assert_allclose(p, qL[1][f], rtol=0)  # static pressure from the interior
    speed = np.linalg.norm(v, axis=0)
    np.testing.assert_allclose(v / speed, np.repeat(d[:, None], len(f), 1), atol=1e-15)
    T = p / (rho * R)
    m2 = speed ** 2 / (GAMMA * p / rho)
    fac = 1 + 0.5 * (GAMMA - 1) * m2
    np.testing.assert_allclose(T * fac, ti, rtol=1e-13)
    np.testing.assert_allclose(p * fac ** (GAMMA / (GAMMA - 1)), pi, rtol=1e-13)


# ---------------------------------------------------------------------------
# GPU parity

TOL64 = 1e-12


def _gpu_parity(mk, steps=3, tol=TOL64, precision=S.PREC_FP64):
    from typhon_amd.solver import Solver
    m, s, g, u = mk()
    s = dict(s, precision=precision)
    gpu = Solver(m, s, g)
    gpu.set_state(u)
    orc = K.make_oracle(m, s, g)
    orc.set_cons(u)
    dt = orc.timestep(s["cfl"])
    assert abs(gpu.timestep() - dt) <= 1e-13 * dt
    nv = m.dim + 2
    err = K.rel_err(gpu.calc_rhs(dt)[:nv], orc.calc_rhs(dt)[:nv])
    assert err <= tol, f"residual rel err {err:.3e}"
    for _ in range(steps):
        gpu.rk_step(dt)
        orc.rk_step(dt)
    err = K.rel_err(gpu.get_state()[:nv], orc.get_cons()[:nv])
    assert err <= tol, f"state rel err {err:.3e}"


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["isoth", "adiab", "flux"])
@pytest.mark.parametrize("dis", [S.DIS_CELLAVG2, S.DIS_CELLFULL])
def test_gpu_couette_walls(kind, dis):
    _gpu_parity(lambda: couette(kind, q=-40.0, jitter=0.15, dis=dis, visc=S.VISC_SUTH, mu=1.68e-5, noise=0.01))


@pytest.mark.gpu
@pytest.mark.parametrize("second", [False, True])
def test_gpu_subsonic_channel_ns(second):
    _gpu_parity(lambda: subsonic_channel(second=second))


@pytest.mark.gpu
@pytest.mark.parametrize("post", [S.POSTLIM_MONOTONIC0, S.POSTLIM_MONOTONIC1, S.POSTLIM_MONOTONIC2])
def test_gpu_monotonic_postlimiters_at_walls(post):
    """the ghost sides of isothermal, adiabatic, subsonic-inlet and outlet faces (MONOTONIC1 reads
    the ghost cell's own extrapolation with the boundary map of its gradient)"""
    def mk():
        m, s, g, u = subsonic_channel(second=True)
        return m, dict(s, limiter="S", postlimiter=post), g, u
    _gpu_parity(mk)


@pytest.mark.gpu
@pytest.mark.parametrize("dis,why", [(S.DIS_COMPACT, "COMPACT"), (S.DIS_FACECENTERED, "FACE-AVERAGE"),
                                     (S.DIS_FACEPENALTY, "PENALTY")])
def test_gpu_unrun_viscous_discretisations_refused(dis, why):
    """COMPACT stops in the reference's NS velocity-gradient interpolation; FACE-AVERAGE and PENALTY
    need face gradients, which the reference offers only with its spectral-volume method: the
    scheme is refused and the previous one stays"""
    from typhon_amd.solver import Solver
    from typhon_amd._lib import TyError
    m, s, g, u = subsonic_channel(second=True)
    gpu = Solver(m, s, g)
    gpu.set_state(u)
    with pytest.raises(TyError, match=why):
        gpu.set_scheme(dict(s, sch_dis=dis), g)
    r = gpu.calc_rhs(1e-6)
    assert np.isfinite(r).all()


@pytest.mark.gpu
def test_gpu_subsonic_channel_euler():
    _gpu_parity(lambda: subsonic_channel(second=True, viscous=False))


@pytest.mark.gpu
def test_gpu_walls_3d_hex():
    """3D hexes, periodic in x and z, isothermal wall at y = 0 and heat-flux wall at y = L"""
    def mk():
        m0, s, g, u = K.tgv_case(8, walls=True)
        L = m0.vtx[:, 1].max() - m0.vtx[:, 1].min()
        bocos = [dict(family="IMIN", type=M.BC_PERIOD, partner="IMAX"),
                 dict(family="IMAX", type=M.BC_PERIOD, partner="IMIN"),
                 dict(family="KMIN", type=M.BC_PERIOD, partner="KMAX"),
                 dict(family="KMAX", type=M.BC_PERIOD, partner="KMIN"),
                 dict(family="JMIN", type=M.BC_WALL_ISOTH, param=[290.0, 0, 0, 0, 3.0, 0, 1.0]),
                 dict(family="JMAX", type=M.BC_WALL_FLUX, param=[-2.0e3, 0, 0, 0, -3.0, 0, 0])]
        m = M.box_hex(8, L, 0.0, 3, periodic=(True, False, True), bocos=bocos)
        return m, s, g, u
    _gpu_parity(mk)


@pytest.mark.gpu
def test_gpu_walls_mixed_precision():
    _gpu_parity(lambda: subsonic_channel(second=True), tol=2e-5, precision=S.PREC_MIXED, steps=2)


@pytest.mark.gpu
def test_gpu_adjoint_rejects_walls():
    from typhon_amd.solver import Solver
    from typhon_amd._lib import TyError
    m, s, g, u = couette("isoth")
    gpu = Solver(m, s, g)
    gpu.set_state(u)
    with pytest.raises(TyError):
        gpu.residual_adjoint(1e-3, np.ones((5, m.ncell_int)))


@pytest.mark.gpu
@pytest.mark.parametrize("btype", [M.BC_WALL_HCONV, M.BC_WALL_HGEN])
def test_gpu_convection_walls_refused(btype):
    """CONVECTION_WALL / GEN_CONV_WALL: the reference has no NS ghost rule for them
    (MENU_NS.f90:95-113; setboco_ns_hconv.f90:78 stops), so the upload refuses them"""
    from typhon_amd.solver import Solver
    from typhon_amd._lib import TyError
    m, s, g, u = couette("isoth")
    m.bocos[-1].type = btype
    with pytest.raises(TyError, match="convection wall"):
        Solver(m, s, g)


@pytest.mark.gpu
def test_gpu_failed_rescheme_keeps_previous_scheme():
    """set_scheme validates everything before it releases anything: a refused scheme (adiabatic
    walls under the Euler model) leaves the previous scheme and state as they were"""
    from typhon_amd.solver import Solver
    from typhon_amd._lib import TyError
    m, s, g, u = couette("adiab", noise=0.01)
    gpu = Solver(m, s, g)
    gpu.set_state(u)
    before = gpu.get_state()
    with pytest.raises(TyError, match="laminar NS") as e:
        gpu.set_scheme(s, dict(g, typ_fluid=S.EQ_EULER))
    assert e.value.code == 4
    assert np.array_equal(gpu.get_state(), before)
    o = K.make_oracle(m, s, g)
    o.set_cons(before)
    assert K.rel_err(gpu.calc_rhs(1e-3), o.calc_rhs(1e-3)) <= TOL64
    gpu.rk_step(1e-3)
    o.rk_step(1e-3)
    assert K.rel_err(gpu.get_state(), o.get_cons()) <= TOL64


@pytest.mark.gpu
def test_gpu_mixed_context_refuses_residual_derivatives():
    """the residual tangent and adjoint differentiate the FP64 residual, which a mixed-precision
    context does not evaluate"""
    from typhon_amd.solver import Solver
    from typhon_amd._lib import TyError
    m, s, g, u = K.bump_case(6, 3, 3, precision=S.PREC_MIXED)
    gpu = Solver(m, s, g)
    gpu.set_state(u)
    lam = np.ones((5, m.ncell_int))
    with pytest.raises(TyError, match="FP64 schemes only") as e:
        gpu.residual_adjoint(1e-3, lam)
    assert e.value.code == 4
    with pytest.raises(TyError, match="FP64 schemes only"):
        gpu.residual_tangent(1e-3, du=lam)
