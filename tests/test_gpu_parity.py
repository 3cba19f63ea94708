"""Parity of the HIP residual path (through the C-ABI) with the oracle.

FP64 bar: 1e-12 relative (max-norm per field) on residuals and on states after
a few RK steps -- the summation order differs (per-cell slot order on the GPU,
face order in the oracle, colour order in the reference), so results agree to
a few ulps, not bitwise.  Mixed precision (FP32 fluxes/gradients, FP64
accumulation) is checked at the tolerance written in each test.
"""
import numpy as np
import pytest

import cases as K
from typhon_amd import scheme as S
from typhon_amd.solver import Solver
from typhon_amd._lib import TyError

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def _pair(mk):
    m, s, g, u = mk()
    gpu = Solver(m, s, g)
    gpu.set_state(u)
    orc = K.make_oracle(m, s, g)
    orc.set_cons(u)
    return m, s, g, u, gpu, orc


def _residual_parity(mk, tol=TOL64):
    m, s, g, u, gpu, orc = _pair(mk)
    dt_g = gpu.timestep()
    dt_o = orc.timestep(s["cfl"])
    assert abs(dt_g - dt_o) <= 1e-13 * dt_o
    rg = gpu.calc_rhs(dt_o)
    ro = orc.calc_rhs(dt_o)
    nv = m.dim + 2
    err = K.rel_err(rg[:nv], ro[:nv])
    assert err <= tol, f"residual rel err {err:.3e}"
    return gpu, orc


def _step_parity(mk, nsteps=5, tol=TOL64):
    m, s, g, u, gpu, orc = _pair(mk)
    for _ in range(nsteps):
        dt = orc.timestep(s["cfl"])
        gpu.rk_step(dt)
        orc.rk_step(dt)
    nv = m.dim + 2
    err = K.rel_err(gpu.get_state()[:nv], orc.get_cons()[:nv])
    assert err <= tol, f"state rel err after {nsteps} steps {err:.3e}"
    return gpu, orc


# --- first order (config: 2D channel, HLLC, RK2) ------------------------------

def test_channel_first_order_residual():
    _residual_parity(lambda: K.channel_case(32, 16))


def test_channel_first_order_full_size_steps():
    _step_parity(lambda: K.channel_case(256, 128), nsteps=5)


ALL_FLUXES = [S.SCH_HLLC, S.SCH_HLLCB, S.SCH_HLLCK, S.SCH_HLLCKB, S.SCH_HLLE, S.SCH_HLLEB, S.SCH_HLLEK,
              S.SCH_HLLEKB, S.SCH_RUSANOV, S.SCH_ROE]


@pytest.mark.parametrize("sch", ALL_FLUXES)
def test_fluxes_residual(sch):
    _residual_parity(lambda: K.channel_case(24, 12, sch_hyp=sch))


@pytest.mark.parametrize("sch", [S.SCH_HLLCK, S.SCH_HLLCKB, S.SCH_HLLEK, S.SCH_HLLEKB])
def test_kinetic_fluxes_second_order_steps(sch):
    """the kinetic wave speeds (erf/exp of the half-Maxwellian Mach numbers) on 3D tets with MUSCL"""
    _step_parity(lambda: K.bump_case(6, 3, 3, noise=0.05, sch_hyp=sch, method="M", limiter="A",
                                     gradmeth=S.GRAD_LSQ, tps_meth=S.TPS_RK2), nsteps=4)


@pytest.mark.parametrize("sch", [S.SCH_HLLCK, S.SCH_HLLEKB])
def test_kinetic_fluxes_mixed_precision(sch):
    """FP32 fluxes (erff/expf) with FP64 accumulation: the FP64 state to 1e-5 after 5 steps"""
    m, s, g, u = K.channel_case(24, 12, sch_hyp=sch)
    res = {}
    for prec in (S.PREC_FP64, S.PREC_MIXED):
        gpu = Solver(m, dict(s, precision=prec), g)
        gpu.set_state(u)
        for _ in range(5):
            gpu.rk_step(2e-3)
        res[prec] = gpu.get_state()
    assert K.rel_err(res[S.PREC_MIXED][:4], res[S.PREC_FP64][:4]) <= 1e-5


# --- second order reference MUSCL family ---------------------------------------

@pytest.mark.parametrize("method", ["M", "F", "U"])
@pytest.mark.parametrize("lim", ["X", "M", "A", "V", "S", "K", "L"])
def test_muscl_family_channel(method, lim):
    _residual_parity(lambda: K.channel_case(24, 12, method=method, limiter=lim))


@pytest.mark.parametrize("method", ["M", "F", "U"])
def test_muscl_periodic_jittered(method):
    _residual_parity(lambda: K.periodic_quad_case(16, 12, jitter=0.2, method=method, limiter="A"))


@pytest.mark.parametrize("post", [S.POSTLIM_BARTH, S.POSTLIM_SUPERBARTH])
@pytest.mark.parametrize("method", ["M", "U"])
def test_barth_postlimiter(method, post):
    _residual_parity(lambda: K.channel_case(24, 12, method=method, limiter="X", postlimiter=post))
    _residual_parity(lambda: K.vortex_case(4, L=4.0, method=method, limiter="X", postlimiter=post))


MONO = [S.POSTLIM_MONOTONIC0, S.POSTLIM_MONOTONIC1, S.POSTLIM_MONOTONIC2]


@pytest.mark.parametrize("post", MONO)
@pytest.mark.parametrize("method", ["M", "F", "U"])
def test_monotonic_postlimiter_channel(method, post):
    """MONOTONIC0/1/2 on the channel's interior, slip-wall, inlet and outlet faces"""
    _residual_parity(lambda: K.channel_case(24, 12, method=method, limiter="S", postlimiter=post))


@pytest.mark.parametrize("post", MONO)
def test_monotonic_postlimiter_tets_steps(post):
    _step_parity(lambda: K.bump_case(6, 3, 3, noise=0.08, method="M", limiter="A", gradmeth=S.GRAD_LSQ,
                                     postlimiter=post, tps_meth=S.TPS_RK2), nsteps=4)


@pytest.mark.parametrize("post", MONO)
def test_monotonic_postlimiter_periodic(post):
    _residual_parity(lambda: K.vortex_case(4, L=4.0, method="U", limiter="A", postlimiter=post))


def test_monotonic1_mixed_precision():
    _step_parity(lambda: K.channel_case(24, 12, method="M", limiter="S", postlimiter=S.POSTLIM_MONOTONIC1,
                                        precision=S.PREC_MIXED), nsteps=3, tol=2e-5)


def test_weighted_lsq():
    _residual_parity(lambda: K.channel_case(24, 12, method="M", limiter="M", gradmeth=S.GRAD_LSQW))


# --- 3D tets, Venkatakrishnan, RK4 ----------------------------------------------

def test_vortex_venkat_residual():
    _residual_parity(lambda: K.vortex_case(6, L=6.0))


def test_vortex_venkat_steps():
    _step_parity(lambda: K.vortex_case(6, L=6.0), nsteps=3)


def test_vortex_reference_variant_steps():
    # the reference-expressible form of the 3D case: MUSCL + van Albada
    _step_parity(lambda: K.vortex_case(5, L=5.0, limiter="A"), nsteps=2)


@pytest.mark.parametrize("tps", [S.TPS_EXPL, S.TPS_RK2, S.TPS_RK2SSP, S.TPS_RK3SSP, S.TPS_RK4,
                                 S.TPS_LSRK25BB, S.TPS_LSRK26BB, S.TPS_LSRK12BS, S.TPS_LSRK13BS])
def test_rk_methods(tps):
    _step_parity(lambda: K.periodic_quad_case(12, 10, jitter=0.1, tps_meth=tps), nsteps=3)


# --- Navier-Stokes ----------------------------------------------------------------

@pytest.mark.parametrize("dis", [S.DIS_CELLAVG2, S.DIS_CELLFULL])
def test_tgv_viscous_residual(dis):
    _residual_parity(lambda: K.tgv_case(6, sch_dis=dis))


def test_tgv_viscous_steps():
    _step_parity(lambda: K.tgv_case(6), nsteps=3)


def test_viscous_second_order_jittered():
    _residual_parity(lambda: K.tgv_case(5, jitter=0.1, method="M", limiter="W"))


# --- mixed precision ----------------------------------------------------------------

def test_mixed_precision_residual():
    # FP32 flux/gradient arithmetic: residual within 2e-5 relative of the FP64 oracle
    _residual_parity(lambda: K.vortex_case(6, L=6.0, precision=S.PREC_MIXED), tol=2e-5)


def _field_errors(a, b):
    """max-norm relative error of each field of conservative states a vs the reference b:
    rho, rhoE, momentum (as one vector), pressure"""
    def p(U):
        return 0.4 * (U[1] - 0.5 * (U[2:5] ** 2).sum(0) / U[0])
    out = {}
    for k, (x, y) in dict(rho=(a[0], b[0]), rhoE=(a[1], b[1]), mom=(a[2:5], b[2:5]), p=(p(a), p(b))).items():
        out[k] = float(np.abs(x - y).max() / np.abs(y).max())
    return out


@pytest.mark.parametrize("form", [S.PREC_MIXED, S.PREC_MIXED_F32SUM])
def test_mixed_precision_steps(form):
    # FP64 RK accumulation (or, PREC_MIXED_F32SUM, the stage-increment sum kept in FP32 and the FP64
    # state update) and FP64 pressure recovery: every field within 1e-6 relative of the FP64 run
    # after 20 RK4 steps (second-order MUSCL + Venkatakrishnan)
    m, s, g, u = K.vortex_case(6, L=6.0)
    a = Solver(m, s, g)
    s2 = dict(s)
    s2["precision"] = form
    b = Solver(m, s2, g)
    a.set_state(u)
    b.set_state(u)
    for _ in range(20):
        dt = a.timestep()
        a.rk_step(dt)
        b.rk_step(dt)
    err = _field_errors(b.get_state(), a.get_state())
    print("mixed vs fp64 after 20 RK4 steps:", err)
    assert max(err.values()) <= 1e-6, err


def test_f32sum_form_differs_from_mixed_by_the_sum_rounding_only():
    """the two mixed forms run the same FP32 kernels; they differ only in the FP32 rounding of the
    stage-increment sum, a relative 6e-8 of an increment: after 20 RK4 steps they agree far more
    closely than either agrees with FP64, and are not bitwise equal (the FP32 sum is in use)"""
    m, s, g, u = K.vortex_case(6, L=6.0)
    sv = {f: Solver(m, dict(s, precision=f), g) for f in (S.PREC_FP64, S.PREC_MIXED, S.PREC_MIXED_F32SUM)}
    for x in sv.values():
        x.set_state(u)
    for _ in range(20):
        dt = sv[S.PREC_FP64].timestep()
        for x in sv.values():
            x.rk_step(dt)
    st = {f: x.get_state() for f, x in sv.items()}
    for x in sv.values():
        x.close()
    d_forms = max(_field_errors(st[S.PREC_MIXED_F32SUM], st[S.PREC_MIXED]).values())
    d_fp64 = max(_field_errors(st[S.PREC_MIXED], st[S.PREC_FP64]).values())
    print(f"f32sum vs mixed {d_forms:.2e}, mixed vs fp64 {d_fp64:.2e}")
    assert 0.0 < d_forms <= 1e-6 and d_forms <= d_fp64


def test_mixed_precision_recovers_pressure_in_fp64():
    """The mixed form recovers p = (gamma - 1)(rhoE - rho |u|^2 / 2) in FP64 before it rounds the
    stage record to FP32 (calc_varprim_ns.f90:47-52).  At Mach 20 the kinetic energy is about 110x
    the internal energy: recovered in FP32, p would carry about 7e-6 of relative error.  Here the
    y-momentum residual is driven by a 10% pressure variation in y alone (v = 0, uniform rho and u),
    and neighbouring faces differ by about 8% of p, so that error would reach the residual about 12x
    larger; with FP64 recovery only the FP32 flux arithmetic remains."""
    from typhon_amd import mesh as M
    from typhon_amd.solver import prim_to_cons
    m = M.periodic_quad(16, 16, 1.0, 1.0)
    n = m.ncell_int
    y = m.ccent[:n, 1]
    p = (1.0 + 0.1 * np.sin(2 * np.pi * y)) / 1.4
    vel = np.vstack([np.full(n, 20.0), np.zeros(n), np.zeros(n)])
    u = prim_to_cons(np.ones(n), p, vel)
    # the same recovery in FP32 arithmetic: its pressure error (what this test would see)
    f = np.float32
    u32 = u.astype(f)
    p32 = f(0.4) * (u32[1] - f(0.5) * u32[0] * ((u32[2] / u32[0]) ** 2 + (u32[3] / u32[0]) ** 2))
    assert np.abs(p32.astype(np.float64) - p).max() / p.max() > 3e-6
    s = S.default_scheme(sch_hyp=S.SCH_HLLC, tps_meth=S.TPS_RK2, precision=S.PREC_MIXED)
    g = S.default_gas()
    gpu = Solver(m, s, g)
    gpu.set_state(u)
    orc = K.make_oracle(m, s, g)
    orc.set_cons(u)
    dt = orc.timestep(s["cfl"])
    rg, ro = gpu.calc_rhs(dt), orc.calc_rhs(dt)
    err_y = np.abs(rg[3] - ro[3]).max() / np.abs(ro[3]).max()
    print(f"mixed y-momentum residual error at Mach 20: {err_y:.3e}")
    assert err_y <= 1e-5, err_y


# --- driver behaviour ------------------------------------------------------------------

def test_advance_matches_host_loop():
    m, s, g, u = K.vortex_case(5, L=5.0)
    a = Solver(m, s, g)
    b = Solver(m, s, g)
    a.set_state(u)
    b.set_state(u)
    for _ in range(3):
        b.rk_step(b.timestep())
    dt = a.advance(3, want_dt=True)
    assert dt > 0
    assert np.array_equal(a.get_state(), b.get_state())


def test_bitwise_repeatable():
    m, s, g, u = K.vortex_case(5, L=5.0)
    out = []
    for _ in range(2):
        a = Solver(m, s, g)
        a.set_state(u)
        a.advance(3)
        out.append(a.get_state())
    assert np.array_equal(out[0], out[1])


def test_bad_state_is_reported():
    # calc_ns_timestep.f90:89-100: a cell with rhoE < 1/2 rho |u|^2 stops the run
    m, s, g, u = K.channel_case(16, 8)
    u = u.copy()
    u[1, 5] = 0.1 * 0.5 * u[2, 5] ** 2 / u[0, 5]
    a = Solver(m, s, g)
    a.set_state(u)
    with pytest.raises(TyError) as e:
        a.timestep()
    assert e.value.code == 3
