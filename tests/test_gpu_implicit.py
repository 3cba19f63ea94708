"""The matrix-free implicit path on the GPU (block-Jacobi from the reference's HLL Jacobian
assembly, FD Jacobian-vector products, GMRES in FP32 or FP64 inside FP64 refinement):
- diagonal blocks against the oracle's restatement of build_implicit_bdlu (1e-12 relative);
- J v against the oracle's central difference taken with the same step (1e-7 relative), and,
  where the residual is smooth at the base state, against the derivative itself (1e-6);
- (I - J) x = b against a dense numpy solve with J assembled column by column from the
  oracle's residual (tiny meshes), for both inner precisions.

Base states carry 1% additive noise: a uniform freestream sits exactly on the kinks of the
HLLC speed estimates and of the slope limiters (equal wave-speed candidates, zero slopes),
where the residual has no derivative and every finite difference depends on its step.
"""
import numpy as np
import pytest

import cases as K
from typhon_amd import scheme as S
from typhon_amd.solver import Solver, krylov

pytestmark = pytest.mark.gpu

CASES = {
    "channel_o1": lambda: K.channel_case(8, 4),
    "channel_muscl": lambda: K.channel_case(8, 4, method="M", limiter="A"),
    "channel_unlimited": lambda: K.channel_case(8, 4, method="M", limiter="X"),
    "vortex_venkat": lambda: K.vortex_case(3, L=3.0),
    "tgv_viscous": lambda: K.tgv_case(3),
}
# residuals whose central difference at the library's step is within 1e-9 of the derivative
# at the noisy base state (van Albada's curvature near small slopes is not: its FD J v is
# only checked against the oracle's FD with the same step)
SMOOTH = ["channel_o1", "channel_unlimited", "vortex_venkat", "tgv_viscous"]
EPS_MACH = 2.220446049250313e-16


def _oracle_residual(orc, u, dt):
    orc.set_cons(u)
    return orc.calc_rhs(dt)


def _dense_jacobian(orc, u, dt, nv):
    """J = dR/dU by central differences of the oracle residual, unknowns ordered [var][cell]"""
    n = u.shape[1]
    N = nv * n
    J = np.empty((N, N))
    for col in range(N):
        v, c = divmod(col, n)
        h = 1e-6 * max(1.0, abs(u[v, c]))
        up = u.copy()
        up[v, c] += h
        um = u.copy()
        um[v, c] -= h
        J[:, col] = ((_oracle_residual(orc, up, dt) - _oracle_residual(orc, um, dt))[:nv] / (2 * h)).ravel()
    return J


def _setup(name):
    m, s, g, u = CASES[name]()
    nv = m.dim + 2
    rng = np.random.default_rng(7)
    u = u.copy()
    u[:nv] += 1e-2 * np.abs(u[:nv]).max(axis=1, keepdims=True) * rng.uniform(-1, 1, size=u[:nv].shape)
    gpu = Solver(m, s, g)
    gpu.set_state(u)
    orc = K.make_oracle(m, s, g)
    orc.set_cons(u)
    dt = 5.0 * orc.timestep(s["cfl"])  # implicit: several explicit time steps
    return m, s, g, u, gpu, orc, dt


@pytest.mark.parametrize("name", list(CASES))
def test_jacobian_diag_matches_oracle(name):
    m, s, g, u, gpu, orc, dt = _setup(name)
    orc.calc_rhs(dt)
    ref = orc.jac_diag(dt)
    got = gpu.jacobian_diag(dt)
    nv = m.dim + 2
    scale = np.abs(ref[:, :nv, :nv]).max()
    assert np.abs(got[:, :nv, :nv] - ref[:, :nv, :nv]).max() <= 1e-12 * scale


def _direction(u, nv, seed=3):
    rng = np.random.default_rng(seed)
    v = np.zeros_like(u)
    v[:nv] = rng.normal(size=(nv, u.shape[1])) * np.abs(u[:nv]).max(axis=1, keepdims=True) * 1e-2
    return v


def _central(orc, u, v, h, dt):
    return (_oracle_residual(orc, u + h * v, dt) - _oracle_residual(orc, u - h * v, dt)) / (2 * h)


@pytest.mark.parametrize("name", list(CASES))
def test_jvp_matches_oracle_same_step(name):
    """ty_jvp is the central difference with step cbrt(eps)(1 + |u|_rms)/|v|_rms"""
    m, s, g, u, gpu, orc, dt = _setup(name)
    nv = m.dim + 2
    v = _direction(u, nv)
    h = np.cbrt(EPS_MACH) * (1.0 + np.sqrt(np.mean(u[:nv] ** 2))) / np.sqrt(np.mean(v[:nv] ** 2))
    ref = _central(orc, u, v, h, dt)
    got = gpu.jvp(dt, v)
    assert K.rel_err(got[:nv], ref[:nv]) <= 1e-7


@pytest.mark.parametrize("name", SMOOTH)
def test_jvp_approximates_derivative(name):
    m, s, g, u, gpu, orc, dt = _setup(name)
    nv = m.dim + 2
    v = _direction(u, nv)
    ref = _central(orc, u, v, 1e-5, dt)
    got = gpu.jvp(dt, v)
    assert K.rel_err(got[:nv], ref[:nv]) <= 1e-6


@pytest.mark.parametrize("inner", [S.PREC_MIXED, S.PREC_FP64])
@pytest.mark.parametrize("name", SMOOTH)
def test_linear_solve_matches_dense(name, inner):
    m, s, g, u, gpu, orc, dt = _setup(name)
    nv = m.dim + 2
    J = _dense_jacobian(orc, u, dt, nv)
    b = _oracle_residual(orc, u, dt)
    x_ref = np.linalg.solve(np.eye(J.shape[0]) - J, b[:nv].ravel()).reshape(nv, -1)
    x, st = gpu.linear_solve(dt, b, krylov(nkrylov=20, inner_precision=inner, outer_tol=1e-9, max_outer=10))
    assert st["rel_res"] <= 1e-9
    assert K.rel_err(x[:nv], x_ref) <= 1e-6


def test_implicit_step_is_state_plus_solution():
    m, s, g, u, gpu, orc, dt = _setup("tgv_viscous")
    nv = m.dim + 2
    J = _dense_jacobian(orc, u, dt, nv)
    b = _oracle_residual(orc, u, dt)
    x_ref = np.linalg.solve(np.eye(J.shape[0]) - J, b[:nv].ravel()).reshape(nv, -1)
    st = gpu.implicit_step(dt, krylov(nkrylov=30, outer_tol=1e-9))
    assert st["rel_res"] <= 1e-9
    du = gpu.get_state()[:nv] - u[:nv]
    assert K.rel_err(du, x_ref) <= 1e-6


def test_fp32_inner_refines_to_fp64_answer():
    """FP32 Krylov iterations inside FP64 refinement reach the same answer as all-FP64 GMRES
    to the refinement tolerance (1e-8 relative Newton residual)"""
    m, s, g, u = K.tgv_case(12)
    a = Solver(m, s, g)
    a.set_state(u)
    dt = 20.0 * a.timestep()
    b = a.calc_rhs(dt)
    x32, st32 = a.linear_solve(dt, b, krylov(nkrylov=30, inner_precision=S.PREC_MIXED, outer_tol=1e-8))
    x64, st64 = a.linear_solve(dt, b, krylov(nkrylov=30, inner_precision=S.PREC_FP64, outer_tol=1e-8))
    assert st32["rel_res"] <= 1e-8 and st64["rel_res"] <= 1e-8
    assert st32["outer_iters"] >= 2  # FP32 inner alone cannot reach 1e-8: refinement did the rest
    assert K.rel_err(x32[:5], x64[:5]) <= 1e-6


@pytest.mark.gpu
def test_krylov_buffers_regrow_between_solves():
    """a solve with a larger basis, then the other inner precision, then the first again on one
    context (the basis and work vectors are released and re-allocated): every solve converges and
    repeats the first solve's answer bit for bit when its settings come back"""
    m, s, g, u, gpu, orc, dt = _setup(SMOOTH[0])
    b = _oracle_residual(orc, u, dt)
    x1, st1 = gpu.linear_solve(dt, b, krylov(nkrylov=8, inner_precision=S.PREC_FP64, outer_tol=1e-9, max_outer=20))
    x2, st2 = gpu.linear_solve(dt, b, krylov(nkrylov=24, inner_precision=S.PREC_FP64, outer_tol=1e-9, max_outer=20))
    x3, st3 = gpu.linear_solve(dt, b, krylov(nkrylov=24, inner_precision=S.PREC_MIXED, outer_tol=1e-9, max_outer=20))
    x4, st4 = gpu.linear_solve(dt, b, krylov(nkrylov=8, inner_precision=S.PREC_FP64, outer_tol=1e-9, max_outer=20))
    scale = np.abs(x1).max()
    for x in (x2, x3):
        assert np.abs(x - x1).max() <= 1e-7 * scale
    np.testing.assert_array_equal(x4, x1)
