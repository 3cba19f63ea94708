"""The reference's assembled implicit path: the block matrix of
build_implicit_bdlu.f90 over every cell including ghosts, its restarted GMRES (bdlu_gmres.f90), and
one implicit step (implicit_solve.f90 via tstep_implicit.f90).

CPU: the oracle's restatement against what fixes it independently --
  * the GMRES restatement against a dense numpy solve of the same assembled system (tight
    tolerance, large Krylov space), and the reference's stopping rule at the default tolerance;
  * the matrix's structure: ghost rows are the identity, slip-wall and supersonic-outlet faces
    carry no coupling to their ghosts (ns_bocojacobian.f90), interior couplings are the HLL face
    Jacobians times area and dt/V of the row's cell, diagonals of cells away from boundaries are
    the existing implicit diagonal (or_jac_diag).
GPU: the HIP path's implicit step against the oracle's.
"""
import numpy as np
import pytest

import cases as K
from oracle import oracle as OR
from typhon_amd import mesh as M
from typhon_amd import scheme as S


def _dense(m, diag, upper, lower):
    n5 = 5 * m.ncell
    A = np.zeros((n5, n5))
    for i in range(m.ncell):
        A[5 * i:5 * i + 5, 5 * i:5 * i + 5] = diag[i].reshape(5, 5)
    for f in range(m.nface):
        i1, i2 = m.ic1[f], m.ic2[f]
        A[5 * i2:5 * i2 + 5, 5 * i1:5 * i1 + 5] += lower[f].reshape(5, 5)
        A[5 * i1:5 * i1 + 5, 5 * i2:5 * i2 + 5] += upper[f].reshape(5, 5)
    return A


def _cases():
    def channel():
        return K.channel_case(12, 6, jitter=0.1)

    def periodic():
        return K.periodic_quad_case(6, 5, jitter=0.1)

    def tgv():
        return K.tgv_case(4, jitter=0.05)
    return {"channel": channel, "periodic": periodic, "tgv": tgv}


@pytest.mark.parametrize("name", ["channel", "periodic", "tgv"])
def test_bdlu_gmres_matches_dense_solve(name):
    m, s, g, u = _cases()[name]()
    o = K.make_oracle(m, s, g)
    o.set_cons(u)
    dt = o.timestep(s["cfl"]) * 5.0
    o.calc_rhs(dt)
    diag, up, lo = o.bdlu_assemble(dt)
    rhs = o.res_all().T.copy()  # [ncell, 5] block-major
    A = _dense(m, diag, up, lo)
    x = np.linalg.solve(A, rhs.ravel())
    sol, info = o.bdlu_gmres(diag, up, lo, rhs, nkrylov=min(200, 5 * m.ncell), maxres=1e-15, max_it=20)
    assert np.abs(sol.ravel() - x).max() <= 1e-9 * np.abs(x).max()
    # the reference's default tolerance (INV_RES 1e-4, NKRYLOV 50): converged per its own test
    sol2, info2 = o.bdlu_gmres(diag, up, lo, rhs)
    assert info2 >= 0
    x0 = rhs / diag[:, ::6]  # the reference's initial guess rhs / diag(ib, ib)
    r = rhs.ravel() - A @ sol2.ravel()
    assert np.linalg.norm(r) / np.linalg.norm(rhs) < np.abs(x0).sum() * 1e-4


def test_bdlu_structure():
    m, s, g, u = K.channel_case(10, 5, jitter=0.1)
    o = K.make_oracle(m, s, g)
    o.set_cons(u)
    dt = o.timestep(s["cfl"])
    o.calc_rhs(dt)
    diag, up, lo = o.bdlu_assemble(dt)
    eye = np.eye(5).ravel()
    ghosts = np.arange(m.ncell_int, m.ncell)
    np.testing.assert_array_equal(diag[ghosts], np.tile(eye, (len(ghosts), 1)))
    bf = np.arange(m.nface_int, m.nface)
    np.testing.assert_array_equal(lo[bf], 0.0)
    for b in m.bocos:
        if b.type == M.BC_SYM:
            np.testing.assert_array_equal(up[b.iface], 0.0)
        else:  # inlet / outlet: the ghost coupling stays
            assert np.abs(up[b.iface]).max() > 0
    # interior faces: HLL Jacobians at the face states, times area and dt/V of the row's cell
    qL, qR = o.face_states()
    for f in range(0, m.nface_int, 7):
        jl, jr = OR.jac_hll(qL[:, f], qR[:, f], m.fnorm[f])
        a = m.fsurf[f]
        i1, i2 = m.ic1[f], m.ic2[f]
        np.testing.assert_allclose(up[f], (a * jr).ravel() * dt / m.cvol[i1], rtol=1e-13, atol=1e-16)
        np.testing.assert_allclose(lo[f], -(a * jl).ravel() * dt / m.cvol[i2], rtol=1e-13, atol=1e-16)
    # cells with no boundary face: the implicit diagonal without boundary corrections
    jd = o.jac_diag(dt).reshape(m.ncell_int, 25)
    bcells = set(m.ic1[m.nface_int:].tolist())
    inner = np.array([i for i in range(m.ncell_int) if i not in bcells])
    np.testing.assert_allclose(diag[inner], jd[inner], rtol=1e-13, atol=1e-15)


def test_bdlu_step_is_the_gmres_solution():
    m, s, g, u = K.channel_case(12, 6, jitter=0.1)
    o = K.make_oracle(m, s, g)
    o.set_cons(u)
    dt = o.timestep(s["cfl"]) * 5.0
    o.calc_rhs(dt)
    diag, up, lo = o.bdlu_assemble(dt)
    sol, info = o.bdlu_gmres(diag, up, lo, o.res_all().T.copy())
    info2 = o.implicit_bdlu_step(dt)
    assert info2 == info
    np.testing.assert_allclose(o.get_cons(), u + sol[:m.ncell_int].T, rtol=1e-13, atol=1e-13)


def test_bdlu_sym_boundary_jacobian():
    """slip-wall rows: the diagonal gains A jR S dt/V with S the momentum mirror (ns_bocojacobian),
    against the implicit diagonal without boundary corrections (or_jac_diag)"""
    m, s, g, u = K.channel_case(10, 5, jitter=0.1)
    o = K.make_oracle(m, s, g)
    o.set_cons(u)
    dt = o.timestep(s["cfl"])
    o.calc_rhs(dt)
    diag, up, lo = o.bdlu_assemble(dt)
    jd = o.jac_diag(dt).reshape(m.ncell_int, 25)
    qL, qR = o.face_states()
    sym = [b for b in m.bocos if b.type == M.BC_SYM]
    other_b = set()
    for b in m.bocos:
        if b.type != M.BC_SYM:
            other_b.update(m.ic1[b.iface].tolist())
    expect = {}
    for b in sym:
        for f in b.iface:
            c = int(m.ic1[f])
            n = m.fnorm[f]
            S = np.eye(5)
            S[2:, 2:] -= 2.0 * np.outer(n, n)
            jl, jr = OR.jac_hll(qL[:, f], qR[:, f], n)
            expect[c] = expect.get(c, 0.0) + (m.fsurf[f] * (jr @ S)).ravel() * dt / m.cvol[c]
    checked = 0
    for c, e in expect.items():
        if c in other_b:
            continue
        np.testing.assert_allclose(diag[c] - jd[c], e, rtol=1e-10, atol=1e-14)
        checked += 1
    assert checked > 5


def _gpu_cases():
    def bump():
        return K.bump_case(8, 5, 4)
    d = _cases()
    d["bump"] = bump
    # HLL Jacobians with the kinetic fluxes' own wave speeds
    d["channel_hllekb"] = lambda: K.channel_case(12, 6, jitter=0.1, sch_hyp=S.SCH_HLLEKB)
    d["channel_hllck"] = lambda: K.channel_case(12, 6, jitter=0.1, sch_hyp=S.SCH_HLLCK)
    return d


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["channel", "periodic", "tgv", "bump", "channel_hllekb", "channel_hllck"])
def test_gpu_bdlu_step_matches_oracle(name):
    from typhon_amd.solver import Solver
    m, s, g, u = _gpu_cases()[name]()
    gpu = Solver(m, s, g)
    gpu.set_state(u)
    orc = K.make_oracle(m, s, g)
    orc.set_cons(u)
    dt = orc.timestep(s["cfl"]) * 5.0
    nv = m.dim + 2
    for _ in range(2):
        ig = gpu.implicit_bdlu_step(dt)
        io = orc.implicit_bdlu_step(dt)
        assert ig == io and ig >= 0
        err = K.rel_err(gpu.get_state()[:nv], orc.get_cons()[:nv])
        assert err <= 1e-10, f"state rel err {err:.3e}"
