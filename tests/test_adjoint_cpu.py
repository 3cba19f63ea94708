"""Host side of the adjoint (CPU only): binomial checkpointing schedule, the torch restatement
of the mesh geometry, and the autograd plumbing of the RK composition (with a linear stand-in
for the residual, whose exact gradient is known)."""
from math import comb

import numpy as np
import pytest
import torch

import cases as K
from typhon_amd import adjoint as A
from typhon_amd import mesh as M
from typhon_amd import scheme as S


@pytest.mark.parametrize("N,snaps", [(1, 0), (2, 1), (7, 0), (10, 2), (25, 3), (40, 20), (100, 4)])
def test_treeverse_reproduces_the_stored_adjoint(N, snaps):
    """lam_0 for U_{k+1} = A_k U_k equals the product of transposes, with at most `snaps`
    checkpoints held, and no more forward steps than the binomial bound allows"""
    rng = np.random.default_rng(N * 31 + snaps)
    mats = [np.eye(3) + 0.1 * rng.normal(size=(3, 3)) for _ in range(N)]
    U0 = rng.normal(size=3)
    lamN = rng.normal(size=3)
    ref = lamN.copy()
    for k in range(N - 1, -1, -1):
        ref = mats[k].T @ ref
    count = {"adv": 0}

    def advance(U, k):
        count["adv"] += 1
        return mats[k] @ U

    seen = []

    def step_adjoint(U, k, lam):
        # the state handed over must be U_k
        Uk = U0
        for j in range(k):
            Uk = mats[j] @ Uk
        assert np.allclose(U, Uk)
        seen.append(k)
        return mats[k].T @ lam

    stats = {}
    lam0 = A.treeverse(0, N, U0, lamN, snaps, advance, step_adjoint, stats)
    assert np.allclose(lam0, ref, rtol=1e-12, atol=1e-14)
    assert seen == list(range(N - 1, -1, -1))
    assert stats.get("max_live", 0) <= snaps
    if snaps >= 1:
        t = 0
        while comb(snaps + t, snaps) < N:
            t += 1
        assert count["adv"] <= t * N


@pytest.mark.parametrize("kind", ["bump", "quad"])
def test_torch_geometry_matches_mesh_builder(kind):
    if kind == "bump":
        m = M.bump_channel_tets(6, 4, 3)
    else:
        m, *_ = K.channel_case(6, 4)
    vtx = torch.as_tensor(m.vtx)
    Sv = A.face_area_vectors(m, vtx).numpy()
    ref = m.fnorm * m.fsurf[:, None]
    assert np.abs(Sv - ref).max() <= 1e-14 * np.abs(ref).max()
    V = A.cell_volumes(m, vtx).numpy()
    assert np.abs(V - m.cvol[:m.ncell_int]).max() <= 1e-13 * m.cvol.max()


def test_torch_geometry_derivatives_match_differences():
    m = M.bump_channel_tets(5, 4, 3)
    rng = np.random.default_rng(0)
    d = rng.normal(size=m.vtx.shape) * 1e-2
    vtx = torch.as_tensor(m.vtx).clone().requires_grad_(True)
    S = A.face_area_vectors(m, vtx)
    V = A.cell_volumes(m, vtx, S)
    wS = torch.as_tensor(rng.normal(size=S.shape))
    wV = torch.as_tensor(rng.normal(size=V.shape))
    g = torch.autograd.grad((S * wS).sum() + (V * wV).sum(), vtx)[0].numpy()
    h = 1e-5

    def val(x):
        mm = M.bump_channel_tets(5, 4, 3, vtx=x)
        return (mm.fnorm * mm.fsurf[:, None] * wS.numpy()).sum() + (mm.cvol[:mm.ncell_int] * wV.numpy()).sum()
    fd = (val(m.vtx + h * d) - val(m.vtx - h * d)) / (2 * h)
    assert abs((g * d).sum() - fd) <= 1e-7 * abs(fd)


class LinearSolver:
    """stand-in for Solver with R(U) = dt * (M U + b(q)) + geometry terms, for the plumbing"""

    viscous = False

    def __init__(self, n, rng):
        self.n = n
        self.M = rng.normal(size=(5 * n, 5 * n)) * 0.1
        self.state = None

    def set_state(self, U):
        self.state = np.array(U, dtype=np.float64).reshape(5, self.n)

    def calc_rhs(self, dt):
        return dt * (self.M @ self.state.ravel()).reshape(5, self.n)

    def residual_adjoint(self, dt, lam, area_vectors=True, volumes=True):
        gu = dt * (self.M.T @ lam.ravel()).reshape(5, self.n)
        return dict(u=gu, s=np.zeros((2, 3)) if area_vectors else None, vol=np.zeros(self.n) if volumes else None,
                    q=np.zeros((1, 5)))


@pytest.mark.parametrize("method", [S.TPS_RK2, S.TPS_RK4, S.TPS_LSRK25BB])
def test_rk_step_gradient_through_autograd(method):
    rng = np.random.default_rng(method)
    n = 4
    s = LinearSolver(n, rng)
    coef = S.rk_tableau(method)
    dt = 0.3
    U = torch.as_tensor(rng.normal(size=(5, n))).requires_grad_(True)
    Sg = torch.zeros(2, 3, requires_grad=True, dtype=torch.float64)
    Vg = torch.zeros(n, requires_grad=True, dtype=torch.float64)
    qg = torch.zeros(1, 5, requires_grad=True, dtype=torch.float64)
    out = A.rk_step(U, Sg, Vg, qg, s, dt, coef)
    w = torch.as_tensor(rng.normal(size=(5, n)))
    (out * w).sum().backward()
    # the RK step of a linear residual is the polynomial sum_k (dt M)^k c_k applied to U
    R = []
    Us = np.eye(5 * n)
    for st in range(coef.shape[0]):
        R.append(dt * s.M @ Us)
        Us = np.eye(5 * n) + sum(coef[st][j] * R[j] for j in range(st + 1))
    J = Us
    ref = (J.T @ w.numpy().ravel()).reshape(5, n)
    assert np.allclose(U.grad.numpy(), ref, rtol=1e-12, atol=1e-13)
    assert np.allclose(out.detach().numpy().ravel(), J @ U.detach().numpy().ravel(), rtol=1e-12, atol=1e-13)
