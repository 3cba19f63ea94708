"""Discrete adjoint of the first-order explicit residual path, behind torch.autograd.

The reference has no derivative code; this module differentiates the path it does have
(calc_rhs.f90:10 residual, treelevel_rungekutta.f90:50-113 stages) exactly:

* ``Residual`` is a ``torch.autograd.Function`` whose forward is the library's residual
  R(U; S, V, q) = -dt/V sum_f F.A (fixed dt) and whose backward is the hand-written
  adjoint kernel (``ty_residual_adjoint``): gradients w.r.t. the state U [5, ncell_int],
  the face area vectors S [nface, 3], the cell volumes V [ncell_int] and the supersonic
  inlet ghost states q [nboco, 5].
* ``rk_step`` composes residuals with the RK tableau, so torch differentiates the stage
  combination and calls ``Residual.backward`` for every stage.
* ``face_area_vectors`` / ``cell_volumes`` restate the mesh geometry (MESHGEOM.f90:37-331) in
  torch, so a gradient w.r.t. S and V reaches vertex coordinates.
* ``unsteady_gradient`` runs N steps forward and the adjoint backward with binomial
  checkpointing (``treeverse``: at most ``snaps`` stored states, Griewank's schedule).

Scope: first-order schemes in FP64 contexts, inviscid (every flux; state, area-vector, volume
and inlet-state gradients) and laminar Navier-Stokes (HLLC/HLLCB; state, volume and inlet-state
gradients), with symmetry, periodic, supersonic inlet and outlet and subsonic outlet boundaries;
the library refuses the rest.  The time step of every step is an input (held fixed under
differentiation), as a fixed-dt schedule.  ``unsteady_gradient_device`` accepts a mixed-precision
solver: the library then recomputes the trajectory in FP64, so the gradient is that of the FP64
map, and the forward sweep for the objective here runs in FP64 too.
"""
from __future__ import annotations

from math import comb

import numpy as np
import torch


class Residual(torch.autograd.Function):
    """R = residual(U; S, V, q) on the solver's mesh.  S, V and q must hold the geometry
    and inlet states the solver was built with: they are inputs so that autograd routes
    their gradients, the forward value does not read them.  The gradient w.r.t. S is
    computed only when S requires it (not provided for viscous residuals)."""

    @staticmethod
    def forward(ctx, U, S, V, q, solver, dt):
        solver.set_state(U.detach().cpu().numpy())
        R = solver.calc_rhs(dt)
        ctx.solver = solver
        ctx.dt = dt
        ctx.save_for_backward(U)
        return torch.from_numpy(np.ascontiguousarray(R))

    @staticmethod
    def backward(ctx, gR):
        (U,) = ctx.saved_tensors
        s = ctx.solver
        s.set_state(U.detach().cpu().numpy())
        g = s.residual_adjoint(ctx.dt, gR.detach().cpu().numpy(), area_vectors=ctx.needs_input_grad[1],
                               volumes=ctx.needs_input_grad[2])
        t = [None if g[k] is None else torch.from_numpy(np.ascontiguousarray(g[k])) for k in ("u", "s", "vol", "q")]
        return t[0], t[1], t[2], t[3], None, None


def rk_step(U, S, V, q, solver, dt, coef):
    """one RK step, U_s = U0 + sum_{j<s} coef[s-1][j] R_j, U_new = U0 + sum_j coef[n-1][j] R_j
    (treelevel_rungekutta.f90:60-105)"""
    ns = coef.shape[0]
    Rs = []
    Us = U
    for s in range(ns):
        Rs.append(Residual.apply(Us, S, V, q, solver, dt))
        Us = U
        for j in range(s + 1):
            if coef[s][j] != 0.0:
                Us = Us + float(coef[s][j]) * Rs[j]
    return Us


# ---------------------------------------------------------------------------
# geometry in torch (MESHGEOM.f90 calc_ust_face :37-139, calc_ust_midcell/elemvol :148-331)

def face_area_vectors(mesh, vtx):
    """S_f = A_f n_f [nface, 3] from vertex coordinates (torch, differentiable), oriented as
    the mesh's normals (owner -> other)."""
    fv = torch.as_tensor(mesh.fvtx, dtype=torch.long)
    v1, v2 = vtx[fv[:, 0]], vtx[fv[:, 1]]
    if mesh.dim == 2:
        e = v2 - v1
        S = torch.stack([-e[:, 1], e[:, 0], torch.zeros_like(e[:, 0])], dim=1)
    else:
        v3 = vtx[fv[:, 2]]
        S = 0.5 * torch.linalg.cross(v2 - v1, v3 - v1)
        quad = fv[:, 3] >= 0
        if bool(quad.any()):
            v4 = vtx[torch.where(quad, fv[:, 3], fv[:, 0])]
            S2 = 0.5 * torch.linalg.cross(v3 - v1, v4 - v1)
            S = S + torch.where(quad[:, None], S2, torch.zeros_like(S2))
    sign = np.sign(np.einsum("ij,ij->i", S.detach().numpy(), mesh.fnorm))
    return S * torch.as_tensor(sign)[:, None]


def face_centres(mesh, vtx):
    fv = torch.as_tensor(mesh.fvtx, dtype=torch.long)
    if mesh.dim == 2:
        return 0.5 * (vtx[fv[:, 0]] + vtx[fv[:, 1]])
    if (mesh.fvtx[:, 3] >= 0).any():
        raise NotImplementedError("torch face centres: triangular faces only")
    return (vtx[fv[:, 0]] + vtx[fv[:, 1]] + vtx[fv[:, 2]]) / 3.0


def cell_volumes(mesh, vtx, S=None):
    """V_c = kvol sum_f |(c_f - mid_c) . S_f| with mid_c the mean of the cell's face centres
    (pyramids on the faces, MESHGEOM.f90:148-331)"""
    if S is None:
        S = face_area_vectors(mesh, vtx)
    fc = face_centres(mesh, vtx)
    n = mesh.ncell_int
    ic1 = torch.as_tensor(mesh.ic1, dtype=torch.long)
    ic2 = torch.as_tensor(mesh.ic2, dtype=torch.long)
    inner = ic2 < n
    cnt = torch.zeros(n, dtype=vtx.dtype).index_add_(0, ic1, torch.ones_like(ic1, dtype=vtx.dtype))
    cnt = cnt.index_add_(0, ic2[inner], torch.ones(int(inner.sum()), dtype=vtx.dtype))
    mid = torch.zeros(n, 3, dtype=vtx.dtype).index_add_(0, ic1, fc).index_add_(0, ic2[inner], fc[inner])
    mid = mid / cnt[:, None]
    kvol = 0.5 if mesh.dim == 2 else 1.0 / 3.0
    V = torch.zeros(n, dtype=vtx.dtype)
    V = V.index_add_(0, ic1, kvol * torch.abs(((fc - mid[ic1]) * S).sum(1)))
    V = V.index_add_(0, ic2[inner], kvol * torch.abs(((fc[inner] - mid[ic2[inner]]) * S[inner]).sum(1)))
    return V


def wall_force(mesh, U, S, family, axis=0, gamma=1.4):
    """J = sum over the family's faces of p(U of the wall cell) S_axis (pressure force along the
    axis on the wall, S pointing out of the domain)"""
    b = mesh.boco(family)
    f = torch.as_tensor(b.iface, dtype=torch.long)
    c = torch.as_tensor(mesh.ic1[b.iface], dtype=torch.long)
    rho, rhoE, mom = U[0, c], U[1, c], U[2:2 + mesh.dim, c]
    p = (gamma - 1.0) * (rhoE - 0.5 * (mom * mom).sum(0) / rho)
    return (p * S[f, axis]).sum()


def wall_force_x(mesh, U, S, family, gamma=1.4):
    """wall_force along x"""
    return wall_force(mesh, U, S, family, 0, gamma)


# ---------------------------------------------------------------------------
# binomial checkpointing

def treeverse_split(steps: int, snaps: int) -> int:
    """where to place the next checkpoint in a range of ``steps`` steps with ``snaps`` free
    snapshots: the smallest repetition number t with C(snaps + t, snaps) >= steps, then the
    right part gets C(snaps - 1 + t, snaps - 1) steps (Griewank, Optim. Methods Softw. 1 (1992))"""
    if steps < 2:
        return 1
    t = 0
    while comb(snaps + t, snaps) < steps:
        t += 1
    right = comb(snaps - 1 + t, snaps - 1) if snaps >= 1 else 1
    return min(max(1, steps - right), steps - 1)


def treeverse(a: int, b: int, Ua, lam_b, snaps: int, advance, step_adjoint, stats=None):
    """Adjoint of the steps a..b-1: given the state at a (Ua) and the adjoint at b, return the
    adjoint at a.  advance(U, k) -> state after step k; step_adjoint(U_k, k, lam_{k+1}) -> lam_k.
    At most ``snaps`` extra states are held at once."""
    if stats is not None:
        stats["max_live"] = max(stats.get("max_live", 0), stats.get("live", 0))
    if b - a == 1:
        return step_adjoint(Ua, a, lam_b)
    if snaps == 0:
        lam = lam_b
        for k in range(b - 1, a - 1, -1):
            U = Ua
            for j in range(a, k):
                U = advance(U, j)
            lam = step_adjoint(U, k, lam)
        return lam
    m = a + treeverse_split(b - a, snaps)
    U = Ua
    for j in range(a, m):
        U = advance(U, j)
    if stats is not None:
        stats["live"] = stats.get("live", 0) + 1
    lam_m = treeverse(m, b, U, lam_b, snaps - 1, advance, step_adjoint, stats)
    if stats is not None:
        stats["live"] -= 1
    return treeverse(a, m, Ua, lam_m, snaps, advance, step_adjoint, stats)


def unsteady_gradient(solver, mesh, U0, dts, coef, objective, vtx=None, q=None, snaps=20):
    """Gradient of J(U_N, S) after len(dts) RK steps from U0 (fixed dt schedule).

    objective(U_N, S) -> scalar torch tensor.  vtx: torch [nvtx, 3] vertex coordinates (the
    mesh's), q: torch [nboco, 5] inlet ghost states (the scheme's).  Returns
    (J, dJ/dU0 [5, ncell_int], dJ/dvtx [nvtx, 3], dJ/dq [nboco, 5]); dJ/dvtx is None for a
    viscous residual (its flux also depends on the cell and face centres, which the area
    vectors and volumes do not carry)."""
    geometry = not solver.viscous
    vtx = torch.as_tensor(mesh.vtx if vtx is None else vtx, dtype=torch.float64).clone().requires_grad_(geometry)
    nb = max(1, len(mesh.bocos))
    q = torch.zeros(nb, 5, dtype=torch.float64) if q is None else torch.as_tensor(q, dtype=torch.float64)
    q = q.clone().requires_grad_(True)
    S = face_area_vectors(mesh, vtx)
    V = cell_volumes(mesh, vtx, S)
    Sd, Vd = S.detach().requires_grad_(geometry), V.detach().requires_grad_(geometry)
    gS = torch.zeros_like(Sd)
    gV = torch.zeros_like(Vd)
    U0 = torch.as_tensor(U0, dtype=torch.float64)
    N = len(dts)

    def advance(U, k):
        with torch.no_grad():
            return rk_step(U, Sd, Vd, q, solver, float(dts[k]), coef)

    def step_adjoint(U, k, lam):
        nonlocal gS, gV
        Ui = U.detach().clone().requires_grad_(True)
        for t in (Sd, Vd, q):
            t.grad = None
        out = rk_step(Ui, Sd, Vd, q, solver, float(dts[k]), coef)
        torch.autograd.backward(out, lam)
        if geometry:
            gS = gS + Sd.grad
            gV = gV + Vd.grad
        return Ui.grad.detach()

    # forward to U_N for the objective (no checkpoint kept: treeverse recomputes)
    U = U0
    for k in range(N):
        U = advance(U, k)
    UN = U.detach().clone().requires_grad_(True)
    Sj = Sd.detach().clone().requires_grad_(True)
    J = objective(UN, Sj)
    lamN, gS_obj = torch.autograd.grad(J, (UN, Sj), allow_unused=True)
    lamN = torch.zeros_like(UN) if lamN is None else lamN
    qgrad = torch.zeros_like(q)

    def step_adjoint_q(U, k, lam):
        nonlocal qgrad
        r = step_adjoint(U, k, lam)
        if q.grad is not None:
            qgrad = qgrad + q.grad
        return r

    lam0 = treeverse(0, N, U0, lamN, snaps, advance, step_adjoint_q) if N > 0 else lamN
    if not geometry:
        return float(J.detach()), lam0.numpy(), None, qgrad.detach().numpy()
    if gS_obj is not None:
        gS = gS + gS_obj
    # chain S(vtx), V(vtx) -> vtx
    gvtx = torch.autograd.grad((S, V), vtx, (gS, gV), allow_unused=True)[0]
    return float(J.detach()), lam0.numpy(), gvtx.numpy(), qgrad.detach().numpy()


def unsteady_gradient_device(solver, mesh, U0, dts, objective, snaps=20):
    """As unsteady_gradient for dJ/dU0 and dJ/dq, with the reverse sweep on the device
    (ty_unsteady_adjoint: checkpoints and stage states in device memory).  One forward sweep gives
    U_N; the objective and lam_N = dJ/dU_N are evaluated with torch on the host.  Returns
    (J, dJ/dU0 [5, ncell_int], dJ/dq [nboco, 5], stats); the solver's state is U0 on return.
    The objective is evaluated on the FP64 trajectory, the map the library's sweep linearises."""
    UN = torch.as_tensor(solver.advance_fp64(U0, dts)).requires_grad_(True)
    S = torch.as_tensor(mesh.fnorm * mesh.fsurf[:, None])
    J = objective(UN, S)
    (lamN,) = torch.autograd.grad(J, (UN,), allow_unused=True)
    lamN = np.zeros(UN.shape) if lamN is None else lamN.numpy()
    solver.set_state(U0)
    out = solver.unsteady_adjoint(np.asarray(dts, dtype=np.float64), lamN, snaps)
    return float(J.detach()), out["u"], out["q"], out["stats"]
