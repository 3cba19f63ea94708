"""Gradient of the bump-channel wall force, the discrete-adjoint workload: 3D first-order HLLC +
RK2 on the jittered tet channel with the sin^2 bump (63 x 33 x 20 hexes x 6 = 249,480 tets at full
size), J = pressure force along x on the bump wall after ``nsteps`` steps of a fixed time-step
schedule (the nominal run's): 20 steps for the wall vertices, 400 for the inflow state, which
reaches the bump after about 300.

- adjoint: typhon_amd.adjoint.unsteady_gradient (torch.autograd over the library's residual and
  its hand-written adjoint kernel, binomial checkpointing) gives dJ/d(inflow rho, u, p) and
  dJ/d(vertex coordinates); for the inflow alone, unsteady_gradient_device (the library's
  device-resident reverse sweep);
- central differences of FP64 forward runs: the inflow state through the inlet's parameters, and
  seeded wall vertices moved along the wall's outward normal (geometry rebuilt by mesh.moved, the
  mesh uploaded again);
- tangent-linear product along one seeded direction: the library's exact residual tangent
  (ty_residual_tangent) carried through the RK stages, the geometry's tangent from torch.

Used by bench.py (full size: 64 wall vertices; the inflow state) and tests/test_gpu_bump_gradient.py."""
from __future__ import annotations

import dataclasses
import time

import numpy as np
import torch

import cases as K
from typhon_amd import adjoint as A
from typhon_amd import mesh as M
from typhon_amd import scheme as S
from typhon_amd.solver import Solver

INFLOW = (1.0, 0.5 * np.sqrt(K.GAMMA), 1.0)  # rho, u, p


def with_inflow(m, inflow):
    """the mesh with the supersonic inlet's parameters set for the static state (rho, u, p)"""
    rho, u, p = inflow
    par = np.zeros(8)
    v = K.freestream_inlet(rho, p, u / np.sqrt(K.GAMMA * p / rho))
    par[:len(v)] = v
    bocos = [dataclasses.replace(b, param=par.copy()) if b.type == M.BC_INLET_SUP else b for b in m.bocos]
    return dataclasses.replace(m, bocos=bocos)


def wall_vertices(m, nxyz, count, seed=64, lx=3.0):
    """``count`` seeded vertices of the bump (j = 0, 1 < x < 2, 0 < k < nz) and their outward unit
    normals (area-weighted sum of the adjacent wall faces' area vectors).  Off the bump the wall is
    flat and J, a force along x, moves by O(h^2) only: central differences there are round-off."""
    nx, ny, nz = nxyz
    ib = np.arange(1, nx)
    ib = ib[(ib * lx / nx > 1.0) & (ib * lx / nx < 2.0)]
    i, k = np.meshgrid(ib, np.arange(1, nz), indexing="ij")
    cand = ((i.ravel() * (ny + 1)) * (nz + 1) + k.ravel()).astype(np.int64)
    vids = np.sort(np.random.default_rng(seed).choice(cand, size=min(count, cand.size), replace=False))
    b = m.boco("JMIN")
    fv = m.fvtx[b.iface, :3]
    Sf = m.fnorm[b.iface] * m.fsurf[b.iface, None]
    nrm = np.zeros((m.vtx.shape[0], 3))
    for j in range(3):
        np.add.at(nrm, fv[:, j], Sf)
    n = nrm[vids]
    return vids, n / np.linalg.norm(n, axis=1)[:, None]


class BumpGradient:
    def __init__(self, nxyz=(63, 33, 20), nsteps=20, snaps=20):
        self.nxyz = nxyz
        self.m, self.s, self.g, self.U0 = K.bump_case(*nxyz, inflow=INFLOW)
        self.solver = Solver(self.m, self.s, self.g)
        self.coef = S.rk_tableau(self.s["tps_meth"])
        self.snaps = snaps
        self.solver.set_state(self.U0)
        self.dts = []
        for _ in range(nsteps):  # the nominal run's time steps, then held fixed
            dt = self.solver.timestep()
            self.dts.append(dt)
            self.solver.rk_step(dt)
        self.mesh = self.m  # mesh currently uploaded

    def close(self):
        self.solver.close()

    def objective(self, mesh):
        return lambda U, Sx: A.wall_force_x(mesh, U, Sx, "JMIN")

    def J(self, mesh):
        """J of a forward FP64 run on ``mesh`` (same connectivity, other vertices or inlet)"""
        if mesh is not self.mesh:
            self.solver.upload(mesh)
            self.mesh = mesh
        self.solver.set_state(self.U0)
        for dt in self.dts:
            self.solver.rk_step(dt)
        U = torch.as_tensor(self.solver.get_state())
        return float(self.objective(mesh)(U, torch.as_tensor(mesh.fnorm * mesh.fsurf[:, None])))

    def adjoint(self):
        """(J, dJ/d(rho, u, p)_inflow [3], dJ/dvtx [nvtx, 3], seconds)"""
        if self.mesh is not self.m:
            self.solver.upload(self.m)
            self.mesh = self.m
        t0 = time.perf_counter()
        J, _, gvtx, gq = A.unsteady_gradient(self.solver, self.m, self.U0, self.dts, self.coef,
                                             self.objective(self.m), snaps=self.snaps)
        ib = [k for k, b in enumerate(self.m.bocos) if b.type == M.BC_INLET_SUP][0]
        # q = (rho, p, u, v, w) of the inlet's ghost state; the inlet direction is x
        return J, np.array([gq[ib, 0], gq[ib, 2], gq[ib, 1]]), gvtx, time.perf_counter() - t0

    def adjoint_inflow_device(self):
        """(J, dJ/d(rho, u, p)_inflow [3], seconds) from the device-resident reverse sweep
        (ty_unsteady_adjoint; it gives the state and inlet sensitivities, not the vertex ones)"""
        if self.mesh is not self.m:
            self.solver.upload(self.m)
            self.mesh = self.m
        t0 = time.perf_counter()
        J, _, gq, _ = A.unsteady_gradient_device(self.solver, self.m, self.U0, self.dts, self.objective(self.m),
                                                 snaps=self.snaps)
        ib = [k for k, b in enumerate(self.m.bocos) if b.type == M.BC_INLET_SUP][0]
        return J, np.array([gq[ib, 0], gq[ib, 2], gq[ib, 1]]), time.perf_counter() - t0

    def fd_inflow(self, k, h_rel=1e-6):
        h = h_rel * INFLOW[k]
        fp, fm = np.array(INFLOW), np.array(INFLOW)
        fp[k] += h
        fm[k] -= h
        return (self.J(with_inflow(self.m, fp)) - self.J(with_inflow(self.m, fm))) / (2 * h)

    def fd_vertex(self, vid, n, h=1e-6):
        vp, vm = self.m.vtx.copy(), self.m.vtx.copy()
        vp[vid] += h * n
        vm[vid] -= h * n
        return (self.J(M.moved(self.m, vp)) - self.J(M.moved(self.m, vm))) / (2 * h)

    def fd_direction(self, dvtx, dinflow, h=1e-6):
        f0 = np.array(INFLOW)
        Jp = self.J(with_inflow(M.moved(self.m, self.m.vtx + h * dvtx), f0 + h * dinflow))
        Jm = self.J(with_inflow(M.moved(self.m, self.m.vtx - h * dvtx), f0 - h * dinflow))
        return (Jp - Jm) / (2 * h)

    def tangent(self, dvtx, dinflow):
        """dJ along (dvtx [nvtx, 3], d(rho, u, p)_inflow): exact residual tangents through the RK
        stages (U_s = U0 + sum_j coef[s-1][j] R_j, treelevel_rungekutta.f90:60-105)"""
        if self.mesh is not self.m:
            self.solver.upload(self.m)
            self.mesh = self.m
        m, sv = self.m, self.solver
        vtx = torch.as_tensor(m.vtx, dtype=torch.float64)

        def geo(v):
            Sv = A.face_area_vectors(m, v)
            return Sv, A.cell_volumes(m, v, Sv)
        _, (dS, dV) = torch.autograd.functional.jvp(geo, vtx, torch.as_tensor(dvtx, dtype=torch.float64))
        dS, dV = dS.numpy(), dV.numpy()
        ib = [k for k, b in enumerate(m.bocos) if b.type == M.BC_INLET_SUP][0]
        dq = np.zeros((len(m.bocos), 5))
        dq[ib, :3] = (dinflow[0], dinflow[2], dinflow[1])
        U, dU = self.U0.copy(), np.zeros_like(self.U0)
        ns = self.coef.shape[0]
        for dt in self.dts:
            Rs, dRs = [], []
            Us, dUs = U, dU
            for s in range(ns):
                sv.set_state(Us)
                Rs.append(sv.calc_rhs(dt))
                dRs.append(sv.residual_tangent(dt, du=dUs, ds=dS, dvol=dV, dq=dq))
                Us, dUs = U.copy(), dU.copy()
                for j in range(s + 1):
                    if self.coef[s][j] != 0.0:
                        Us += self.coef[s][j] * Rs[j]
                        dUs += self.coef[s][j] * dRs[j]
            U, dU = Us, dUs
        Ut = torch.as_tensor(U).requires_grad_(True)
        St = torch.as_tensor(m.fnorm * m.fsurf[:, None]).requires_grad_(True)
        gU, gS = torch.autograd.grad(self.objective(m)(Ut, St), (Ut, St))
        return float((gU.numpy() * dU).sum() + (gS.numpy() * dS).sum())


def check(nxyz=(63, 33, 20), nsteps=20, nvert=64, inflow=False, h=1e-6, tol=1e-5, log=None):
    """adjoint vs central differences and vs the tangent-linear product along one seeded direction
    through every checked parameter; returns a report dict.  Parameters: ``nvert`` seeded bump-wall
    vertices along the wall normal, and with ``inflow`` the inlet's rho, u, p (they reach the bump
    only after about 300 steps: information from the inlet travels at u + a = 1.77 and the time step
    is 1.8e-3).  Each row has its difference relative to the component (rel_diff) and to the norm
    of the checked gradient (rel_to_norm): a component far smaller than the others (rho at fixed u
    and p, 1.4e-7 against 1e-2) is compared at the accuracy central differences of J reach, which
    is set by J's round-off and its kinks (the min/max of the HLLC wave speeds), not by h alone."""
    t0 = time.perf_counter()
    bg = BumpGradient(nxyz, nsteps)
    try:
        if nvert == 0:  # inlet state only: the device sweep
            J, gin, t_adj = bg.adjoint_inflow_device()
            gvtx = np.zeros_like(bg.m.vtx)
            sweep = "device (ty_unsteady_adjoint)"
        else:
            J, gin, gvtx, t_adj = bg.adjoint()
            sweep = "host (torch.autograd over ty_residual_adjoint)"
        rows = []
        if inflow:
            for k, name in enumerate(("rho", "u", "p")):
                fd = bg.fd_inflow(k, h)
                rows.append(dict(param=f"inflow {name}", adjoint=float(gin[k]), fd=float(fd)))
        vids, nrm = wall_vertices(bg.m, nxyz, nvert)
        for vid, n in zip(vids, nrm):
            fd = bg.fd_vertex(vid, n, h)
            rows.append(dict(param=f"wall vertex {int(vid)} (normal)", adjoint=float(gvtx[vid] @ n), fd=float(fd)))
            if log and len(rows) % 16 == 0:
                log(f"bump gradient: {len(rows)} of {len(vids) + 3 * inflow} differences")
        gnorm = float(np.sqrt(sum(r["adjoint"] ** 2 for r in rows)))
        for r in rows:
            r["rel_diff"] = float(abs(r["adjoint"] - r["fd"]) / max(abs(r["fd"]), 1e-300))
            r["rel_to_norm"] = float(abs(r["adjoint"] - r["fd"]) / max(gnorm, 1e-300))
        # one seeded direction through every checked parameter
        rng = np.random.default_rng(7)
        c = rng.normal(size=len(vids))
        dvtx = np.zeros_like(bg.m.vtx)
        dvtx[vids] = c[:, None] * nrm
        din = rng.normal(size=3) * np.array(INFLOW) if inflow else np.zeros(3)
        ad_dir = float((gvtx * dvtx).sum() + gin @ din)
        tl_dir = bg.tangent(dvtx, din)
        fd_dir = bg.fd_direction(dvtx, din, h)
        worst = max(rows, key=lambda r: r["rel_diff"])
        params = ", ".join(([f"inflow rho, u, p"] if inflow else []) + ([f"{len(vids)} bump-wall vertices"] if nvert else []))
        return dict(workload=f"3D first-order HLLC + RK2, bump channel, {bg.m.ncell_int} tets, J = pressure force "
                             f"along x on the bump wall after {nsteps} steps; parameters: {params}",
                    J=J, steps=nsteps, checks=len(rows), tol=tol,
                    passed=int(sum(bool(r["rel_diff"] <= tol) for r in rows)),
                    max_rel_diff=float(worst["rel_diff"]), worst=worst["param"], h=h, adjoint_s=round(t_adj, 2),
                    reverse_sweep=sweep,
                    gradient_norm=gnorm, max_rel_to_norm=float(max(r["rel_to_norm"] for r in rows)),
                    direction=dict(adjoint=ad_dir, tangent_linear=tl_dir, fd=fd_dir,
                                   rel_diff_tangent=float(abs(ad_dir - tl_dir) / abs(tl_dir)),
                                   rel_diff_fd=float(abs(ad_dir - fd_dir) / abs(fd_dir))),
                    rows=rows, total_s=round(time.perf_counter() - t0, 1))
    finally:
        bg.close()
