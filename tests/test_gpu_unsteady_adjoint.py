"""Device-resident unsteady adjoint (ty_unsteady_adjoint) against the host-driven one.

The host path (typhon_amd.adjoint.unsteady_gradient: residual adjoints per stage through
torch.autograd, checkpoints on the host) is itself checked against central differences of GPU
runs in test_gpu_adjoint.py / test_gpu_adjoint_visc.py.  Here the device sweep must give the same
dJ/dU0 and dJ/d(inlet state) to round-off, for any snapshot budget, with the forward-step count
of binomial checkpointing.
"""
import numpy as np
import pytest
import torch

import cases as K
from typhon_amd import adjoint as A
from typhon_amd import scheme as S
from typhon_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _viscous(m, s, g, u, mu=0.02):
    return m, dict(s, sch_dis=S.DIS_CELLFULL, gradmeth=S.GRAD_LSQ), \
        dict(g, typ_fluid=S.EQ_NSLAM, typ_visc=S.VISC_DYNCST, visc_dyn=mu), u


def _schedule(gpu, U0, N):
    dts, U = [], U0
    for _ in range(N):
        gpu.set_state(U)
        dt = gpu.timestep()
        dts.append(dt)
        gpu.rk_step(dt)
        U = gpu.get_state()
    gpu.set_state(U0)
    return np.array(dts)


CASES = {
    "bump_euler_rk2": lambda: K.bump_case(8, 4, 3),
    "bump_ns_rk2": lambda: _viscous(*K.bump_case(8, 4, 3)),
    "bump_euler_rk4": lambda: K.bump_case(8, 4, 3, tps_meth=S.TPS_RK4),
    "bump_ns_lsrk25": lambda: _viscous(*K.bump_case(8, 4, 3, tps_meth=S.TPS_LSRK25BB)),
}


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name", list(CASES))
def test_device_sweep_matches_host_sweep(name):
    m, s, g, U0 = CASES[name]()
    gpu = Solver(m, s, g)
    N = 9
    dts = _schedule(gpu, U0, N)
    coef = S.rk_tableau(s["tps_meth"])
    obj = lambda U, Sx: A.wall_force_x(m, U, Sx, "JMIN")  # noqa: E731
    J_h, gU_h, _, gq_h = A.unsteady_gradient(gpu, m, U0, dts, coef, obj, snaps=2)
    J_d, gU_d, gq_d, st = A.unsteady_gradient_device(gpu, m, U0, dts, obj, snaps=2)
    assert abs(J_d - J_h) <= 1e-12 * abs(J_h)
    assert _rel(gU_d, gU_h) <= 1e-10, _rel(gU_d, gU_h)
    assert _rel(gq_d, gq_h) <= 1e-10, _rel(gq_d, gq_h)
    assert st["adjoint_steps"] == N and st["max_snapshots"] <= 2
    # the state is U0 again
    assert np.array_equal(gpu.get_state()[:, :5], np.asarray(U0)[:, :5])


def _count_forward(N, snaps):
    stats = {"fwd": 0}

    def advance(U, k):
        stats["fwd"] += 1
        return U

    A.treeverse(0, N, 0, 0, snaps, advance, lambda U, k, lam: lam)
    return stats["fwd"]


def test_snapshot_budget_changes_work_not_result():
    m, s, g, U0 = _viscous(*K.bump_case(8, 4, 3))
    gpu = Solver(m, s, g)
    N = 13
    dts = _schedule(gpu, U0, N)
    rng = np.random.default_rng(2)
    lamN = np.zeros((5, m.ncell_int))
    lamN[:5] = rng.normal(size=(5, m.ncell_int))
    ref = None
    for snaps in (0, 1, 3, 20):
        gpu.set_state(U0)
        out = gpu.unsteady_adjoint(dts, lamN, snaps)
        st = out["stats"]
        assert st["adjoint_steps"] == N
        assert st["max_snapshots"] <= snaps
        assert st["forward_steps"] == _count_forward(N, snaps), (snaps, st)
        if ref is None:
            ref = out
        else:
            assert _rel(out["u"], ref["u"]) <= 1e-13
            assert _rel(out["q"], ref["q"]) <= 1e-13


def test_mixed_context_gives_the_fp64_gradient():
    """a mixed-precision context recomputes the trajectory in FP64, so its gradient is the FP64
    map's: the same as an FP64 context's, with min(snaps, N - 1) checkpoint buffers"""
    m, s, g, U0 = _viscous(*K.bump_case(8, 4, 3))
    N = 9
    res = {}
    for prec in (S.PREC_FP64, S.PREC_MIXED):
        gpu = Solver(m, dict(s, precision=prec), g)
        dts = _schedule(gpu, U0, N) if prec == S.PREC_FP64 else res[S.PREC_FP64][2]
        J, gU, gq, st = A.unsteady_gradient_device(gpu, m, U0, dts, lambda U, Sx: A.wall_force_x(m, U, Sx, "JMIN"),
                                                   snaps=3)
        assert st["snapshots_allocated"] == 3 and st["max_snapshots"] <= 3
        assert gpu.scheme["precision"] == prec
        res[prec] = (J, gU, dts, gq)
        gpu.close()
    assert abs(res[S.PREC_MIXED][0] - res[S.PREC_FP64][0]) <= 1e-14 * abs(res[S.PREC_FP64][0])
    assert _rel(res[S.PREC_MIXED][1], res[S.PREC_FP64][1]) <= 1e-13
    assert _rel(res[S.PREC_MIXED][3], res[S.PREC_FP64][3]) <= 1e-13


def test_snapshot_buffers_bounded_by_steps():
    m, s, g, U0 = K.bump_case(8, 4, 3)
    gpu = Solver(m, s, g)
    dts = _schedule(gpu, U0, 3)
    lamN = np.random.default_rng(3).normal(size=(5, m.ncell_int))
    for N, snaps, alloc in ((3, 20, 2), (1, 20, 0), (3, 1, 1)):
        gpu.set_state(U0)
        st = gpu.unsteady_adjoint(dts[:N], lamN, snaps)["stats"]
        assert st["snapshots_allocated"] == alloc, (N, snaps, st)
        assert st["max_snapshots"] <= alloc
