"""The discrete-adjoint workload at full size: 249,480-tet bump channel, first-order HLLC + RK2,
J = pressure force along x on the bump wall after 20 steps.  dJ/d(inflow rho, u, p) and dJ/d(wall
vertex along the wall normal) from the checkpointed adjoint against central differences of FP64
forward runs (h = 1e-6, agreement 1e-5), and against the tangent-linear product along one seeded
direction.  bench.py runs the same check over 64 seeded wall vertices."""
import pytest

import bump_gradient as BG

pytestmark = pytest.mark.gpu


def test_full_size_gradient_matches_differences_and_tangent():
    rep = BG.check(nvert=4)
    print(f"bump gradient: {rep['passed']}/{rep['checks']} within {rep['tol']:g}, worst {rep['worst']} "
          f"{rep['max_rel_diff']:.2e}; direction {rep['direction']}", flush=True)
    for r in rep["rows"]:
        assert r["rel_diff"] <= 1e-5, r
    d = rep["direction"]
    assert d["rel_diff_tangent"] <= 1e-9, d
    assert d["rel_diff_fd"] <= 1e-5, d
