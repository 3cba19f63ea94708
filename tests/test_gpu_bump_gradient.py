"""The discrete-adjoint workload at full size: 249,480-tet bump channel, first-order HLLC + RK2,
J = pressure force along x on the bump wall.  dJ/d(wall vertex along the wall normal) after 20
steps, and dJ/d(inflow rho, u, p) after 400 steps (the inflow reaches the bump after about 300),
from the checkpointed adjoint against central differences of FP64 forward runs (h = 1e-6, agreement
1e-5; rho, whose sensitivity is 7e4 times smaller than u's and p's, against the gradient's norm),
and against the tangent-linear product along one seeded direction.  bench.py runs the vertex check
over 64 seeded wall vertices."""
import pytest

import bump_gradient as BG

pytestmark = pytest.mark.gpu


def _report(rep):
    print(f"bump gradient: {rep['passed']}/{rep['checks']} within {rep['tol']:g}, worst {rep['worst']} "
          f"{rep['max_rel_diff']:.2e}; direction {rep['direction']}", flush=True)


def test_full_size_wall_vertex_gradient_matches_differences_and_tangent():
    rep = BG.check(nvert=4)
    _report(rep)
    for r in rep["rows"]:
        assert r["rel_diff"] <= 1e-5, r
    d = rep["direction"]
    assert d["rel_diff_tangent"] <= 1e-9, d
    assert d["rel_diff_fd"] <= 1e-5, d


def test_full_size_inflow_gradient_matches_differences_and_tangent():
    rep = BG.check(nsteps=400, nvert=0, inflow=True)
    _report(rep)
    rows = {r["param"]: r for r in rep["rows"]}
    for name in ("inflow u", "inflow p"):
        assert abs(rows[name]["adjoint"]) > 1e-3, rows[name]
        assert rows[name]["rel_diff"] <= 1e-5, rows[name]
    # rho at fixed u, p moves J 7e4 times less than u or p: compared against the gradient's norm
    assert rows["inflow rho"]["rel_to_norm"] <= 1e-8, rows["inflow rho"]
    d = rep["direction"]
    assert d["rel_diff_tangent"] <= 1e-9, d
    assert d["rel_diff_fd"] <= 1e-5, d
