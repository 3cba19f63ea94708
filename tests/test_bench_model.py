"""The algorithmic byte models bench.py prices its rooflines with: the per-cell figures of the
counting rule for each workload (each array once per pass, int32 indices, FP64 or the mixed form)."""
import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


@pytest.mark.parametrize("nc, nf, nb, kw, total, residual", [
    # 2D quad channel 256 x 128, first order, RK2
    (32768, 65920, 768, dict(dim=2, nv=4, rk4=False, gradients=False), 156.3, 108.3),
    # 6M periodic tets, MUSCL + LSQ, RK4
    (6_000_000, 12_000_000, 0, dict(), 732.0, 592.0),
    # 250k-tet bump channel, first order, RK2
    (250_000, 510_000, 20_000, dict(rk4=False, gradients=False), 193.3, 133.3),
    # 256^3 hexes NS (gradients for the viscous flux), RK4
    (16_777_216, 50_331_648, 0, dict(), 812.0, 672.0),
])
def test_fp64_bytes_per_cell(nc, nf, nb, kw, total, residual):
    bm = bench.byte_model(nc, nf, nb, **kw)
    assert bm["total"] / nc == pytest.approx(total, abs=0.05)
    assert bm["residual_only"] / nc == pytest.approx(residual, abs=0.05)


@pytest.mark.parametrize("nc, nf, nb, kw, total, residual", [
    (32768, 65920, 768, dict(dim=2, nv=4, rk4=False, gradients=False), 116.2, 76.2),
    (6_000_000, 12_000_000, 0, dict(), 485.0, 360.0),
    (4_096_000, 12_288_000, 0, dict(), None, 416.0),  # the implicit workload's JVP residual
    (16_777_216, 50_331_648, 0, dict(), 541.0, 416.0),
    # the FP32 stage-increment sum, RK4: 5 x (10 + 1.5 x 4 + 0.75 x 4) = 95 B per cell per launch
    (6_000_000, 12_000_000, 0, dict(f32sum=True), 455.0, 360.0),
    (16_777_216, 50_331_648, 0, dict(f32sum=True), 511.0, 416.0),
])
def test_mixed_bytes_per_cell(nc, nf, nb, kw, total, residual):
    bm = bench.byte_model(nc, nf, nb, mixed=True, **kw)
    if total is not None:
        assert bm["total"] / nc == pytest.approx(total, abs=0.05)
    assert bm["residual_only"] / nc == pytest.approx(residual, abs=0.05)


def test_roofline_object():
    r = bench.flux_roofline(8e9, {"flux_ms": 10.0, "flux_n": 10}, "model")
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert r["achieved"] == pytest.approx(8000.0) and r["frac"] == pytest.approx(1.0)
