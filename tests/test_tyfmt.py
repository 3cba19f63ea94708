"""TYPHON XBIN (.tym) reader/writer against the reference's own mesh files (copied as data into
tests/golden/meshes/, from NRG/COMMON/): section structure, round trip, and the solver mesh built
from a file against the synthetic layout the NRG replays use."""
import os

import numpy as np
import pytest

import nrg
from typhon_amd import mesh as M
from typhon_amd import tyfmt as T

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = os.path.join(HERE, "golden", "meshes")


def path(name):
    return os.path.join(MESHES, name)


@pytest.mark.parametrize("name,nvtx,ncell,nbf,version", [("quad20x20.tym", 441, 400, 80, 2),
                                                         ("quad20x10.tym", 231, 200, 60, 1),
                                                         ("mxlay2x20quad-h50.tym", 63, 40, 44, 2)])
def test_read_reference_meshes(name, nvtx, ncell, nbf, version):
    v, secs = T.read_xbin(path(name))
    assert v == version
    t = T.read_tym(path(name))
    assert t.vtx.shape == (nvtx, 3)
    assert t.cells[T.ELEM_QUAD4].shape == (ncell, 4)
    assert sum(b.shape[0] for b in t.bfaces) == nbf
    assert sorted(t.marks) == ["IMAX", "IMIN", "JMAX", "JMIN"]
    # every boundary face is marked exactly once
    allm = np.sort(np.concatenate(list(t.marks.values())))
    np.testing.assert_array_equal(allm, np.arange(nbf))
    assert np.all(t.vtx[:, 2] == 0.0)


def test_write_read_round_trip(tmp_path):
    v, secs = T.read_xbin(path("quad20x20.tym"))
    out = tmp_path / "copy.tym"
    T.write_xbin(out, secs, version=v)
    assert out.read_bytes() == open(path("quad20x20.tym"), "rb").read()


def test_quad20x20_matches_the_synthetic_nrg_mesh():
    """the NRG RK replays build quad20x20 from its description (uniform unit square, periodic in
    x and y); the file's own vertices and connectivity give the same geometry"""
    bocos = [dict(family="IMIN", type=M.BC_PERIOD, partner="IMAX", period=(1.0, 0.0, 0.0)),
             dict(family="IMAX", type=M.BC_PERIOD, partner="IMIN", period=(-1.0, 0.0, 0.0)),
             dict(family="JMIN", type=M.BC_PERIOD, partner="JMAX", period=(0.0, 1.0, 0.0)),
             dict(family="JMAX", type=M.BC_PERIOD, partner="JMIN", period=(0.0, -1.0, 0.0))]
    mf = T.tym_ustmesh(path("quad20x20.tym"), bocos)
    ms, _, _, _ = nrg.setup("RK2", from_file=False)
    assert mf.ncell_int == ms.ncell_int == 400
    assert mf.nface == ms.nface
    # same cell centres (as sets) and volumes
    cf = np.round(mf.ccent[:400, :2], 12)
    cs = np.round(ms.ccent[:400, :2], 12)
    of, os_ = np.lexsort(cf.T), np.lexsort(cs.T)
    np.testing.assert_allclose(cf[of], cs[os_], atol=1e-13)
    np.testing.assert_allclose(mf.cvol[:400][of], ms.cvol[:400][os_], rtol=1e-13)
    # every periodic face has a partner whose centre is one period away
    for b in mf.bocos:
        d = mf.fcent[b.ipart] - mf.fcent[b.iface]
        assert np.allclose(np.abs(d).sum(axis=1), 1.0)


def test_stretched_mixing_layer_mesh():
    """mxlay2x20quad-h50: 2 x 20 quads on [0,1] x [-50,50], y clustered at 0 by
    y = 50 sinh((2s - 1) asinh 50) / 50 ... i.e. the file's own grid lines; periodic in x"""
    bocos = [dict(family="IMIN", type=M.BC_PERIOD, partner="IMAX", period=(1.0, 0.0, 0.0)),
             dict(family="IMAX", type=M.BC_PERIOD, partner="IMIN", period=(-1.0, 0.0, 0.0)),
             dict(family="JMIN", type=M.BC_SYM), dict(family="JMAX", type=M.BC_SYM)]
    m = T.tym_ustmesh(path("mxlay2x20quad-h50.tym"), bocos)
    assert m.ncell_int == 40
    assert abs(m.cvol[:40].sum() - 100.0) < 1e-10
    ys = np.unique(np.round(m.vtx[:, 1], 12))
    assert ys.size == 21 and ys[0] == -50.0 and ys[-1] == 50.0
    # sinh stretching: symmetric about 0, finest spacing at the centre
    np.testing.assert_allclose(ys, -ys[::-1], atol=1e-12)
    dy = np.diff(ys)
    assert dy[9] < dy[0]
