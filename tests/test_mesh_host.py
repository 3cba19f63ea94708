"""Host mesh build (typhon_amd/mesh.py): periodic pairing, boundary grouping and the errors the
build stops on.  The geometry itself is pinned in test_oracle.py (closed forms and the oracle's C
restatement of MESHGEOM.f90); the device build is compared with this one in test_gpu_meshbuild.py."""
import numpy as np
import pytest

from typhon_amd import mesh as M


@pytest.mark.parametrize("mk, L", [(lambda: M.box_hex(6, 2.0, jitter=0.1), 2.0),
                                   (lambda: M.box_tets(4, 3.0), 3.0),
                                   (lambda: M.box_hex(5, 1.0, periodic=(True, False, True)), 1.0),
                                   (lambda: M.periodic_quad(7, 5, lx=1.0, ly=1.0, jitter=0.1), 1.0)])
def test_periodic_partners_are_mutual_translates(mk, L):
    m = mk()
    fpart = np.full(m.nface, -1)
    nper = 0
    for b in m.bocos:
        if b.type == M.BC_PERIOD:
            assert b.ipart is not None and len(b.ipart) == len(b.iface)
            fpart[b.iface] = b.ipart
            nper += len(b.iface)
    assert nper > 0
    per = np.nonzero(fpart >= 0)[0]
    assert np.array_equal(fpart[fpart[per]], per)
    # partners carry the same vertices up to one period along one axis (vertex means: the centre
    # of a warped face depends on its vertex order)
    nvf = 2 if m.dim == 2 else (4 if m.fvtx[0, 3] >= 0 else 3)
    vmean = m.vtx[m.fvtx[:, :nvf]].mean(axis=1)
    d = np.sort(np.abs(vmean[fpart[per]] - vmean[per]), axis=1)
    assert np.allclose(d, [0.0, 0.0, L], atol=1e-12 * L)
    # ghost cells of periodic faces sit across the face from the partner's interior cell
    g = m.ic2[per]
    assert np.allclose(m.cvol[g], m.cvol[m.ic1[fpart[per]]], rtol=0, atol=0)


def test_boundary_faces_grouped_by_condition_in_spec_order():
    m = M.channel_quad(8, 4)
    start = m.nface_int
    for b in m.bocos:
        assert np.array_equal(b.iface, np.arange(start, start + len(b.iface)))
        start += len(b.iface)
    assert start == m.nface
    # internal faces: left cell below right cell, right cells of boundary faces are their ghosts
    assert (m.ic1[:m.nface_int] < m.ic2[:m.nface_int]).all()
    assert np.array_equal(m.ic2[m.nface_int:], m.ncell_int + np.arange(m.nface - m.nface_int))


def _tri_pair_fam(fv, fc):
    return np.array(["WALL"] * fv.shape[0], dtype=object)


def test_non_manifold_mesh_is_refused():
    vtx = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, -1, 0], [1, 1, 0]], dtype=np.float64)
    cells = np.array([[0, 1, 2], [1, 0, 3], [0, 1, 4]])
    with pytest.raises(ValueError, match="non-manifold"):
        M.build_ustmesh(2, vtx, cells, "tri", _tri_pair_fam, [dict(family="WALL", type=M.BC_SYM)])


def test_boundary_face_without_family_is_refused():
    vtx = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, -1, 0]], dtype=np.float64)
    cells = np.array([[0, 1, 2], [1, 0, 3]])

    def fam(fv, fc):
        out = np.array(["WALL"] * fv.shape[0], dtype=object)
        out[0] = None
        return out
    with pytest.raises(ValueError, match="without a boundary condition family"):
        M.build_ustmesh(2, vtx, cells, "tri", fam, [dict(family="WALL", type=M.BC_SYM)])


def test_two_triangles_by_hand():
    """faces in order of first appearance, internal first; areas, centroids and normals by hand"""
    vtx = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, -1, 0]], dtype=np.float64)
    cells = np.array([[0, 1, 2], [1, 0, 3]])
    m = M.build_ustmesh(2, vtx, cells, "tri", _tri_pair_fam, [dict(family="WALL", type=M.BC_SYM)])
    assert (m.nface, m.nface_int, m.ncell_int, m.ncell) == (5, 1, 2, 6)
    assert m.fvtx[:, :2].tolist() == [[0, 1], [1, 2], [2, 0], [0, 3], [3, 1]]
    assert m.ic1.tolist() == [0, 0, 0, 1, 1] and m.ic2.tolist() == [1, 2, 3, 4, 5]
    assert np.allclose(m.cvol[:2], [0.5, 0.5], rtol=1e-15)
    assert np.allclose(m.ccent[:2], [[1 / 3, 1 / 3, 0], [1 / 3, -1 / 3, 0]], atol=1e-15)
    # the internal face's normal points from cell 0 (y > 0) to cell 1 (y < 0)
    assert np.allclose(m.fnorm[0], [0.0, -1.0, 0.0]) and m.fsurf[0] == 1.0
