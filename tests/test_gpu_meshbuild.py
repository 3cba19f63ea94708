"""Mesh preprocessing on the device (ty_meshbuild_*: the faces of create_face_connect,
MESHCONNECT.f90:34-163, and the geometry of calc_meshgeom, MESHGEOM.f90:421-461) against the host
build (typhon_amd/mesh.py, numpy).

The host build is the checker here: the meshes it builds feed the residual tests that match the
reference's NRG golden histories, and tests/test_oracle.py checks its geometry against closed forms
(test_geometry_closed_forms) and against the oracle's C restatement of MESHGEOM.f90
(test_meshgeom_c_restatement_matches_builder).  Bar: bit for bit on every array -- face order, facecell, facevtex, boundary faces and periodic
partners, and every geometric array (the device evaluates the host's expressions in the host's
order, without FMA contraction).  At the 256^3-hex size the host build takes minutes, so there the
device build is checked through properties that fix it: volumes add up to the box, every cell is
closed (its outward area vectors sum to zero), normals point from the left to the right cell, every
cell has six faces and periodic partners are translates of each other."""
import ctypes as C
import os
import time

import numpy as np
import pytest

from typhon_amd import _lib
from typhon_amd import mesh as M
from typhon_amd import tyfmt

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def tri_channel(nx=14, ny=9, lx=1.4, ly=0.9, jitter=0.2, seed=5, device=None):
    """2D channel of 2 nx ny jittered triangles (each quad split along alternating diagonals)"""
    h = lx / nx
    i, j = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="ij")
    i, j = i.ravel(), j.ravel()
    vid = i * (ny + 1) + j
    inner = (i > 0) & (i < nx) & (j > 0) & (j < ny)
    dx = np.where(inner, jitter * h * M.uniform_pm1(seed, vid, 0), 0.0)
    dy = np.where(inner, jitter * h * M.uniform_pm1(seed, vid, 1), 0.0)
    vtx = np.stack([i * h + dx, j * (ly / ny) + dy, np.zeros_like(dx)], axis=1)
    ci, cj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    ci, cj = ci.ravel(), cj.ravel()
    v00, v10 = ci * (ny + 1) + cj, (ci + 1) * (ny + 1) + cj
    v11, v01 = v10 + 1, v00 + 1
    alt = (ci + cj) % 2 == 0
    t1 = np.where(alt[:, None], np.stack([v00, v10, v11], 1), np.stack([v00, v10, v01], 1))
    t2 = np.where(alt[:, None], np.stack([v00, v11, v01], 1), np.stack([v10, v11, v01], 1))
    cells = np.stack([t1, t2], axis=1).reshape(-1, 3)

    def fam(fv, fc):
        out = np.empty(fv.shape[0], dtype=object)
        out[np.isclose(fc[:, 0], 0.0)] = "IMIN"
        out[np.isclose(fc[:, 0], lx)] = "IMAX"
        out[np.isclose(fc[:, 1], 0.0)] = "JMIN"
        out[np.isclose(fc[:, 1], ly)] = "JMAX"
        return out
    bocos = [dict(family="IMIN", type=M.BC_INLET_SUP, param=[1.0, 1.0, 2.0, 1.0, 0.0, 0.0]),
             dict(family="IMAX", type=M.BC_OUTLET_SUP), dict(family="JMIN", type=M.BC_SYM),
             dict(family="JMAX", type=M.BC_SYM)]
    return M.build_ustmesh(2, vtx, cells, "tri", fam, bocos, None, device)


def _tym_mixing_layer(device=None):
    """the reference's mxlay2x20quad-h50.tym: periodic in x, moving adiabatic walls"""
    bocos = [dict(family="IMIN", type=M.BC_PERIOD, partner="IMAX", period=(1.0, 0.0, 0.0)),
             dict(family="IMAX", type=M.BC_PERIOD, partner="IMIN", period=(-1.0, 0.0, 0.0)),
             dict(family="JMIN", type=M.BC_WALL_ADIAB, param=[0.0, 0.0, 0.0, 0.0, -1.0, 0.0, 0.0]),
             dict(family="JMAX", type=M.BC_WALL_ADIAB, param=[0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0])]
    return tyfmt.tym_ustmesh(os.path.join(HERE, "golden", "meshes", "mxlay2x20quad-h50.tym"), bocos, device=device)


def _tym_periodic(device=None):
    bocos = [dict(family="IMIN", type=M.BC_PERIOD, partner="IMAX", period=(1.0, 0.0, 0.0)),
             dict(family="IMAX", type=M.BC_PERIOD, partner="IMIN", period=(-1.0, 0.0, 0.0)),
             dict(family="JMIN", type=M.BC_PERIOD, partner="JMAX", period=(0.0, 1.0, 0.0)),
             dict(family="JMAX", type=M.BC_PERIOD, partner="JMIN", period=(0.0, -1.0, 0.0))]
    return tyfmt.tym_ustmesh(os.path.join(HERE, "golden", "meshes", "quad20x20.tym"), bocos, device=device)


MAKERS = {
    "tri channel": tri_channel,
    "quad channel": lambda device=None: M.channel_quad(24, 12, device=device),
    "periodic quad": lambda device=None: M.periodic_quad(10, 6, jitter=0.1, device=device),
    "quad20x20.tym periodic": _tym_periodic,
    "mxlay2x20quad-h50.tym walls": _tym_mixing_layer,
    "periodic tets": lambda device=None: M.box_tets(6, device=device),
    "walled tets": lambda device=None: M.box_tets(5, periodic=False, device=device),
    "bump tets": lambda device=None: M.bump_channel_tets(12, 6, 4, device=device),
    "periodic hexes": lambda device=None: M.box_hex(8, jitter=0.1, device=device),
    "channel hexes": lambda device=None: M.box_hex(7, jitter=0.1, periodic=(True, False, True), device=device),
    "hexes 40^3": lambda device=None: M.box_hex(40, jitter=0.05, device=device),
}

ARRAYS = ("ic1", "ic2", "fvtx", "fnorm", "fcent", "fsurf", "ccent", "cvol", "vtx")


def assert_same_mesh(a, b):
    for k in ("dim", "ncell_int", "ncell", "nface", "nface_int"):
        assert getattr(a, k) == getattr(b, k), k
    for k in ARRAYS:
        x, y = getattr(a, k), getattr(b, k)
        assert x.shape == y.shape, k
        assert x.dtype == y.dtype, k
        assert np.array_equal(x, y), (k, np.abs(x - y).max())
    assert len(a.bocos) == len(b.bocos)
    for p, q in zip(a.bocos, b.bocos):
        assert (p.family, p.type) == (q.family, q.type)
        assert np.array_equal(p.iface, q.iface) and np.array_equal(p.param, q.param)
        assert (p.ipart is None) == (q.ipart is None)
        if p.ipart is not None:
            assert np.array_equal(p.ipart, q.ipart)


@pytest.mark.parametrize("name", list(MAKERS))
def test_device_build_matches_host_bitwise(name):
    print(f"\nmesh build: {name}", flush=True)
    host = MAKERS[name]()
    dev = MAKERS[name](device=0)
    assert_same_mesh(host, dev)


def test_device_build_errors():
    """the argument checks (all on the host before any launch) and the non-manifold report"""
    lib = _lib.load()
    mb = C.c_void_p()
    assert lib.ty_meshbuild_create(C.byref(mb), 0) == 0
    i32 = lambda a: a.ctypes.data_as(_lib.I32P)  # noqa: E731
    f64 = lambda a: a.ctypes.data_as(_lib.F64P)  # noqa: E731
    nf, nfi = C.c_int32(), C.c_int32()
    try:
        vtx = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0], [1, 1, 0]], dtype=np.float64)
        out = np.zeros(64, dtype=np.int32)
        dout = np.zeros(64)
        assert lib.ty_meshbuild_finish(mb, i32(out), i32(out), i32(out), f64(dout), f64(dout), f64(dout)) == 1
        assert b"first" in lib.ty_meshbuild_error(mb)
        # vertex number outside 1..nvtex
        ev = np.ascontiguousarray(np.array([[1, 2, 7]], dtype=np.int32).T)
        assert lib.ty_meshbuild_faces(mb, 2, 13, 1, i32(ev), 6, f64(vtx), C.byref(nf), C.byref(nfi)) == 1
        assert b"outside" in lib.ty_meshbuild_error(mb)
        # element type not offered in 2D
        assert lib.ty_meshbuild_faces(mb, 2, 101, 1, i32(ev), 6, f64(vtx), C.byref(nf), C.byref(nfi)) == 4
        # three triangles on the edge (1, 2)
        ev = np.ascontiguousarray(np.array([[1, 2, 3], [2, 1, 5], [1, 2, 6]], dtype=np.int32).T)
        assert lib.ty_meshbuild_faces(mb, 2, 13, 3, i32(ev), 6, f64(vtx), C.byref(nf), C.byref(nfi)) == 1
        assert b"non-manifold" in lib.ty_meshbuild_error(mb)
        # two triangles: 5 faces, 1 internal; a boundary order that is not a permutation
        ev = np.ascontiguousarray(np.array([[1, 2, 3], [2, 1, 5]], dtype=np.int32).T)
        assert lib.ty_meshbuild_faces(mb, 2, 13, 2, i32(ev), 6, f64(vtx), C.byref(nf), C.byref(nfi)) == 0
        assert (nf.value, nfi.value) == (5, 1)
        bfv = np.zeros((4, 4), dtype=np.int32)
        bc = np.zeros((4, 3))
        assert lib.ty_meshbuild_boundary(mb, i32(bfv), f64(bc)) == 0
        # boundary faces in order of first appearance: (2,3), (3,1) of the first triangle, then (1,5), (5,2)
        assert bfv[:2].T.tolist() == [[2, 3], [3, 1], [1, 5], [5, 2]]
        assert np.array_equal(bc, 0.5 * (vtx[bfv[0] - 1] + vtx[bfv[1] - 1]))
        fc = np.zeros(10, dtype=np.int32)
        fv = np.zeros(20, dtype=np.int32)
        face = np.zeros(35)
        cen = np.zeros(6)
        vol = np.zeros(2)
        bad = np.array([1, 2, 2, 4], dtype=np.int32)
        assert lib.ty_meshbuild_finish(mb, i32(bad), i32(fc), i32(fv), f64(face), f64(cen), f64(vol)) == 1
        assert b"permutation" in lib.ty_meshbuild_error(mb)
        border = np.array([4, 3, 2, 1], dtype=np.int32)
        assert lib.ty_meshbuild_finish(mb, i32(border), i32(fc), i32(fv), f64(face), f64(cen), f64(vol)) == 0
        # internal face (1,2) between elements 1 and 2, then the boundary faces in the given order
        assert fc.reshape(2, 5).T.tolist() == [[1, 2], [2, 0], [2, 0], [1, 0], [1, 0]]
        assert fv.reshape(4, 5)[:2].T.tolist() == [[1, 2], [5, 2], [1, 5], [3, 1], [2, 3]]
        assert np.allclose(vol, [0.5, 0.5], rtol=1e-15)
        assert np.allclose(cen.reshape(2, 3), [[1 / 3, 1 / 3, 0], [1 / 3, -1 / 3, 0]], rtol=1e-15, atol=1e-15)
    finally:
        lib.ty_meshbuild_destroy(mb)


def _topology_checks(m, L):
    ni, nfi = m.ncell_int, m.nface_int
    assert (m.cvol[:ni] > 0).all()
    # six faces per cell
    cnt = np.bincount(m.ic1, minlength=ni)[:ni] + np.bincount(m.ic2[:nfi], minlength=ni)
    assert (cnt == 6).all()
    # closed cells: sum of outward area vectors = 0 (to round-off of the face areas)
    h = L / round(ni ** (1 / 3))
    sf = m.fnorm * m.fsurf[:, None]
    for d in range(3):
        s = np.bincount(m.ic1, weights=sf[:, d], minlength=m.ncell)[:ni]
        s -= np.bincount(m.ic2[:nfi], weights=sf[:nfi, d], minlength=ni)[:ni]
        assert np.abs(s).max() <= 1e-12 * h * h, d
    # normals from the left cell to the right cell (ghost centres of periodic faces included)
    v12 = m.ccent[m.ic2] - m.ccent[m.ic1]
    assert (np.einsum("ij,ij->i", v12, m.fnorm) > 0).all()
    # periodic partners: translates by L along one axis (their vertex means: the face centres of
    # warped faces depend on the vertex order, which differs between partners), and each other's
    # partners
    vmean = m.vtx[m.fvtx].mean(axis=1)
    fpart = np.full(m.nface, -1)
    for b in m.bocos:
        d = vmean[b.ipart] - vmean[b.iface]
        assert np.allclose(np.sort(np.abs(d), axis=1), [0.0, 0.0, L], atol=1e-12 * L)
        fpart[b.iface] = b.ipart
    assert (fpart[nfi:] >= 0).all()
    assert np.array_equal(fpart[fpart[nfi:]], np.arange(nfi, m.nface))


def test_full_size_hex_box_properties():
    """the 256^3-hex periodic box (16,777,216 cells, 50,331,648 faces) built on the device: cubes
    (every face, centre and volume known in closed form), then the jittered box (warped faces)"""
    n, L = 256, 2 * np.pi
    h = L / n
    t0 = time.perf_counter()
    m = M.box_hex(n, L, jitter=0.0, device=0)
    print(f"\n256^3 device mesh build {time.perf_counter() - t0:.1f} s", flush=True)
    ni, nfi = m.ncell_int, m.nface_int
    assert (ni, m.nface, nfi) == (n ** 3, 3 * n ** 3 + 3 * n * n, 3 * n ** 3 - 3 * n * n)
    # the vertices i h carry round-off of order n ulp(h): volumes and areas to n x 2e-15
    assert np.allclose(m.cvol[:ni], h ** 3, rtol=5e-13, atol=0)
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    cc = (np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1) + 0.5) * h
    assert np.abs(m.ccent[:ni] - cc).max() <= 1e-13 * L
    assert np.allclose(m.fsurf, h * h, rtol=5e-13, atol=0)
    assert np.allclose(np.sort(np.abs(m.fnorm), axis=1), [0.0, 0.0, 1.0], rtol=0, atol=1e-15)
    _topology_checks(m, L)
    del m, cc, i, j, k
    m = M.box_hex(n, L, jitter=0.05, device=0)
    # pyramid volumes of warped faces (calc_ust_elemvol) add up to the box only approximately
    assert abs(m.cvol[:m.ncell_int].sum() - L ** 3) <= 1e-7 * L ** 3
    _topology_checks(m, L)
