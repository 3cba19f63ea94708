"""Unstructured mesh in Typhon's own form, built on the host with numpy.

This is the host side of the drop-in: it produces exactly the arrays the
reference holds after ``init_world`` (SOURCE/MAIN/main.f90:95) and hands them to
the C-ABI once.  Geometry restates CFDTOOLS/Mesh/MESHGEOM.f90:37-461, face
ordering follows MESHCONNECT.f90:538-548 (internal faces first, then boundary
faces grouped by boundary condition), ghost-cell geometry follows
SOURCE/ZONE/init_ustboco_ghostcell.f90:65-73, init_ustboco_ghostface.f90:53-54
and init_connect_grid.f90:226-231.

All indices here are 0-based; ``ic2`` of a boundary face is its ghost cell
(ncell_int + k).  The C-ABI converts to the reference's 1-based arrays.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

# boundary condition codes (SOURCE/MODCOM/VARCOM.f90:75-93)
BC_SYM = 10
BC_PERIOD = 11
BC_WALL_ADIAB = 20
BC_WALL_ISOTH = 21
BC_WALL_FLUX = 22
BC_WALL_HCONV = 23  # CONVECTION_WALL and GEN_CONV_WALL: refused for NS (MENU_NS.f90:95-113)
BC_WALL_HGEN = 24
BC_INLET_SUB = 30
BC_INLET_SUP = 31
BC_OUTLET_SUB = 35
BC_OUTLET_SUP = 36

# ghost-cell geometry kind per BC type (EQNS/MENU_NS.f90:91-114, PARAM/MENU_BOCO.f90:127-136)
_GHOSTCELL = {BC_SYM, BC_WALL_ADIAB, BC_WALL_ISOTH, BC_WALL_FLUX}


@dataclass
class Boco:
    """One boundary condition: ``umesh%boco(ib)`` plus its resolved parameters."""

    family: str
    type: int
    iface: np.ndarray                     # 0-based face indices
    ipart: np.ndarray | None = None       # periodic: 0-based partner face per face
    param: np.ndarray = field(default_factory=lambda: np.zeros(8))


@dataclass
class UstMesh:
    dim: int
    ncell_int: int
    ncell: int
    nface: int
    nface_int: int
    ic1: np.ndarray      # int32 [nface]
    ic2: np.ndarray      # int32 [nface]
    fnorm: np.ndarray    # [nface,3]
    fcent: np.ndarray    # [nface,3]
    fsurf: np.ndarray    # [nface]
    ccent: np.ndarray    # [ncell,3]
    cvol: np.ndarray     # [ncell]
    bocos: list
    vtx: np.ndarray      # [nvtx,3]
    fvtx: np.ndarray     # int32 [nface,4], -1 padded

    def boco(self, family: str) -> Boco:
        for b in self.bocos:
            if b.family == family:
                return b
        raise KeyError(family)


# ---------------------------------------------------------------------------
# seeded counter-based RNG: SplitMix64 keyed by (seed, entity id, draw)

def _splitmix64(x: np.ndarray) -> np.ndarray:
    x = (x + np.uint64(0x9E3779B97F4A7C15)).astype(np.uint64)
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)).astype(np.uint64)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)).astype(np.uint64)
    return z ^ (z >> np.uint64(31))


def uniform_pm1(seed: int, ids: np.ndarray, draw: int) -> np.ndarray:
    """Uniform [-1, 1) from SplitMix64(seed*2^32 ^ (3*id + draw))."""
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) << np.uint64(32)) ^ (ids.astype(np.uint64) * np.uint64(3) + np.uint64(draw))
        z = _splitmix64(key)
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


# ---------------------------------------------------------------------------
# local faces of each element type (vertex order defines the geometry split)

_LOCAL_FACES = {
    "quad": [(0, 1), (1, 2), (2, 3), (3, 0)],
    "tri": [(0, 1), (1, 2), (2, 0)],
    "tet": [(0, 2, 1), (0, 1, 3), (1, 2, 3), (0, 3, 2)],
    "hex": [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)],
}


def _extract_faces(cells: np.ndarray, etype: str):
    """Unique faces of an element mesh (create_facevtex.f90:15-220).

    Returns fvtx [nf,4] (first occurrence's vertex order), cell1, cell2 (-1 when
    boundary), in order of first appearance (cell-major)."""
    lf = _LOCAL_FACES[etype]
    nloc = len(lf)
    nvf = len(lf[0])
    ncell = cells.shape[0]
    fv = np.empty((ncell, nloc, 4), dtype=np.int64)
    fv[:] = -1
    for k, f in enumerate(lf):
        for j, v in enumerate(f):
            fv[:, k, j] = cells[:, v]
    fv = fv.reshape(ncell * nloc, 4)
    owner = np.repeat(np.arange(ncell, dtype=np.int64), nloc)
    srt = np.sort(fv[:, :nvf], axis=1)
    nvtx = int(cells.max()) + 1
    bits = int(np.ceil(np.log2(max(nvtx, 2))))
    first = np.ones(srt.shape[0], dtype=bool)
    if bits * nvf <= 63:
        key = np.zeros(srt.shape[0], dtype=np.int64)
        for j in range(nvf):
            key = (key << bits) | srt[:, j]
        order = np.argsort(key, kind="stable")
        ks = key[order]
        first[1:] = ks[1:] != ks[:-1]
    else:  # too many vertices to pack the sorted ids into one int64: sort on the columns
        order = np.lexsort(srt.T[::-1])
        ss = srt[order]
        first[1:] = np.any(ss[1:] != ss[:-1], axis=1)
    grp = np.cumsum(first) - 1
    nuniq = int(grp[-1]) + 1
    # first occurrence (lowest local index in cell-major order, stable sort)
    fidx = order[first]
    c1 = owner[fidx]
    c2 = np.full(nuniq, -1, dtype=np.int64)
    second = ~first
    c2[grp[second]] = owner[order[second]]
    if np.bincount(grp, minlength=nuniq).max() > 2:
        raise ValueError("non-manifold mesh: a face is shared by more than two cells")
    # order of first appearance
    app = np.argsort(fidx, kind="stable")
    return fv[fidx[app]], c1[app], c2[app]


def _periodic_keys(bfvtx: np.ndarray, nvf: int, wrap: np.ndarray) -> np.ndarray:
    """one id per set of wrapped vertex ids of each boundary face: periodic partners share it"""
    ws = np.sort(wrap[bfvtx[:, :nvf]], axis=1)
    if ws.shape[0] == 0:
        return np.zeros(0, dtype=np.int64)
    _, inv = np.unique(ws, axis=0, return_inverse=True)
    return inv.ravel()


# ---------------------------------------------------------------------------
# geometry: CFDTOOLS/Mesh/MESHGEOM.f90

def _dot(a, b):
    """row-wise dot product of [n,3] arrays, summed x, y, z in that order"""
    return a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1] + a[:, 2] * b[:, 2]


def _face_geometry(dim: int, vtx: np.ndarray, fvtx: np.ndarray):
    """calc_ust_face (MESHGEOM.f90:37-139)."""
    nf = fvtx.shape[0]
    v1 = vtx[fvtx[:, 0]]
    v2 = vtx[fvtx[:, 1]]
    if dim == 2:
        e = v2 - v1
        nrm = np.stack([-e[:, 1], e[:, 0], np.zeros(nf)], axis=1)
        surf = np.sqrt(nrm[:, 0] * nrm[:, 0] + nrm[:, 1] * nrm[:, 1] + nrm[:, 2] * nrm[:, 2])
        cen = .5 * (v1 + v2)
    else:
        v3 = vtx[fvtx[:, 2]]
        a = v2 - v1
        b = v3 - v1
        n1 = .5 * np.cross(a, b)
        quad = fvtx[:, 3] >= 0
        nrm = n1.copy()
        cen = (v1 + v2 + v3) / 3.0
        if quad.any():
            v4 = vtx[np.where(quad, fvtx[:, 3], fvtx[:, 0])]
            s1 = np.sqrt(_dot(n1, n1))
            n2 = .5 * np.cross(b, v4 - v1)
            s2 = np.sqrt(_dot(n2, n2))
            cg2 = (v1 + v3 + v4) / 3.0
            qc = (s1[:, None] * cen + s2[:, None] * cg2) / (s1 + s2)[:, None]
            nrm = np.where(quad[:, None], n1 + n2, n1)
            cen = np.where(quad[:, None], qc, cen)
        surf = np.sqrt(_dot(nrm, nrm))
    return nrm / surf[:, None], cen, surf


def _cell_geometry(dim, ncell_int, ic1, ic2, fnorm, fcent, fsurf):
    """calc_ust_midcell / calc_ust_elemvol / calc_ust_cell (MESHGEOM.f90:148-331)."""
    inner = (ic2 >= 0) & (ic2 < ncell_int)
    nbf = np.bincount(ic1, minlength=ncell_int) + np.bincount(ic2[inner], minlength=ncell_int)
    mid = np.zeros((ncell_int, 3))
    for d in range(3):
        mid[:, d] = (np.bincount(ic1, weights=fcent[:, d], minlength=ncell_int)
                     + np.bincount(ic2[inner], weights=fcent[inner, d], minlength=ncell_int))
    mid /= nbf[:, None]
    kcg, kvol = (1. / 3.0, 1. / 2.0) if dim == 2 else (1. / 4.0, 1. / 3.0)
    vol = np.zeros(ncell_int)
    cen = np.zeros((ncell_int, 3))
    for side, cells, sel in ((0, ic1, slice(None)), (1, ic2[inner], inner)):
        kh = fcent[sel] - mid[cells]
        h = np.abs(_dot(kh, fnorm[sel]))
        v = kvol * h * fsurf[sel]
        cg = fcent[sel] - kcg * kh
        vol += np.bincount(cells, weights=v, minlength=ncell_int)
        for d in range(3):
            cen[:, d] += np.bincount(cells, weights=v * cg[:, d], minlength=ncell_int)
    cen /= vol[:, None]
    return cen, vol


ELEMTYPE = {"tri": 13, "quad": 24, "tet": 101, "hex": 141}  # ELEMVTEX.f90 element codes

# wall-clock seconds of the phases of the last build_ustmesh call (for benchmarks)
LAST_BUILD: dict = {}


def _boco_groups(face_family, bocos_spec, bfvtx, bcent):
    """boundary faces (in order of first appearance) of each boco of the spec, as index arrays"""
    fam = np.asarray(face_family(bfvtx, bcent))
    groups = [np.nonzero(fam == spec["family"])[0] for spec in bocos_spec]
    if sum(len(g) for g in groups) != bfvtx.shape[0]:
        raise ValueError("boundary faces without a boundary condition family")
    return groups


def _host_faces(dim, vtx, cells, etype, face_family, bocos_spec):
    """faces in the final order and geometry, on the host (numpy)"""
    fvtx, c1, c2 = _extract_faces(cells, etype)
    ncell_int = cells.shape[0]
    internal = c2 >= 0
    fn, fc, fs = _face_geometry(dim, vtx, fvtx)
    bidx = np.nonzero(~internal)[0]
    groups = _boco_groups(face_family, bocos_spec, fvtx[bidx], fc[bidx])
    # internal faces first, then boundary faces grouped by boco in spec order
    order = np.concatenate([np.nonzero(internal)[0]] + [bidx[g] for g in groups])
    fvtx = fvtx[order].astype(np.int32)
    c1 = c1[order]
    c2 = c2[order]
    fn, fc, fs = fn[order], fc[order], fs[order]
    nface_int = int(internal.sum())
    # calc_ust_checkface (MESHGEOM.f90:362-410): ic1 < ic2 for internal faces
    swap = (c2 >= 0) & (c1 > c2)
    t = c1[swap].copy()
    c1[swap] = c2[swap]
    c2[swap] = t
    ccent_i, cvol_i = _cell_geometry(dim, ncell_int, c1, c2, fn, fc, fs)
    v12 = np.where((c2 >= 0)[:, None], ccent_i[np.maximum(c2, 0)], fc) - ccent_i[c1]
    flip = _dot(v12, fn) < 0.0
    fn[flip] = -fn[flip]
    return fvtx, c1, c2, fn, fc, fs, ccent_i, cvol_i, nface_int, [len(g) for g in groups]


def _device_faces(device, dim, vtx, cells, etype, face_family, bocos_spec):
    """the same arrays from the library's device mesh build (ty_meshbuild_*); no host fallback"""
    import ctypes as C
    from . import _lib
    lib = _lib.load()
    mb = C.c_void_p()
    rc = lib.ty_meshbuild_create(C.byref(mb), int(device))
    if rc:
        raise _lib.TyError(rc, "ty_meshbuild_create")

    def check(rc, what):
        if rc:
            raise _lib.TyError(rc, f"{what}: {lib.ty_meshbuild_error(mb).decode()}")
    i32 = lambda a: a.ctypes.data_as(_lib.I32P)  # noqa: E731
    f64 = lambda a: a.ctypes.data_as(_lib.F64P)  # noqa: E731
    try:
        ncell = cells.shape[0]
        ev = np.ascontiguousarray(cells.T, dtype=np.int32) + 1  # elemvtex(1:ncell, 1:nv), 1-based
        vt = np.ascontiguousarray(vtx, dtype=np.float64)
        nf, nfi = C.c_int32(), C.c_int32()
        t0 = time.perf_counter()
        check(lib.ty_meshbuild_faces(mb, dim, ELEMTYPE[etype], ncell, i32(ev), vt.shape[0], f64(vt), C.byref(nf),
                                     C.byref(nfi)), "ty_meshbuild_faces")
        LAST_BUILD["device_faces"] = time.perf_counter() - t0
        del ev
        nface, nface_int = nf.value, nfi.value
        nb = nface - nface_int
        bfv = np.empty((4, nb), dtype=np.int32)
        bc = np.empty((nb, 3))
        check(lib.ty_meshbuild_boundary(mb, i32(bfv), f64(bc)), "ty_meshbuild_boundary")
        groups = _boco_groups(face_family, bocos_spec, bfv.T - 1, bc)
        border = (np.concatenate(groups) + 1).astype(np.int32) if nb else np.zeros(1, dtype=np.int32)
        facecell = np.empty((2, nface), dtype=np.int32)
        facevtex = np.empty((4, nface), dtype=np.int32)
        face = np.empty((nface, 7))
        cen = np.empty((ncell, 3))
        vol = np.empty(ncell)
        t0 = time.perf_counter()
        check(lib.ty_meshbuild_finish(mb, i32(border), i32(facecell), i32(facevtex), f64(face), f64(cen), f64(vol)),
              "ty_meshbuild_finish")
        LAST_BUILD["device_finish"] = time.perf_counter() - t0
    finally:
        lib.ty_meshbuild_destroy(mb)
    c1 = facecell[0].astype(np.int64) - 1
    c2 = facecell[1].astype(np.int64) - 1
    fvtx = np.ascontiguousarray(facevtex.T) - 1
    fn = np.ascontiguousarray(face[:, 0:3])
    fc = np.ascontiguousarray(face[:, 3:6])
    fs = np.ascontiguousarray(face[:, 6])
    return fvtx, c1, c2, fn, fc, fs, cen, vol, nface_int, [len(g) for g in groups]


def build_ustmesh(dim: int, vtx: np.ndarray, cells: np.ndarray, etype: str,
                  face_family, bocos_spec: list, wrap: np.ndarray | None = None, device: int | None = None) -> UstMesh:
    """Build Typhon's ``st_ustmesh`` from an element mesh.

    face_family(fvtx_boundary, centres) -> array of family names per boundary face.
    bocos_spec: list of dicts {family, type, param[, partner]} in boundary order;
    a periodic boco names its partner family, paired through ``wrap``.
    device: None builds faces and geometry on the host (numpy); a GPU index builds them there
    (ty_meshbuild_*: the same arrays bit for bit).  Boundary grouping, periodic pairing and ghost
    cells (boundary-sized work) stay on the host either way.
    """
    t0 = time.perf_counter()
    LAST_BUILD.clear()
    if device is None:
        out = _host_faces(dim, vtx, cells, etype, face_family, bocos_spec)
    else:
        out = _device_faces(device, dim, vtx, cells, etype, face_family, bocos_spec)
    LAST_BUILD["faces_and_geometry"] = time.perf_counter() - t0
    fvtx, c1, c2, fn, fc, fs, ccent_i, cvol_i, nface_int, sizes = out
    ncell_int = cells.shape[0]
    nface = fvtx.shape[0]
    # ghost cells, one per boundary face, in boco order (init_connect_grid.f90:82-88)
    nghost = nface - nface_int
    ncell = ncell_int + nghost
    ccent = np.zeros((ncell, 3))
    cvol = np.zeros(ncell)
    ccent[:ncell_int] = ccent_i
    cvol[:ncell_int] = cvol_i
    c2[nface_int:] = ncell_int + np.arange(nghost)
    pkey = None
    if wrap is not None:
        pkey = _periodic_keys(fvtx[nface_int:], len(_LOCAL_FACES[etype][0]), wrap)
    spans = {}
    start = nface_int
    for spec, n in zip(bocos_spec, sizes):
        spans[spec["family"]] = np.arange(start, start + n, dtype=np.int32)
        start += n
    bocos = []
    for spec in bocos_spec:
        faces = spans[spec["family"]]
        btype = spec["type"]
        ipart = None
        if btype == BC_PERIOD:
            # partner face: the face of the partner family with the same wrapped vertex set
            pf = spans[spec["partner"]]
            kp = pkey[pf - nface_int]
            srt = np.argsort(kp, kind="stable")
            pos = np.minimum(np.searchsorted(kp[srt], pkey[faces - nface_int]), max(len(pf) - 1, 0))
            if len(pf) != len(faces) or not np.array_equal(kp[srt][pos], pkey[faces - nface_int]):
                raise ValueError(f"periodic boco {spec['family']}: faces without a partner in {spec['partner']}")
            ipart = pf[srt[pos]].astype(np.int32)
        param = np.zeros(8)
        p = np.asarray(spec.get("param", []), dtype=np.float64)
        param[:p.size] = p
        bocos.append(Boco(spec["family"], btype, faces, ipart, param))
    _ghost_geometry(ncell_int, nface_int, c1, fn, fc, ccent, cvol, bocos)
    LAST_BUILD["total"] = time.perf_counter() - t0
    return UstMesh(dim, ncell_int, ncell, nface, nface_int, c1.astype(np.int32), c2.astype(np.int32),
                   fn, fc, fs, ccent, cvol, bocos, vtx, fvtx)


def _ghost_geometry(ncell_int, nface_int, c1, fn, fc, ccent, cvol, bocos):
    """ghost cell centres and volumes, one ghost per boundary face (init_ustboco_ghostcell.f90:46-82,
    periodic: init_connect_grid.f90:226-231); ccent/cvol hold the interior cells and are completed
    in place"""
    for b in bocos:
        faces = b.iface
        ghost = ncell_int + (faces - nface_int)
        inner = c1[faces]
        if b.type == BC_PERIOD:
            icper = c1[b.ipart]
            cvol[ghost] = cvol[icper]
            ccent[ghost] = fc[faces] + (ccent[icper] - fc[b.ipart])
        elif b.type == BC_SYM:
            gc = ccent[inner]
            n = fn[faces]
            k = 2.0 * np.einsum("ij,ij->i", fc[faces] - gc, n)
            cvol[ghost] = cvol[inner]
            ccent[ghost] = gc + k[:, None] * n
        elif b.type in _GHOSTCELL:
            cvol[ghost] = cvol[inner]
            ccent[ghost] = 2.0 * fc[faces] - ccent[inner]
        else:
            cvol[ghost] = 0.0
            ccent[ghost] = fc[faces]


def moved(m: UstMesh, vtx: np.ndarray) -> UstMesh:
    """The mesh ``m`` with its vertices moved to ``vtx`` (same numbering and connectivity): face,
    cell and ghost geometry recomputed as build_ustmesh does, each face keeping its orientation
    (the displacements are assumed small enough not to turn a face over)."""
    vtx = np.asarray(vtx, dtype=np.float64)
    fn, fc, fs = _face_geometry(m.dim, vtx, m.fvtx)
    fn = np.where((np.einsum("ij,ij->i", fn, m.fnorm) < 0.0)[:, None], -fn, fn)
    ccent = np.zeros((m.ncell, 3))
    cvol = np.zeros(m.ncell)
    ccent[:m.ncell_int], cvol[:m.ncell_int] = _cell_geometry(m.dim, m.ncell_int, m.ic1, m.ic2, fn, fc, fs)
    _ghost_geometry(m.ncell_int, m.nface_int, m.ic1, fn, fc, ccent, cvol, m.bocos)
    return UstMesh(m.dim, m.ncell_int, m.ncell, m.nface, m.nface_int, m.ic1, m.ic2, fn, fc, fs, ccent, cvol,
                   m.bocos, vtx, m.fvtx)


# ---------------------------------------------------------------------------
# synthetic meshes

def channel_quad(nx=256, ny=128, lx=1.0, ly=0.5, jitter=0.1, seed=1, bocos=None, device=None):
    """2D channel of nx*ny quads, interior vertices jittered by +-jitter*h.

    Families IMIN/IMAX (x=0/lx) and JMIN/JMAX (y=0/ly)."""
    hx, hy = lx / nx, ly / ny
    i, j = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="ij")
    vid = (i * (ny + 1) + j).ravel()
    x = (i * hx).ravel().astype(np.float64)
    y = (j * hy).ravel().astype(np.float64)
    interior = ((i > 0) & (i < nx) & (j > 0) & (j < ny)).ravel()
    x = x + np.where(interior, jitter * hx * uniform_pm1(seed, vid, 0), 0.0)
    y = y + np.where(interior, jitter * hy * uniform_pm1(seed, vid, 1), 0.0)
    vtx = np.stack([x, y, np.zeros_like(x)], axis=1)
    ci, cj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    v0 = ci * (ny + 1) + cj
    cells = np.stack([v0, v0 + (ny + 1), v0 + (ny + 2), v0 + 1], axis=-1).reshape(-1, 4)
    wrap = None
    if bocos is None:
        bocos = [dict(family="JMIN", type=BC_SYM), dict(family="JMAX", type=BC_SYM),
                 dict(family="IMIN", type=BC_SYM), dict(family="IMAX", type=BC_SYM)]
    if any(b["type"] == BC_PERIOD for b in bocos):
        wi = np.where(i == nx, 0, i)
        wj = np.where(j == ny, 0, j)
        wrap = (wi * (ny + 1) + wj).ravel()

    def fam(fv, fc):
        iv = fv[:, 0] // (ny + 1)
        jv = fv[:, 0] % (ny + 1)
        iw = fv[:, 1] // (ny + 1)
        jw = fv[:, 1] % (ny + 1)
        out = np.empty(fv.shape[0], dtype=object)
        out[(iv == 0) & (iw == 0)] = "IMIN"
        out[(iv == nx) & (iw == nx)] = "IMAX"
        out[(jv == 0) & (jw == 0)] = "JMIN"
        out[(jv == ny) & (jw == ny)] = "JMAX"
        return out

    return build_ustmesh(2, vtx, cells, "quad", fam, bocos, wrap, device)


def _axes(periodic):
    """per-axis periodic flags from a bool or a 3-sequence"""
    if isinstance(periodic, (bool, np.bool_)):
        return (bool(periodic),) * 3
    return tuple(bool(p) for p in periodic)


def _box_vertices(n, L, jitter, seed, periodic):
    axes = _axes(periodic)
    if any(axes) and n < 3:
        # partner faces are found through vertex ids taken modulo n: at n = 2 two different
        # faces of one side carry the same set of wrapped ids
        raise ValueError("a periodic box needs n >= 3")
    h = L / n
    i, j, k = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    if all(axes):
        wi, wj, wk = i % n, j % n, k % n
        rid = (wi * n + wj) * n + wk
        dx = jitter * h * uniform_pm1(seed, rid, 0)
        dy = jitter * h * uniform_pm1(seed, rid, 1)
        dz = jitter * h * uniform_pm1(seed, rid, 2)
        wrap = ((wi * (n + 1) + wj) * (n + 1) + wk)
    elif any(axes):
        # periodic along some axes: ids wrap there; walls (the other axes' end planes) stay flat
        wi, wj, wk = (idx % n if ax else idx for idx, ax in zip((i, j, k), axes))
        wrap = (wi * (n + 1) + wj) * (n + 1) + wk
        inner = np.ones(i.shape, dtype=bool)
        for idx, ax in zip((i, j, k), axes):
            if not ax:
                inner &= (idx > 0) & (idx < n)
        dx = np.where(inner, jitter * h * uniform_pm1(seed, wrap, 0), 0.0)
        dy = np.where(inner, jitter * h * uniform_pm1(seed, wrap, 1), 0.0)
        dz = np.where(inner, jitter * h * uniform_pm1(seed, wrap, 2), 0.0)
    else:
        vid = (i * (n + 1) + j) * (n + 1) + k
        inner = (i > 0) & (i < n) & (j > 0) & (j < n) & (k > 0) & (k < n)
        dx = np.where(inner, jitter * h * uniform_pm1(seed, vid, 0), 0.0)
        dy = np.where(inner, jitter * h * uniform_pm1(seed, vid, 1), 0.0)
        dz = np.where(inner, jitter * h * uniform_pm1(seed, vid, 2), 0.0)
        wrap = None
    vtx = np.stack([i * h + dx, j * h + dy, k * h + dz], axis=1)
    return vtx, wrap


def _box_bocos(periodic):
    """periodic pairs along the periodic axes, slip walls (SYMMETRY) on the others"""
    out = []
    for ax, (lo, hi) in zip(_axes(periodic), (("IMIN", "IMAX"), ("JMIN", "JMAX"), ("KMIN", "KMAX"))):
        if ax:
            out += [dict(family=lo, type=BC_PERIOD, partner=hi), dict(family=hi, type=BC_PERIOD, partner=lo)]
        else:
            out += [dict(family=lo, type=BC_SYM), dict(family=hi, type=BC_SYM)]
    return out


def _box_family(n):
    def fam(fv, fc):
        vi = fv[:, :3] // ((n + 1) * (n + 1))
        vj = (fv[:, :3] // (n + 1)) % (n + 1)
        vk = fv[:, :3] % (n + 1)
        out = np.empty(fv.shape[0], dtype=object)
        for name, arr, val in (("IMIN", vi, 0), ("IMAX", vi, n), ("JMIN", vj, 0),
                               ("JMAX", vj, n), ("KMIN", vk, 0), ("KMAX", vk, n)):
            out[np.all(arr == val, axis=1)] = name
        return out
    return fam


def box_tets(n=100, L=10.0, jitter=0.05, seed=2, periodic=True, bocos=None, device=None):
    """n^3 hexes on [0,L]^3, each split into 6 tets about the main diagonal
    (Kuhn/Freudenthal split, conforming between neighbours)."""
    vtx, wrap = _box_vertices(n, L, jitter, seed, periodic)
    ci, cj, ck = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()

    def v(di, dj, dk):
        return ((ci + di) * (n + 1) + (cj + dj)) * (n + 1) + (ck + dk)
    h = [v(0, 0, 0), v(1, 0, 0), v(1, 1, 0), v(0, 1, 0), v(0, 0, 1), v(1, 0, 1), v(1, 1, 1), v(0, 1, 1)]
    kuhn = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]
    cells = np.stack([np.stack([h[a], h[b], h[c], h[d]], axis=1) for a, b, c, d in kuhn], axis=1).reshape(-1, 4)
    if bocos is None:
        bocos = _box_bocos(periodic)
    return build_ustmesh(3, vtx, cells, "tet", _box_family(n), bocos, wrap, device)


def bump_channel_vertices(nx=63, ny=33, nz=20, lx=3.0, ly=1.0, lz=0.6, jitter=0.05, seed=4, amp=0.05):
    """Vertices of the bump channel: a (nx+1)(ny+1)(nz+1) grid on [0,lx]x[0,ly]x[0,lz] whose
    bottom wall follows y_w = amp sin^2(pi (x - 1)) for 1 <= x <= 2 (grid lines stretched
    between the bump and y = ly), interior vertices jittered by +-jitter h (seeded).
    Vertex id = (i (ny+1) + j) (nz+1) + k."""
    hx, hy, hz = lx / nx, ly / ny, lz / nz
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    vid = (i * (ny + 1) + j) * (nz + 1) + k
    inner = (i > 0) & (i < nx) & (j > 0) & (j < ny) & (k > 0) & (k < nz)
    x = i * hx + np.where(inner, jitter * hx * uniform_pm1(seed, vid, 0), 0.0)
    y = j * hy + np.where(inner, jitter * hy * uniform_pm1(seed, vid, 1), 0.0)
    z = k * hz + np.where(inner, jitter * hz * uniform_pm1(seed, vid, 2), 0.0)
    yw = np.where((x >= 1.0) & (x <= 2.0), amp * np.sin(np.pi * (x - 1.0)) ** 2, 0.0)
    y = yw + y * (ly - yw) / ly
    return np.stack([x, y, z], axis=1)


def bump_channel_tets(nx=63, ny=33, nz=20, lx=3.0, ly=1.0, lz=0.6, jitter=0.05, seed=4, amp=0.05, bocos=None,
                      vtx=None, device=None):
    """Tet channel with a bottom-wall bump (6 Kuhn tets per hex): families IMIN/IMAX (x = 0 / lx),
    JMIN (bump wall) / JMAX, KMIN/KMAX.  ``vtx`` replaces the generated vertex coordinates (same
    numbering), e.g. with displaced wall vertices."""
    if vtx is None:
        vtx = bump_channel_vertices(nx, ny, nz, lx, ly, lz, jitter, seed, amp)
    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()

    def v(di, dj, dk):
        return ((ci + di) * (ny + 1) + (cj + dj)) * (nz + 1) + (ck + dk)
    h = [v(0, 0, 0), v(1, 0, 0), v(1, 1, 0), v(0, 1, 0), v(0, 0, 1), v(1, 0, 1), v(1, 1, 1), v(0, 1, 1)]
    kuhn = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]
    cells = np.stack([np.stack([h[a], h[b], h[c], h[d]], axis=1) for a, b, c, d in kuhn], axis=1).reshape(-1, 4)

    def fam(fv, fc):
        vi = fv[:, :3] // ((ny + 1) * (nz + 1))
        vj = (fv[:, :3] // (nz + 1)) % (ny + 1)
        vk = fv[:, :3] % (nz + 1)
        out = np.empty(fv.shape[0], dtype=object)
        for name, arr, val in (("IMIN", vi, 0), ("IMAX", vi, nx), ("JMIN", vj, 0),
                               ("JMAX", vj, ny), ("KMIN", vk, 0), ("KMAX", vk, nz)):
            out[np.all(arr == val, axis=1)] = name
        return out
    if bocos is None:
        bocos = [dict(family=f, type=BC_SYM) for f in ("IMIN", "IMAX", "JMIN", "JMAX", "KMIN", "KMAX")]
    return build_ustmesh(3, np.asarray(vtx, dtype=np.float64), cells, "tet", fam, bocos, None, device)


def box_hex(n=16, L=2 * np.pi, jitter=0.0, seed=3, periodic=True, bocos=None, device=None):
    """n^3 hexes on [0,L]^3; periodic: bool or per-axis (x, y, z) flags (slip walls elsewhere)."""
    vtx, wrap = _box_vertices(n, L, jitter, seed, periodic)
    ci, cj, ck = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()

    def v(di, dj, dk):
        return ((ci + di) * (n + 1) + (cj + dj)) * (n + 1) + (ck + dk)
    cells = np.stack([v(0, 0, 0), v(1, 0, 0), v(1, 1, 0), v(0, 1, 0),
                      v(0, 0, 1), v(1, 0, 1), v(1, 1, 1), v(0, 1, 1)], axis=1)
    if bocos is None:
        bocos = _box_bocos(periodic)
    return build_ustmesh(3, vtx, cells, "hex", _box_family(n), bocos, wrap, device)


def periodic_quad(nx=20, ny=20, lx=1.0, ly=1.0, jitter=0.0, seed=1, device=None):
    """2D doubly periodic quad mesh (the NRG TIME/RUNGEKUTTA layout)."""
    bocos = [dict(family="IMIN", type=BC_PERIOD, partner="IMAX"),
             dict(family="IMAX", type=BC_PERIOD, partner="IMIN"),
             dict(family="JMIN", type=BC_PERIOD, partner="JMAX"),
             dict(family="JMAX", type=BC_PERIOD, partner="JMIN")]
    hx, hy = lx / nx, ly / ny
    i, j = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="ij")
    wi = np.where(i == nx, 0, i)
    wj = np.where(j == ny, 0, j)
    rid = (wi * ny + wj).ravel()
    x = (i * hx).ravel() + jitter * hx * uniform_pm1(seed, rid, 0)
    y = (j * hy).ravel() + jitter * hy * uniform_pm1(seed, rid, 1)
    vtx = np.stack([x, y, np.zeros_like(x)], axis=1)
    ci, cj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    v0 = ci * (ny + 1) + cj
    cells = np.stack([v0, v0 + (ny + 1), v0 + (ny + 2), v0 + 1], axis=-1).reshape(-1, 4)
    wrap = (wi * (ny + 1) + wj).ravel()

    def fam(fv, fc):
        iv, jv = fv[:, 0] // (ny + 1), fv[:, 0] % (ny + 1)
        iw, jw = fv[:, 1] // (ny + 1), fv[:, 1] % (ny + 1)
        out = np.empty(fv.shape[0], dtype=object)
        out[(iv == 0) & (iw == 0)] = "IMIN"
        out[(iv == nx) & (iw == nx)] = "IMAX"
        out[(jv == 0) & (jw == 0)] = "JMIN"
        out[(jv == ny) & (jw == ny)] = "JMAX"
        return out

    return build_ustmesh(2, vtx, cells, "quad", fam, bocos, wrap, device)
