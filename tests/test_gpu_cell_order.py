"""The device cell orders of ty_mesh_upload.

Reverse Cuthill-McKee runs on the device level by level (ty_order.hip); it must give exactly the
order of the serial algorithm, restated here in plain Python: a pseudo-peripheral start (a
breadth-first search from the lowest-numbered unordered cell, then from the lowest-degree cell of
its last level -- the first in queue order, and only if its degree is below the start's), then a
breadth-first queue in which each cell appends its not yet queued neighbours stably sorted by
degree, one pass per connected component, the queue reversed.  The graph is the interior faces in
face order, then the merged periodic pairs (first face of each pair in face order), adjacency lists
in that edge order.  Checked on 2D/3D, periodic, walled, disconnected and degenerate meshes, and at
the 256^3-hex size against the serial algorithm's bandwidth and validity."""
import numpy as np
import pytest

from typhon_amd import mesh as M
from typhon_amd import scheme as S
from typhon_amd.solver import Solver

def graph_edges(m):
    """(a, b) edges of the interior-cell graph, in the order ty_mesh_upload lists them"""
    e1 = np.stack([m.ic1[:m.nface_int], m.ic2[:m.nface_int]], axis=1)
    fpart = np.full(m.nface, -1, dtype=np.int64)
    for b in m.bocos:
        if b.ipart is not None:
            fpart[b.iface] = b.ipart
    f = np.arange(m.nface_int, m.nface)
    p = fpart[f]
    sel = (p >= 0) & (f < p)
    e2 = np.stack([m.ic1[f[sel]], m.ic1[p[sel]]], axis=1)
    return np.concatenate([e1, e2]).astype(np.int64)


def rcm_serial(n, edges):
    adj = [[] for _ in range(n)]
    for a, b in np.asarray(edges).tolist():
        adj[a].append(b)
        adj[b].append(a)
    deg = [len(x) for x in adj]
    done = [False] * n

    def far_end(s):
        lvl = {s: 0}
        q = [s]
        h = 0
        while h < len(q):
            u = q[h]
            h += 1
            for v in adj[u]:
                if done[v] or v in lvl:
                    continue
                lvl[v] = lvl[u] + 1
                q.append(v)
        last = lvl[q[-1]]
        best = s
        for v in q:
            if lvl[v] == last and deg[v] < deg[best]:
                best = v
        return best

    order = []
    scan = 0
    while len(order) < n:
        while done[scan]:
            scan += 1
        s = far_end(far_end(scan))
        head = len(order)
        order.append(s)
        done[s] = True
        while head < len(order):
            u = order[head]
            head += 1
            nb = []
            for v in adj[u]:
                if not done[v]:
                    done[v] = True
                    nb.append(v)
            nb.sort(key=lambda v: deg[v])  # stable: adjacency order among equal degrees
            order.extend(nb)
    return order[::-1]


def two_boxes():
    """two disjoint hex boxes (two graph components), slip walls everywhere"""
    a = M.box_hex(4, 1.0, periodic=False)
    nv = a.vtx.shape[0]
    n = 4
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()

    def v(di, dj, dk):
        return ((i + di) * (n + 1) + (j + dj)) * (n + 1) + (k + dk)
    cells = np.stack([v(0, 0, 0), v(1, 0, 0), v(1, 1, 0), v(0, 1, 0), v(0, 0, 1), v(1, 0, 1), v(1, 1, 1), v(0, 1, 1)], 1)
    vtx = np.concatenate([a.vtx, a.vtx + [3.0, 0.0, 0.0]])
    cells = np.concatenate([cells + nv, cells])  # the second box's cells first

    def fam(fv, fc):
        return np.array(["WALL"] * fv.shape[0], dtype=object)
    return M.build_ustmesh(3, vtx, cells, "hex", fam, [dict(family="WALL", type=M.BC_SYM)])


MESHES = {
    "channel quads": lambda: M.channel_quad(24, 12),
    "periodic quads": lambda: M.periodic_quad(10, 6, jitter=0.1),
    "periodic quads 3 x 3": lambda: M.periodic_quad(3, 3),
    "periodic tets": lambda: M.box_tets(6),
    "walled tets": lambda: M.box_tets(5, periodic=False),
    "bump tets": lambda: M.bump_channel_tets(12, 6, 4),
    "periodic hexes": lambda: M.box_hex(8, jitter=0.1),
    "channel hexes": lambda: M.box_hex(7, periodic=(True, False, True)),
    "two boxes": two_boxes,
    "quad20x20.tym": lambda: _tym("quad20x20.tym"),
}


def _tym(name):
    import os
    from typhon_amd import tyfmt
    here = os.path.dirname(os.path.abspath(__file__))
    bocos = [dict(family="IMIN", type=M.BC_PERIOD, partner="IMAX", period=(1.0, 0.0, 0.0)),
             dict(family="IMAX", type=M.BC_PERIOD, partner="IMIN", period=(-1.0, 0.0, 0.0)),
             dict(family="JMIN", type=M.BC_PERIOD, partner="JMAX", period=(0.0, 1.0, 0.0)),
             dict(family="JMAX", type=M.BC_PERIOD, partner="JMIN", period=(0.0, -1.0, 0.0))]
    return tyfmt.tym_ustmesh(os.path.join(here, "golden", "meshes", name), bocos)


def _solver(m, order):
    return Solver(m, S.default_scheme(), S.default_gas(), cell_order=order)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(MESHES))
def test_device_rcm_is_the_serial_order(name):
    print(f"\ncell order: {name}", flush=True)
    m = MESHES[name]()
    sv = _solver(m, "rcm")
    try:
        got = sv.cell_order()
    finally:
        sv.close()
    want = rcm_serial(m.ncell_int, graph_edges(m))
    assert np.array_equal(got, np.array(want, dtype=np.int32))


def _bandwidth(m, order):
    pos = np.empty(m.ncell_int, dtype=np.int64)
    pos[order] = np.arange(m.ncell_int)
    e = graph_edges(m)
    return int(np.abs(pos[e[:, 0]] - pos[e[:, 1]]).max())


@pytest.mark.gpu
def test_device_rcm_at_full_size():
    """the 256^3-hex NS box (periodic in x and z, slip walls in y): a permutation with the bandwidth
    the serial order has on the same box at 8^3, 16^3 and 24^3 -- n^2, one plane of cells"""
    n = 256
    m = M.box_hex(n, 1.0, periodic=(True, False, True), device=0)
    sv = _solver(m, "rcm")
    try:
        order = sv.cell_order()
    finally:
        sv.close()
    assert np.array_equal(np.sort(order), np.arange(m.ncell_int))
    bw = _bandwidth(m, order)
    print(f"\n256^3 RCM bandwidth {bw}", flush=True)
    assert bw == n * n


def test_small_boxes_rcm_bandwidth():
    """the serial restatement itself on the boxes the full-size test extrapolates from"""
    for n in (8, 16):
        m = M.box_hex(n, 1.0, periodic=(True, False, True))
        assert _bandwidth(m, np.array(rcm_serial(m.ncell_int, graph_edges(m)))) == n * n
