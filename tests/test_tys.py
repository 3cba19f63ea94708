"""TYPHON solution files (.tys) and the restart initialisation (INIT TYPE=TYPHON/TYS).

The reference ships no .tys file, so the layout is checked byte by byte against the XBIN section
format (XBIN_IO.f90/XBIN_DATA.f90) and the order typhonwrite_sol writes (TYFMT_SOL.f90:30-72); the
mesh part is the reference's own .tym content; the restart is checked by continuing a run from the
file against the uninterrupted run (oracle on the CPU here, the library in the GPU test)."""
import os
import struct

import numpy as np
import pytest

import cases as K
import nrg
from typhon_amd import tyfmt as T

HERE = os.path.dirname(os.path.abspath(__file__))
Q20 = os.path.join(HERE, "golden", "meshes", "quad20x20.tym")
Q10 = os.path.join(HERE, "golden", "meshes", "quad20x10.tym")


def _fields(n, seed=0):
    rng = np.random.default_rng(seed)
    return 1.0 + rng.random(n), 1e5 * (1.0 + rng.random(n)), rng.normal(size=(n, 3))


def test_tys_section_bytes(tmp_path):
    """header, mesh sections, SOLUTION and the three quantity sections, decoded by hand"""
    n = 480  # 400 cells + 80 ghosts of quad20x20
    rho, p, vel = _fields(n)
    f = tmp_path / "sol.tys"
    T.write_tys(f, T.mesh_sections(Q20), rho, p, vel)
    b = f.read_bytes()
    assert b[:4] == b"XBIN" and b[4:6] == bytes([0, 2]) and struct.unpack("<i", b[6:10])[0] == 0x01020304
    pos = 10

    def head():
        nonlocal pos
        dtype, = struct.unpack_from("<h", b, pos)
        name = b[pos + 2:pos + 34].decode().strip()
        group, user, isz, rsz, npar = struct.unpack_from("<5h", b, pos + 34)
        pos += 44
        par = list(struct.unpack_from(f"<{npar}i", b, pos))
        pos += 4 * npar
        slen, = struct.unpack_from("<h", b, pos)
        pos += 2 + slen
        assert (group, isz, rsz) == (0, 4, 8)
        return dtype, name, user, par

    assert head() == (1, "FILE_HEADER", 1, [3, 20, 1, 1, 11])
    assert b[pos:pos + 3] == b"END"
    pos += 3
    # the mesh sections are the .tym file's own bytes after its FILE_HEADER
    tym = open(Q20, "rb").read()
    hdr_len = 10 + 44 + 4 * 5 + 2 + 3
    mesh_bytes = tym[hdr_len:]
    assert b[pos:pos + len(mesh_bytes)] == mesh_bytes
    pos += len(mesh_bytes)
    assert head() == (1, "SOLUTION", 30, [2, n, n, 1, 2, 1])
    assert b[pos:pos + 3] == b"END"
    pos += 3
    for name, qid, arr in (("RHO", 1, rho), ("PS", 2, p)):
        assert head() == (20, name, 31, [qid])
        first, nelem, dim = struct.unpack_from("<3i", b, pos)
        assert (first, nelem, dim) == (1, n, 0)
        pos += 12
        assert np.array_equal(np.frombuffer(b, "<f8", n, pos), arr)
        pos += 8 * n
        assert b[pos:pos + 3] == b"END"
        pos += 3
    assert head() == (20, "V", 32, [101])
    first, nelem, dim = struct.unpack_from("<3i", b, pos)
    assert (first, nelem, dim) == (1, n, 3)
    pos += 12
    # Fortran sol(1:3, 1:nelem): x, y, z of each element in turn
    assert np.array_equal(np.frombuffer(b, "<f8", 3 * n, pos).reshape(n, 3), vel)
    pos += 24 * n
    assert b[pos:] == b"END"


def test_format1_mesh_header_is_converted():
    secs = T.mesh_sections(Q10)
    h = secs[0]
    assert (h.name, h.usertype) == ("USTMESH", T.TY_MESHDEF)
    assert h.params == [10, 0, 60, 231, 1, 1, 0, 4]
    raw = T.read_xbin(Q10)[1]
    assert [s.name for s in secs] == [s.name for s in raw[1:]]


def test_read_tys_round_trip_and_versions(tmp_path):
    n = 480
    rho, p, vel = _fields(n, 1)
    f = tmp_path / "a.tys"
    T.write_tys(f, T.mesh_sections(Q20), rho, p, vel)
    sol = T.read_tys(f)
    assert (sol.version, sol.soltype, sol.ncell, sol.ndof) == (3, T.SOL_CELL, n, 1)
    assert np.array_equal(sol.scalars[T.QS_DENSITY], rho)
    assert np.array_equal(sol.scalars[T.QS_PRESSURE], p)
    assert np.array_equal(sol.vectors[T.QV_VELOCITY], vel)
    assert [s.name for s in sol.mesh_sections] == [s.name for s in T.mesh_sections(Q20)]
    # a format-2 file: SOLUTION with (nelem, nscal, nvect) only
    _, secs = T.read_xbin(f)
    secs[0].params = [2, 20, 1, 1, 11]
    k = [i for i, s in enumerate(secs) if s.name == "SOLUTION"][0]
    secs[k].params = [n, 2, 1]
    T.write_xbin(tmp_path / "b.tys", secs)
    sol2 = T.read_tys(tmp_path / "b.tys")
    assert (sol2.version, sol2.soltype, sol2.ncell, sol2.ndof) == (2, T.SOL_CELL, n, 1)
    assert np.array_equal(sol2.vectors[T.QV_VELOCITY], vel)
    # errors the reference's reader stops on
    secs[k].params = [n, 2]
    T.write_xbin(tmp_path / "c.tys", secs)
    with pytest.raises(ValueError, match="3 parameters"):
        T.read_tys(tmp_path / "c.tys")
    secs[k].params = [n, 2, 1]
    secs[0].params = [2, 10, 1, 0, 11]
    T.write_xbin(tmp_path / "d.tys", secs)
    with pytest.raises(ValueError, match="missing solution"):
        T.read_tys(tmp_path / "d.tys")


def test_restart_state_is_calc_varcons():
    """rhoE = p/(gamma-1) + rho|u|^2/2 and rho u from the file's interior elements"""
    m, s, g, u = nrg.setup_mesh_case("PER2D-TRANS")
    rho, p, vel = T.solution_fields(u)
    assert np.allclose(p, 100000 / 14., rtol=1e-14)
    cons = np.vstack([rho, p / 0.4 + 0.5 * rho * (vel * vel).sum(1), (rho[:, None] * vel).T])
    assert np.allclose(cons, u, rtol=1e-14, atol=0)


def test_restart_continues_the_run(tmp_path):
    """PER2D-TRANS (quad20x20.tym, periodic, Rusanov, RK3-SSP): 20 steps, write the state as a .tys,
    restart from it and run 20 more = 40 steps straight (the prim/cons round trip of the file is the
    only difference, as in the reference, which stores primitive variables)"""
    m, s, g, u = nrg.setup_mesh_case("PER2D-TRANS")
    o = K.make_oracle(m, s, g)
    o.set_cons(u)
    dts = []
    for _ in range(40):
        dt = o.timestep(s["cfl"])
        dts.append(dt)
        o.rk_step(dt)
    straight = o.get_cons()
    o.set_cons(u)
    for dt in dts[:20]:
        o.rk_step(dt)
    f = tmp_path / "restart.tys"
    T.write_state_tys(f, Q20, m, o.get_cons())
    u20 = T.restart_state(f, m.ncell_int)
    assert np.abs(u20 - o.get_cons()).max() <= 1e-15 * np.abs(o.get_cons()).max() * 10
    o.set_cons(u20)
    for dt in dts[20:]:
        o.rk_step(dt)
    assert K.rel_err(o.get_cons(), straight) <= 1e-13


@pytest.mark.gpu
def test_gpu_restart_continues_the_run(tmp_path):
    """the same restart through the library: 20 device steps, .tys, restart, 20 more = 40 straight
    device steps, and = the oracle's uninterrupted run"""
    from typhon_amd.solver import Solver
    m, s, g, u = nrg.setup_mesh_case("PER2D-TRANS")
    gpu = Solver(m, s, g)
    gpu.set_state(u)
    dts = []
    for _ in range(40):
        dt = gpu.timestep()
        dts.append(dt)
        gpu.rk_step(dt)
    straight = gpu.get_state()
    gpu.set_state(u)
    for dt in dts[:20]:
        gpu.rk_step(dt)
    f = tmp_path / "restart.tys"
    T.write_state_tys(f, Q20, m, gpu.get_state())
    gpu.set_state(T.restart_state(f, m.ncell_int))
    for dt in dts[20:]:
        gpu.rk_step(dt)
    assert K.rel_err(gpu.get_state(), straight) <= 1e-13
    o = K.make_oracle(m, s, g)
    o.set_cons(u)
    for dt in dts:
        o.rk_step(dt)
    assert K.rel_err(straight, o.get_cons()) <= 1e-12
    gpu.close()
