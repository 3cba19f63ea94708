"""The reference's golden probe histories (NRG/TIME/RUNGEKUTTA/*/RHOERR.tmon, NRG/NS/NUM/FLUX/VISCOUS,
NRG/MESH/ELEM/TYM-QUAD, NRG/MESH/CON/PER2D-TRANS) replayed by the oracle
(CPU) and by the HIP path (GPU).  The files hold 8 significant digits; the check is 2e-7 relative
per cycle over every cycle.  On the GPU the probes are the cases' own QUANTITY strings evaluated
by the device probes (ty_probe_eval)."""
import numpy as np
import pytest

import cases as K
import nrg

RTOL = 2e-7


@pytest.mark.parametrize("case", list(nrg.CASES))
def test_oracle_matches_nrg_golden(case):
    m, s, g, u = nrg.setup(case)
    o = K.make_oracle(m, s, g)
    o.set_cons(u)
    res = nrg.replay(o.rk_step, lambda rem: o.timestep(s["cfl"], rem), lambda: o.get_cons()[0], m)
    gold = nrg.golden(case)
    np.testing.assert_allclose(res[:, 1], gold[:, 1], rtol=1e-7)
    np.testing.assert_allclose(res[:, 2], gold[:, 2], rtol=RTOL)


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(nrg.CASES))
def test_gpu_matches_nrg_golden(case):
    from typhon_amd.solver import Solver
    m, s, g, u = nrg.setup(case)
    gpu = Solver(m, s, g)
    gpu.set_state(u)
    pid = gpu.define_probe(nrg.RK_PROBE, "VOL_AVERAGE")
    res = nrg.replay_probes(gpu.rk_step, lambda rem: gpu.timestep(rem), lambda t: (gpu.probe(pid, t),), 100, 0.1)
    gold = nrg.golden(case)
    np.testing.assert_allclose(res[:, 2], gold[:, 2], rtol=RTOL)


@pytest.mark.parametrize("case", list(nrg.VISC_CASES))
def test_oracle_matches_nrg_visc_golden(case):
    """NRG/NS/NUM/FLUX/VISCOUS/{CELLFULL,CELLAVG}: Stokes layer between moving adiabatic walls,
    VXERR and DELTA histories over 10 cycles (8 significant digits)."""
    m, s, g, u = nrg.setup_visc(case)
    o = K.make_oracle(m, s, g)
    o.set_cons(u)
    res = nrg.replay_probes(o.rk_step, lambda rem: o.timestep(s["cfl"], rem),
                            lambda t: nrg.visc_probes(m, o.get_cons(), t), 10, 10.0)
    np.testing.assert_allclose(res[:, 2], nrg.golden_visc(case, "VXERR")[:, 2], rtol=RTOL)
    np.testing.assert_allclose(res[:, 3], nrg.golden_visc(case, "DELTA")[:, 2], rtol=RTOL)


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(nrg.VISC_CASES))
def test_gpu_matches_nrg_visc_golden(case):
    from typhon_amd.solver import Solver
    m, s, g, u = nrg.setup_visc(case)
    gpu = Solver(m, s, g)
    gpu.set_state(u)
    ids = [gpu.define_probe(q) for q in nrg.VISC_PROBES]
    res = nrg.replay_probes(gpu.rk_step, lambda rem: gpu.timestep(rem), lambda t: gpu.probe(ids, t), 10, 10.0)
    np.testing.assert_allclose(res[:, 2], nrg.golden_visc(case, "VXERR")[:, 2], rtol=RTOL)
    np.testing.assert_allclose(res[:, 3], nrg.golden_visc(case, "DELTA")[:, 2], rtol=RTOL)


@pytest.mark.parametrize("case", list(nrg.MESH_CASES))
def test_oracle_matches_nrg_mesh_golden(case):
    """NRG/MESH/ELEM/TYM-QUAD (first-order HLLC, slip walls, PAVG) and NRG/MESH/CON/PER2D-TRANS
    (Rusanov, periodic, RHODEV) on the reference's own .tym meshes."""
    probe, ncyc, dur = nrg.MESH_CASES[case]
    m, s, g, u = nrg.setup_mesh_case(case)
    o = K.make_oracle(m, s, g)
    o.set_cons(u)
    res = nrg.replay_probes(o.rk_step, lambda rem: o.timestep(s["cfl"], rem),
                            lambda t: nrg.mesh_case_probe(case, m, o.get_cons()), ncyc, dur)
    np.testing.assert_allclose(res[:, 2], nrg.golden_mesh(case)[:, 2], rtol=RTOL)


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(nrg.MESH_CASES))
def test_gpu_matches_nrg_mesh_golden(case):
    from typhon_amd.solver import Solver
    probe, ncyc, dur = nrg.MESH_CASES[case]
    m, s, g, u = nrg.setup_mesh_case(case)
    gpu = Solver(m, s, g)
    gpu.set_state(u)
    pid = gpu.define_probe(nrg.MESH_PROBES[case], "VOL_AVERAGE")
    res = nrg.replay_probes(gpu.rk_step, lambda rem: gpu.timestep(rem), lambda t: (gpu.probe(pid, t),), ncyc, dur)
    np.testing.assert_allclose(res[:, 2], nrg.golden_mesh(case)[:, 2], rtol=RTOL)
