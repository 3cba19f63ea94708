"""ty_probe_eval_noalloc (the probe evaluation with its partials in a fixed per-GPU workspace)
against ty_probe_eval: bit for bit, on two calls in a row, the second on a smaller mesh (another
context sharing the GPU's workspace, so the second call starts from the first call's leftovers),
and the limit check on the probe count."""
import numpy as np
import pytest

import cases as K

EXPRS = ["RHO", "PS", "sqr(V_X-0.5)+V_Y", "(RHO-1.25)^2", "erf(x)*PS"]
TYPES = ["VOL_AVERAGE", "VOL_MIN", "VOL_MAX"]


def _solver(case):
    from typhon_amd.solver import Solver
    m, s, g, u = case
    gpu = Solver(m, s, g)
    gpu.set_state(u)
    ids = [gpu.define_probe(e, t) for e in EXPRS for t in TYPES]
    return gpu, ids


@pytest.mark.gpu
def test_noalloc_probes_match_the_allocating_entry_point():
    from typhon_amd import _lib
    big, ids_b = _solver(K.channel_case(96, 48, seed=4))
    small, ids_s = _solver(K.bump_case(4, 2, 2, noise=0.05))
    assert big.ni > 10 * small.ni
    for gpu, ids, t in ((big, ids_b, 0.25), (small, ids_s, 0.5)):
        got = gpu.probe(ids, t, noalloc=True)
        want = gpu.probe(ids, t)
        assert np.array_equal(got, want), (got, want)
    # the first context again, after the smaller one has used the workspace
    assert np.array_equal(big.probe(ids_b, 0.25, noalloc=True), big.probe(ids_b, 0.25))
    # more probes than the workspace holds: refused before any launch
    with pytest.raises(_lib.TyError, match="exceed"):
        big.probe([ids_b[0]] * 65, 0.0, noalloc=True)
    assert big.probe([ids_b[0]] * 64, 0.0, noalloc=True).shape == (64,)
    big.close()
    small.close()
