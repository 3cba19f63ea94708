"""Volume probes: the FCT expression language (FCT_PARSER.f90, FCT_EVAL.f90, FCT_MATH.f90) and
VOL_AVERAGE / VOL_MIN / VOL_MAX (PROBECALC.f90, prb_zone_vol.f90).

CPU: the oracle's parser and evaluator (oracle/fct.py) against what fixes them independently --
  * hand-worked expressions whose value follows from the grammar's rules (splitting at the LAST
    operator of the lowest priority found, so - and / associate to the left and so does ^; a
    leading - is the opposite; exponent signs of numbers are not operators; sign is inv);
  * random fully bracketed trees against a direct evaluation of the tree;
  * the NRG cases' QUANTITY strings against the probes written out by hand in numpy, which
    replay the reference's golden .tmon histories (test_nrg_golden.py);
  * the errors the reference raises (unmatched brackets, operators at the ends, comparisons).
GPU: ty_probe_define / ty_probe_eval against the oracle on seeded states of 2D and 3D meshes for
every function and operator and the three probe types; repeated calls give the same bits;
definition errors come back as TY_ERR_ARG.
"""
import math

import numpy as np
import pytest

import cases as K
import nrg
from oracle import fct
from typhon_amd import mesh as M

HAND = [
    ("1-2-3", -4.0), ("8/4/2", 1.0), ("2^3^2", 64.0), ("-2^2", -4.0), ("2*3+4*5", 26.0),
    ("1e-3*2", 0.002), ("2.5E+2-1", 249.0), ("1d2/4", 25.0), ("((1+2))*3", 9.0), ("+3", 3.0), ("-(2)", -2.0),
    ("sqr(3)", 9.0), ("sqrt(16)", 4.0), ("inv(4)", 0.25), ("sign(4)", 0.25), ("sign(-2)", -0.5),
    ("step(-1)", 0.0), ("step(0)", 1.0), ("ramp(-2)", 0.0), ("ramp(3)", 3.0), ("log(1000)", 3.0),
    ("ln(_e)", 1.0), ("sin(_pi/2)", 1.0), ("abs(-1.5)", 1.5), ("exp(0)", 1.0), ("10-2*3^2", -8.0),
    ("2-3+4", 3.0), ("6/2*3", 9.0), ("1-(2-3)", 2.0), (".5*4", 2.0), ("2.*3", 6.0), (" 1 + 2 ", 3.0),
    ("SQR(2)", 4.0), ("erf(0)+erfc(0)", 1.0), ("atan(1)*4", math.pi), ("cosh(0)-sinh(0)+tanh(0)", 1.0),
]


@pytest.mark.parametrize("expr,value", HAND)
def test_hand_worked_expressions(expr, value):
    assert abs(fct.evaluate(fct.parse(expr), {}) - value) <= 1e-15 * max(1.0, abs(value))


def test_variables_and_precedence_with_names():
    env = {"x": 2.0, "v_x": 3.0, "rho": 0.5}
    assert fct.evaluate(fct.parse("x-v_x*rho"), env) == 0.5
    assert fct.evaluate(fct.parse("-x^2"), env) == -4.0
    # a name that merely contains the letters of a function is a variable
    assert fct.evaluate(fct.parse("rho*sqr(x)"), env) == 2.0
    with pytest.raises(fct.FctError):
        fct.evaluate(fct.parse("foo+1"), env)


@pytest.mark.parametrize("bad", ["(1+2", "1+2)", "1+", "*2", "1<2", "1==1", "2&&3"])
def test_errors(bad):
    with pytest.raises(fct.FctError):
        fct.evaluate(fct.parse(bad), {})


def _random_tree(rng, depth):
    if depth == 0 or rng.random() < 0.25:
        if rng.random() < 0.5:
            return ("cst", float(rng.integers(1, 9)) + float(rng.integers(0, 4)) / 4), None
        name = rng.choice(["x", "y", "rho"])
        return ("var", name), None
    if rng.random() < 0.3:
        f = rng.choice(["sqr", "abs", "exp", "sin", "cos", "atan", "inv", "ramp", "step", "erf"])
        return ("fn", f), _random_tree(rng, depth - 1)
    op = rng.choice(["+", "-", "*", "/"])
    return ("op", op), (_random_tree(rng, depth - 1), _random_tree(rng, depth - 1))


def _render(t):
    (kind, val), sub = t
    if kind == "cst":
        return repr(val)
    if kind == "var":
        return val
    if kind == "fn":
        return f"{val}({_render(sub)})"
    return f"({_render(sub[0])}){val}({_render(sub[1])})"


def _direct(t, env):
    (kind, val), sub = t
    if kind == "cst":
        return val
    if kind == "var":
        return env[val]
    if kind == "fn":
        a = _direct(sub, env)
        return {"sqr": lambda x: x * x, "abs": abs, "exp": math.exp, "sin": math.sin, "cos": math.cos,
                "atan": math.atan, "inv": lambda x: 1.0 / x, "ramp": lambda x: max(x, 0.0),
                "step": lambda x: 1.0 if x >= 0 else 0.0, "erf": math.erf}[val](a)
    a, b = _direct(sub[0], env), _direct(sub[1], env)
    return {"+": a + b, "-": a - b, "*": a * b, "/": a / b}[val]


def test_random_bracketed_trees():
    rng = np.random.default_rng(7)
    env = {"x": 0.3, "y": -1.7, "rho": 2.25}
    for _ in range(300):
        t = _random_tree(rng, 4)
        try:
            want = _direct(t, env)
        except (ZeroDivisionError, OverflowError):
            continue
        got = fct.evaluate(fct.parse(_render(t)), env)
        assert got == want or abs(got - want) <= 1e-14 * abs(want), _render(t)


def test_nrg_quantities_match_the_hand_written_probes():
    m, s, g, u = nrg.setup("RK4")
    rng = np.random.default_rng(1)
    u = u * (1 + 0.01 * rng.standard_normal(u.shape))
    for t in (0.0, 0.0137):
        got = fct.vol_probe(nrg.RK_PROBE, fct.VOL_AVERAGE, m.ccent, m.cvol, u, t, 1.4)
        assert abs(got - nrg.rhoerr(m, u[0], t)) <= 1e-13 * nrg.rhoerr(m, u[0], t)
    m, s, g, u = nrg.setup_visc("CELLAVG")
    for t in (0.0, 3.0):
        want = nrg.visc_probes(m, u, t)
        got = [fct.vol_probe(q, fct.VOL_AVERAGE, m.ccent, m.cvol, u, t, 1.4) for q in nrg.VISC_PROBES]
        np.testing.assert_allclose(got, want, rtol=1e-13)
    for case in nrg.MESH_CASES:
        m, s, g, u = nrg.setup_mesh_case(case)
        got = fct.vol_probe(nrg.MESH_PROBES[case], fct.VOL_AVERAGE, m.ccent, m.cvol, u, 0.0, g["gamma"])
        assert abs(got - nrg.mesh_case_probe(case, m, u)[0]) <= 1e-13 * abs(got)


def test_min_max_and_average_of_a_field():
    m, s, g, u = K.channel_case(6, 3, seed=2)
    n = m.ncell_int
    rho = u[0]
    assert fct.vol_probe("RHO", fct.VOL_MIN, m.ccent, m.cvol, u, 0, 1.4) == rho.min()
    assert fct.vol_probe("RHO", fct.VOL_MAX, m.ccent, m.cvol, u, 0, 1.4) == rho.max()
    v = m.cvol[:n]
    assert abs(fct.vol_probe("RHO", fct.VOL_AVERAGE, m.ccent, m.cvol, u, 0, 1.4) - (rho * v).sum() / v.sum()) < 1e-15


# ---------------------------------------------------------------------------
# device probes

EXPRS = [
    "RHO", "PS", "V", "V_X", "V_Y", "V_Z", "x", "y", "z", "t", "2.5", "-RHO", "+PS",
    "sqr(RHO-(1+sqr(sin(2*_pi*(x-50*t))*sin(2*_pi*y))/2))", "(RHO-1.25)^2", "sqr(V_X-(erf(y/2/sqrt(t+1.))))",
    "2*50*(1-sqr(V_X))", "inv(RHO)+sign(PS)", "exp(-x)*ln(PS)*log(RHO)", "tan(y)+cos(x)-sin(z)",
    "sinh(V_X)*cosh(V_Y)/tanh(1+RHO)", "asin(V_X/V)+acos(0.3)+atan(V_Y)", "asinh(x)+acosh(1+RHO)+atanh(0.5*V_X/V)",
    "erfc(RHO)+abs(V_Y)+step(V_X-0.5)+ramp(y-0.5)", "1e-3*PS^1.5-2.5E+2/RHO", "RHO*V^2/2+PS/0.4",
]
TYPES = {"VOL_AVERAGE": fct.VOL_AVERAGE, "VOL_MIN": fct.VOL_MIN, "VOL_MAX": fct.VOL_MAX}


def _probe_cases():
    m2, s2, g2, u2 = K.channel_case(24, 12, seed=5)
    u2 = u2.copy()
    u2[3] = 0.02 * np.random.default_rng(3).standard_normal(m2.ncell_int) * u2[0]
    m3, s3, g3, u3 = K.bump_case(6, 3, 3, noise=0.05)
    return {"quad": (m2, s2, g2, u2), "tets": (m3, s3, g3, u3)}


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["quad", "tets"])
def test_gpu_probes_match_oracle(case):
    from typhon_amd.solver import Solver
    m, s, g, u = _probe_cases()[case]
    gpu = Solver(m, s, g)
    gpu.set_state(u)
    ids, want, scale = [], [], []
    t = 0.37
    env = fct.probe_env(m.ccent, u, t, g["gamma"])
    v = m.cvol[:m.ncell_int]
    for e in EXPRS:
        f = np.broadcast_to(fct.evaluate(fct.parse(e), env), v.shape)
        assert np.isfinite(f).all(), e
        for name, code in TYPES.items():
            ids.append(gpu.define_probe(e, name))
            want.append(fct.vol_probe(e, code, m.ccent, m.cvol, u, t, g["gamma"]))
            scale.append((np.abs(f) * v).sum() / v.sum() if code == fct.VOL_AVERAGE else abs(want[-1]))
    got = gpu.probe(ids, t)
    want = np.array(want)
    # FP64 on both sides; the cell values differ by the last bits of the transcendental functions
    # and of calc_varprim's operation order, the averages also by the order of summation (relative
    # to the average of |f|, as the sums of signed fields cancel)
    err = np.abs(got - want) / np.maximum(np.array(scale), 1e-300)
    assert err.max() <= 1e-13, (EXPRS[int(err.argmax()) // 3], got[err.argmax()], want[err.argmax()])
    # fixed-order reductions: the same bits on every call
    assert np.array_equal(gpu.probe(ids, t), got)


@pytest.mark.gpu
def test_gpu_probe_follows_the_state():
    from typhon_amd.solver import Solver
    m, s, g, u = K.channel_case(16, 8, seed=1)
    gpu = Solver(m, s, g)
    gpu.set_state(u)
    pid = gpu.define_probe("RHO", "VOL_MAX")
    assert gpu.probe(pid) == u[0].max()
    gpu.rk_step(gpu.timestep())
    U = gpu.get_state()
    assert gpu.probe(pid) == U[0].max()
    # a re-scheme keeps the probes
    gpu.set_scheme(dict(s, cfl=0.5), g)
    gpu.set_state(u)
    assert gpu.probe([pid, pid])[1] == u[0].max()


@pytest.mark.gpu
@pytest.mark.parametrize("bad", ["1<2", "RHO&&PS", "TS", "(RHO", "RHO+", "*RHO"])
def test_gpu_probe_definition_errors(bad):
    from typhon_amd import _lib
    from typhon_amd.solver import Solver
    m, s, g, u = K.channel_case(4, 2)
    gpu = Solver(m, s, g)
    with pytest.raises(_lib.TyError):
        gpu.define_probe(bad)
    with pytest.raises(_lib.TyError):
        gpu.define_probe("RHO", 7)
    gpu.set_state(u)
    with pytest.raises(_lib.TyError):
        gpu.probe(3)
