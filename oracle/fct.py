"""CPU restatement of the reference's FCT expressions and volume probes.

TEST INFRASTRUCTURE ONLY: tests/ use it as the checker of the device probes; the product
(typhon_amd/) never imports it.

  parse      -- string_to_node, CFDTOOLS/Math/FCT_PARSER.f90:113-340, with index_oper
                (FCT_PARSER.f90:350-405) and is_real (CFDTOOLS/General/STRING.f90:265-316)
  evaluate   -- fct_node_eval, CFDTOOLS/Math/FCT_EVAL.f90:170-330, with the functions of
                CFDTOOLS/Math/FCT_MATH.f90 (sign evaluates as inv, FCT_EVAL.f90:297-298)
  vol_probe  -- prb_vol_calc / prb_vol_add, CFDTOOLS/Mesh/PROBECALC.f90:41-252, on the primitive
                fields of calc_varprim (SOURCE/EQNS/calc_varprim_ns.f90:47-52)

A tree is a tuple: ("cst", value) | ("var", name) | ("fn", name, arg) | ("op", symbol, left, right).
"""
from __future__ import annotations

from math import erf as _erf
from math import erfc as _erfc

import numpy as np

BINARY = ["&&", "||", "!=", "==", "<", ">", "<=", ">=", "+", "-", "*", "/", "^"]  # lowest priority first
FUNCTIONS = ["inv", "sqr", "sqrt", "exp", "ln", "log", "sin", "cos", "tan", "sinh", "cosh", "tanh", "asin",
             "acos", "atan", "asinh", "acosh", "atanh", "erf", "erfc", "abs", "sign", "step", "ramp"]


class FctError(ValueError):
    pass


def _index_oper(s: str, op: str) -> int:
    """first 1-based index of op in s outside brackets, skipping the sign of a number's exponent; 0: none"""
    ls = len(s)
    iop = s.find(op) + 1
    ibk1 = s.find("(") + 1 or ls + 1
    if iop == 0:
        return 0
    if iop == ibk1:
        return iop
    if iop < ibk1:
        if 3 <= iop < ls and s[iop - 1] in "-+" and s[iop - 2] in "eE" and s[iop - 3] in ".0123456789" \
                and s[iop] in "0123456789":
            r = _index_oper(s[iop:], op)
            return 0 if r == 0 else iop + r
        return iop
    rb = _index_oper(s[ibk1:], ")")
    if rb == 0:
        raise FctError(f"did not find ending bracket in {s!r}")
    ibk2 = ibk1 + rb
    r = _index_oper(s[ibk2:], op)
    return 0 if r == 0 else ibk2 + r


def _figures(s: str) -> bool:
    return all(ch.isdigit() for ch in s)


def _is_real(w: str) -> bool:
    n = len(w)
    if n == 0:
        return False
    iexp = next((i + 1 for i, ch in enumerate(w) if ch in "dDeE"), n + 1)
    idot = w.find(".") + 1 or iexp
    afterdot = False
    if idot + 1 <= iexp - 1:
        if not _figures(w[idot:iexp - 1]):
            return False
        afterdot = True
    if iexp == n:
        return False
    if iexp + 1 <= n:
        i = iexp + 1
        if i != n and w[i - 1] in "+-":
            i += 1
        if not _figures(w[i - 1:]):
            return False
    i = 2 if w[0] in "+-" else 1
    if idot - 1 >= i:
        return _figures(w[i - 1:idot - 1])
    return afterdot


def _parse(s: str):
    ls = len(s)
    if ls == 0:
        raise FctError("empty expression")
    depth = 0
    for ch in s:
        depth += (ch == "(") - (ch == ")")
        if depth < 0:
            break
    if depth != 0:
        raise FctError(f"unmatched bracket in {s!r}")
    if s[0] == "(" and _index_oper(s[1:], ")") == ls - 1:
        return _parse(s[1:-1])
    if _is_real(s):
        return ("cst", float(s.replace("d", "e").replace("D", "e")))
    for op in BINARY:
        last, pos = 0, 1
        while pos:
            pos = _index_oper(s[last:], op)
            last += pos
        if last == 0:
            continue
        if last == ls:
            raise FctError(f"binary operator at the end of {s!r}")
        if last == 1:
            if op == "+":
                return _parse(s[1:])
            if op == "-":
                return ("fn", "opp", _parse(s[1:]))
            raise FctError(f"binary operator at the start of {s!r}")
        return ("op", op, _parse(s[:last - 1]), _parse(s[last - 1 + len(op):]))
    for f in FUNCTIONS:
        if s.startswith(f + "("):
            return ("fn", f, _parse(s[len(f):]))
    if s.startswith("_pi"):
        return ("cst", np.pi)
    if s.startswith("_e"):
        return ("cst", np.e)
    if not (s[0].isalpha() or s[0] == "_"):
        raise FctError(f"bad variable name {s!r}")
    return ("var", s)


def parse(expr: str):
    """string_to_node on the space-stripped, lower-cased expression"""
    return _parse("".join(expr.split()).lower())


_FN = {
    "opp": np.negative, "inv": lambda x: 1.0 / x, "sqr": lambda x: x ** 2, "sqrt": np.sqrt, "exp": np.exp,
    "ln": np.log, "log": np.log10, "sin": np.sin, "cos": np.cos, "tan": np.tan, "sinh": np.sinh, "cosh": np.cosh,
    "tanh": np.tanh, "asin": np.arcsin, "acos": np.arccos, "atan": np.arctan, "asinh": np.arcsinh,
    "acosh": np.arccosh, "atanh": np.arctanh, "erf": np.vectorize(_erf, otypes=[float]),
    "erfc": np.vectorize(_erfc, otypes=[float]), "abs": np.abs, "sign": lambda x: 1.0 / x,
    "step": lambda x: np.where(x >= 0.0, 1.0, 0.0), "ramp": lambda x: np.where(x >= 0.0, x, 0.0),
}


def evaluate(tree, env: dict):
    """fct_node_eval: env maps lower-case names to scalars or arrays"""
    kind = tree[0]
    if kind == "cst":
        return tree[1]
    if kind == "var":
        if tree[1] not in env:
            raise FctError(f"unknown variable {tree[1]!r}")
        return env[tree[1]]
    if kind == "fn":
        return _FN[tree[1]](np.asarray(evaluate(tree[2], env), dtype=np.float64))
    a = np.asarray(evaluate(tree[2], env), dtype=np.float64)
    b = np.asarray(evaluate(tree[3], env), dtype=np.float64)
    op = tree[1]
    if op == "+":
        return a + b
    if op == "-":
        return a - b
    if op == "*":
        return a * b
    if op == "/":
        return a / b
    if op == "^":
        return a ** b
    raise FctError(f"operator {op} is not evaluated")


def probe_env(ccent, cons, t, gamma):
    """prb_vol_calc's variables from the conservative state [nv][n] (calc_varprim) and centres [n][3]"""
    cons = np.asarray(cons, dtype=np.float64)
    n = cons.shape[1]
    rho = cons[0]
    vel = np.zeros((3, n))
    vel[:cons.shape[0] - 2] = cons[2:] / rho
    ps = (gamma - 1.0) * (cons[1] - 0.5 * rho * (vel ** 2).sum(0))
    cc = np.asarray(ccent, dtype=np.float64)[:n]
    return {"rho": rho, "ps": ps, "v": np.sqrt((vel ** 2).sum(0)), "v_x": vel[0], "v_y": vel[1], "v_z": vel[2],
            "x": cc[:, 0], "y": cc[:, 1], "z": cc[:, 2], "t": t}


VOL_AVERAGE, VOL_MIN, VOL_MAX = 21, 22, 23


def vol_probe(expr, kind, ccent, cvol, cons, t, gamma):
    """prb_vol_calc: the probe over the interior cells (one block: prb_vol_add once)"""
    env = probe_env(ccent, cons, t, gamma)
    n = env["rho"].size
    f = np.broadcast_to(np.asarray(evaluate(parse(expr), env), dtype=np.float64), (n,))
    if kind == VOL_MIN:
        return float(min(np.finfo(float).max, f.min()))
    if kind == VOL_MAX:
        return float(max(-np.finfo(float).max, f.max()))
    v = np.asarray(cvol, dtype=np.float64)[:n]
    return float((f * v).sum() / v.sum())
