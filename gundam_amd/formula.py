"""Event-selection formulas of the YAML configuration (``applyCondition``,
``selectionCutFormula``), evaluated over the columns of an event store.

The reference hands these strings to ROOT's TTreeFormula / TFormula (with
``[var]`` naming a leaf).  This is a small evaluator of the same expression
language subset: numbers, ``[var]`` or bare variable names, + - * / %,
comparisons, && || ! (and Python's and/or/not), parentheses, and the
functions abs, sqrt, exp, log, sin, cos, tan, pow, min, max.  Nothing else is
accepted (no attribute access, no calls of other names), and it never calls
Python's eval.
"""
from __future__ import annotations

import ast
import re
from typing import Callable, Dict

import numpy as np


class FormulaError(ValueError):
    pass


_FUNCS: Dict[str, Callable] = {
    "abs": np.abs, "sqrt": np.sqrt, "exp": np.exp, "log": np.log, "sin": np.sin,
    "cos": np.cos, "tan": np.tan, "pow": np.power, "min": np.minimum, "max": np.maximum,
}


def _to_python(src: str) -> str:
    if "~" in src:
        raise FormulaError(f"unsupported operator '~' in {src!r}")
    s = re.sub(r"\[([A-Za-z_][A-Za-z0-9_.]*)\]", lambda m: "__v_" + m.group(1).replace(".", "__dot__"), src)
    s = s.replace("&&", " and ").replace("||", " or ")
    # TFormula's '!' is a unary operator that binds tighter than arithmetic
    # and comparisons ('!a == b' is '(!a) == b'); Python's unary '~' has that
    # precedence (Python's 'not' binds looser than '=='), and the evaluator
    # reads it as logical not
    s = re.sub(r"!(?!=)", "~", s)
    return s


def _name(n: str) -> str:
    return n[4:].replace("__dot__", ".") if n.startswith("__v_") else n


def evaluate(src: str, column: Callable[[str], np.ndarray], n: int) -> np.ndarray:
    """Value of the formula for each of n events; column(name) -> [n]."""
    try:
        tree = ast.parse(_to_python(src.strip()) or "1", mode="eval")
    except SyntaxError as e:
        raise FormulaError(f"cannot parse {src!r}: {e.msg}") from None

    def ev(node):
        if isinstance(node, ast.Expression):
            return ev(node.body)
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)) \
                and not isinstance(node.value, bool):
            return np.full(n, float(node.value))
        if isinstance(node, ast.Name):
            return np.asarray(column(_name(node.id)), dtype=np.float64)
        if isinstance(node, ast.UnaryOp):
            v = ev(node.operand)
            if isinstance(node.op, ast.USub):
                return -v
            if isinstance(node.op, ast.UAdd):
                return v
            if isinstance(node.op, (ast.Not, ast.Invert)):
                return (v == 0).astype(np.float64)
        if isinstance(node, ast.BinOp):
            a, b = ev(node.left), ev(node.right)
            ops = {ast.Add: np.add, ast.Sub: np.subtract, ast.Mult: np.multiply,
                   ast.Div: np.divide, ast.Mod: np.fmod, ast.Pow: np.power}
            for t, f in ops.items():
                if isinstance(node.op, t):
                    with np.errstate(all="ignore"):
                        return f(a, b)
        if isinstance(node, ast.BoolOp):
            vals = [ev(v) != 0 for v in node.values]
            f = np.logical_and if isinstance(node.op, ast.And) else np.logical_or
            out = vals[0]
            for v in vals[1:]:
                out = f(out, v)
            return out.astype(np.float64)
        if isinstance(node, ast.Compare):
            left = ev(node.left)
            out = np.ones(n, dtype=bool)
            cmp = {ast.Lt: np.less, ast.LtE: np.less_equal, ast.Gt: np.greater,
                   ast.GtE: np.greater_equal, ast.Eq: np.equal, ast.NotEq: np.not_equal}
            for op, rnode in zip(node.ops, node.comparators):
                right = ev(rnode)
                f = next((f for t, f in cmp.items() if isinstance(op, t)), None)
                if f is None:
                    break
                out &= f(left, right)
                left = right
            else:
                return out.astype(np.float64)
        if isinstance(node, ast.Call) and isinstance(node.func, ast.Name) \
                and node.func.id in _FUNCS and not node.keywords:
            with np.errstate(all="ignore"):
                return _FUNCS[node.func.id](*[ev(a) for a in node.args])
        raise FormulaError(f"unsupported element in {src!r}: {ast.dump(node)[:60]}")

    return np.broadcast_to(ev(tree), (n,)).astype(np.float64)


def select(src, column, n) -> np.ndarray:
    """Boolean mask: the formula is nonzero (an empty formula selects all)."""
    if src is None or str(src).strip() == "":
        return np.ones(n, dtype=bool)
    return evaluate(str(src), column, n) != 0
