"""Custom single-table scaling: the caller's code strings, per source level.

volumential's ``NearFieldFromCSR`` accepts ``kernel_scaling_code`` and
``kernel_displacement_code`` (volumential/list1.py:198-205, 389-404): C
expressions spliced into its generated list-1 kernel, with table level 0
for every source box.  ``BOX`` in them names the source box (``sbox``).
The scaling may use ``sbox_extent`` (root_extent / 2^level), ``sbox_level``,
``root_extent`` and ``table_root_extent``; the displacement may further use
``scaling`` (the value of the scaling code) and
``mode_nmlz[table_lev, sid]``, the normaliser of the source mode.

This library has no run-time code generation.  Both strings depend on the
source box only through its level, so they are evaluated here, once per
level, into ``level_scaling[l]`` and ``level_disp[l]``; the kernels apply
``T * level_scaling[l] + level_disp[l] * mode_nmlz[0, s]``.  The
displacement must therefore be linear in ``mode_nmlz[...]`` with no other
term (checked: it is evaluated at three normaliser values).  Expressions are
parsed with Python's ``ast`` and evaluated over numbers only: arithmetic,
comparisons, ``a if c else b``, the math functions below and the names
above; anything else (attributes, other names, other calls) is refused.
Evaluation is in double precision (``/`` is always a float division).
"""

from __future__ import annotations

import ast
import math

import numpy as np

from . import _capi

MAX_LEVELS = _capi.MAX_LEVELS

_FUNCS = {
    "log": math.log, "log2": math.log2, "log10": math.log10, "exp": math.exp,
    "sqrt": math.sqrt, "pow": math.pow, "fabs": math.fabs, "abs": abs,
    "sin": math.sin, "cos": math.cos, "tan": math.tan, "atan": math.atan,
    "min": min, "max": max, "floor": math.floor, "ceil": math.ceil,
}
_CONSTS = {"PI": math.pi, "M_PI": math.pi, "pi": math.pi}
_BINOPS = {
    ast.Add: lambda a, b: a + b, ast.Sub: lambda a, b: a - b,
    ast.Mult: lambda a, b: a * b, ast.Div: lambda a, b: a / b,
    ast.Pow: lambda a, b: a ** b, ast.Mod: math.fmod,
}
_CMPS = {
    ast.Lt: lambda a, b: a < b, ast.LtE: lambda a, b: a <= b,
    ast.Gt: lambda a, b: a > b, ast.GtE: lambda a, b: a >= b,
    ast.Eq: lambda a, b: a == b, ast.NotEq: lambda a, b: a != b,
}


class _Eval:
    def __init__(self, names, nmlz=None):
        self.names = names
        self.nmlz = nmlz  # value of mode_nmlz[...] (displacement only)

    def __call__(self, node):
        meth = getattr(self, "_" + type(node).__name__, None)
        if meth is None:
            raise ValueError(f"unsupported syntax in scaling code: {type(node).__name__}")
        return meth(node)

    def _Expression(self, n):
        return self(n.body)

    def _Constant(self, n):
        if isinstance(n.value, bool) or not isinstance(n.value, (int, float)):
            raise ValueError(f"unsupported constant in scaling code: {n.value!r}")
        return float(n.value)

    def _Name(self, n):
        if n.id in self.names:
            return float(self.names[n.id])
        if n.id in _CONSTS:
            return _CONSTS[n.id]
        raise ValueError(f"unknown name {n.id!r} in scaling code")

    def _BinOp(self, n):
        op = _BINOPS.get(type(n.op))
        if op is None:
            raise ValueError(f"unsupported operator in scaling code: {type(n.op).__name__}")
        return float(op(self(n.left), self(n.right)))

    def _UnaryOp(self, n):
        v = self(n.operand)
        if isinstance(n.op, ast.USub):
            return -v
        if isinstance(n.op, ast.UAdd):
            return v
        if isinstance(n.op, ast.Not):
            return float(not v)
        raise ValueError("unsupported unary operator in scaling code")

    def _Compare(self, n):
        left = self(n.left)
        for op, right_node in zip(n.ops, n.comparators):
            f = _CMPS.get(type(op))
            if f is None:
                raise ValueError("unsupported comparison in scaling code")
            right = self(right_node)
            if not f(left, right):
                return 0.0
            left = right
        return 1.0

    def _BoolOp(self, n):
        vals = [self(v) for v in n.values]
        return float(all(vals) if isinstance(n.op, ast.And) else any(vals))

    def _IfExp(self, n):
        return self(n.body) if self(n.test) else self(n.orelse)

    def _Call(self, n):
        if not isinstance(n.func, ast.Name) or n.func.id not in _FUNCS or n.keywords:
            raise ValueError("unsupported call in scaling code")
        return float(_FUNCS[n.func.id](*[self(a) for a in n.args]))

    def _Subscript(self, n):
        if not (isinstance(n.value, ast.Name) and n.value.id == "mode_nmlz"):
            raise ValueError("only mode_nmlz[table_lev, sid] may be subscripted in scaling code")
        if self.nmlz is None:
            raise ValueError("mode_nmlz is only available to the displacement code")
        return self.nmlz


def _parse(code):
    if not isinstance(code, str):
        raise TypeError("scaling and displacement codes must be strings")
    # the reference splices BOX -> the source box name
    text = " ".join(code.replace("BOX", "sbox").split())
    try:
        return ast.parse(text, mode="eval")
    except SyntaxError as exc:
        raise ValueError(f"cannot parse scaling code {code!r}: {exc.msg}") from None


def level_arrays(scaling_code, displacement_code, root_extent, table_root_extent):
    """``(level_scaling, level_disp)``, float64 arrays of ``MAX_LEVELS``:
    the scaling code's value and the displacement code's coefficient of
    ``mode_nmlz[table_lev, sid]`` for a source box of each level."""
    sc, dc = _parse(scaling_code), _parse(displacement_code)
    scal = np.zeros(MAX_LEVELS)
    disp = np.zeros(MAX_LEVELS)
    for lev in range(MAX_LEVELS):
        names = {"sbox_extent": float(root_extent) / 2.0 ** lev, "sbox_level": float(lev),
                 "root_extent": float(root_extent), "table_root_extent": float(table_root_extent),
                 "table_lev": 0.0}
        try:
            s = _Eval(names)(sc)
            names["scaling"] = s
            d0, d1, d2 = (_Eval(names, v)(dc) for v in (0.0, 1.0, 2.0))
        except (ArithmeticError, TypeError) as exc:
            raise ValueError(f"scaling or displacement code fails at source level {lev}: "
                             f"{exc}") from None
        if not (math.isfinite(s) and math.isfinite(d1) and math.isfinite(d2)):
            raise ValueError(f"scaling or displacement code is not finite at source level {lev}")
        if d0 != 0.0:
            raise NotImplementedError(
                "kernel_displacement_code must be a multiple of mode_nmlz[table_lev, sid] "
                "(a term without it is not supported)")
        if abs(d2 - 2.0 * d1) > 1e-12 * max(abs(d1), 1e-300):
            raise NotImplementedError(
                "kernel_displacement_code must be linear in mode_nmlz[table_lev, sid]")
        scal[lev], disp[lev] = s, d1
    return scal, disp
