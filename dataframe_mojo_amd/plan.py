"""Compiles a LazyFrame of the form source -> filter* -> group_by -> agg into
the device query: the predicate program (leaves + truth table, as
dataframe_metal/predicate_program.mojo:1-71 compiles positive AND/OR
predicates), the key list and the aggregate list.

NOT is pushed into the leaves with Kleene semantics: NOT(a AND b) is
NOT a OR NOT b, and a negated comparison stays "not true" on a null value.
After that the formula is monotone in the leaves' is-true bits, so a 256-bit
truth table over those bits is the SQL filter (false and null dropped,
docs/semantics.md:207-209).

Anything outside the device path raises `Unsupported` before any device
work, as the accelerator contract asks (docs/design/accelerator-contract.md).
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass, field

from . import _lib
from .expr import Expr
from .frame import FLOAT64, INT32, INT64, STRING, LazyFrame

_CMP = {"eq": _lib.EQ, "ne": _lib.NE, "lt": _lib.LT, "le": _lib.LE, "gt": _lib.GT, "ge": _lib.GE}
_FLIP = {"eq": "eq", "ne": "ne", "lt": "gt", "le": "ge", "gt": "lt", "ge": "le"}
_AGG = {
    "len": _lib.AGG_LEN,
    "count": _lib.AGG_COUNT,
    "sum": _lib.AGG_SUM,
    "min": _lib.AGG_MIN,
    "max": _lib.AGG_MAX,
    "mean": _lib.AGG_MEAN,
    "var": _lib.AGG_VAR,
    "std": _lib.AGG_STD,
}
DTYPE_CODE = {INT32: _lib.DFR_INT32, INT64: _lib.DFR_INT64, FLOAT64: _lib.DFR_FLOAT64, STRING: _lib.DFR_DICT}


class Unsupported(_lib.Unsupported):
    def __init__(self, message: str):
        super().__init__(_lib.DFR_UNSUPPORTED, message)


@dataclass
class Leaf:
    column: str
    op: int
    flags: int = 0
    literal: int = 0  # int64 value or double bits
    modulus: int = 0
    column2: str | None = None  # the right-hand column of a column comparison (LEAF_COLUMN)


@dataclass
class Program:
    leaves: list = field(default_factory=list)
    truth: list = field(default_factory=lambda: [0] * 8)


@dataclass
class ValueSpec:
    """An aggregate's expression as a device value program: computed by
    dfr_column_compute into the column `name` of DevicePlan.columns."""

    name: str
    inputs: list  # input column names (program input i = inputs[i])
    ops: list  # (op, arg, dtype code, literal bits), postfix
    conditions: list  # Programs whose leaves name inputs
    dtype: str  # the result type


@dataclass
class DevicePlan:
    keys: list
    columns: list  # names of the query's input columns
    program: Program
    aggs: list  # (op, column index or -1, ddof, min_count)
    agg_names: list
    maintain_order: bool
    values: list = field(default_factory=list)  # ValueSpec of each derived column

    @property
    def output_names(self) -> list:
        return list(self.keys) + list(self.agg_names)


def _f64_bits(x: float) -> int:
    return struct.unpack("<q", struct.pack("<d", float(x)))[0]


def _is_int(v) -> bool:
    return isinstance(v, int) and not isinstance(v, bool)


class _Compiler:
    def __init__(self, frame):
        self.frame = frame
        self.leaves: list[Leaf] = []

    def dtype(self, name: str) -> str:
        try:
            return self.frame.column(name).dtype
        except KeyError:
            raise Unsupported(f"no column named {name!r}") from None

    # returns a formula: ("leaf", i) | ("and", a, b) | ("or", a, b) | ("const", bool)
    def formula(self, e: Expr, neg: bool):
        k = e.kind
        if k == "not":
            return self.formula(e.args[0], not neg)
        if k in ("and", "or"):
            a = self.formula(e.args[0], neg)
            b = self.formula(e.args[1], neg)
            op = k if not neg else ("or" if k == "and" else "and")
            return (op, a, b)
        if k == "lit" and isinstance(e.args[0], bool):
            return ("const", e.args[0] != neg)
        if k in ("is_null", "is_not_null"):
            target = e.args[0]
            if target.kind != "col":
                raise Unsupported("is_null of an expression")
            op = _lib.IS_NULL if (k == "is_null") != neg else _lib.IS_NOT_NULL
            return self.leaf(Leaf(target.args[0], op))
        if k == "cmp":
            return self.comparison(e, neg)
        if k == "between":
            x, lo, hi, closed = e.args
            op_lo = "ge" if closed in ("both", "left") else "gt"
            op_hi = "le" if closed in ("both", "right") else "lt"
            return self.formula(Expr("and", (Expr("cmp", (op_lo, x, lo)), Expr("cmp", (op_hi, x, hi)))), neg)
        if k == "is_in":
            return self.membership(e, neg)
        if k == "col" and self.dtype(e.args[0]) not in (INT32, INT64, FLOAT64, STRING):
            raise Unsupported("boolean columns are not on the device path")
        raise Unsupported(f"predicate {e!r} is not on the device path")

    def membership(self, e: Expr, neg: bool):
        """is_in as an OR of equality leaves, one per distinct non-null
        candidate that can equal a value of the column (an integer column
        never equals a fractional candidate); a null input makes every leaf
        null, so the OR stays null. With no candidate left the result is
        false, or null for a null input (negated: is_not_null)."""
        x, values = e.args
        if x.kind != "col":
            raise Unsupported("is_in of an expression")
        dt = self.dtype(x.args[0])
        cands = []
        for v in values:
            if v is None:
                continue  # never matches
            if isinstance(v, bool):
                raise Unsupported("is_in with a boolean candidate")
            if dt == STRING:
                if not isinstance(v, str):
                    raise Unsupported("dictionary columns take string candidates")
            elif isinstance(v, str):
                raise Unsupported(f"numeric column is_in with {v!r}")
            elif dt != FLOAT64 and isinstance(v, float):
                if not math.isfinite(v) or v != int(v):
                    continue  # no integer equals it
                v = int(v)
            elif dt == FLOAT64:
                v = float(v)
            if v not in cands:
                cands.append(v)
        if not cands:
            if neg:
                return self.leaf(Leaf(x.args[0], _lib.IS_NOT_NULL))
            return ("const", False)
        f = None
        for v in cands:
            term = Expr("cmp", ("eq", x, Expr("lit", (v,))))
            f = term if f is None else Expr("or", (f, term))
        return self.formula(f, neg)

    def leaf(self, leaf: Leaf):
        self.leaves.append(leaf)
        if len(self.leaves) > _lib.MAX_LEAVES:
            raise Unsupported(f"more than {_lib.MAX_LEAVES} predicate leaves")
        return ("leaf", len(self.leaves) - 1)

    def comparison(self, e: Expr, neg: bool):
        op, a, b = e.args
        if a.kind == "lit" and b.kind != "lit":
            a, b, op = b, a, _FLIP[op]
        if a.kind == "col" and b.kind == "col":
            # column <op> column: two integer columns (compared by value) or
            # two Float64 columns; null on either side is null
            na, nb = a.args[0], b.args[0]
            ta, tb = self.dtype(na), self.dtype(nb)
            ints = ta in (INT32, INT64) and tb in (INT32, INT64)
            if not (ints or (ta == FLOAT64 and tb == FLOAT64)):
                raise Unsupported(f"comparison of {ta} with {tb} columns is not on the device path")
            flags = _lib.LEAF_COLUMN | (_lib.LEAF_NOT if neg else 0)
            return self.leaf(Leaf(na, _CMP[op], flags, 0, 0, nb))
        if b.kind != "lit":
            raise Unsupported("comparisons of expressions other than columns and literals are not on the device path")
        value = b.args[0]
        flags = _lib.LEAF_NOT if neg else 0
        modulus = 0
        if a.kind == "mod":
            base, m = a.args
            if base.kind != "col" or m.kind != "lit" or not _is_int(m.args[0]):
                raise Unsupported("only column % integer literal")
            a = base
            modulus = int(m.args[0])
            flags |= _lib.LEAF_MOD
        if a.kind != "col":
            raise Unsupported(f"comparison of {a!r}")
        name = a.args[0]
        dt = self.dtype(name)
        if value is None:
            # a comparison with null is null: never true, also when negated
            return ("const", False)
        if dt == STRING:
            if modulus or op not in ("eq", "ne") or not isinstance(value, str):
                raise Unsupported("dictionary columns compare by (in)equality with a string")
            labels = self.frame.column(name).column.labels
            code = labels.index(value) if value in labels else -1
            return self.leaf(Leaf(name, _CMP[op], flags, code))
        if dt == FLOAT64:
            if modulus:
                raise Unsupported("Float64 modulo is not on the device path")
            if not isinstance(value, (int, float)) or isinstance(value, bool):
                raise Unsupported(f"Float64 column compared with {value!r}")
            return self.leaf(Leaf(name, _CMP[op], flags, _f64_bits(float(value))))
        # integer column
        if isinstance(value, float):
            if not math.isfinite(value) or value != int(value):
                raise Unsupported("integer column compared with a fractional Float literal")
            value = int(value)
        if not _is_int(value):
            raise Unsupported(f"integer column compared with {value!r}")
        if not (-(1 << 63) <= value < (1 << 63)):
            raise Unsupported("literal outside Int64")
        return self.leaf(Leaf(name, _CMP[op], flags, int(value), modulus))


# ------------------------------------------------------------ value programs
#
# Types follow the reference's promotion (docs/expressions.md:57-84): two
# integer operands compute in the wider (Int32 with Int32 stays Int32,
# checked at 32 bits); an integer with Float64 is Float64; `/` is Float64. A
# bare literal takes the other operand's type when it fits, else Int32, or
# Int64 when it needs more; a float literal beside an integer is Float64; a
# literal with a literal is Int64 (Float64 for a float).

_RANK = {INT32: 0, INT64: 1, FLOAT64: 2}
_VAL_CODE = {INT32: _lib.DFR_INT32, INT64: _lib.DFR_INT64, FLOAT64: _lib.DFR_FLOAT64}
_VAL_OP = {"add": _lib.VAL_ADD, "sub": _lib.VAL_SUB, "mul": _lib.VAL_MUL, "div": _lib.VAL_DIV}


def _is_bare_lit(e: Expr) -> bool:
    return e.kind == "lit"


def _lit_own_type(v) -> str:
    if isinstance(v, float):
        return FLOAT64
    return INT32 if -(1 << 31) <= v < (1 << 31) else INT64


class _ValueCompiler:
    def __init__(self, frame, name: str):
        self.frame = frame
        self.name = name
        self.inputs: list[str] = []
        self.ops: list = []
        self.conditions: list = []

    def input(self, name: str) -> int:
        if name not in self.inputs:
            self.inputs.append(name)
        return self.inputs.index(name)

    def type_of(self, e: Expr) -> str:
        """The node's type; "int?"/"float?"/"null" for a bare literal."""
        k = e.kind
        if k == "col":
            try:
                dt = self.frame.column(e.args[0]).dtype
            except KeyError:
                raise Unsupported(f"no column named {e.args[0]!r}") from None
            if dt not in _RANK:
                raise Unsupported(f"{dt} columns are not values on the device path")
            return dt
        if k == "lit":
            v = e.args[0]
            if v is None:
                return "null"
            if isinstance(v, bool) or not isinstance(v, (int, float)):
                raise Unsupported(f"literal {v!r} is not a numeric value")
            if isinstance(v, int) and not (-(1 << 63) <= v < (1 << 63)):
                raise Unsupported("literal outside Int64")
            return "float?" if isinstance(v, float) else "int?"
        if k == "arith":
            op, a, b = e.args
            if op == "div":
                self.type_of(a), self.type_of(b)
                return FLOAT64
            return self._combine(a, b)
        if k in ("neg", "abs"):
            t = self.type_of(e.args[0])
            return {"int?": INT64, "float?": FLOAT64, "null": INT64}.get(t, t)
        if k == "case":
            return self._combine(e.args[1], e.args[2])
        raise Unsupported(f"{e!r} is not a value on the device path")

    def _combine(self, a: Expr, b: Expr) -> str:
        ta, tb = self.type_of(a), self.type_of(b)
        return self._bind_pair(a, ta, b, tb)[2]

    def _bind_pair(self, a, ta, b, tb):
        """(type of a, type of b, node type) after literal binding."""
        def bound(e, t, other):
            if t in _RANK:
                return t
            if t == "null":
                return other if other in _RANK else None
            if t == "float?":
                return FLOAT64
            # an untyped integer literal
            if other == FLOAT64 or other == "float?":
                return FLOAT64
            if other in (INT32, INT64) and _is_bare_lit(e):
                v = e.args[0]
                lo, hi = (-(1 << 31), 1 << 31) if other == INT32 else (-(1 << 63), 1 << 63)
                return other if lo <= v < hi else _lit_own_type(v)
            return INT64
        ba, bb = bound(a, ta, tb), bound(b, tb, ta)
        if ba is None and bb is None:
            ba = bb = INT64
        ba = ba or bb
        bb = bb or ba
        node = ba if _RANK[ba] >= _RANK[bb] else bb
        return ba, bb, node

    def emit(self, e: Expr, t: str):
        """Postfix ops of `e` computed as type `t` (its bound type)."""
        k = e.kind
        if k == "col":
            self.ops.append((_lib.VAL_COLUMN, self.input(e.args[0]), _VAL_CODE[self.type_of(e)], 0))
            return
        if k == "lit":
            v = e.args[0]
            if v is None:
                self.ops.append((_lib.VAL_NULL, 0, _VAL_CODE[t], 0))
            elif t == FLOAT64:
                self.ops.append((_lib.VAL_LITERAL, 0, _VAL_CODE[t], _f64_bits(float(v))))
            else:
                self.ops.append((_lib.VAL_LITERAL, 0, _VAL_CODE[t], int(v)))
            return
        if k == "arith":
            op, a, b = e.args
            ba, bb, node = self._bind_pair(a, self.type_of(a), b, self.type_of(b))
            if op == "div":
                node = FLOAT64
            self.emit(a, ba)
            self.emit(b, bb)
            self.ops.append((_VAL_OP[op], 0, _VAL_CODE[node], 0))
            return
        if k in ("neg", "abs"):
            x = e.args[0]
            tx = self.type_of(x)
            bt = {"int?": INT64, "float?": FLOAT64, "null": INT64}.get(tx, tx)
            self.emit(x, bt)
            self.ops.append((_lib.VAL_NEG if k == "neg" else _lib.VAL_ABS, 0, _VAL_CODE[bt], 0))
            return
        if k == "case":
            cond, then, other = e.args
            bt, bo, node = self._bind_pair(then, self.type_of(then), other, self.type_of(other))
            if len(self.conditions) == _lib.MAX_CONDITIONS:
                raise Unsupported(f"more than {_lib.MAX_CONDITIONS} CASE conditions in one value")
            prog = compile_predicate(self.frame, [cond])
            for leaf in prog.leaves:
                self.input(leaf.column)
                if leaf.column2 is not None:
                    self.input(leaf.column2)
            self.conditions.append(prog)
            ci = len(self.conditions) - 1
            self.emit(then, bt)
            self.emit(other, bo)
            self.ops.append((_lib.VAL_CASE, ci, _VAL_CODE[node], 0))
            return
        raise Unsupported(f"{e!r} is not a value on the device path")

    def compile(self, e: Expr, filter_program: Program) -> ValueSpec:
        t = self.type_of(e)
        t = {"int?": INT64, "float?": FLOAT64, "null": INT64}.get(t, t)
        self.emit(e, t)
        for leaf in filter_program.leaves:
            self.input(leaf.column)
            if leaf.column2 is not None:
                self.input(leaf.column2)
        if len(self.ops) > _lib.MAX_VALUE_OPS:
            raise Unsupported(f"a value program has at most {_lib.MAX_VALUE_OPS} nodes")
        if len(self.inputs) > 8:
            raise Unsupported("a value program reads at most 8 columns (its filter's included)")
        depth = peak = 0
        for op, _, _, _ in self.ops:
            depth += 1 if op in (_lib.VAL_COLUMN, _lib.VAL_LITERAL, _lib.VAL_NULL) else (
                0 if op in (_lib.VAL_NEG, _lib.VAL_ABS) else -1)
            peak = max(peak, depth)
        if peak > 8:
            raise Unsupported("a value program needs more than 8 stack entries")
        return ValueSpec(self.name, self.inputs, self.ops, self.conditions, t)


def _eval(f, bits: int) -> bool:
    t = f[0]
    if t == "leaf":
        return bool((bits >> f[1]) & 1)
    if t == "const":
        return f[1]
    if t == "and":
        return _eval(f[1], bits) and _eval(f[2], bits)
    return _eval(f[1], bits) or _eval(f[2], bits)


def compile_predicate(frame, filters) -> Program:
    comp = _Compiler(frame)
    formula = ("const", True)
    for f in filters:
        formula = ("and", formula, comp.formula(f, False))
    prog = Program(leaves=comp.leaves)
    n = len(comp.leaves)
    for m in range(1 << n):
        if _eval(formula, m):
            prog.truth[m >> 5] |= 1 << (m & 31)
    return prog


def compile_plan(plan: LazyFrame) -> DevicePlan:
    if not plan.grouped:
        raise Unsupported("the device path runs filter -> group_by -> agg plans")
    frame = plan.source
    if not plan.keys:
        raise Unsupported("group_by needs at least one key")
    for k in plan.keys:
        dt = frame.column(k).dtype
        if dt == FLOAT64 and len(plan.keys) > 1:
            # a Float64 key groups by its canonical 64-bit word (-0.0 == 0.0,
            # one NaN group; group_by.mojo:319-329), the whole word range
            raise Unsupported("a Float64 group key is the only key on the device path")
    program = compile_predicate(frame, plan.filters)
    columns: list[str] = []

    def colindex(name: str) -> int:
        frame.column(name)
        if name not in columns:
            columns.append(name)
        return columns.index(name)

    for leaf in program.leaves:
        colindex(leaf.column)
        if leaf.column2 is not None:
            colindex(leaf.column2)
    aggs, names = [], []
    values: list[ValueSpec] = []
    value_exprs: list[str] = []
    for e in plan.aggs:
        if not isinstance(e, Expr) or e.kind != "agg":
            raise Unsupported(f"{e!r} is not a reduction")
        op, target, kw = e.args
        kw = dict(kw)
        if target is None:
            aggs.append((_lib.AGG_LEN, -1, 0, 0))
        elif target.kind != "col":
            # an arithmetic or CASE value: a derived column (one per distinct
            # expression) computed on the device for the kept rows
            vname = next((v.name for v, ex in zip(values, value_exprs) if ex == repr(target)), None)
            if vname is None:
                vname = f"\x00value{len(values)}"
                values.append(_ValueCompiler(frame, vname).compile(target, program))
                value_exprs.append(repr(target))
            columns.append(vname) if vname not in columns else None
            aggs.append((_AGG[op], columns.index(vname), int(kw.get("ddof", 1)), int(kw.get("min_count", 0))))
        else:
            name = target.args[0]
            dt = frame.column(name).dtype
            if dt == STRING and op not in ("count", "len"):
                raise Unsupported(f"{op} of a String column is not on the device path")
            aggs.append((_AGG[op], colindex(name), int(kw.get("ddof", 1)), int(kw.get("min_count", 0))))
        names.append(e.output_name())
    out = list(plan.keys) + names
    if len(set(out)) != len(out):
        raise ValueError(f"duplicate output names {out}")
    return DevicePlan(list(plan.keys), columns, program, aggs, names, plan.maintain_order, values)
