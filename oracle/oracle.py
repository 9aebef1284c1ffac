"""TEST INFRASTRUCTURE ONLY — the CPU oracle for the filtered group-by path.

Loaded only by tests/, __graft_entry__.smoke() and bench.py's CPU baseline,
as the checker; the product (dataframe_mojo_amd) never imports it.

Two independent halves:
  * `evaluate` restates the reference's filter semantics in numpy: Kleene
    AND/OR/NOT, comparisons with null are null, IEEE float comparisons,
    `%` with the divisor's sign and modulo-by-zero null
    (docs/semantics.md:207-209; docs/expressions.md:91-94). It walks the
    expression tree directly and shares nothing with the product's
    predicate compiler.
  * `dfr_oracle.c` (libdforacle.so) restates the CPU sink's grouping,
    per-row state updates and finish (file:line citations in its header).

`collect(lazyframe)` returns the same DataFrame shape the provider returns.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from dataframe_mojo_amd.expr import Expr
from dataframe_mojo_amd.frame import FLOAT64, INT32, INT64, STRING, Column, DataFrame, Series

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdforacle.so")

_CODES = {INT32: 1, INT64: 2, FLOAT64: 3, STRING: 4}
_AGG = {"len": 0, "count": 1, "sum": 2, "min": 3, "max": 4, "mean": 5, "var": 6, "std": 7}


class OracleError(RuntimeError):
    pass


class _Col(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int32), ("pad", ctypes.c_int32), ("values", ctypes.c_void_p), ("valid", ctypes.c_void_p)]


class _Agg(ctypes.Structure):
    _fields_ = [("op", ctypes.c_int32), ("col", ctypes.c_int32), ("ddof", ctypes.c_int32), ("pad", ctypes.c_int32), ("min_count", ctypes.c_int64)]


_lib = None


def build() -> str:
    import subprocess

    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i64 = ctypes.c_void_p, ctypes.c_int64
        L.dfo_groupby.argtypes = [i64, ctypes.c_int, vp, vp, ctypes.c_int, vp, ctypes.c_int, vp, vp, vp, vp, vp, vp, vp, ctypes.c_char_p, ctypes.c_int]
        L.dfo_groupby.restype = ctypes.c_int
        L.dfo_groupby_parallel.argtypes = [ctypes.c_int] + L.dfo_groupby.argtypes
        L.dfo_groupby_parallel.restype = ctypes.c_int
        L.dfo_dd_moments.argtypes = [i64, vp, vp, vp, i64, ctypes.c_int, vp, vp, vp, vp]
        _lib = L
    return _lib


# ------------------------------------------------------------ filter semantics


def _column_eval(frame, name):
    c = frame.column(name).column
    if c.dtype == STRING:
        labels = np.array(c.labels + [""], dtype=object)
        vals = labels[np.minimum(c.values.astype(np.int64), len(c.labels))]
        return vals, c.is_valid().copy(), STRING
    return c.values, c.is_valid().copy(), c.dtype


def evaluate(frame, e):
    """(values, valid, dtype) of an expression over every row of `frame`."""
    n = frame.height()
    k = e.kind
    if k == "col":
        return _column_eval(frame, e.args[0])
    if k == "lit":
        v = e.args[0]
        if v is None:
            return np.zeros(n, dtype=np.int64), np.zeros(n, dtype=bool), INT64
        if isinstance(v, bool):
            return np.full(n, v), np.ones(n, dtype=bool), "bool"
        if isinstance(v, str):
            return np.full(n, v, dtype=object), np.ones(n, dtype=bool), STRING
        if isinstance(v, float):
            return np.full(n, v, dtype=np.float64), np.ones(n, dtype=bool), FLOAT64
        return np.full(n, v, dtype=np.int64), np.ones(n, dtype=bool), INT64
    if k == "cmp":
        op, a, b = e.args
        av, am, at = evaluate(frame, a)
        bv, bm, bt = evaluate(frame, b)
        if FLOAT64 in (at, bt):
            av, bv = av.astype(np.float64), bv.astype(np.float64)
        with np.errstate(invalid="ignore"):
            f = {"eq": np.equal, "ne": np.not_equal, "lt": np.less, "le": np.less_equal, "gt": np.greater, "ge": np.greater_equal}[op]
            v = np.asarray(f(av, bv), dtype=bool)
        return v, am & bm, "bool"
    if k == "mod":
        av, am, _ = evaluate(frame, e.args[0])
        bv, bm, _ = evaluate(frame, e.args[1])
        av = av.astype(np.int64)
        bv = bv.astype(np.int64)
        zero = bv == 0
        safe = np.where(zero, 1, bv)
        with np.errstate(all="ignore"):
            r = np.remainder(av, safe)  # takes the divisor's sign
        return r, am & bm & ~zero, INT64
    if k in ("and", "or"):
        av, am, _ = evaluate(frame, e.args[0])
        bv, bm, _ = evaluate(frame, e.args[1])
        a_t, a_f = am & av.astype(bool), am & ~av.astype(bool)
        b_t, b_f = bm & bv.astype(bool), bm & ~bv.astype(bool)
        if k == "and":
            t, f = a_t & b_t, a_f | b_f
        else:
            t, f = a_t | b_t, a_f & b_f
        return t, t | f, "bool"
    if k == "not":
        av, am, _ = evaluate(frame, e.args[0])
        return ~av.astype(bool), am, "bool"
    if k == "is_null":
        _, am, _ = evaluate(frame, e.args[0])
        return ~am, np.ones(n, dtype=bool), "bool"
    if k == "is_not_null":
        _, am, _ = evaluate(frame, e.args[0])
        return am.copy(), np.ones(n, dtype=bool), "bool"
    if k == "is_in":
        # EQ per candidate with nulls filled false, OR, then the input's nulls
        # kept (docs/design/scalar-membership-chunks.md:49-51;
        # docs/expressions.md:445-448)
        av, am, at = evaluate(frame, e.args[0])
        hit = np.zeros(n, dtype=bool)
        for c in e.args[1]:
            if c is None:
                continue
            if at == STRING:
                hit |= av == c
            elif at == FLOAT64:
                with np.errstate(invalid="ignore"):
                    hit |= av == float(c)
            elif isinstance(c, float):
                # an integer equals a float candidate only when it is integral
                if np.isfinite(c) and c == int(c):
                    hit |= av.astype(np.int64) == int(c)
            else:
                hit |= av.astype(np.int64) == int(c)
        return hit & am, am.copy(), "bool"
    if k == "between":
        x, lo, hi, closed = e.args
        op_lo = "ge" if closed in ("both", "left") else "gt"
        op_hi = "le" if closed in ("both", "right") else "lt"
        return evaluate(frame, Expr("and", (Expr("cmp", (op_lo, x, lo)), Expr("cmp", (op_hi, x, hi)))))
    raise OracleError(f"oracle cannot evaluate {e!r}")


# ------------------------------------------------------------ values
#
# Aggregate values (arithmetic, CASE), restated from docs/expressions.md:
# 50-100 (nulls propagate; integer + - * and unary -/abs checked at the
# operand width, Int32 computed wider and range-checked; `/` is Float64;
# promotion: the wider integer, an integer with a float is the float; a
# bare literal takes the other operand's type when it fits, else Int32, or
# Int64 when it needs more; two literals are Int64) and the reference's
# error texts (kernels/binary.mojo:108-122 "Int64 sum overflow", "Int64
# subtraction overflow", "Int64 multiplication overflow"; :415-421 "<dtype>
# <operation> overflow" below 64 bits; :1620-1635 negation and abs). Only
# rows a node is evaluated on can fault: the filter's kept rows, within a
# CASE branch the rows that take it (docs/apple-metal.md:711-725; a null
# condition takes ELSE), and only where the operands are valid. Exact
# Python integers decide every overflow.

_BOUNDS = {INT32: (-(1 << 31), (1 << 31) - 1), INT64: (-(1 << 63), (1 << 63) - 1)}
_RANKS = {INT32: 0, INT64: 1, FLOAT64: 2}
_WORD = {("add", INT64): "Int64 sum overflow", ("sub", INT64): "Int64 subtraction overflow",
         ("mul", INT64): "Int64 multiplication overflow", ("neg", INT64): "Int64 negation overflow",
         ("abs", INT64): "Int64 abs overflow", ("add", INT32): "int32 addition overflow",
         ("sub", INT32): "int32 subtraction overflow", ("mul", INT32): "int32 multiplication overflow",
         ("neg", INT32): "int32 negation overflow", ("abs", INT32): "int32 abs overflow"}


class ValueFault(OracleError):
    """A checked overflow: the reference raises its text."""


def _vtype(frame, e):
    """A node's type, or ("lit", value) for a bare numeric literal, or
    "null" for a null literal."""
    if e.kind == "col":
        return frame.column(e.args[0]).dtype
    if e.kind == "lit":
        return "null" if e.args[0] is None else ("lit", e.args[0])
    if e.kind == "arith":
        op, a, b = e.args
        return FLOAT64 if op == "div" else _pair_types(frame, a, b)[2]
    if e.kind in ("neg", "abs"):
        t = _vtype(frame, e.args[0])
        return _alone(t)
    if e.kind == "case":
        return _pair_types(frame, e.args[1], e.args[2])[2]
    raise OracleError(f"oracle cannot type {e!r}")


def _alone(t):
    if isinstance(t, tuple):
        return FLOAT64 if isinstance(t[1], float) else INT64
    return INT64 if t == "null" else t


def _lit_as(v, other):
    """A literal's type beside an operand of type `other`."""
    if isinstance(v, float) or other == FLOAT64 or (isinstance(other, tuple) and isinstance(other[1], float)):
        return FLOAT64
    if other in _BOUNDS:
        lo, hi = _BOUNDS[other]
        if lo <= v <= hi:
            return other
    if isinstance(other, tuple) or other == "null":
        return INT64
    return INT32 if _BOUNDS[INT32][0] <= v <= _BOUNDS[INT32][1] else INT64


def _pair_types(frame, a, b):
    ta, tb = _vtype(frame, a), _vtype(frame, b)
    ba = _lit_as(ta[1], tb) if isinstance(ta, tuple) else ta
    bb = _lit_as(tb[1], ta) if isinstance(tb, tuple) else tb
    if ba == "null" and bb == "null":
        ba = bb = INT64
    ba = bb if ba == "null" else ba
    bb = ba if bb == "null" else bb
    return ba, bb, (ba if _RANKS[ba] >= _RANKS[bb] else bb)


def _as_type(vals, t):
    if t == FLOAT64:
        return np.array([float(v) for v in vals], dtype=np.float64)
    return np.array([int(v) for v in vals], dtype=object)


def evaluate_value(frame, e, rows: np.ndarray, as_type=None):
    """(values, valid, type) of a value expression on the rows `rows`
    selects (other rows: invalid). Raises ValueFault for a checked overflow
    on an evaluated row with valid operands."""
    n = frame.height()
    k = e.kind
    t = as_type or _alone(_vtype(frame, e))
    if k == "col":
        c = frame.column(e.args[0]).column
        vals = c.values.astype(np.float64) if t == FLOAT64 else c.values.astype(object)
        return vals, c.is_valid() & rows, t
    if k == "lit":
        v = e.args[0]
        if v is None:
            return np.zeros(n, dtype=np.float64 if t == FLOAT64 else object), np.zeros(n, dtype=bool), t
        return (np.full(n, float(v)) if t == FLOAT64 else np.full(n, int(v), dtype=object)), rows.copy(), t
    if k == "arith":
        op, a, b = e.args
        ta, tb, node = _pair_types(frame, a, b)
        if op == "div":
            node = FLOAT64
        av, am, _ = evaluate_value(frame, a, rows, ta)
        bv, bm, _ = evaluate_value(frame, b, rows, tb)
        ok = am & bm
        if node == FLOAT64:
            av, bv = av.astype(np.float64), bv.astype(np.float64)
            with np.errstate(all="ignore"):
                r = {"add": np.add, "sub": np.subtract, "mul": np.multiply, "div": np.true_divide}[op](av, bv)
            return r, ok, node
        r = np.array([0] * n, dtype=object)
        lo, hi = _BOUNDS[node]
        for i in np.nonzero(ok)[0]:
            x, y = int(av[i]), int(bv[i])
            z = x + y if op == "add" else (x - y if op == "sub" else x * y)
            if z < lo or z > hi:
                raise ValueFault(_WORD[(op, node)])
            r[i] = z
        return r, ok, node
    if k in ("neg", "abs"):
        xv, xm, _ = evaluate_value(frame, e.args[0], rows, t)
        if t == FLOAT64:
            return (-xv if k == "neg" else np.abs(xv)), xm, t
        r = np.array([0] * n, dtype=object)
        lo, _ = _BOUNDS[t]
        for i in np.nonzero(xm)[0]:
            x = int(xv[i])
            if x == lo:
                raise ValueFault(_WORD[(k, t)])
            r[i] = -x if (k == "neg" or x < 0) else x
        return r, xm, t
    if k == "case":
        cond, then, other = e.args
        tt, to, node = _pair_types(frame, then, other)
        cv, cm, _ = evaluate(frame, cond)
        take = cm & cv.astype(bool)
        tv, tm, _ = evaluate_value(frame, then, rows & take, tt)
        ov, om, _ = evaluate_value(frame, other, rows & ~take, to)
        tv, ov = _as_type(tv, node), _as_type(ov, node)
        return np.where(take, tv, ov), np.where(take, tm, om) & rows, node
    raise OracleError(f"oracle cannot evaluate the value {e!r}")


def value_column(frame, e, keep: np.ndarray) -> Column:
    vals, valid, t = evaluate_value(frame, e, keep)
    if t == FLOAT64:
        out = np.where(valid, vals.astype(np.float64), 0.0)
    else:
        out = np.array([int(v) if ok else 0 for v, ok in zip(vals, valid)], dtype=np.int64)
        if t == INT32:
            out = out.astype(np.int32)
    return Column(t, out, valid)


def kept_rows(frame, filters) -> np.ndarray:
    keep = np.ones(frame.height(), dtype=bool)
    for f in filters:
        v, m, _ = evaluate(frame, f)
        keep &= m & v.astype(bool)
    return keep


# ------------------------------------------------------------ grouping


def _ocol(column: Column, hold: list) -> _Col:
    vals = np.ascontiguousarray(column.values)
    valid = None if column.valid is None else np.ascontiguousarray(column.valid.astype(np.uint8))
    hold.extend([vals, valid])
    c = _Col()
    c.dtype = _CODES[column.dtype]
    c.values = vals.ctypes.data
    c.valid = valid.ctypes.data if valid is not None else None
    return c


def _out_dtype(op: str, dt: str) -> str:
    if op in ("len", "count"):
        return INT64
    if op in ("mean", "var", "std"):
        return FLOAT64
    return dt


def groupby(frame, keep: np.ndarray, keys: list, aggs: list, maintain_order: bool = True, threads: int = 1):
    """Runs the C oracle. `aggs`: (op, column name or None, ddof, min_count).
    Returns (DataFrame pieces as (name, Column) lists, first_rows). With
    threads > 1 the parallel restatement runs (dfo_groupby_parallel: the
    same states, bit for bit); its groups are put in first-row order when
    `maintain_order` asks for it."""
    n = frame.height()
    hold: list = []
    kcols = (_Col * max(1, len(keys)))(*[_ocol(frame.column(k).column, hold) for k in keys])
    names = []
    for _, c, _, _ in aggs:
        if c is not None and c not in names:
            names.append(c)
    vcols = (_Col * max(1, len(names)))(*[_ocol(frame.column(c).column, hold) for c in names])
    oaggs = (_Agg * max(1, len(aggs)))()
    for i, (op, c, ddof, mc) in enumerate(aggs):
        oaggs[i].op = _AGG[op]
        oaggs[i].col = names.index(c) if c is not None else 0
        oaggs[i].ddof = ddof
        oaggs[i].min_count = mc
    keep8 = np.ascontiguousarray(keep.astype(np.uint8))
    cap = max(n, 1)
    groups = ctypes.c_int64(0)
    first = np.empty(cap, dtype=np.int64)
    kv = np.empty(max(1, len(keys)) * cap, dtype=np.int64)
    kval = np.empty(max(1, len(keys)) * cap, dtype=np.uint8)
    ab = np.empty(max(1, len(aggs)) * cap, dtype=np.uint64)
    aval = np.empty(max(1, len(aggs)) * cap, dtype=np.uint8)
    err = ctypes.create_string_buffer(256)
    args = (
        n, len(keys), ctypes.addressof(kcols), keep8.ctypes.data, len(names), ctypes.addressof(vcols), len(aggs),
        ctypes.addressof(oaggs), ctypes.addressof(groups), first.ctypes.data, kv.ctypes.data, kval.ctypes.data,
        ab.ctypes.data, aval.ctypes.data, err, len(err),
    )
    rc = lib().dfo_groupby_parallel(threads, *args) if threads > 1 else lib().dfo_groupby(*args)
    if rc != 0:
        raise OracleError(err.value.decode())
    G = groups.value
    if threads > 1 and maintain_order:
        order = np.argsort(first[:G], kind="stable")
        first[:G] = first[:G][order]
        for arr, count in ((kv, len(keys)), (kval, len(keys)), (ab, len(aggs)), (aval, len(aggs))):
            for i in range(count):
                arr[i * cap : i * cap + G] = arr[i * cap : i * cap + G][order]
    out = []
    for i, k in enumerate(keys):
        src = frame.column(k).column
        vals = kv[i * cap : i * cap + G]
        valid = kval[i * cap : i * cap + G].astype(bool)
        if src.dtype == STRING:
            out.append((k, Column(STRING, vals.astype(np.uint32), valid, src.labels)))
        elif src.dtype == FLOAT64:
            out.append((k, Column(FLOAT64, vals.view(np.float64).copy(), valid)))
        else:
            out.append((k, Column(src.dtype, vals.astype(np.int32 if src.dtype == INT32 else np.int64), valid)))
    agg_cols = []
    for i, (op, c, _, _) in enumerate(aggs):
        dt = _out_dtype(op, frame.column(c).column.dtype if c is not None else INT64)
        bits = ab[i * cap : i * cap + G]
        valid = aval[i * cap : i * cap + G].astype(bool)
        if dt == FLOAT64:
            vals = bits.view(np.float64).copy()
        elif dt == INT32:
            vals = bits.view(np.int64).astype(np.int32)
        else:
            vals = bits.view(np.int64).copy()
        agg_cols.append(Column(dt, vals, valid))
    return out, agg_cols, first[:G].copy()


def collect(plan, threads: int = 1):
    """The oracle's answer for a LazyFrame: filter -> group_by -> agg (on
    `threads` workers when more than one)."""
    frame = plan.source
    keep = kept_rows(frame, plan.filters)
    specs, names = [], []
    derived = []
    for e in plan.aggs:
        op, target, kw = e.args
        kw = dict(kw)
        name = None if target is None else (target.args[0] if target.kind == "col" else None)
        if target is not None and target.kind != "col":
            # an aggregate's value: computed for the kept rows, then grouped
            name = f"\x00value{len(derived)}"
            derived.append(Series(name, value_column(frame, target, keep)))
        specs.append((op, name, int(kw.get("ddof", 1)), int(kw.get("min_count", 0))))
        names.append(e.output_name())
    if derived:
        frame = DataFrame([frame.column(c) for c in frame.columns] + derived)
    key_cols, agg_cols, first = groupby(frame, keep, list(plan.keys), specs, plan.maintain_order, threads)
    series = [Series(k, c) for k, c in key_cols] + [Series(nm, c) for nm, c in zip(names, agg_cols)]
    return DataFrame(series), first


def dd_moments(gid: np.ndarray, x: np.ndarray, valid, G: int, ddof: int = 1):
    """Double-double per-group sum (hi + lo), count and two-pass variance."""
    n = len(x)
    gid = np.ascontiguousarray(gid, dtype=np.int64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    v8 = None if valid is None else np.ascontiguousarray(np.asarray(valid, dtype=np.uint8))
    hi = np.empty(max(G, 1))
    lo = np.empty(max(G, 1))
    cnt = np.empty(max(G, 1), dtype=np.int64)
    var = np.empty(max(G, 1))
    lib().dfo_dd_moments(n, gid.ctypes.data, x.ctypes.data, v8.ctypes.data if v8 is not None else None, G, ddof,
                         hi.ctypes.data, lo.ctypes.data, cnt.ctypes.data, var.ctypes.data)
    return hi[:G], lo[:G], cnt[:G], var[:G]
