"""Value programs: arithmetic and CASE inside aggregates
(docs/apple-metal.md:104-120, 635-650, 711-725; docs/expressions.md:50-100).

CPU: the oracle's value semantics are pinned by the reference's documented
cases (promotion, literal binding, checked overflow texts of
kernels/binary.mojo, null propagation, CASE with a null condition, faults
only on evaluated rows) and by a row-by-row scalar evaluator with exact
Python integers on random expressions; the product's compiler types its
nodes as documented. GPU: random filtered group-bys over value expressions
equal the oracle (integers exact, Float64 within 1e-12 relative), and
checked overflows raise the oracle's text as data faults."""
import math

import numpy as np
import pytest

from dataframe_mojo_amd import Column, DataFrame, Series, _lib, col, len_, lit, when
from dataframe_mojo_amd.plan import Unsupported, compile_plan
from oracle import oracle
from oracle.oracle import ValueFault

I32_MAX, I32_MIN = (1 << 31) - 1, -(1 << 31)
I64_MAX, I64_MIN = (1 << 63) - 1, -(1 << 63)


def _frame(a32, b64, x=None, valid_a=None):
    n = len(a32)
    cols = [Series("k", Column.int64(np.arange(n) % 3)), Series("a", Column.int32(a32, valid_a)),
            Series("b", Column.int64(b64))]
    cols.append(Series("x", Column.float64(x if x is not None else np.linspace(-1, 1, n))))
    return DataFrame(cols)


def _value(frame, e, keep=None):
    keep = np.ones(frame.height(), dtype=bool) if keep is None else keep
    vals, valid, t = oracle.evaluate_value(frame, e, keep)
    return [(int(v) if t != "float64" else float(v)) if ok else None for v, ok in zip(vals, valid)], t


# --------------------------------------------------------------- documented cases


def test_int32_stays_int32_and_is_checked_at_32_bits():
    f = _frame([I32_MAX, 1], [0, 0])
    with pytest.raises(ValueFault, match="^int32 addition overflow$"):
        _value(f, col("a") + 1)
    # a literal that does not fit Int32 makes the node Int64 (docs: `col("u8") + 300` is Int32)
    vals, t = _value(f, col("a") + 3_000_000_000)
    assert t == "int64" and vals == [I32_MAX + 3_000_000_000, 3_000_000_001]
    vals, t = _value(f, col("a") * col("b"))
    assert t == "int64" and vals == [0, 0]


def test_int64_overflow_texts():
    f = _frame([0, 0], [I64_MAX, I64_MIN])
    with pytest.raises(ValueFault, match="^Int64 sum overflow$"):
        _value(f, col("b") + 1)
    with pytest.raises(ValueFault, match="^Int64 subtraction overflow$"):
        _value(f, col("b") - 1)
    with pytest.raises(ValueFault, match="^Int64 multiplication overflow$"):
        _value(f, col("b") * 2)
    with pytest.raises(ValueFault, match="^Int64 negation overflow$"):
        _value(f, -col("b"))
    with pytest.raises(ValueFault, match="^Int64 abs overflow$"):
        _value(f, col("b").abs())
    f2 = _frame([I32_MIN, 0], [0, 0])
    with pytest.raises(ValueFault, match="^int32 negation overflow$"):
        _value(f2, -col("a"))


def test_division_is_float64_and_ieee():
    f = _frame([1, -1, 0, 7], [2, 0, 0, 2])
    vals, t = _value(f, col("a") / col("b"))
    assert t == "float64" and vals[0] == 0.5 and vals[1] == -math.inf and math.isnan(vals[2]) and vals[3] == 3.5
    vals, t = _value(f, col("a") + 0.5)
    assert t == "float64" and vals[0] == 1.5


def test_nulls_propagate_and_mask_overflow():
    f = _frame([I32_MAX, 5], [1, 2], valid_a=np.array([False, True]))
    vals, _ = _value(f, col("a") * 2 + col("b"))
    assert vals == [None, 12]


def test_case_null_condition_takes_else_and_only_the_taken_branch_faults():
    f = _frame([1, 2, 3], [I64_MAX, 10, 20], valid_a=np.array([True, False, True]))
    # a > 1: row 0 false, row 1 null, row 2 true
    vals, t = _value(f, when(col("a") > 1).then(col("b") * 2).otherwise(col("b") - 1))
    assert vals == [I64_MAX - 1, 9, 40]
    # the untaken branch of row 0 would overflow: no fault
    vals, _ = _value(f, when(col("a") > 1).then(col("b") * 2).otherwise(lit(None)))
    assert vals == [None, None, 40]
    with pytest.raises(ValueFault):
        _value(f, when(col("a") < 2).then(col("b") * 2).otherwise(0))
    # a chain: when/when/otherwise
    vals, _ = _value(f, when(col("a") > 2).then(100).when(col("a") > 0).then(10).otherwise(-1))
    assert vals == [10, -1, 100]


def test_filtered_out_rows_never_fault():
    f = _frame([0, 0], [I64_MAX, 3])
    vals, _ = _value(f, col("b") * 2, keep=np.array([False, True]))
    assert vals == [None, 6]


def test_product_compiler_types_nodes_as_documented():
    f = _frame([1], [1])
    lf = f.group_by("k").agg((col("a") * 2).sum().alias("s32"), (col("a") + 3_000_000_000).sum().alias("s64"),
                             (col("a") / 2).sum().alias("f"), (col("a") + col("b")).sum().alias("w"))
    plan = compile_plan(lf)
    types = [v.ops[-1][2] for v in plan.values]
    assert types == [_lib.DFR_INT32, _lib.DFR_INT64, _lib.DFR_FLOAT64, _lib.DFR_INT64]
    with pytest.raises(Unsupported):
        compile_plan(f.group_by("k").agg((col("a") + lit("s")).sum()))


# --------------------------------------------------------------- brute force


def _scalar(frame, e, i):
    """Row i's value by direct recursion with exact integers: (value, valid,
    type); raises ValueFault. Independent of the vectorised oracle."""
    k = e.kind
    if k == "col":
        c = frame.column(e.args[0]).column
        v = c.values[i].item()
        return v, bool(c.is_valid()[i]), c.dtype
    if k == "lit":
        v = e.args[0]
        return v, v is not None, None
    if k == "arith":
        op, a, b = e.args
        av, ao, at = _scalar(frame, a, i)
        bv, bo, bt = _scalar(frame, b, i)
        t = oracle._pair_types(frame, a, b)[2] if op != "div" else "float64"
        if not (ao and bo):
            return None, False, t
        if t == "float64":
            x, y = float(av), float(bv)
            if op == "div":
                return (x / y if y != 0 else (math.copysign(math.inf, x) * math.copysign(1, y) if x != 0 and not math.isnan(x) else math.nan)), True, t
            return {"add": x + y, "sub": x - y, "mul": x * y}[op], True, t
        z = {"add": av + bv, "sub": av - bv, "mul": av * bv}[op]
        lo, hi = (I32_MIN, I32_MAX) if t == "int32" else (I64_MIN, I64_MAX)
        if not lo <= z <= hi:
            raise ValueFault("overflow")
        return z, True, t
    if k == "case":
        cond, then, other = e.args
        cv, cm, _ = oracle.evaluate(frame, cond)
        return _scalar(frame, then if (cm[i] and cv[i]) else other, i)
    raise AssertionError(k)


def _random_expr(rng, depth=0):
    r = rng.random()
    if depth > 2 or r < 0.3:
        choice = rng.integers(0, 4)
        return [col("a"), col("b"), col("x"), lit(int(rng.integers(-5, 6)))][choice]
    if r < 0.8:
        op = ["add", "sub", "mul"][rng.integers(0, 3)]
        from dataframe_mojo_amd.expr import Expr
        return Expr("arith", (op, _random_expr(rng, depth + 1), _random_expr(rng, depth + 1)))
    return when(col("b") > int(rng.integers(-3, 4))).then(_random_expr(rng, depth + 1)).otherwise(
        _random_expr(rng, depth + 1))


def test_vectorised_oracle_equals_the_scalar_evaluator_on_random_values():
    rng = np.random.default_rng(11)
    n = 300
    big = rng.random(n) < 0.1
    f = _frame(np.where(big, rng.integers(-I32_MAX, I32_MAX, n), rng.integers(-9, 10, n)).astype(np.int32),
               np.where(big, rng.integers(I64_MIN // 2, I64_MAX // 2, n), rng.integers(-9, 10, n)),
               rng.normal(size=n), valid_a=rng.random(n) > 0.2)
    for _ in range(60):
        e = _random_expr(rng)
        try:
            vec = _value(f, e)[0]
            vec_fault = None
        except ValueFault as exc:
            vec, vec_fault = None, exc
        scal, fault = [], None
        for i in range(n):
            try:
                v, ok, _ = _scalar(f, e, i)
                scal.append(v if ok else None)
            except ValueFault as exc:
                fault = exc
                break
        assert (vec_fault is None) == (fault is None), e
        if fault is None:
            for a, b in zip(vec, scal):
                assert (a is None) == (b is None), e
                if a is not None:
                    assert a == b or (isinstance(a, float) and math.isclose(a, b, rel_tol=0, abs_tol=0)), (e, a, b)


# --------------------------------------------------------------- device


def _gpu_frame(rng, n):
    return DataFrame([
        Series("k", Column.int64(rng.integers(0, 40, n), rng.random(n) > 0.02)),
        Series("a", Column.int32(rng.integers(-1000, 1000, n), rng.random(n) > 0.1)),
        Series("b", Column.int64(rng.integers(-10**6, 10**6, n))),
        Series("x", Column.float64(rng.normal(size=n), rng.random(n) > 0.05)),
    ])


def _close(got, exp, n_keys=0):
    """Equal frames; with n_keys, rows compared in key order (an unordered
    group-by's row order is unspecified)."""
    g, e = got.to_dict(), exp.to_dict()
    assert list(g) == list(e)
    if n_keys:
        def order(d):
            rows = sorted(zip(*d.values()), key=lambda r: tuple((v is None, v or 0) for v in r[:n_keys]))
            return {name: [r[i] for r in rows] for i, name in enumerate(d)}
        g, e = order(g), order(e)
    for name in e:
        assert len(g[name]) == len(e[name]), name
        for a, b in zip(g[name], e[name]):
            if isinstance(b, float) and a is not None:
                assert (math.isnan(a) and math.isnan(b)) or abs(a - b) <= 1e-12 * max(1.0, abs(b)), (name, a, b)
            else:
                assert a == b, (name, a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_value_aggregates_match_the_oracle(seed):
    rng = np.random.default_rng(seed)
    f = _gpu_frame(rng, 20000)
    lf = (f.filter(col("b") > -500_000).group_by("k", maintain_order=True).agg(
        len_(),
        (col("a") * 2 + 1).sum().alias("s32"),
        (col("a") * col("b") - 7).sum().alias("s64"),
        (col("x") * col("a") + 0.5).mean().alias("fx"),
        (col("b") / 3).max().alias("dv"),
        when(col("a") > 0).then(col("b") * 2).otherwise(-col("b")).min().alias("case1"),
        when(col("x").is_null()).then(lit(None)).when(col("a") < -500).then(col("a").abs()).otherwise(0)
        .sum().alias("case2"),
        when(col("a") > 0).then(col("x") * 2.0).otherwise(col("b")).var().alias("cvar"),
    ))
    got = lf.collect(engine="accel")
    exp = oracle.collect(lf)[0]
    _close(got, exp)


@pytest.mark.gpu
def test_value_overflow_is_a_data_fault_with_the_reference_text():
    n = 5000
    rng = np.random.default_rng(5)
    b = rng.integers(-10, 10, n)
    b[4321] = I64_MAX // 2 + 1
    f = DataFrame([Series("k", Column.int64(np.arange(n) % 7)), Series("b", Column.int64(b)),
                   Series("a", Column.int32(rng.integers(0, 5, n)))])
    lf = f.group_by("k").agg((col("b") * 2).sum().alias("s"))
    with pytest.raises(ValueFault) as ref:
        oracle.collect(lf)
    with pytest.raises(_lib.DataFault) as got:
        lf.collect(engine="accel")
    assert str(got.value).endswith(str(ref.value)), (got.value, ref.value)
    # the same row filtered out, or in an untaken branch: no fault
    lf2 = f.filter(col("b") < 100).group_by("k").agg((col("b") * 2).sum().alias("s"))
    _close(lf2.collect(engine="accel"), oracle.collect(lf2)[0], n_keys=1)
    lf3 = f.group_by("k", maintain_order=True).agg(
        when(col("b") < 100).then(col("b") * 2).otherwise(0).sum().alias("s"))
    _close(lf3.collect(engine="accel"), oracle.collect(lf3)[0])


@pytest.mark.gpu
def test_malformed_value_program_is_invalid_before_device_work():
    from dataframe_mojo_amd import runtime

    ctx = runtime.context(0)
    colm = ctx.upload(Column.int64(np.arange(10)))
    vp = _lib.ValueProgram()
    vp.n_ops = 1
    vp.ops[0].op = _lib.VAL_ADD  # pops an empty stack
    vp.ops[0].dtype = _lib.DFR_INT64
    import ctypes

    arr = (ctypes.c_void_p * 1)(colm.handle)
    h, e = ctypes.c_void_p(), ctypes.c_void_p()
    st = ctx.lib.dfr_column_compute(ctx.handle, ctypes.byref(vp), arr, 1, ctypes.byref(h), ctypes.byref(e))
    assert st == _lib.DFR_INVALID
    ctx.lib.dfr_free(e)
    colm.release()
