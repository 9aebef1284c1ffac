"""Pins the CPU oracle (oracle/dfr_oracle.c + oracle/oracle.py) against the
reference's known answers, closed forms, brute force and pyarrow, so that
the GPU parity tests check against something trustworthy.

Known answers used:
  README.md sales example (east 100, west 200)
  tests/execution_modes/test_metal_grouped_reductions.mojo:81-157 (a wide
    sum past Int64 then cancelled, a null group)
  tests/kernel_equivalence/test_state_merge.mojo:252 (overflow deferred to
    the final total), :605-639 (zero ties, NaN)
  tests/support/aggregate_table_reference.mojo:27-90 (closed-form
    generators: group_of, narrow, wide, real and their validity)
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

from dataframe_mojo_amd import Column, DataFrame, Series, col, datagen, len_
from oracle import oracle

INT64_MAX = (1 << 63) - 1

HERE = os.path.dirname(os.path.abspath(__file__))


def run(lf):
    return oracle.collect(lf)[0].to_dict()


def test_readme_sales_example():
    df = DataFrame([
        Series("region", Column.string(["east", "west", "east", "west"])),
        Series("amount", Column.int64([100, 80, -20, 120])),
    ])
    out = run(df.filter(col("amount") > 0).group_by("region", maintain_order=True).agg(col("amount").sum().alias("gross_sales")))
    assert out == {"region": ["east", "west"], "gross_sales": [100, 200]}


def test_wide_sum_cancels_and_null_group():
    # the literal finish of test_metal_grouped_reductions.mojo:81-157, as rows
    df = DataFrame([
        Series("k", Column.int64([0, 9, 0, 0, 9], [False, True, False, False, True])),
        Series("v", Column.int64([INT64_MAX, 3, 7, -INT64_MAX, 3])),
    ])
    out = run(df.lazy().group_by("k").agg(
        col("v").sum().alias("s"), col("v").count().alias("n"), col("v").mean().alias("avg"),
        col("v").min().alias("lo"), col("v").max().alias("hi")))
    assert out["k"] == [None, 9]
    assert out["s"] == [7, 6]
    assert out["n"] == [3, 2]
    assert out["avg"][0] == 7 / 3
    assert out["lo"][0] == -INT64_MAX and out["hi"][0] == INT64_MAX


def test_int64_overflow_only_on_final_total():
    ok = DataFrame([Series("k", Column.int64([1, 1, 1])), Series("v", Column.int64([INT64_MAX, 1, -1]))])
    assert run(ok.lazy().group_by("k").agg(col("v").sum()))["v"] == [INT64_MAX]
    bad = DataFrame([Series("k", Column.int64([1, 1])), Series("v", Column.int64([INT64_MAX, 1]))])
    with pytest.raises(oracle.OracleError, match="Int64 expression sum overflow"):
        run(bad.lazy().group_by("k").agg(col("v").sum()))
    # an invalid group (below min_count) never raises
    assert run(bad.lazy().group_by("k").agg(col("v").sum(min_count=3)))["v"] == [None]


def test_int32_sum_stays_int32_and_raises():
    df = DataFrame([Series("k", Column.int64([1, 1])), Series("v", Column.int32([2**31 - 1, 1]))])
    with pytest.raises(oracle.OracleError, match="int32 expression sum overflow"):
        run(df.lazy().group_by("k").agg(col("v").sum()))
    df2 = DataFrame([Series("k", Column.int64([1, 1])), Series("v", Column.int32([2**31 - 1, -5]))])
    out, _ = oracle.collect(df2.lazy().group_by("k").agg(col("v").sum()))
    assert out.column("v").dtype == "int32" and out.to_dict()["v"] == [2**31 - 6]


def test_zero_ties_keep_first_and_nan_rules():
    # test_state_merge.mojo:605-639: ties keep the first value, -0.0 == 0.0
    df = DataFrame([
        Series("k", Column.int64([1, 1, 2, 2, 3, 3, 4])),
        Series("x", Column.float64([0.0, -0.0, -0.0, 0.0, math.nan, 1.5, math.nan])),
    ])
    out = run(df.lazy().group_by("k", maintain_order=True).agg(col("x").min().alias("mn"), col("x").max().alias("mx")))
    assert math.copysign(1, out["mn"][0]) == 1 and math.copysign(1, out["mx"][0]) == 1
    assert math.copysign(1, out["mn"][1]) == -1 and math.copysign(1, out["mx"][1]) == -1
    assert out["mn"][2] == 1.5 and out["mx"][2] == 1.5  # NaN skipped
    assert math.isnan(out["mn"][3]) and math.isnan(out["mx"][3])  # only NaN


def test_empty_group_sum_zero_and_null_rules():
    df = DataFrame([
        Series("k", Column.int64([1, 1, 2])),
        Series("x", Column.float64([1.0, 2.0, 5.0], [False, False, True])),
    ])
    out = run(df.lazy().group_by("k", maintain_order=True).agg(
        col("x").sum().alias("s"), col("x").mean().alias("m"), col("x").var().alias("v"),
        col("x").min().alias("mn"), col("x").count().alias("c"), len_()))
    assert out["s"] == [0.0, 5.0]  # empty sum is 0, valid
    assert out["m"] == [None, 5.0]
    assert out["v"] == [None, None]  # n - ddof <= 0
    assert out["mn"] == [None, 5.0]
    assert out["c"] == [0, 1] and out["len"] == [2, 1]


# --- closed-form generators of tests/support/aggregate_table_reference.mojo


def group_of(row, groups):
    return (row * 7919) % groups


def narrow(row):
    return (row * 104729) % 1999 - 1000


def narrow_valid(row):
    return row % 17 != 3


def wide(row):
    return (row % 977) * 1_000_000_007 - 400_000_000_000


def wide_valid(row):
    return row % 13 != 4


def real(row):
    return (row % 211) / 8 - 13


def real_valid(row):
    return row % 19 != 9


def int_key(g):
    return g * 1_000_003 - 5_000_000


def int_key_valid(g):
    return not (g % 31 == 7 and g % 29 != 5)


@pytest.mark.parametrize("rows,groups", [(1, 1), (9, 3), (65, 7), (1000, 64), (4099, 33)])
def test_aggregate_table_closed_forms(rows, groups):
    r = range(rows)
    gs = [group_of(i, groups) for i in r]
    df = DataFrame([
        Series("i", Column.int64([int_key(g) for g in gs], [int_key_valid(g) for g in gs])),
        Series("x", Column.int32([narrow(i) for i in r], [narrow_valid(i) for i in r])),
        Series("w", Column.int64([wide(i) for i in r], [wide_valid(i) for i in r])),
        Series("y", Column.float64([real(i) for i in r], [real_valid(i) for i in r])),
    ])
    lf = df.lazy().group_by("i", maintain_order=True).agg(
        col("x").sum().alias("xs"), col("w").sum().alias("ws"), col("w").min().alias("wmin"),
        col("w").max().alias("wmax"), col("x").count().alias("xc"), col("y").sum().alias("ys"),
        col("y").mean().alias("ym"), col("y").var().alias("yv"), col("w").mean().alias("wm"), len_())
    got = run(lf)
    # brute force in Python integers and Fractions
    model = {}
    order = []
    for i in r:
        g = gs[i]
        key = int_key(g) if int_key_valid(g) else None
        if key not in model:
            model[key] = {"xs": 0, "ws": 0, "w": [], "xc": 0, "y": [], "len": 0}
            order.append(key)
        m = model[key]
        m["len"] += 1
        if narrow_valid(i):
            m["xs"] += narrow(i)
            m["xc"] += 1
        if wide_valid(i):
            m["ws"] += wide(i)
            m["w"].append(wide(i))
        if real_valid(i):
            m["y"].append(real(i))
    assert got["i"] == order
    for j, key in enumerate(order):
        m = model[key]
        assert got["xs"][j] == m["xs"] and got["ws"][j] == m["ws"] and got["xc"][j] == m["xc"]
        assert got["len"][j] == m["len"]
        assert got["wmin"][j] == (min(m["w"]) if m["w"] else None)
        assert got["wmax"][j] == (max(m["w"]) if m["w"] else None)
        if m["w"]:
            assert got["wm"][j] == float(Fraction(m["ws"]) / len(m["w"])) or abs(got["wm"][j] - m["ws"] / len(m["w"])) <= 1e-12 * abs(m["ws"] / len(m["w"]))
        ys = m["y"]
        exact_sum = math.fsum(ys)
        assert abs(got["ys"][j] - exact_sum) <= 1e-12 * max(1.0, abs(exact_sum))
        if ys:
            assert abs(got["ym"][j] - exact_sum / len(ys)) <= 1e-12 * max(1.0, abs(exact_sum / len(ys)))
        if len(ys) > 1:
            fr = [Fraction(v) for v in ys]
            mu = sum(fr) / len(fr)
            var = float(sum((v - mu) ** 2 for v in fr) / (len(fr) - 1))
            assert abs(got["yv"][j] - var) <= 1e-9 * max(1.0, var)  # test_numeric_variance.mojo:10-12
        else:
            assert got["yv"][j] is None


def test_pyarrow_independent_integer_aggregates():
    pa = pytest.importorskip("pyarrow")
    rng = np.random.default_rng(7)
    n = 20000
    k = rng.integers(0, 300, n)
    v = rng.integers(-10**12, 10**12, n)
    vm = rng.random(n) > 0.1
    df = DataFrame([Series("k", Column.int64(k)), Series("v", Column.int64(v, vm))])
    got = oracle.collect(df.lazy().group_by("k").agg(
        col("v").sum().alias("s"), col("v").count().alias("c"), col("v").min().alias("mn"), col("v").max().alias("mx")))[0].to_dict()
    t = pa.table({"k": k, "v": pa.array(v, mask=~vm)})
    ref = t.group_by("k", use_threads=False).aggregate([("v", "sum"), ("v", "count"), ("v", "min"), ("v", "max")]).to_pydict()
    exp = {kk: (s, c, mn, mx) for kk, s, c, mn, mx in zip(ref["k"], ref["v_sum"], ref["v_count"], ref["v_min"], ref["v_max"])}
    for kk, s, c, mn, mx in zip(got["k"], got["s"], got["c"], got["mn"], got["mx"]):
        assert exp[kk] == (s, c, mn, mx)


def test_filter_kleene_semantics():
    df = DataFrame([
        Series("a", Column.int64([1, -2, 3, 7, 9], [True, True, False, True, True])),
        Series("b", Column.int64([5, 5, 5, 0, -3])),
    ])
    keep = lambda e: oracle.kept_rows(df, [e]).tolist()  # noqa: E731
    assert keep(col("a") > 0) == [True, False, False, True, True]
    assert keep(~(col("a") > 0)) == [False, True, False, False, False]  # null stays not true
    assert keep((col("a") > 0) | (col("b") == 5)) == [True, True, True, True, True]
    assert keep(~((col("a") > 0) & (col("b") == 5))) == [False, True, False, True, True]
    assert keep(col("a").is_null()) == [False, False, True, False, False]
    # modulo takes the divisor's sign; modulo by zero is null
    assert keep((col("a") % 3) == 1) == [True, True, False, True, False]
    assert keep((col("a") % -4) == -3) == [True, False, False, False, True]
    assert keep((col("a") % col("b")) == 0) == [False, False, False, False, True]


def test_membership_and_range_semantics():
    # docs/expressions.md:445-450: a null candidate never matches, a null
    # input stays null, an empty list gives false; is_between is the Kleene
    # AND of its two comparisons, `closed` picking the inclusive bounds
    df = DataFrame([
        Series("a", Column.int64([1, 2, 3, 4, 5], [True, True, False, True, True])),
        Series("x", Column.float64([0.5, -0.0, 2.0, float("nan"), 3.0])),
        Series("s", Column.string(["p", "q", None, "p", "r"])),
    ])
    keep = lambda e: oracle.kept_rows(df, [e]).tolist()  # noqa: E731
    assert keep(col("a").is_in([2, 4, None])) == [False, True, False, True, False]
    assert keep(~col("a").is_in([2, 4, None])) == [True, False, False, False, True]  # null input stays null
    assert keep(col("a").is_in([])) == [False] * 5
    assert keep(~col("a").is_in([])) == [True, True, False, True, True]
    assert keep(col("a").is_in([2.0, 4.5])) == [False, True, False, False, False]  # by value
    assert keep(col("x").is_in([0.0, float("nan"), 3])) == [False, True, False, False, True]  # IEEE equality
    assert keep(col("s").is_in(["p", "zz"])) == [True, False, False, True, False]
    assert keep(col("a").is_between(2, 4)) == [False, True, False, True, False]
    assert keep(col("a").is_between(2, 4, closed="none")) == [False, False, False, False, False]
    assert keep(col("a").is_between(2, 4, closed="left")) == [False, True, False, False, False]
    assert keep(col("a").is_between(2, 4, closed="right")) == [False, False, False, True, False]
    # a null bound: null AND c is false where c is false, else null
    assert keep(col("a").is_between(None, 3)) == [False] * 5
    assert keep(~col("a").is_between(None, 3)) == [False, False, False, True, True]


def test_membership_matches_brute_force():
    rng = np.random.default_rng(5)
    n = 400
    a = rng.integers(-6, 6, n)
    av = rng.random(n) > 0.2
    df = DataFrame([Series("a", Column.int64(a, av))])
    for cands in ([], [None], [0], [1, -3, 5, None], [2.0, 2.5, -6], list(range(-8, 8))):
        for neg in (False, True):
            e = col("a").is_in(cands)
            got = oracle.kept_rows(df, [~e if neg else e]).tolist()
            want = []
            for i in range(n):
                if not av[i]:
                    want.append(False)  # null stays null: never kept
                    continue
                hit = any(c is not None and float(c) == float(a[i]) for c in cands)
                want.append(hit != neg)
            assert got == want, (cands, neg)
        for lo, hi in ((-2, 3), (3, -2), (0, 0)):
            for closed in ("both", "left", "right", "none"):
                got = oracle.kept_rows(df, [col("a").is_between(lo, hi, closed)]).tolist()
                lo_ok = (lambda v: v >= lo) if closed in ("both", "left") else (lambda v: v > lo)
                hi_ok = (lambda v: v <= hi) if closed in ("both", "right") else (lambda v: v < hi)
                assert got == [bool(av[i] and lo_ok(a[i]) and hi_ok(a[i])) for i in range(n)]


def test_golden_groupby_1m_fixture():
    with open(os.path.join(HERE, "golden", "groupby_1m.json")) as f:
        gold = json.load(f)
    cols = datagen.groupby_1m(gold["rows"], gold["seed"])
    assert cols["key"][:16].tolist() == gold["key_head"]
    assert cols["amount"][:16].tolist() == gold["amount_head"]
    df = DataFrame([Series("key", Column.int64(cols["key"])), Series("amount", Column.int64(cols["amount"]))])
    out, first = oracle.collect(df.filter(col("amount") > 0).group_by("key", maintain_order=True).agg(col("amount").sum(), len_()))
    d = out.to_dict()
    assert d["key"] == gold["keys"] and d["amount"] == gold["sum"] and d["len"] == gold["len"]
    assert first.tolist() == gold["first_rows"]


def test_dd_moments_closed_form():
    # x = 1e8 + k for k in 0..9: mean 1e8 + 4.5, var (ddof 1) = 55/6
    x = np.array([1e8 + k for k in range(10)] + [3.0, 4.0])
    gid = np.array([0] * 10 + [1, 1])
    hi, lo, cnt, var = oracle.dd_moments(gid, x, None, 2)
    assert cnt.tolist() == [10, 2]
    assert hi[0] + lo[0] == 1e9 + 45
    assert abs(var[0] - 55 / 6) < 1e-12 and var[1] == 0.5


def test_float_keys_canonical_groups_keep_first_value():
    """Float64 keys group by canonical words (-0.0 == 0.0, every NaN one
    group) and report the first row's own value (group_by.mojo:319-329)."""
    import math

    df = DataFrame([
        Series("k", Column.float64([-0.0, 0.0, math.nan, 1.5, -math.nan, 0.0, 1.5, 2.0], [True] * 7 + [False])),
        Series("v", Column.int64([1, 2, 3, 4, 5, 6, 7, 8])),
    ])
    out, first = oracle.collect(df.lazy().group_by("k", maintain_order=True).agg(col("v").sum(), len_()))
    k = out.column("k").column
    assert list(first) == [0, 2, 3, 7]
    assert np.signbit(k.values[0]) and k.values[0] == 0.0  # the first row's -0.0
    assert math.isnan(k.values[1]) and k.values[2] == 1.5 and not k.is_valid()[3]
    assert out.to_dict()["v"] == [9, 8, 11, 8] and out.to_dict()["len"] == [3, 2, 2, 1]


@pytest.mark.parametrize("threads", [2, 3, 8])
@pytest.mark.parametrize("keys", [["k"], ["k", "s"], ["f"]])
def test_parallel_restatement_equals_sequential(threads, keys):
    """dfo_groupby_parallel (workers over contiguous row ranges, per-worker
    partitions by key hash, partition-parallel folding and finish) gives
    every group the same state as the sequential dfo_groupby, bit for bit,
    Float64 sums and moments included, because each partition folds its
    rows in ascending row order; in first-row order when ordered."""
    rng = np.random.default_rng(threads * 7 + len(keys))
    n = 60000
    df = DataFrame([
        Series("k", Column.int64(rng.integers(-300, 300, n), rng.random(n) > 0.02)),
        Series("s", Column.int32(rng.integers(0, 7, n))),
        Series("f", Column.float64(rng.choice([0.0, -0.0, 1.5, np.nan, -2.0], n), rng.random(n) > 0.05)),
        Series("v", Column.int64(rng.integers(-10**12, 10**12, n), rng.random(n) > 0.1)),
        Series("x", Column.float64(rng.normal(size=n) * 1e3, rng.random(n) > 0.05)),
    ])
    aggs = [len_(), col("v").sum().alias("v_sum"), col("v").min().alias("v_min"), col("x").sum().alias("x_sum"),
            col("x").mean().alias("x_mean"), col("x").var().alias("x_var"), col("x").max().alias("x_max"),
            col("v").count().alias("v_count")]
    for mo in (True, False):
        lf = df.filter(col("v") > -5 * 10**11).group_by(*keys, maintain_order=mo).agg(aggs)
        seq, fs = oracle.collect(lf)
        par, fp = oracle.collect(lf, threads=threads)
        if not mo:
            assert sorted(fp) == sorted(fs)
            o = np.argsort(fp, kind="stable")
            fp = fp[o]
            par = DataFrame([Series(s.name, Column(s.column.dtype, s.column.values[o], s.column.is_valid()[o],
                                                   s.column.labels)) for s in par._series])
        assert np.array_equal(fs, fp)
        for name in seq.columns:
            a, b = seq.column(name).column, par.column(name).column
            assert np.array_equal(a.is_valid(), b.is_valid()), name
            assert np.array_equal(a.values.view(np.uint8), b.values.view(np.uint8)), name


def test_parallel_restatement_raises_the_same_overflow():
    df = DataFrame([Series("k", Column.int64([1, 1, 2])), Series("v", Column.int64([INT64_MAX, 1, 5]))])
    lf = df.lazy().group_by("k").agg(col("v").sum())
    with pytest.raises(oracle.OracleError, match="Int64 expression sum overflow"):
        oracle.collect(lf, threads=4)


# --- Float64 min/max between distinct values, by brute force


def _brute_extrema(keys, x, valid):
    """Per key in first-row order: (min, max) by the reference's rules
    (dataframe/aggregation/function.mojo:1143-1235 _scatter, reducer.mojo
    MIN/MAX; tests/kernel_equivalence/test_state_merge.mojo:605-639): NaN
    values are skipped unless every valid value is NaN (then NaN), no valid
    value gives null, and a tie keeps the first value in row order (so the
    sign of a zero is the first zero's). Python's min/max keep the first of
    equal elements, which is that rule."""
    groups: dict = {}
    for i, k in enumerate(keys):
        groups.setdefault(int(k), [])
        if valid[i]:
            groups[int(k)].append(float(x[i]))
    out = {}
    for k, vals in groups.items():
        if not vals:
            out[k] = (None, None)
            continue
        real = [v for v in vals if not math.isnan(v)]
        if not real:
            out[k] = (math.nan, math.nan)
            continue
        out[k] = (min(real), max(real))
    return out


@pytest.mark.parametrize("seed", range(6))
def test_float_min_max_brute_force(seed):
    rng = np.random.default_rng(100 + seed)
    n = 4000
    keys = rng.integers(0, 60, n)
    x = rng.normal(size=n) * 10.0 ** rng.integers(-3, 4, n)
    # signed zeros, NaNs, infinities and repeated values, in every group
    pick = rng.random(n)
    x[pick < 0.05] = 0.0
    x[(pick >= 0.05) & (pick < 0.10)] = -0.0
    x[(pick >= 0.10) & (pick < 0.13)] = math.nan
    x[(pick >= 0.13) & (pick < 0.14)] = math.inf
    x[(pick >= 0.14) & (pick < 0.15)] = -math.inf
    x[(pick >= 0.15) & (pick < 0.25)] = 1.25
    x[keys == 7] = math.nan  # a group of NaNs only
    valid = rng.random(n) > 0.1
    valid[keys == 9] = False  # a group with no valid value
    df = DataFrame([Series("k", Column.int64(keys)), Series("x", Column.float64(x, valid))])
    out = run(df.lazy().group_by("k", maintain_order=True).agg(col("x").min().alias("mn"), col("x").max().alias("mx")))
    exp = _brute_extrema(keys, x, valid)
    assert out["k"] == list(exp)  # first-row order
    for i, k in enumerate(out["k"]):
        for got, want in ((out["mn"][i], exp[k][0]), (out["mx"][i], exp[k][1])):
            if want is None:
                assert got is None, k
            elif math.isnan(want):
                assert math.isnan(got), k
            else:
                assert np.float64(got).tobytes() == np.float64(want).tobytes(), (k, got, want)


def test_float_min_max_distinct_values_not_swapped():
    df = DataFrame([Series("k", Column.int64([1, 1, 1, 2, 2])),
                    Series("x", Column.float64([3.5, -2.25, 7.0, -0.0, 1e-300]))])
    out = run(df.lazy().group_by("k", maintain_order=True).agg(col("x").min().alias("mn"), col("x").max().alias("mx")))
    assert out["mn"] == [-2.25, -0.0] and out["mx"] == [7.0, 1e-300]
    assert math.copysign(1, out["mn"][1]) == -1


# --- var/std at ddof 0 and 1 against exact two-pass references


def _variance_frame(n, nulls):
    """The columns of tests/public_contracts/test_numeric_variance.mojo:14-88
    (step = row * 7919 % 2001; Float64 1e6 + step / 8, Int32 100000 + step,
    Int64 step; null every row % 7 == 3 when asked), and the reference's
    expected variance of `step` from its kept values."""
    step = (np.arange(n, dtype=np.int64) * 7919) % 2001
    valid = np.ones(n, dtype=bool) if not nulls else (np.arange(n) % 7 != 3)
    df = DataFrame([
        Series("g", Column.int64(np.zeros(n, dtype=np.int64))),
        Series("d", Column.float64(1e6 + step / 8, valid)),
        Series("i", Column.int32((100_000 + step).astype(np.int32), valid)),
        Series("u", Column.int64(step, valid)),
    ])
    return df, [Fraction(int(s)) for s, ok in zip(step, valid) if ok]


@pytest.mark.parametrize("n", [0, 1, 2, 9, 127, 128, 129, 257, 1000, 70_001])
@pytest.mark.parametrize("nulls", [False, True])
@pytest.mark.parametrize("ddof", [0, 1])
def test_variance_matches_two_pass_reference(n, nulls, ddof):
    df, kept = _variance_frame(n, nulls)
    if n == 0:
        return  # no group at all
    lf = df.lazy().group_by("g").agg(
        [col(c).var(ddof=ddof).alias("v" + c) for c in "diu"] + [col(c).std(ddof=ddof).alias("s" + c) for c in "diu"])
    out = run(lf)
    defined = len(kept) > ddof
    if defined:
        m = sum(kept) / len(kept)
        unit = sum((v - m) ** 2 for v in kept) / (len(kept) - ddof)
    for c, scale in (("d", Fraction(1, 64)), ("i", Fraction(1)), ("u", Fraction(1))):
        v, s = out["v" + c][0], out["s" + c][0]
        if not defined:
            assert v is None and s is None
            continue
        want = unit * scale
        # the reference's own bound (1e-9, its _close) and the exact value's
        assert abs(Fraction(v) - want) <= Fraction(1, 10**9) * max(1, want)
        assert abs(Fraction(v) - want) <= Fraction(1, 10**12) * max(1, want), (c, v, float(want))
        assert abs(s - math.sqrt(float(want))) <= 1e-12 * max(1.0, math.sqrt(float(want))), (c, s)


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("ddof", [0, 1, 3])
def test_random_variance_against_exact(seed, ddof):
    rng = np.random.default_rng(200 + seed)
    n = 3000
    keys = rng.integers(0, 25, n)
    x = rng.normal(size=n) * 10.0 ** rng.integers(-2, 3, n) + rng.choice([0.0, 50.0], n)
    valid = rng.random(n) > 0.05
    df = DataFrame([Series("k", Column.int64(keys)), Series("x", Column.float64(x, valid))])
    out = run(df.lazy().group_by("k", maintain_order=True).agg(
        col("x").var(ddof=ddof).alias("v"), col("x").std(ddof=ddof).alias("s")))
    for i, k in enumerate(out["k"]):
        vals = [Fraction(float(v)) for v, ok, kk in zip(x, valid, keys) if ok and kk == k]
        if len(vals) - ddof <= 0:
            assert out["v"][i] is None and out["s"][i] is None
            continue
        m = sum(vals) / len(vals)
        want = sum((v - m) ** 2 for v in vals) / (len(vals) - ddof)
        assert abs(Fraction(out["v"][i]) - want) <= Fraction(1, 10**12) * max(1, want), (k, out["v"][i], float(want))
        assert abs(out["s"][i] - math.sqrt(float(want))) <= 1e-12 * max(1.0, math.sqrt(float(want)))
