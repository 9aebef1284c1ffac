"""Float64 sums and variances whose column holds values of very different
magnitudes: no group may lose precision to another group's large values,
to a value in a row the filter drops, or to a value outside a slice.

Each column's fixed-point accumulators are chosen from its value range
(128 to 256 bits; see the query compiler): sums are exact whenever some
width up to 256 bits is, variances are within 2^-84 absolute, and a query
that no width serves raises Unsupported before it runs. The checks here are
against exact rational arithmetic (the two-pass model of
tests/public_contracts/test_numeric_variance.mojo:14-88, taken exactly) at
a relative 1e-12, and against the oracle with the reference's assert_close
(tests/support/aggregate_table_reference.mojo:292-297).
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from dataframe_mojo_amd import Column, DataFrame, Series, _lib, col
from dataframe_mojo_amd.plan import compile_plan
from dataframe_mojo_amd.runtime import result_frame
from oracle import oracle

from test_gpu_parity import TUNE_DEFAULTS, assert_same, both

pytestmark = pytest.mark.gpu


def exact_groups(keys, values, valid=None):
    """key -> list of exact (Fraction) values."""
    out: dict = {}
    for i, (k, v) in enumerate(zip(keys, values)):
        if valid is not None and not valid[i]:
            continue
        out.setdefault(int(k), []).append(Fraction(float(v)))
    return out


def exact_var(vals, ddof):
    n = len(vals)
    if n - ddof <= 0:
        return None
    m = sum(vals) / n
    return sum((v - m) ** 2 for v in vals) / (n - ddof)


def rel_ok(got, exp: Fraction, tol=1e-12):
    if exp == 0:
        return got == 0
    return abs(Fraction(got) - exp) <= Fraction(tol) * abs(exp)


def by_key(frame, key="k"):
    d = frame.to_dict()
    return {k: {c: d[c][i] for c in d} for i, k in enumerate(d[key])}


def test_huge_value_does_not_erase_another_groups_sum():
    # 1e45 in group 0; {1, 2, 3} in group 1: the sum needs 256 bits and is
    # exact; the variance of such a column fits no width and is refused
    # before it runs
    df = DataFrame([Series("k", Column.int64([0, 1, 1, 1, 0])),
                    Series("x", Column.float64([1e45, 1.0, 2.0, 3.0, -2.5e44]))])
    got = both(df.lazy().group_by("k", maintain_order=True).agg(
        col("x").sum().alias("s"), col("x").mean().alias("m"), col("x").count().alias("n")))
    d = got.to_dict()
    assert d["s"] == [1e45 - 2.5e44, 6.0] and d["m"] == [(1e45 - 2.5e44) / 2, 2.0]
    with pytest.raises(_lib.Unsupported, match="variance"):
        df.lazy().group_by("k").agg(col("x").var()).collect()


def test_wide_magnitudes_keep_small_groups_exact(gpu_ctx):
    # group 0 at ~1e9, group 1 at ~1e-3 (variance ~1e-6), group 2 at ~1e-9
    rng = np.random.default_rng(3)
    n = 3000
    k = rng.integers(0, 3, n)
    x = rng.normal(size=n) * np.array([1e9, 1e-3, 1e-9])[k]
    valid = rng.random(n) > 0.02
    df = DataFrame([Series("k", Column.int64(k)), Series("x", Column.float64(x, valid))])
    lf = df.lazy().group_by("k", maintain_order=True).agg(
        col("x").sum().alias("s"), col("x").mean().alias("m"), col("x").var().alias("v"),
        col("x").std().alias("sd"), col("x").var(ddof=0).alias("v0"), col("x").std(ddof=0).alias("sd0"))
    got = both(lf)
    exact = exact_groups(k, x, valid)
    for key, row in by_key(got).items():
        vals = exact[key]
        # sums and means: exact, rounded once
        assert row["s"] == float(sum(vals)), (key, row["s"])
        assert rel_ok(row["m"], sum(vals) / len(vals)), key
        for name, ddof in (("v", 1), ("v0", 0)):
            assert rel_ok(row[name], exact_var(vals, ddof)), (key, name, row[name])
        for name, ddof in (("sd", 1), ("sd0", 0)):
            assert abs(row[name] - math.sqrt(exact_var(vals, ddof))) <= 1e-12 * row[name], (key, name)


def test_int64_near_2_62_beside_small_ints():
    # the variance of an integer column goes through doubles; values near
    # +-2^62 need a wider sum of squares, and must not zero the small group
    big = (1 << 62) - np.arange(1000, dtype=np.int64) * 977
    small = np.arange(1, 101, dtype=np.int64)
    neg = -(1 << 62) + np.arange(50, dtype=np.int64) * 1000003
    v = np.concatenate([big, small, neg])
    k = np.concatenate([np.zeros(1000, np.int64), np.ones(100, np.int64), np.full(50, 2, np.int64)])
    df = DataFrame([Series("k", Column.int64(k)), Series("v", Column.int64(v))])
    # (no sum: group 0's total leaves Int64, a data fault)
    lf = df.lazy().group_by("k", maintain_order=True).agg(
        col("v").var().alias("var"), col("v").std().alias("sd"), col("v").mean().alias("m"))
    got = lf.collect()
    rows = by_key(got)
    # the reference takes each integer to Float64 first (its Welford state)
    exact = exact_groups(k, v.astype(np.float64))
    for key in (0, 1, 2):
        assert rel_ok(rows[key]["var"], exact_var(exact[key], 1)), (key, rows[key]["var"])
    assert rows[1]["var"] == pytest.approx(841.6666666666666, rel=1e-15)
    # the small group also against the oracle (whose Welford is accurate
    # there); the means come from exact integer totals
    exp, _ = oracle.collect(lf)
    erows = by_key(exp)
    assert abs(rows[1]["var"] - erows[1]["var"]) <= 1e-12 * erows[1]["var"]
    for key in (0, 1, 2):
        total = sum(int(t) for t in v[k == key])
        assert rows[key]["m"] == float(Fraction(total) / int((k == key).sum())) or rel_ok(
            rows[key]["m"], Fraction(total, int((k == key).sum())), 1e-15)


def test_outlier_in_a_filtered_out_row():
    # a 1e300 sentinel in rows the filter drops: the kept rows' range sets
    # the accumulators (otherwise every sum would be truncated to zero)
    rng = np.random.default_rng(5)
    n = 20000
    a = rng.integers(-100, 100, n)
    x = rng.normal(size=n)
    x[a == -100] = 1e300
    x[a == -99] = -1e300
    df = DataFrame([Series("k", Column.int64(rng.integers(0, 40, n))), Series("a", Column.int64(a)),
                    Series("x", Column.float64(x))])
    lf = df.filter(col("a") > -99).group_by("k", maintain_order=True).agg(
        col("x").sum().alias("s"), col("x").mean().alias("m"), col("x").var().alias("v"), col("x").std().alias("sd"))
    got = both(lf)
    exact = exact_groups(df.column("k").column.values[a > -99], x[a > -99])
    for key, row in by_key(got).items():
        assert row["s"] == float(sum(exact[key]))
        assert rel_ok(row["v"], exact_var(exact[key], 1))


def test_outlier_in_a_slices_parent(gpu_ctx):
    # the slice's statistics are its parent's: its own window is measured
    # before the accumulators are chosen
    rng = np.random.default_rng(6)
    n = 50000
    k = rng.integers(0, 64, n)
    x = rng.normal(size=n) * 1e-3
    x[:10] = 1e300
    x[-3:] = -1e300
    parent = {"k": Column.int64(k), "x": Column.float64(x)}
    dev = {name: gpu_ctx.upload(c) for name, c in parent.items()}
    off, length = 10, n - 13
    host = DataFrame([Series("k", Column.int64(k[off:off + length])), Series("x", Column.float64(x[off:off + length]))])
    lf = host.lazy().group_by("k", maintain_order=True).agg(
        col("x").sum().alias("s"), col("x").var().alias("v"), col("x").mean().alias("m"))
    plan = compile_plan(lf)
    sl = {name: dev[name].slice(off, length) for name in dev}
    res = gpu_ctx.groupby(plan, [sl[c] for c in plan.keys], [sl[c] for c in plan.columns])
    got = result_frame(res, plan, [sl[c] for c in plan.keys])
    res.release()
    exp, _ = oracle.collect(lf)
    assert_same(got, exp, True)
    exact = exact_groups(k[off:off + length], x[off:off + length])
    for key, row in by_key(got).items():
        assert row["s"] == float(sum(exact[key]))
        assert rel_ok(row["v"], exact_var(exact[key], 1))


WIDE_SETTINGS = [
    {}, {"jit": 2}, {"agg_mode": 0}, {"lds_mode": 0}, {"two_level": 1}, {"two_level": 1, "jit": 2},
    {"direct_max_slots": 0, "lds_mode": 0, "agg_mode": 0},
]


@pytest.mark.parametrize("setting", WIDE_SETTINGS)
def test_wide_accumulators_through_every_path(setting, gpu_ctx):
    # values from 1e-20 to 1e20 in every group: 256-bit sums and sums of
    # squares through pass A's LDS table, spilled records, pass C, the
    # second level, the one-pass atomic kernel and HBM merges
    rng = np.random.default_rng(8)
    n = 300_000
    k = rng.integers(0, 3000, n)
    x = rng.normal(size=n) * 10.0 ** rng.integers(-20, 21, n)
    valid = rng.random(n) > 0.05
    df = DataFrame([Series("k", Column.int64(k)), Series("a", Column.int64(rng.integers(-10, 10, n))),
                    Series("x", Column.float64(x, valid))])
    lf = df.filter(col("a") != 0).group_by("k", maintain_order=True).agg(
        col("x").sum().alias("s"), col("x").mean().alias("m"), col("x").var().alias("v"), col("x").count().alias("n"))
    try:
        for key, value in setting.items():
            gpu_ctx.tune(key, value)
        got = lf.collect()
    finally:
        for key, value in TUNE_DEFAULTS.items():
            gpu_ctx.tune(key, value)
    exp, _ = oracle.collect(lf)
    assert_same(got, exp, True)
    keep = df.column("a").column.values != 0
    exact = exact_groups(k[keep], x[keep], valid[keep])
    rows = by_key(got)
    for key in list(rows)[:200]:
        assert rows[key]["s"] == float(sum(exact[key])), key
        assert rel_ok(rows[key]["v"], exact_var(exact[key], 1)), key


def test_batch_and_wide_accumulators_agree(gpu_ctx):
    # the batch kernel folds the same wide states: bitwise equal to one query
    from dataframe_mojo_amd.runtime import PreparedBatch

    rng = np.random.default_rng(9)
    jobs, frames = [], []
    for q in range(6):
        n = 20000 + 1000 * q
        k = rng.integers(0, 50 + 100 * q, n)
        x = rng.normal(size=n) * 10.0 ** rng.integers(-15, 16, n)
        df = DataFrame([Series("k", Column.int64(k)), Series("x", Column.float64(x))])
        lf = df.lazy().group_by("k").agg(col("x").sum().alias("s"), col("x").var().alias("v"))
        plan = compile_plan(lf)
        dev = {"k": gpu_ctx.upload(df.column("k").column), "x": gpu_ctx.upload(df.column("x").column)}
        jobs.append((plan, [dev[c] for c in plan.keys], [dev[c] for c in plan.columns]))
        frames.append(lf)
    results = PreparedBatch(jobs).run(gpu_ctx)
    for (plan, keys, cols), lf, res in zip(jobs, frames, results):
        assert not isinstance(res, Exception), res
        got = result_frame(res, plan, keys)
        one = gpu_ctx.groupby(plan, keys, cols)
        alone = result_frame(one, plan, keys)
        a, b = by_key(got), by_key(alone)
        assert a.keys() == b.keys()
        for key in a:
            for name in ("s", "v"):
                assert np.float64(a[key][name]).tobytes() == np.float64(b[key][name]).tobytes()
        exp, _ = oracle.collect(lf)
        assert_same(got, exp, False)
        res.release()
        one.release()


# --- per-group error against a compensated (double-double) reference


def _dd_errors(keys_kept, x_kept, valid_kept, got_rows, exp_rows, sum_name, var_name):
    """Per group: |result - dd| / max(1, |dd|) of the sum and the variance,
    for the device result and for the oracle's; returns the maxima."""
    uniq, gid = np.unique(keys_kept, return_inverse=True)
    hi, lo, cnt, var = oracle.dd_moments(gid, x_kept, valid_kept, len(uniq), ddof=1)
    err = {"gpu_sum": 0.0, "gpu_var": 0.0, "cpu_sum": 0.0, "cpu_var": 0.0}
    for g, key in enumerate(uniq.tolist()):
        dd_sum = hi[g] + lo[g]
        for who, rows in (("gpu", got_rows), ("cpu", exp_rows)):
            row = rows[key]
            err[who + "_sum"] = max(err[who + "_sum"], abs(row[sum_name] - dd_sum) / max(1.0, abs(dd_sum)))
            if cnt[g] > 1:
                err[who + "_var"] = max(err[who + "_var"], abs(row[var_name] - var[g]) / max(1.0, abs(var[g])))
    return err


def test_double_double_error_report_zipf(gpu_ctx):
    """The 200M-row Zipf workload's columns (a 4M-row prefix): each group's
    Float64 sum and variance from the device and from the oracle, against
    double-double sums and two-pass variances; both within the budget, and
    the device (exact sums) no worse than the oracle's sequential sums."""
    from dataframe_mojo_amd import datagen

    n = 4_000_000
    cols = datagen.zipf_200m_device(n, seed=0)
    key = cols["key"].numpy(n)
    a = cols["a"].numpy(n)
    x = cols["x"].numpy(n)
    xv = datagen.valid_from_words(cols["x.valid"], n)
    host = DataFrame([Series("key", Column.int32(key)), Series("a", Column.int64(a)),
                      Series("x", Column.float64(x, xv))])
    lf = host.filter(col("a") > 0).group_by("key").agg(col("x").sum().alias("s"), col("x").var().alias("v"))
    plan = compile_plan(lf)
    dev = {"key": gpu_ctx.wrap_array(cols["key"], length=n), "a": gpu_ctx.wrap_array(cols["a"], length=n),
           "x": gpu_ctx.wrap_array(cols["x"], cols["x.valid"], length=n)}
    res = gpu_ctx.groupby(plan, [dev[c] for c in plan.keys], [dev[c] for c in plan.columns])
    got = result_frame(res, plan, [dev[c] for c in plan.keys])
    res.release()
    exp, _ = oracle.collect(lf)
    keep = a > 0
    err = _dd_errors(key[keep], x[keep], xv[keep], by_key(got, "key"), by_key(exp, "key"), "s", "v")
    print("double-double error, zipf prefix:", err)
    assert max(err.values()) <= 1e-12
    assert err["gpu_sum"] <= err["cpu_sum"] + 1e-16


def test_double_double_error_report_dict(gpu_ctx):
    """The same report on the 500M-row dictionary workload's columns (a
    2M-row prefix; lognormal values)."""
    from dataframe_mojo_amd import datagen

    n = 2_000_000
    cols = datagen.dict_500m_device(n, seed=0)
    k1 = cols["k1"].numpy(n)
    k2 = cols["k2"].numpy(n)
    v = cols["v"].numpy(n)
    vv = datagen.valid_from_words(cols["v.valid"], n)
    labels = datagen.dict_labels()
    host = DataFrame([Series("k1", Column.dictionary(k1, labels)), Series("k2", Column.int64(k2)),
                      Series("v", Column.float64(v, vv))])
    lf = host.filter(col("k2") % 3 != 0).group_by("k1", "k2").agg(col("v").sum().alias("s"),
                                                                   col("v").var().alias("var"))
    plan = compile_plan(lf)
    dev = {"k1": gpu_ctx.wrap_array(cols["k1"], labels=labels, dict_codes=True, length=n),
           "k2": gpu_ctx.wrap_array(cols["k2"], length=n), "v": gpu_ctx.wrap_array(cols["v"], cols["v.valid"], length=n)}
    res = gpu_ctx.groupby(plan, [dev[c] for c in plan.keys], [dev[c] for c in plan.columns])
    got = result_frame(res, plan, [dev[c] for c in plan.keys])
    res.release()
    exp, _ = oracle.collect(lf)
    keep = k2 % 3 != 0
    comb = k1.astype(np.int64) * 10_000_000 + k2

    def rows_by_pair(frame):
        d = frame.to_dict()
        lab = {s: i for i, s in enumerate(labels)}
        return {lab[a] * 10_000_000 + b: {"s": d["s"][i], "var": d["var"][i]} for i, (a, b) in enumerate(zip(d["k1"], d["k2"]))}

    err = _dd_errors(comb[keep], v[keep], vv[keep], rows_by_pair(got), rows_by_pair(exp), "s", "var")
    print("double-double error, dict prefix:", err)
    assert max(err.values()) <= 1e-12
