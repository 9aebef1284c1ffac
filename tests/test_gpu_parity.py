"""Parity of the device path (libdfrocm through the provider boundary) with
the CPU oracle on the same seeded inputs.

Bar: keys, group order (under maintain_order), nulls, error texts and every
integer aggregate bit-exact; Float64 sum/mean/var/std within
|a - e| <= 1e-12 * max(1, |e|) (the reference's assert_close,
tests/support/aggregate_table_reference.mojo:292-297); Float64 min/max
bit-exact, including the sign of zero.
"""
import math

import numpy as np
import pytest

from dataframe_mojo_amd import Column, DataFrame, Series, col, datagen, len_, provider, runtime
from dataframe_mojo_amd import _lib
from oracle import oracle

pytestmark = pytest.mark.gpu

INT64_MAX = (1 << 63) - 1
FTOL = 1e-12


def close(a, e, tol=FTOL):
    if a is None or e is None:
        return a is None and e is None
    if isinstance(e, float) or isinstance(a, float):
        if math.isnan(e):
            return math.isnan(a)
        if math.isinf(e):
            return a == e
        return abs(a - e) <= tol * max(1.0, abs(e))
    return a == e


def assert_same(got, exp, ordered, exact_float=(), n_keys=1):
    assert got.columns == exp.columns, (got.columns, exp.columns)
    grows, erows = got.rows(), exp.rows()
    assert len(grows) == len(erows), (len(grows), len(erows))
    if not ordered:
        def sk(r):
            return tuple((v is None, 0 if v is None else v) for v in r[:n_keys])

        grows, erows = sorted(grows, key=sk), sorted(erows, key=sk)
    for gr, er in zip(grows, erows):
        for name, a, b in zip(got.columns, gr, er):
            if name in exact_float and isinstance(b, float):
                assert (np.float64(a).tobytes() == np.float64(b).tobytes()) or (math.isnan(a) and math.isnan(b)), (name, a, b)
            else:
                assert close(a, b), (name, a, b, gr, er)


def both(lf, exact_float=()):
    got = lf.collect(engine="accel")
    exp, _ = oracle.collect(lf)
    assert_same(got, exp, lf.maintain_order, exact_float, len(lf.keys))
    return got


def test_readme_sales_example():
    df = DataFrame([
        Series("region", Column.string(["east", "west", "east", "west"])),
        Series("amount", Column.int64([100, 80, -20, 120])),
    ])
    out = (df.filter(col("amount") > 0).group_by("region", maintain_order=True)
           .agg(col("amount").sum().alias("gross_sales")).collect(engine="accel"))
    assert out.to_dict() == {"region": ["east", "west"], "gross_sales": [100, 200]}


def test_wide_sum_cancels_and_null_group():
    df = DataFrame([
        Series("k", Column.int64([0, 9, 0, 0, 9], [False, True, False, False, True])),
        Series("v", Column.int64([INT64_MAX, 3, 7, -INT64_MAX, 3])),
    ])
    out = both(df.lazy().group_by("k", maintain_order=True).agg(
        col("v").sum().alias("s"), col("v").count().alias("n"), col("v").mean().alias("avg"),
        col("v").min().alias("lo"), col("v").max().alias("hi")))
    d = out.to_dict()
    assert d["k"] == [None, 9] and d["s"] == [7, 6] and d["avg"][0] == 7 / 3


def test_groupby_1m_golden():
    import json
    import os

    with open(os.path.join(os.path.dirname(__file__), "golden", "groupby_1m.json")) as f:
        gold = json.load(f)
    cols = datagen.groupby_1m(gold["rows"], gold["seed"])
    df = DataFrame([Series("key", Column.int64(cols["key"])), Series("amount", Column.int64(cols["amount"]))])
    out = df.filter(col("amount") > 0).group_by("key", maintain_order=True).agg(col("amount").sum(), len_()).collect()
    d = out.to_dict()
    assert d["key"] == gold["keys"] and d["amount"] == gold["sum"] and d["len"] == gold["len"]


def test_int64_overflow_is_a_data_fault():
    bad = DataFrame([Series("k", Column.int64([1, 1])), Series("v", Column.int64([INT64_MAX, 1]))])
    with pytest.raises(_lib.DataFault, match="Int64 expression sum overflow"):
        bad.lazy().group_by("k").agg(col("v").sum()).collect()
    ok = DataFrame([Series("k", Column.int64([1, 1, 1])), Series("v", Column.int64([INT64_MAX, 1, -1]))])
    assert ok.lazy().group_by("k").agg(col("v").sum()).collect().to_dict()["v"] == [INT64_MAX]
    i32 = DataFrame([Series("k", Column.int64([1, 1])), Series("v", Column.int32([2**31 - 1, 1]))])
    with pytest.raises(_lib.DataFault, match="int32 expression sum overflow"):
        i32.lazy().group_by("k").agg(col("v").sum()).collect()


def test_zero_ties_and_nan():
    df = DataFrame([
        Series("k", Column.int64([1, 1, 2, 2, 3, 3, 4, 5])),
        Series("x", Column.float64([0.0, -0.0, -0.0, 0.0, math.nan, 1.5, math.nan, math.inf])),
    ])
    both(df.lazy().group_by("k", maintain_order=True).agg(
        col("x").min().alias("mn"), col("x").max().alias("mx"), col("x").sum().alias("s"),
        col("x").mean().alias("m"), col("x").var().alias("v")), exact_float=("mn", "mx"))


def test_empty_and_all_filtered():
    df = DataFrame([Series("k", Column.int64([1, 2, 3])), Series("v", Column.int64([-1, -2, -3]))])
    out = df.filter(col("v") > 0).group_by("k").agg(col("v").sum()).collect()
    assert out.height() == 0
    empty = DataFrame([Series("k", Column.int64([])), Series("v", Column.int64([]))])
    assert empty.lazy().group_by("k").agg(col("v").sum()).collect().height() == 0


def random_frame(n, seed, key_card=50):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=n)
    x[rng.random(n) < 0.01] = np.nan
    x[rng.random(n) < 0.01] = 0.0
    x[rng.random(n) < 0.005] = -0.0
    labels = [f"lbl{i}" for i in range(17)] + ["lbl3"]
    return DataFrame([
        Series("k", Column.int64(rng.integers(-key_card, key_card, n), rng.random(n) > 0.03)),
        Series("k32", Column.int32(rng.integers(0, 7, n))),
        Series("s", Column.dictionary(rng.integers(0, len(labels), n), labels, rng.random(n) > 0.05)),
        Series("a", Column.int64(rng.integers(-10**6, 10**6, n), rng.random(n) > 0.1)),
        Series("b", Column.int32(rng.integers(-1000, 1000, n))),
        Series("x", Column.float64(x, rng.random(n) > 0.05)),
        Series("w", Column.int64(rng.integers(-(2**50), 2**50, n))),
    ])


ALL_AGGS = [
    len_(), col("a").sum().alias("a_sum"), col("a").count().alias("a_cnt"), col("a").min().alias("a_min"),
    col("a").max().alias("a_max"), col("a").mean().alias("a_mean"), col("a").var().alias("a_var"),
    col("b").sum().alias("b_sum"), col("b").min().alias("b_min"), col("b").std(ddof=0).alias("b_std0"),
    col("x").sum().alias("x_sum"), col("x").mean().alias("x_mean"), col("x").var().alias("x_var"),
    col("x").std().alias("x_std"), col("x").min().alias("x_min"), col("x").max().alias("x_max"),
    col("x").count().alias("x_cnt"), col("w").sum().alias("w_sum"), col("s").count().alias("s_cnt"),
    col("a").sum(min_count=3).alias("a_sum3"),
]

FILTERS = [
    [],
    [col("a") > 0],
    [(col("b") % 3 != 0) | col("x").is_null()],
    [~((col("x") > 0.25) & (col("s") == "lbl3"))],
    [col("a").is_not_null(), col("x") <= 1.0],
    [col("k").is_in([3, 17, None, 40.0]) | col("s").is_in(["lbl3", "lbl5"]), ~col("x").is_between(-1.0, 0.0)],
    [(col("a") > col("b")) | ~(col("k32") <= col("k")), col("x") == col("x")],
]

KEYS = [["k"], ["k32"], ["s"], ["k", "s"], ["k32", "k", "s"]]


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("fi", range(len(FILTERS)))
@pytest.mark.parametrize("ki", range(len(KEYS)))
def test_random_queries(seed, fi, ki):
    df = random_frame(20000 + 977 * seed, seed)
    for mo in (True, False):
        lf = df.lazy()
        for f in FILTERS[fi]:
            lf = lf.filter(f)
        lf = lf.group_by(*KEYS[ki], maintain_order=mo).agg(ALL_AGGS)
        both(lf, exact_float=("x_min", "x_max"))


class tuned:
    """Tune settings for a block; the context's previous values (built-in or
    from a calibration profile) come back afterwards."""

    def __init__(self, ctx, setting):
        self.ctx, self.setting = ctx, setting

    def __enter__(self):
        self.saved = {k: self.ctx.tune_get(k) for k in self.setting}
        for k, v in self.setting.items():
            self.ctx.tune(k, v)

    def __exit__(self, *exc):
        for k, v in self.saved.items():
            self.ctx.tune(k, v)


TUNE_DEFAULTS = {"lds_mode": -1, "lds_bytes": 65536, "direct_max_slots": 1 << 26, "hash_slots": 0, "block": 512,
                 "wg_per_cu": 2, "lds_probe": 4, "scan_unroll": 1, "agg_mode": 1, "two_level": -1,
                 "max_parts": 4096, "part_lds_bytes": 131072, "sub_lds_bytes": 49152, "sub_block": 512, "jit": 1,
                 "scan_prefetch": 1, "lds_odd_stride": 0, "part_odd_stride": 1, "part_block": 1024, "scan_block": 1024,
                 "scan_wg_per_cu": 1, "scan_lds_bytes": 159744, "two_level_parts": 512,
                 "scatter_slices": 1, "scatter_chunk": 16384, "scatter_stage": 1, "jit_wait": 1,
                 "sub_wg_per_cu": -1, "sub_fill": 50, "hot_keys": -1, "hot_rows": 1 << 24, "hot_domain": 1 << 24,
                 "hot_sample": 1 << 19, "hot_share": 30, "rec_pack": 1}


@pytest.mark.parametrize("setting", [
    {"lds_mode": 0}, {"lds_mode": 2}, {"lds_mode": 2, "lds_bytes": 8192}, {"direct_max_slots": 0},
    {"direct_max_slots": 0, "lds_mode": 0}, {"direct_max_slots": 0, "hash_slots": 64},
    {"block": 256, "wg_per_cu": 1, "agg_mode": 0}, {"scan_block": 1024, "scan_wg_per_cu": 4}, {"lds_probe": 1},
    {"scan_unroll": 2}, {"scan_unroll": 4}, {"agg_mode": 0},
    {"two_level": 1}, {"two_level": 1, "lds_mode": 0}, {"two_level": 1, "scan_unroll": 4},
    # one first-level partition, many small second-level tables
    {"two_level": 1, "max_parts": 1, "sub_lds_bytes": 4096, "sub_block": 64},
    {"two_level": 1, "sub_lds_bytes": 98304, "sub_block": 512},
    {"two_level": 1, "two_level_parts": 64}, {"two_level": 1, "two_level_parts": 2, "jit": 2},
    # second-level workgroups: one per sub-partition, or looping over many;
    # sub-partitions nearly full (capacity retries) or sparse
    {"two_level": 1, "sub_wg_per_cu": 0}, {"two_level": 1, "sub_wg_per_cu": 1, "jit": 2},
    {"two_level": 1, "sub_fill": 90, "sub_lds_bytes": 4096}, {"two_level": 1, "sub_fill": 10, "sub_wg_per_cu": 3},
    # pass A compiled at run time for the query's shape
    {"jit": 2}, {"jit": 2, "lds_mode": 0}, {"jit": 2, "two_level": 1}, {"jit": 2, "scan_unroll": 4},
    # its loads pipelined (the default) or not, over ranges of odd and even tile counts
    {"jit": 2, "scan_prefetch": 0}, {"jit": 2, "scan_unroll": 2, "scan_block": 256},
    {"jit": 2, "scan_block": 64, "scan_wg_per_cu": 1},
    {"lds_odd_stride": 1}, {"lds_odd_stride": 1, "jit": 2, "two_level": 1},
    {"part_odd_stride": 0}, {"part_odd_stride": 0, "jit": 2, "two_level": 1},
    {"part_block": 512}, {"part_block": 64, "jit": 2},
    # pass A's launch shape and LDS budget (bucket counts that are not powers of two)
    {"scan_block": 256, "scan_wg_per_cu": 2, "scan_lds_bytes": 40000},
    {"scan_block": 512, "scan_wg_per_cu": 4, "scan_lds_bytes": 65536, "jit": 2},
    {"scan_block": 64, "scan_lds_bytes": 12000, "lds_mode": 2},
    # the scatter's work split and chunk (slices that must divide 5 waves: one)
    {"scatter_slices": 2}, {"scatter_slices": 16, "scatter_chunk": 8192},
    {"scatter_slices": 4, "scan_block": 320, "two_level": 1},
    {"scatter_stage": 0}, {"scatter_stage": 0, "two_level": 1, "scatter_slices": 4},
    {"scatter_stage": 1, "jit": 2, "scan_block": 128, "two_level": 1},
    # pass A tables seeded with sampled heavy-hitter keys
    {"hot_keys": 1, "lds_mode": 2}, {"hot_keys": 1, "lds_mode": 2, "jit": 2, "hot_share": 0},
    # every record value in its own word
    {"rec_pack": 0}, {"rec_pack": 0, "two_level": 1, "jit": 2},
])
def test_table_modes_give_identical_results(setting, gpu_ctx):
    df = random_frame(30000, 5, key_card=400)
    lf = df.filter(col("a") > -500000).group_by("k", "s", maintain_order=True).agg(ALL_AGGS)
    base = lf.collect()
    with tuned(gpu_ctx, setting):
        got = lf.collect()
    # bitwise identical: every state is order independent
    for name in base.columns:
        a, b = base.column(name).column, got.column(name).column
        assert np.array_equal(a.values.view(np.uint8), b.values.view(np.uint8)), name
        assert np.array_equal(a.is_valid(), b.is_valid()), name
    exp, _ = oracle.collect(lf)
    assert_same(got, exp, True, ("x_min", "x_max"), 2)


def test_device_generator_matches_host_twin():
    n = 100003
    u = datagen.device_uniform(n, 11, 3, -1000, 1000).numpy()
    assert np.array_equal(u, datagen.uniform(11, 3, n, -1000, 1000))
    v = datagen.device_validity(n, 11, 4, 0.05)
    assert np.array_equal(datagen.valid_from_words(v, n), datagen.validity_bits(11, 4, n, 0.05))
    z = datagen.device_zipf(n, 11, 5, 1.1, 1000).numpy()
    assert np.array_equal(z, datagen.zipf(11, 5, n, datagen.zipf_cdf(1.1, 1000)))


def _zipf_frame_from_device(cols, n):
    """Host copies of the first n rows of device workload columns."""
    return DataFrame([
        Series("key", Column.int32(cols["key"].numpy(n))),
        Series("a", Column.int64(cols["a"].numpy(n))),
        Series("x", Column.float64(cols["x"].numpy(n), datagen.valid_from_words(cols["x.valid"], n))),
        Series("y", Column.int64(cols["y"].numpy(n))),
    ])


ZIPF_AGGS = [
    col("y").sum().alias("y_sum"), col("y").count().alias("y_cnt"), col("y").min().alias("y_min"),
    col("y").max().alias("y_max"), col("x").sum().alias("x_sum"), col("x").mean().alias("x_mean"),
    col("x").var().alias("x_var"), col("x").min().alias("x_min"),
]


def test_zipf_workload_prefix_matches_oracle(gpu_ctx):
    """The 200M-row Zipf workload's generator and query, checked against the
    oracle on a 3M-row prefix of the device-resident columns (wrapped in
    place, zero copy)."""
    from dataframe_mojo_amd.plan import compile_plan
    from dataframe_mojo_amd.runtime import result_frame

    n_total, n = 4_000_000, 3_000_000
    cols = datagen.zipf_200m_device(n_total, seed=0)
    host = _zipf_frame_from_device(cols, n)
    lf = host.filter(col("a") > 0).group_by("key", maintain_order=True).agg(ZIPF_AGGS)
    dplan = compile_plan(lf)
    dev = {
        "key": gpu_ctx.wrap_array(cols["key"], length=n),
        "a": gpu_ctx.wrap_array(cols["a"], length=n),
        "x": gpu_ctx.wrap_array(cols["x"], cols["x.valid"], length=n),
        "y": gpu_ctx.wrap_array(cols["y"], length=n),
    }
    res = gpu_ctx.groupby(dplan, [dev[k] for k in dplan.keys], [dev[c] for c in dplan.columns])
    got = result_frame(res, dplan, [dev[k] for k in dplan.keys])
    exp, first = oracle.collect(lf)
    assert_same(got, exp, True, ("x_min",))
    assert np.array_equal(res.first_rows(), first)


@pytest.mark.parametrize("ki", range(len(KEYS)))
def test_two_level_random_queries(ki, gpu_ctx):
    """The second partition level (dense group tables) against the oracle,
    ordered and unordered, on every key kind."""
    df = random_frame(40000, 7, key_card=5000)
    try:
        gpu_ctx.tune("two_level", 1)
        for mo in (True, False):
            lf = df.lazy().filter(col("a") > -800000).group_by(*KEYS[ki], maintain_order=mo).agg(ALL_AGGS)
            both(lf, exact_float=("x_min", "x_max"))
    finally:
        gpu_ctx.tune("two_level", -1)


@pytest.mark.parametrize("jit", [1, 2])
def test_zipf_second_level_unordered_matches_oracle(jit, gpu_ctx):
    """The Zipf query (eight aggregates, variance included) through the
    second partition level, unordered: with the runtime-compiled pass C
    (jit 2) its fused output rows would keep part of the pass in scratch
    memory, so the groups go to dense tables and the finalize instead; the
    precompiled pass C (jit 1 at this size) writes the rows itself. Both
    against the oracle, and bit for bit against the one-level result."""
    from dataframe_mojo_amd.plan import compile_plan
    from dataframe_mojo_amd.runtime import result_frame

    n = 2_000_000
    cols = datagen.zipf_200m_device(n, seed=1)
    host = _zipf_frame_from_device(cols, n)
    lf = host.filter(col("a") > 0).group_by("key").agg(ZIPF_AGGS)
    dplan = compile_plan(lf)
    dev = {
        "key": gpu_ctx.wrap_array(cols["key"], length=n),
        "a": gpu_ctx.wrap_array(cols["a"], length=n),
        "x": gpu_ctx.wrap_array(cols["x"], cols["x.valid"], length=n),
        "y": gpu_ctx.wrap_array(cols["y"], length=n),
    }
    keys, inputs = [dev[k] for k in dplan.keys], [dev[c] for c in dplan.columns]
    one = gpu_ctx.groupby(dplan, keys, inputs)
    try:
        gpu_ctx.tune("two_level", 1)
        gpu_ctx.tune("jit", jit)
        res = gpu_ctx.groupby(dplan, keys, inputs)
    finally:
        gpu_ctx.tune("two_level", -1)
        gpu_ctx.tune("jit", 1)
    got = result_frame(res, dplan, keys)
    exp, _ = oracle.collect(lf)
    assert_same(got, exp, False, ("x_min",))
    ka, kb = res.key(0)[1], one.key(0)[1]
    oa, ob = np.argsort(ka, kind="stable"), np.argsort(kb, kind="stable")
    assert np.array_equal(ka[oa], kb[ob])
    for j in range(len(ZIPF_AGGS)):
        _, va, ma = res.agg(j)
        _, vb, mb = one.agg(j)
        assert np.array_equal(ma[oa], mb[ob]) and np.array_equal(va[oa].view(np.uint8), vb[ob].view(np.uint8)), j
    res.release()
    one.release()


def _dict_frame_from_device(cols, n):
    return DataFrame([
        Series("k1", Column.dictionary(cols["k1"].numpy(n).astype(np.int64), datagen.dict_labels())),
        Series("k2", Column.int64(cols["k2"].numpy(n))),
        Series("v", Column.float64(cols["v"].numpy(n), datagen.valid_from_words(cols["v.valid"], n))),
    ])


@pytest.mark.parametrize("two_level", [-1, 1])
def test_dict_workload_prefix_matches_oracle(two_level, gpu_ctx):
    """The 500M-row dictionary workload's generator and query (two keys, a
    modulo filter, nearly one row per group), checked against the oracle on
    a 2M-row prefix wrapped in place, with and without the second partition
    level."""
    from dataframe_mojo_amd.plan import compile_plan
    from dataframe_mojo_amd.runtime import result_frame

    n = 2_000_000
    cols = datagen.dict_500m_device(n, seed=0)
    host = _dict_frame_from_device(cols, n)
    aggs = [col("v").count().alias("v_count"), col("v").sum().alias("v_sum"), col("v").mean().alias("v_mean")]
    lf = host.filter(col("k2") % 3 != 0).group_by("k1", "k2", maintain_order=True).agg(aggs)
    dplan = compile_plan(lf)
    dev = {
        "k1": gpu_ctx.wrap_array(cols["k1"], labels=datagen.dict_labels(), dict_codes=True, length=n),
        "k2": gpu_ctx.wrap_array(cols["k2"], length=n),
        "v": gpu_ctx.wrap_array(cols["v"], cols["v.valid"], length=n),
    }
    try:
        gpu_ctx.tune("two_level", two_level)
        res = gpu_ctx.groupby(dplan, [dev[k] for k in dplan.keys], [dev[c] for c in dplan.columns])
    finally:
        gpu_ctx.tune("two_level", -1)
    got = result_frame(res, dplan, [dev[k] for k in dplan.keys])
    exp, first = oracle.collect(lf)
    assert_same(got, exp, True, (), 2)
    assert np.array_equal(res.first_rows(), first)


def test_dict_workload_size_independent_properties(gpu_ctx):
    """At 60M rows (past one partition level's capacity, so the second level
    runs by itself): the group count equals the number of distinct kept
    keys, the lengths sum to the kept rows, the counts to the kept valid
    rows, and each group's key satisfies the filter."""
    from dataframe_mojo_amd.plan import compile_plan

    n = 60_000_000
    cols = datagen.dict_500m_device(n, seed=3)
    schema = DataFrame([Series("k1", Column.dictionary([], [])), Series("k2", Column.int64([])),
                        Series("v", Column.float64([]))])
    lf = schema.filter(col("k2") % 3 != 0).group_by("k1", "k2").agg(len_(), col("v").count().alias("n"))
    dplan = compile_plan(lf)
    dev = {
        "k1": gpu_ctx.wrap_array(cols["k1"], labels=datagen.dict_labels(), dict_codes=True, length=n),
        "k2": gpu_ctx.wrap_array(cols["k2"], length=n),
        "v": gpu_ctx.wrap_array(cols["v"], cols["v.valid"], length=n),
    }
    res = gpu_ctx.groupby(dplan, [dev[k] for k in dplan.keys], [dev[c] for c in dplan.columns])
    k1, k2 = cols["k1"].numpy(n).astype(np.int64), cols["k2"].numpy(n)
    valid = datagen.valid_from_words(cols["v.valid"], n)
    kept = (k2 % 3) != 0
    packed = np.unique(k1[kept] * 10_000_000 + k2[kept])
    assert res.groups == len(packed)
    _, g1, _ = res.key(0)
    _, g2, _ = res.key(1)
    assert np.array_equal(np.sort(g1.astype(np.int64) * 10_000_000 + g2), packed)
    assert (g2 % 3 != 0).all()
    _, lens, _ = res.agg(0)
    _, cnts, _ = res.agg(1)
    assert int(lens.sum()) == int(kept.sum())
    assert int(cnts.sum()) == int((kept & valid).sum())


@pytest.mark.parametrize("ki,fi", [(ki, fi) for ki in range(len(KEYS)) for fi in (1, 2, 3)] + [(3, 6)])
def test_jit_scan_matches_precompiled(ki, fi, gpu_ctx):
    """Pass A compiled at run time for the query's shape gives the
    precompiled kernel's results bit for bit, and the oracle's."""
    df = random_frame(25000, 11 + fi, key_card=300)
    lf = df.lazy()
    for f in FILTERS[fi]:
        lf = lf.filter(f)
    lf = lf.group_by(*KEYS[ki], maintain_order=True).agg(ALL_AGGS)
    try:
        gpu_ctx.tune("jit", 0)
        base = lf.collect()
        before = gpu_ctx.observe()["jit_launches"]
        gpu_ctx.tune("jit", 2)
        got = lf.collect()
        # pass A and pass C both ran as runtime-compiled kernels
        assert gpu_ctx.observe()["jit_launches"] >= before + 2
    finally:
        gpu_ctx.tune("jit", 1)
    for name in base.columns:
        a, b = base.column(name).column, got.column(name).column
        assert np.array_equal(a.values.view(np.uint8), b.values.view(np.uint8)), name
        assert np.array_equal(a.is_valid(), b.is_valid()), name
    exp, _ = oracle.collect(lf)
    assert_same(got, exp, True, ("x_min", "x_max"), len(KEYS[ki]))


@pytest.mark.parametrize("two_level", [-1, 1])
@pytest.mark.parametrize("width", [2, 3, 4])
def test_staged_scatter_record_widths(width, two_level, gpu_ctx):
    """The staged scatters (records of 2, 3 and 4 words sorted through LDS,
    at the first level and, with two levels, the second) give the gathering
    scatters' results bit for bit, and the oracle's. One-wave pass A
    workgroups make every spill segment span several staging rounds."""
    from test_gpu_batch import frames_equal_bitwise
    rng = np.random.default_rng(width)
    n = 3_000_000
    df = DataFrame([Series("k", Column.int64(rng.integers(0, 300000, n))),
                    Series("v", Column.int64(rng.integers(-1000, 1000, n))),
                    Series("x", Column.float64(rng.normal(size=n), rng.random(n) > 0.1)),
                    Series("y", Column.int32(rng.integers(-50, 50, n)))])
    aggs = [col("v").sum(), col("x").sum().alias("x_sum"), col("y").min().alias("y_min")][: width - 1]
    lf = df.filter(col("v") > -900).group_by("k").agg(aggs)
    got = []
    try:
        gpu_ctx.tune("scan_block", 64)
        gpu_ctx.tune("two_level", two_level)
        gpu_ctx.tune("rec_pack", 0)  # (records of `width` words)
        for stage in (0, 1):
            gpu_ctx.tune("scatter_stage", stage)
            got.append(lf.collect(engine="accel"))
    finally:
        for k, v in TUNE_DEFAULTS.items():
            gpu_ctx.tune(k, v)
    frames_equal_bitwise(got[0], got[1], 1)
    exp, _ = oracle.collect(lf)
    assert_same(got[1], exp, False, (), 1)


@pytest.mark.parametrize("mo", [True, False])
@pytest.mark.parametrize("setting", [{}, {"agg_mode": 0}, {"jit": 2}, {"two_level": 1}])
def test_float_keys(mo, setting, gpu_ctx):
    """A Float64 key: canonical groups (-0.0 == 0.0, NaNs together), the
    first row's own value reported, nulls one group; against the oracle,
    with the key's bits compared exactly."""
    rng = np.random.default_rng(21)
    n = 30000
    base = rng.choice(np.array([0.0, -0.0, 1.5, -2.25, np.nan, np.inf, -np.inf, 3e300, 5e-324]), n)
    base = np.where(rng.random(n) < 0.4, rng.integers(-50, 50, n) / 4.0, base)
    df = DataFrame([Series("k", Column.float64(base, rng.random(n) > 0.03)),
                    Series("v", Column.int64(rng.integers(-1000, 1000, n))),
                    Series("x", Column.float64(rng.normal(size=n)))])
    lf = df.filter(col("v") > -800).group_by("k", maintain_order=mo).agg(
        len_(), col("v").sum().alias("v_sum"), col("x").mean().alias("x_mean"), col("x").min().alias("x_min"))
    try:
        for k, v in setting.items():
            gpu_ctx.tune(k, v)
        got = lf.collect(engine="accel")
    finally:
        for k, v in TUNE_DEFAULTS.items():
            gpu_ctx.tune(k, v)
    exp, _ = oracle.collect(lf)
    if not mo:
        # order is unspecified: match rows by the key's canonical word
        def canon(c):
            v = c.values.copy()
            v[np.isnan(v)] = np.nan
            w = v.view(np.uint64).copy()
            w[v == 0] = 0
            w[np.isnan(v)] = 0x7FF8000000000000
            w[~c.is_valid()] = 0
            return np.lexsort((w, ~c.is_valid()))  # nulls last; 5e-324's word is 1
        ga, ea = canon(got.column("k").column), canon(exp.column("k").column)
        got = DataFrame([Series(s.name, Column(s.column.dtype, s.column.values[ga], s.column.is_valid()[ga]))
                         for s in got._series])
        exp = DataFrame([Series(s.name, Column(s.column.dtype, s.column.values[ea], s.column.is_valid()[ea]))
                         for s in exp._series])
    gk, ek = got.column("k").column, exp.column("k").column
    assert np.array_equal(gk.is_valid(), ek.is_valid())
    m = ek.is_valid()
    assert np.array_equal(gk.values[m].view(np.uint64), ek.values[m].view(np.uint64))
    assert_same(got, exp, True, ("x_min",), 1)


def test_result_buffers_as_device_residences(gpu_ctx):
    """Each output buffer through the residence callbacks, called by address
    as storage/residence.mojo calls _CopyToHost/_ReleaseDevice, after the
    result handle itself is released: the same bytes as the direct copies."""
    from dataframe_mojo_amd.plan import compile_plan

    df = random_frame(5000, 3)
    lf = df.filter(col("a") > 0).group_by("k", "s", maintain_order=True).agg(
        col("x").sum().alias("xs"), col("b").sum().alias("bs"), col("a").min().alias("amin"), len_())
    plan = compile_plan(lf)
    dev = {name: gpu_ctx.upload(df.column(name).column) for name in {*plan.keys, *plan.columns}}
    res = gpu_ctx.groupby(plan, [dev[k] for k in plan.keys], [dev[c] for c in plan.columns])
    G = res.groups
    want = []
    for kind, count, get in ((_lib.DFR_BUFFER_KEY, len(plan.keys), res.key), (_lib.DFR_BUFFER_AGG, len(plan.aggs), res.agg)):
        for i in range(count):
            _, vals, valid = get(i)
            want.append((kind, i, vals.tobytes(), np.packbits(valid, bitorder="little").tobytes()))
    first = res.first_rows().tobytes()
    held = [(w, res.residence(w[0], w[1], _lib.DFR_BUFFER_VALUES), res.residence(w[0], w[1], _lib.DFR_BUFFER_VALIDITY))
            for w in want]
    fr = res.residence(_lib.DFR_BUFFER_FIRST_ROWS, 0, _lib.DFR_BUFFER_VALUES)
    res.release()  # the residences keep the device memory
    for (kind, i, vals, bits), rv, rb in held:
        assert rv.backend == "amd" and rv.bytes == len(vals)
        assert rv.host().tobytes() == vals, (kind, i)
        assert rb.bytes == (G + 7) // 8 and rb.host().tobytes() == bits, (kind, i)
        rv.release()
        rb.release()
    assert fr.host().tobytes() == first
    fr.release()


@pytest.mark.parametrize("kind", ["int64", "int32", "float64", "dict", "float_device"])
@pytest.mark.parametrize("card", [1, 37, 5000, 300_000])
def test_distinct_estimate_bounds_the_groups(kind, card, gpu_ctx):
    """The hand-off sketch's distinct estimate is an upper bound within a
    few percent of the truth (it sizes hash tables: an underestimate costs
    a capacity retry, an overestimate memory)."""
    rng = np.random.default_rng(card)
    n = 1_000_000
    k = rng.integers(0, card, n)
    if kind == "float_device":
        arr = datagen.device_float_keys(n, 5, 77, card)
        col_ = gpu_ctx.wrap_array(arr)
        true = len(np.unique(datagen.float_keys(5, 77, n, card)))
    else:
        c = {"int64": Column.int64(k), "int32": Column.int32(k), "float64": Column.float64(k * 0.125),
             "dict": Column.dictionary(k % 1000, [f"l{i}" for i in range(1000)])}[kind]
        col_ = gpu_ctx.upload(c)
        true = len(np.unique(k if kind != "dict" else k % 1000))
    est = col_.distinct()
    # (HyperLogLog with 4096 registers: 1.6% standard error; the estimate adds three of them and 16)
    assert true <= est <= 1.12 * true + 24, (kind, card, true, est)


def _skewed_series(n, seed, domain=200_000):
    """Zipf(1.1)-skewed Int64 keys over [0, domain) with nulls, and value
    columns whose Float64 magnitudes span 1e-8 to 1e20 (wide fixed-point
    accumulators)."""
    rng = np.random.default_rng(seed)
    ranks = np.arange(1, domain + 1, dtype=np.float64)
    p = ranks ** -1.1
    keys = rng.choice(domain, size=n, p=p / p.sum())
    w = rng.normal(size=n) * 10.0 ** rng.uniform(-8, 20, n)
    return [Series("k", Column.int64(keys, rng.random(n) > 0.01)),
            Series("a", Column.int64(rng.integers(-1000, 1000, n))),
            Series("x", Column.float64(rng.normal(size=n), rng.random(n) > 0.05)),
            Series("w", Column.float64(w)),
            Series("y", Column.int32(rng.integers(-1000, 1000, n)))]


HOT_AGGS = [col("y").sum().alias("y_sum"), col("y").count().alias("y_cnt"), col("y").min().alias("y_min"),
            col("y").max().alias("y_max"), col("x").sum().alias("x_sum"), col("x").mean().alias("x_mean"),
            col("x").var().alias("x_var"), col("x").min().alias("x_min"), col("x").max().alias("x_max"),
            col("w").sum().alias("w_sum"), col("w").std().alias("w_std"), len_().alias("n")]


@pytest.mark.parametrize("setting", [{}, {"jit": 2}, {"scan_block": 256, "scan_wg_per_cu": 3},
                                     {"hot_sample": 4096, "hot_share": 0}, {"lds_odd_stride": 1, "jit": 2},
                                     {"scan_lds_bytes": 30000},
                                     # the key column's hand-off list instead of a per-query sample
                                     {"hot_rows": 1000}, {"hot_rows": 1000, "jit": 2, "hot_share": 0}])
@pytest.mark.parametrize("mo", [False, True])
def test_hot_key_seeding(setting, mo, gpu_ctx):
    """Pass A tables seeded with the sampled heavy hitters of skewed keys
    (their states merged densely, apart from the partitions) give the
    unseeded results bit for bit, and the oracle's."""
    from test_gpu_batch import frames_equal_bitwise
    df = DataFrame(_skewed_series(1_500_000, 21))
    lf = df.filter(col("a") > -600).group_by("k", maintain_order=mo).agg(HOT_AGGS)
    try:
        gpu_ctx.tune("hot_keys", 0)
        base = lf.collect(engine="accel")
        for k, v in setting.items():
            gpu_ctx.tune(k, v)
        gpu_ctx.tune("hot_keys", 1)
        before = gpu_ctx.observe()["hot_keys"]
        got = lf.collect(engine="accel")
        seeded = gpu_ctx.observe()["hot_keys"] - before
    finally:
        for k, v in TUNE_DEFAULTS.items():
            gpu_ctx.tune(k, v)
    assert seeded > 0  # the skew left heavy hitters to seed
    frames_equal_bitwise(base, got, 1)
    if mo or not setting:
        exp, _ = oracle.collect(lf)
        assert_same(got, exp, mo, ("x_min", "x_max"), 1)


def test_hot_key_seeding_multi_key_and_filtered_hot_keys(gpu_ctx):
    """Two keys (dictionary codes and Int64) packed into one domain, and a
    filter that drops the hottest keys' rows: seeding follows the kept rows."""
    from test_gpu_batch import frames_equal_bitwise
    rng = np.random.default_rng(5)
    n = 1_200_000
    codes = rng.integers(0, 7, n).astype(np.int32)
    df = DataFrame(_skewed_series(n, 22, domain=5000) +
                   [Series("d", Column.dictionary(codes, [f"c{i}" for i in range(7)]))])
    lf = df.filter((col("k") > 3) & (col("a") > -900)).group_by("d", "k").agg(HOT_AGGS)
    try:
        gpu_ctx.tune("hot_keys", 0)
        base = lf.collect(engine="accel")
        gpu_ctx.tune("hot_keys", 1)
        before = gpu_ctx.observe()["hot_keys"]
        got = lf.collect(engine="accel")
        seeded = gpu_ctx.observe()["hot_keys"] - before
    finally:
        for k, v in TUNE_DEFAULTS.items():
            gpu_ctx.tune(k, v)
    assert seeded > 0
    frames_equal_bitwise(base, got, 2)
    exp, _ = oracle.collect(lf)
    assert_same(got, exp, False, ("x_min", "x_max"), 2)


def test_hot_key_seeding_dictionary_key_list(gpu_ctx):
    """A dictionary key column's hand-off heavy-hitter list seeds pass A
    (skewed codes over 3000 labels), with the unseeded results bit for bit."""
    from test_gpu_batch import frames_equal_bitwise
    rng = np.random.default_rng(9)
    n, labels = 1_000_000, 3000
    p = np.arange(1, labels + 1, dtype=np.float64) ** -1.2
    codes = rng.choice(labels, size=n, p=p / p.sum()).astype(np.int32)
    df = DataFrame([Series("d", Column.dictionary(codes, [f"s{i}" for i in range(labels)], rng.random(n) > 0.02)),
                    Series("a", Column.int64(rng.integers(-1000, 1000, n))),
                    Series("x", Column.float64(rng.normal(size=n), rng.random(n) > 0.05)),
                    Series("y", Column.int32(rng.integers(-1000, 1000, n)))])
    lf = df.filter(col("a") > 0).group_by("d").agg(HOT_AGGS[:9])
    try:
        gpu_ctx.tune("hot_keys", 0)
        base = lf.collect(engine="accel")
        gpu_ctx.tune("hot_keys", 1)
        gpu_ctx.tune("hot_rows", 1000)
        gpu_ctx.tune("lds_mode", 2)
        before = gpu_ctx.observe()["hot_keys"]
        got = lf.collect(engine="accel")
        seeded = gpu_ctx.observe()["hot_keys"] - before
    finally:
        for k, v in TUNE_DEFAULTS.items():
            gpu_ctx.tune(k, v)
    assert seeded > 0
    frames_equal_bitwise(base, got, 1)
    exp, _ = oracle.collect(lf)
    assert_same(got, exp, False, ("x_min", "x_max"), 1)


@pytest.mark.parametrize("two_level", [-1, 1])
@pytest.mark.parametrize("mo", [False, True])
def test_narrow_values_ride_in_the_key_word(two_level, mo, gpu_ctx):
    """Spilled records carry narrow integer values in the key word's free
    bits (offsets from the column minimum, negative minimums included);
    null slots holding values far outside the valid range are masked. The
    results equal the one-word-per-value records bit for bit, and the
    oracle's."""
    from test_gpu_batch import frames_equal_bitwise
    rng = np.random.default_rng(31)
    n = 2_000_000
    v_valid = rng.random(n) > 0.1
    v = rng.integers(-700, 300, n)
    v[~v_valid] = 1 << 40  # a null slot's value: never read
    df = DataFrame([Series("k", Column.int64(rng.integers(-5000, 200000, n))),
                    Series("v", Column.int64(v, v_valid)),
                    Series("y", Column.int32(rng.integers(-(1 << 31), (1 << 31) - 1, n))),
                    Series("z", Column.int64(rng.integers(-(1 << 62), 1 << 62, n))),
                    Series("u", Column.int32(np.full(n, 7, dtype=np.int32))),
                    Series("x", Column.float64(rng.normal(size=n), rng.random(n) > 0.05))])
    aggs = [col("v").sum().alias("v_sum"), col("v").min().alias("v_min"), col("v").var().alias("v_var"),
            col("y").max().alias("y_max"), col("y").min().alias("y_min"), col("z").max().alias("z_max"),
            col("u").sum().alias("u_sum"), col("x").sum().alias("x_sum"), len_().alias("n")]
    lf = df.filter(col("y") > -(1 << 30)).group_by("k", maintain_order=mo).agg(aggs)
    got = []
    try:
        gpu_ctx.tune("two_level", two_level)
        for pack in (0, 1):
            gpu_ctx.tune("rec_pack", pack)
            got.append(lf.collect(engine="accel"))
    finally:
        for k, val in TUNE_DEFAULTS.items():
            gpu_ctx.tune(k, val)
    frames_equal_bitwise(got[0], got[1], 1)
    exp, _ = oracle.collect(lf)
    assert_same(got[1], exp, mo, (), 1)


def _sorted_rows(frame, n_keys):
    """Every column's (values bits, validity) with rows in key order."""
    cols = [frame.column(c).column for c in frame.columns]
    keys = []
    for c in reversed(cols[:n_keys]):
        raw = c.values.view(np.int64) if c.values.dtype.kind == "f" else c.values.astype(np.int64)
        keys.append(np.where(c.is_valid(), raw, np.iinfo(np.int64).max))
        keys.append(~c.is_valid())
    order = np.lexsort(keys)
    return [(name, c.values[order].view(np.uint8), c.is_valid()[order]) for name, c in zip(frame.columns, cols)]


@pytest.mark.parametrize("setting", [
    {}, {"jit": 2}, {"hot_keys": 1, "lds_mode": 2}, {"hot_keys": 1, "lds_mode": 2, "jit": 2},
    {"rec_pack": 0}, {"part_block": 256}, {"lds_mode": 0},
    # one partition whose LDS table cannot hold the groups: the dense pass C
    # raises its capacity flag and the query runs again through the table
    {"max_parts": 1, "part_lds_bytes": 4096},
])
def test_unordered_one_level_dense_slots(setting, gpu_ctx):
    """An unordered one-level query whose pass C (and the seeded keys'
    merge) puts each group at the next dense slot equals the same query
    through the key-addressed table, occupancy scan and compaction, bit for
    bit, and the oracle."""
    df = random_frame(40000, 9, key_card=3000)
    lf = df.filter(col("a") > -700000).group_by("k", "k32", maintain_order=False).agg(ALL_AGGS)
    with tuned(gpu_ctx, dict(setting, fuse_one_level=0)):
        unfused = lf.collect()
    with tuned(gpu_ctx, dict(setting, fuse_one_level=1)):
        fused = lf.collect()
    assert fused.height() == unfused.height()
    for (name, a, ma), (_, b, mb) in zip(_sorted_rows(fused, 2), _sorted_rows(unfused, 2)):
        assert np.array_equal(ma, mb), name
        assert np.array_equal(a.reshape(len(ma), -1)[ma], b.reshape(len(mb), -1)[mb]), name
    exp, _ = oracle.collect(lf)
    assert_same(fused, exp, False, ("x_min", "x_max"), 2)
