"""The batch entry point (dfr_groupby_batch): many queries in one persistent
launch give what each query gives alone through dfr_groupby, bit for bit,
and the oracle agrees with both.

Bar: the same groups, keys, validity and values (Float64 bit patterns
included) as the one-launch-per-query path; group order is unspecified
without maintain_order, so rows are compared sorted by key.
"""
import numpy as np
import pytest

from dataframe_mojo_amd import Column, DataFrame, Series, _lib, col, len_
from dataframe_mojo_amd.plan import compile_plan
from dataframe_mojo_amd.runtime import result_frame
from oracle import oracle

from test_gpu_parity import assert_same, TUNE_DEFAULTS

pytestmark = pytest.mark.gpu


def frame(n, seed, card):
    rng = np.random.default_rng(seed)
    return DataFrame([
        Series("k", Column.int64(rng.integers(0, max(card, 1), n), rng.random(n) > 0.02)),
        Series("v", Column.int64(rng.integers(-10**6, 10**6, n))),
        Series("x", Column.float64(rng.normal(size=n), rng.random(n) > 0.05)),
        Series("s", Column.dictionary(rng.integers(0, 9, n), [f"c{i}" for i in range(9)])),
    ])


AGG_SETS = [
    [len_(), col("v").sum().alias("v_sum"), col("v").max().alias("v_max"), col("x").sum().alias("x_sum"),
     col("x").mean().alias("x_mean")],
    [col("x").var().alias("x_var"), col("x").min().alias("x_min"), col("v").count().alias("v_n")],
    [col("v").mean().alias("v_mean"), col("v").min().alias("v_min"), col("x").std().alias("x_std")],
]


def make_jobs(gpu_ctx, count, seed):
    """`count` queries over frames of uneven rows and cardinalities."""
    rng = np.random.default_rng(seed)
    jobs, lfs = [], []
    for i in range(count):
        n = int(10 ** rng.uniform(2, 5))
        card = int(10 ** rng.uniform(0, 4))
        df = frame(n, 1000 * seed + i, card)
        thr = int(rng.integers(-10**6, 10**6))
        keys = ["k"] if i % 3 else ["k", "s"]
        lf = df.filter(col("v") > thr).group_by(*keys).agg(AGG_SETS[i % len(AGG_SETS)])
        plan = compile_plan(lf)
        dev = {name: gpu_ctx.upload(df.column(name).column) for name in {*plan.keys, *plan.columns}}
        jobs.append((plan, [dev[k] for k in plan.keys], [dev[c] for c in plan.columns]))
        lfs.append(lf)
    return jobs, lfs


def frames_equal_bitwise(a, b, n_keys):
    ra, rb = a.rows(), b.rows()
    assert len(ra) == len(rb)

    def sk(r):
        return tuple((v is None, 0 if v is None else v) for v in r[:n_keys])

    ia = sorted(range(len(ra)), key=lambda i: sk(ra[i]))
    ib = sorted(range(len(rb)), key=lambda i: sk(rb[i]))
    for name in a.columns:
        ca, cb = a.column(name).column, b.column(name).column
        va, vb = ca.values[ia], cb.values[ib]
        ma, mb = ca.is_valid()[ia], cb.is_valid()[ib]
        assert np.array_equal(ma, mb), name
        assert np.array_equal(va[ma].view(np.uint8), vb[mb].view(np.uint8)), name


@pytest.mark.parametrize("setting", [{}, {"batch_graph": 1}, {"batch_tile": 2048}, {"block": 256, "lds_bytes": 16384},
                                     {"batch_fill": 0}, {"batch_fill": 8, "batch_graph": 1},
                                     {"batch_part_rows": 1000, "batch_streams": 1},
                                     {"batch_part_rows": 1000, "batch_streams": 8, "batch_graph": 1},
                                     {"batch_part_rows": 0}, {"batch_part_rows": 1000},
                                     {"batch_part_rows": 1000, "block": 256, "lds_bytes": 16384}])
def test_batch_equals_one_launch_per_query(setting, gpu_ctx):
    jobs, lfs = make_jobs(gpu_ctx, 40, 1)
    try:
        for k, v in setting.items():
            gpu_ctx.tune(k, v)
        batch = gpu_ctx.groupby_batch(jobs)
    finally:
        for k, v in {**TUNE_DEFAULTS, "batch_graph": 0, "batch_tile": 131072, "batch_fill": 1, "batch_streams": 4,
                     "batch_part_rows": 1 << 22}.items():
            gpu_ctx.tune(k, v)
    for (plan, keys, cols), got in zip(jobs, batch):
        assert not isinstance(got, Exception), got
        one = gpu_ctx.groupby(plan, keys, cols)
        frames_equal_bitwise(result_frame(got, plan, keys), result_frame(one, plan, keys), len(plan.keys))
        got.release()
        one.release()


def test_batch_never_waits_for_a_compile_with_jit_wait_0(gpu_ctx):
    # a serving lane's setting: queries that would be routed to the
    # partitioned passes use runtime-compiled kernels only when they exist
    jobs, lfs = make_jobs(gpu_ctx, 40, 3)
    try:
        for k, v in {"batch_part_rows": 1000, "jit": 2, "jit_wait": 0}.items():
            gpu_ctx.tune(k, v)
        before = gpu_ctx.observe()["jit_compiles"]
        batch = gpu_ctx.groupby_batch(jobs)
        assert gpu_ctx.observe()["jit_compiles"] == before
    finally:
        for k, v in {**TUNE_DEFAULTS, "batch_part_rows": 1 << 22}.items():
            gpu_ctx.tune(k, v)
    for (plan, keys, cols), got in zip(jobs, batch):
        assert not isinstance(got, Exception), got
        one = gpu_ctx.groupby(plan, keys, cols)
        frames_equal_bitwise(result_frame(got, plan, keys), result_frame(one, plan, keys), len(plan.keys))
        got.release()
        one.release()


def test_batch_matches_oracle(gpu_ctx):
    jobs, lfs = make_jobs(gpu_ctx, 12, 2)
    for (plan, keys, cols), got, lf in zip(jobs, gpu_ctx.groupby_batch(jobs), lfs):
        exp, _ = oracle.collect(lf)
        assert_same(result_frame(got, plan, keys), exp, False, ("x_min",), len(plan.keys))


def test_batch_reports_each_query_status(gpu_ctx):
    good = frame(5000, 3, 20)
    bad = DataFrame([Series("k", Column.int64([1, 1])), Series("v", Column.int64([2**63 - 1, 1])),
                     Series("x", Column.float64([0.0, 1.0])), Series("s", Column.string(["a", "b"]))])
    lfs = [
        good.lazy().group_by("k").agg(col("v").sum()),
        bad.lazy().group_by("k").agg(col("v").sum()),
        good.lazy().group_by("k", maintain_order=True).agg(col("v").sum()),
        good.filter(col("v") > 0).group_by("k").agg(col("x").mean()),
    ]
    jobs = []
    for lf in lfs:
        plan = compile_plan(lf)
        df = lf.source
        dev = {name: gpu_ctx.upload(df.column(name).column) for name in {*plan.keys, *plan.columns}}
        jobs.append((plan, [dev[k] for k in plan.keys], [dev[c] for c in plan.columns]))
    out = gpu_ctx.groupby_batch(jobs)
    assert isinstance(out[1], _lib.DataFault) and "Int64 expression sum overflow" in str(out[1])
    assert isinstance(out[2], _lib.Unsupported)
    for i in (0, 3):
        plan, keys, _ = jobs[i]
        exp, _ = oracle.collect(lfs[i])
        assert_same(result_frame(out[i], plan, keys), exp, False, (), 1)


def test_empty_batch_and_empty_queries(gpu_ctx):
    assert gpu_ctx.groupby_batch([]) == []
    df = frame(1000, 4, 10)
    lf = df.filter(col("v") > 10**7).group_by("k").agg(col("v").sum())
    plan = compile_plan(lf)
    dev = {name: gpu_ctx.upload(df.column(name).column) for name in {*plan.keys, *plan.columns}}
    (r,) = gpu_ctx.groupby_batch([(plan, [dev[k] for k in plan.keys], [dev[c] for c in plan.columns])])
    assert r.groups == 0


@pytest.mark.parametrize("sizes", [[(300_000, 50)], [(300_000, 200_000)], [(1_000, 80_000), (20_000, 150_000)],
                                   [(1_000_000, 3_000), (500, 1), (2_000, 120_000), (150_000, 40)]])
def test_few_rows_cut_fine_and_wide_tables_spread(sizes, gpu_ctx):
    # a batch of few rows is cut into items of fewer rows than batch_tile
    # (down to one per workgroup of the grid, never below a query's LDS
    # table); tables of up to 2e5 groups are set up, scanned and compacted
    # in spans of 2048 slots, several workgroups per query
    jobs, lfs = [], []
    for i, (n, card) in enumerate(sizes):
        df = frame(n, 77 + i, card)
        lf = df.filter(col("v") > -900_000).group_by("k").agg(AGG_SETS[i % len(AGG_SETS)])
        plan = compile_plan(lf)
        dev = {name: gpu_ctx.upload(df.column(name).column) for name in {*plan.keys, *plan.columns}}
        jobs.append((plan, [dev[k] for k in plan.keys], [dev[c] for c in plan.columns]))
        lfs.append(lf)
    for (plan, keys, cols), got, lf in zip(jobs, gpu_ctx.groupby_batch(jobs), lfs):
        assert not isinstance(got, Exception), got
        one = gpu_ctx.groupby(plan, keys, cols)
        frames_equal_bitwise(result_frame(got, plan, keys), result_frame(one, plan, keys), len(plan.keys))
        got.release()
        one.release()
    plan, keys, cols = jobs[-1]
    exp, _ = oracle.collect(lfs[-1])
    (got,) = gpu_ctx.groupby_batch([jobs[-1]])
    assert_same(result_frame(got, plan, keys), exp, False, ("x_min",), len(plan.keys))
