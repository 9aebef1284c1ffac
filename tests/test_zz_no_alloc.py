"""dfr_groupby_into, the entry point that allocates nothing: through the C
ABI, into device buffers the caller allocated, it gives dfr_groupby's
result bit for bit (keys, validity, every aggregate's bits; under
maintain_order the same row order and first rows), on two calls in a row,
the second with a smaller input (the shared scratch keeps the first call's
contents), and it refuses a query whose group table exceeds its scratch."""
import numpy as np
import pytest

from dataframe_mojo_amd import Column, DataFrame, Series, _lib, col, datagen, len_
from dataframe_mojo_amd.plan import compile_plan
from dataframe_mojo_amd.runtime import result_frame

pytestmark = pytest.mark.gpu

WIDTH = {_lib.DFR_INT32: 4, _lib.DFR_DICT: 4, _lib.DFR_INT64: 8, _lib.DFR_FLOAT64: 8}
NP = {_lib.DFR_INT32: np.int32, _lib.DFR_DICT: np.uint32, _lib.DFR_INT64: np.int64, _lib.DFR_FLOAT64: np.float64}


def frame(n, seed, card):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=n)
    x[rng.random(n) < 0.01] = np.nan
    return DataFrame([
        Series("k", Column.int64(rng.integers(-card, card, n), rng.random(n) > 0.03)),
        Series("s", Column.dictionary(rng.integers(0, 7, n), [f"c{i}" for i in range(7)])),
        Series("a", Column.int64(rng.integers(-10**6, 10**6, n), rng.random(n) > 0.1)),
        Series("b", Column.int32(rng.integers(-1000, 1000, n))),
        Series("x", Column.float64(x, rng.random(n) > 0.05)),
    ])


AGGS = [len_(), col("a").sum().alias("a_sum"), col("a").min().alias("a_min"), col("a").max().alias("a_max"),
        col("a").mean().alias("a_mean"), col("b").sum().alias("b_sum"), col("x").sum().alias("x_sum"),
        col("x").var().alias("x_var"), col("x").min().alias("x_min"), col("x").count().alias("x_n"),
        col("a").var().alias("a_var")]


def run_both(ctx, df, keys, maintain_order, capacity=None):
    lf = df.filter(col("b") > -900).group_by(*keys, maintain_order=maintain_order).agg(AGGS)
    plan = compile_plan(lf)
    dev = {name: ctx.upload(df.column(name).column) for name in {*plan.keys, *plan.columns}}
    kc, cc = [dev[k] for k in plan.keys], [dev[c] for c in plan.columns]
    ref = ctx.groupby(plan, kc, cc)
    kd = [ref.key(i)[0] for i in range(len(plan.keys))]
    ad = [ref.agg(i)[0] for i in range(len(plan.aggs))]
    cap = capacity if capacity is not None else ref.groups + 5
    hold = []

    def buf(dtype):
        v = datagen.DeviceArray(max(cap, 1), NP[dtype])
        m = datagen.DeviceArray(cap // 64 + 1, np.uint64)
        hold.append((v, m))
        return v, m

    kb = [buf(d) for d in kd]
    ab = [buf(d) for d in ad]
    first = datagen.DeviceArray(max(cap, 1), np.int64) if maintain_order else None
    g = ctx.groupby_into(plan, kc, cc, (cap, [(v.ptr, m.ptr) for v, m in kb], [(v.ptr, m.ptr) for v, m in ab],
                                        first.ptr if first is not None else 0))
    assert g == ref.groups
    exp = result_frame(ref, plan, kc)
    got_cols = []
    for (v, m), d in zip(kb + ab, kd + ad):
        vals = v.numpy(g)
        bits = np.unpackbits(m.numpy(g // 64 + 1).view(np.uint8), bitorder="little")[:g].astype(bool)
        got_cols.append((vals, bits))
    exp_cols = [(exp.column(c).column.values, exp.column(c).column.is_valid()) for c in exp.columns]
    if maintain_order:
        order_g = order_e = np.arange(g)
        assert np.array_equal(first.numpy(g), ref.first_rows())
    else:
        def order_of(cols):
            ks = []
            for vals, bits in reversed(cols[:len(plan.keys)]):
                ks.append(np.where(bits, vals.astype(np.int64), np.iinfo(np.int64).max))
                ks.append(~bits)
            return np.lexsort(ks)
        order_g, order_e = order_of(got_cols), order_of(exp_cols)
    for name, (gv, gm), (ev, em) in zip(exp.columns, got_cols, exp_cols):
        assert np.array_equal(gm[order_g], em[order_e]), name
        a, b = gv[order_g][gm[order_g]], ev[order_e][em[order_e]]
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), name
    ref.release()
    for d in dev.values():
        d.release()
    return g


@pytest.mark.parametrize("maintain_order", [False, True])
@pytest.mark.parametrize("keys", [["k"], ["k", "s"]])
def test_no_alloc_equals_groupby_twice_second_smaller(keys, maintain_order, gpu_ctx):
    g1 = run_both(gpu_ctx, frame(200_000, 1, 3000), keys, maintain_order)
    g2 = run_both(gpu_ctx, frame(30_000, 2, 300), keys, maintain_order)
    assert g2 < g1


def test_no_alloc_hash_table_keys(gpu_ctx):
    # a key span past the direct tables: the open-addressing table in scratch
    rng = np.random.default_rng(4)
    n = 50_000
    df = DataFrame([Series("k", Column.int64(rng.integers(0, 4000, n) * 10**12)),
                    Series("s", Column.dictionary(rng.integers(0, 3, n), ["a", "b", "c"])),
                    Series("a", Column.int64(rng.integers(-10**6, 10**6, n))),
                    Series("b", Column.int32(rng.integers(-1000, 1000, n))),
                    Series("x", Column.float64(rng.normal(size=n)))])
    run_both(gpu_ctx, df, ["k"], False)
    run_both(gpu_ctx, df, ["k"], True)


def test_no_alloc_rejects_what_exceeds_its_scratch_or_capacity(gpu_ctx):
    rng = np.random.default_rng(5)
    n = 4_000_000
    # ~4M distinct keys: the group table alone is past 64 MiB
    df = DataFrame([Series("k", Column.int64(rng.permutation(n).astype(np.int64) * 3)),
                    Series("s", Column.dictionary(np.zeros(n, np.int64), ["a"])),
                    Series("a", Column.int64(rng.integers(-10**6, 10**6, n))),
                    Series("b", Column.int32(rng.integers(-1000, 1000, n))),
                    Series("x", Column.float64(rng.normal(size=n)))])
    with pytest.raises(_lib.DfrError, match="no-alloc scratch"):
        run_both(gpu_ctx, df, ["k"], False)
    small = frame(10_000, 6, 500)
    with pytest.raises(_lib.DfrError, match="exceed the output capacity"):
        run_both(gpu_ctx, small, ["k"], False, capacity=10)
