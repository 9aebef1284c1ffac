"""The benchmark configurations at their full sizes: what bench.py times is
what these tests check.

  zipf_200m   200M rows, unordered, default tuning (the runtime-compiled
              kernels), against the oracle's 16-worker restatement (which
              is bit-identical to its sequential one)
  dict_500m   500M rows: size-independent properties (group count and keys
              against the distinct kept keys, lengths and counts against
              the kept rows, means against sums)
  batch_4096  4096 queries of the batched distribution, graph capture on
              and off: every query's result equal to its own dfr_groupby
              call bit for bit, and a sample against the oracle
  serve       Poisson arrivals over tables with Int64, Float64 and
              dictionary String keys: each served result equal to the query
              alone, bit for bit, and a sample against the oracle

Bar (as the parity tests): keys, validity and integer aggregates exact;
Float64 aggregates within |a - e| <= 1e-12 * max(1, |e|) of the oracle;
batched and served results bit-identical to the one-query path.
"""
import numpy as np
import pytest

from dataframe_mojo_amd import Column, DataFrame, Series, col, datagen, len_
from dataframe_mojo_amd.plan import compile_plan
from dataframe_mojo_amd.runtime import PreparedBatch, Server, result_frame
from oracle import oracle


pytestmark = pytest.mark.gpu

FTOL = 1e-12


def sorted_columns(frame, n_keys):
    """Columns of a result frame as numpy (values, valid), rows sorted by
    the key columns (nulls last)."""
    cols = [frame.column(c).column for c in frame.columns]
    keys = []
    for c in reversed(cols[:n_keys]):
        # (Float64 keys by their bits: a group's key is its first row's value
        # on both sides, NaN and -0.0 included)
        raw = c.values.view(np.int64) if c.values.dtype.kind == "f" else c.values.astype(np.int64)
        keys.append(np.where(c.is_valid(), raw, np.iinfo(np.int64).max))
        keys.append(~c.is_valid())
    order = np.lexsort(keys) if keys else np.arange(frame.height())
    return [(frame.columns[i], c.values[order], c.is_valid()[order]) for i, c in enumerate(cols)]


def assert_frames_close(got, exp, n_keys, exact=()):
    assert got.columns == exp.columns
    assert got.height() == exp.height(), (got.height(), exp.height())
    for (name, a, ma), (_, e, me) in zip(sorted_columns(got, n_keys), sorted_columns(exp, n_keys)):
        assert np.array_equal(ma, me), name
        a, e = a[ma], e[me]
        if a.dtype.kind == "f" and name not in exact:
            both_nan = np.isnan(a) & np.isnan(e)
            ok = both_nan | (np.abs(a - e) <= FTOL * np.maximum(1.0, np.abs(e)))
            assert ok.all(), (name, int((~ok).sum()), a[~ok][:5], e[~ok][:5])
        else:
            assert np.array_equal(a.view(np.uint8), e.view(np.uint8)), name


def float_sum_budget(keys_kept, x_kept, valid_kept, uniq):
    """Per group (in `uniq` key order): the reference's forward error budget
    of a sequential Float64 sum, gamma_{2n+1} * sum|x| with
    gamma_k = k eps / (1 - k eps), eps = 2^-52
    (tests/kernel_equivalence/test_state_merge.mojo:363-430); and n."""
    gid = np.searchsorted(uniq, keys_kept)
    w = np.where(valid_kept, np.abs(x_kept), 0.0)
    mag = np.bincount(gid, weights=w, minlength=len(uniq))
    n = np.bincount(gid, weights=valid_kept.astype(np.float64), minlength=len(uniq))
    k = 2 * n + 1
    eps = 2.220446049250313e-16
    return k * eps / (1 - k * eps) * mag, n


def test_zipf_200m_full_size_unordered(gpu_ctx):
    """Integers exact against the oracle; Float64 sums against double-double
    sums (the device's are exact, so within an ulp) and against the oracle
    within the reference's forward error budget of the oracle's sequential
    sums (at 200M rows a group holds up to 12M values, and the oracle's
    sums there drift past 1e-12 of the exact ones); variances within 1e-12
    of both the oracle and two-pass double-double variances."""
    from bench import host_zipf, wrap_zipf, zipf_query

    n = 200_000_000
    cols = datagen.zipf_200m_device(n, seed=0)
    dev = wrap_zipf(gpu_ctx, cols, n)
    q = zipf_query()
    plan = compile_plan(q)
    before = gpu_ctx.observe()["jit_launches"]
    res = gpu_ctx.groupby(plan, [dev[k] for k in plan.keys], [dev[c] for c in plan.columns])
    assert gpu_ctx.observe()["jit_launches"] > before  # the runtime-compiled kernels ran, as in the bench
    got = result_frame(res, plan, [dev[k] for k in plan.keys])
    res.release()
    host = host_zipf(cols, n)
    del cols, dev
    lf = type(q)(host, q.filters, q.keys, q.aggs, q.maintain_order, True)
    exp, _ = oracle.collect(lf, threads=16)
    assert got.height() > 900_000
    float_aggs = ("x_sum", "x_mean", "x_var")
    g = {name: (a, m) for name, a, m in sorted_columns(got, 1)}
    e = {name: (a, m) for name, a, m in sorted_columns(exp, 1)}
    for name in got.columns:
        assert np.array_equal(g[name][1], e[name][1]), name
        if name not in float_aggs:
            assert np.array_equal(g[name][0].view(np.uint8), e[name][0].view(np.uint8)), name
    keep = host.column("a").column.values > 0
    key = host.column("key").column.values[keep]
    x = host.column("x").column.values[keep]
    xv = host.column("x").column.is_valid()[keep]
    uniq = g["key"][0]
    assert np.array_equal(uniq, np.unique(key))
    gid = np.searchsorted(uniq, key)
    hi, lo, cnt, dd_var = oracle.dd_moments(gid, x, xv, len(uniq), ddof=1)
    dd_sum = hi + lo
    budget, n = float_sum_budget(key, x, xv, uniq)
    gs, es = g["x_sum"][0], e["x_sum"][0]
    # the device's sums: exact, rounded once
    assert (np.abs(gs - dd_sum) <= 2.220446049250313e-16 * np.abs(dd_sum)).all()
    tol = np.maximum(1e-12 * np.maximum(1.0, np.abs(es)), budget)
    assert (np.abs(gs - es) <= tol).all()
    gm, em = g["x_mean"][0], e["x_mean"][0]
    nz = n > 0
    assert (np.abs(gm[nz] - em[nz]) <= tol[nz] / n[nz] + 1e-12 * np.maximum(1.0, np.abs(em[nz]))).all()
    gv, ev = g["x_var"][0], e["x_var"][0]
    mv = g["x_var"][1]
    assert (np.abs(gv[mv] - ev[mv]) <= 1e-12 * np.maximum(1.0, np.abs(ev[mv]))).all()
    assert (np.abs(gv[mv] - dd_var[mv]) <= 1e-12 * np.maximum(1.0, np.abs(dd_var[mv]))).all()


def test_dict_500m_full_size_properties(gpu_ctx):
    from bench import dict_query, wrap_dict

    n = 500_000_000
    cols = datagen.dict_500m_device(n, seed=0)
    dev = wrap_dict(gpu_ctx, cols, n)
    plan = compile_plan(dict_query())
    res = gpu_ctx.groupby(plan, [dev[k] for k in plan.keys], [dev[c] for c in plan.columns])
    _, g1, m1 = res.key(0)
    _, g2, m2 = res.key(1)
    _, cnt, _ = res.agg(0)
    _, vsum, _ = res.agg(1)
    _, vmean, mmean = res.agg(2)
    G = res.groups
    res.release()
    assert m1.all() and m2.all()
    k2 = cols["k2"].numpy(n)
    kept = (k2 % 3) != 0
    packed = cols["k1"].numpy(n).astype(np.int64)[kept] * 10_000_000 + k2[kept]
    del k2
    valid = datagen.valid_from_words(cols["v.valid"], n)[kept]
    v = cols["v"].numpy(n)[kept]
    del cols, dev
    # the groups are exactly the distinct kept keys
    uniq = np.unique(packed)
    assert G == len(uniq) > 3e8
    gp = g1.astype(np.int64) * 10_000_000 + g2
    gp.sort()
    assert np.array_equal(gp, uniq)
    assert (g2 % 3 != 0).all()
    # counts sum to the kept valid values; sums to their total (positive
    # values, so no cancellation); means are sums over counts
    assert int(cnt.sum()) == int(valid.sum())
    total = float(np.sum(v[valid]))
    assert abs(float(np.sum(vsum)) - total) <= 1e-9 * total
    nz = cnt > 0
    assert np.array_equal(mmean, nz)
    assert np.allclose(vmean[nz], vsum[nz] / cnt[nz], rtol=1e-15, atol=0)


def _batch_jobs(gpu_ctx, params, cols):
    from bench import batch_query

    jobs = []
    for p, c in zip(params, cols):
        n = p["rows"]
        dev = {"key": gpu_ctx.wrap_array(c["key"], length=n), "value": gpu_ctx.wrap_array(c["value"], length=n),
               "x": gpu_ctx.wrap_array(c["x"], c["x.valid"], length=n)}
        plan = compile_plan(batch_query(p["threshold"]))
        jobs.append((plan, [dev[k] for k in plan.keys], [dev[k] for k in plan.columns]))
    return jobs


def _result_arrays(res, n_keys, n_aggs):
    """A one-key result's (values, valid) arrays, rows in key order: by
    validity, then by the key's bits (so the null group, a zero key and
    every Float64 key bit pattern sort apart)."""
    _, kv, km = res.key(0)
    bits = kv.view(np.uint32 if kv.itemsize == 4 else np.uint64)
    order = np.lexsort((bits, ~km))
    out = [(kv[order], km[order])]
    for a in range(n_aggs):
        _, vals, valid = res.agg(a)
        out.append((vals[order], valid[order]))
    return out


@pytest.mark.parametrize("graph", [0, 1])
def test_batch_4096_equals_one_launch_per_query(graph, gpu_ctx):
    params = datagen.batch_4096_params(4096, seed=0)
    cols = datagen.batch_4096_device(params, seed=0)
    jobs = _batch_jobs(gpu_ctx, params, cols)
    try:
        gpu_ctx.tune("batch_graph", graph)
        batched = PreparedBatch(jobs).run(gpu_ctx)
    finally:
        gpu_ctx.tune("batch_graph", 0)
    checked = 0
    for i, (job, b) in enumerate(zip(jobs, batched)):
        assert not isinstance(b, Exception), (i, b)
        alone = gpu_ctx.groupby(*job)
        assert b.groups == alone.groups, i
        na = len(job[0].aggs)
        for (va, ma), (vb, mb) in zip(_result_arrays(b, 1, na), _result_arrays(alone, 1, na)):
            assert np.array_equal(ma, mb), i
            assert np.array_equal(va[ma].view(np.uint8), vb[mb].view(np.uint8)), i
        alone.release()
        b.release()
        checked += 1
    assert checked == 4096
    # a sample against the oracle (host copies of those queries' columns)
    from bench import batch_query

    for i in range(0, 4096, 512):
        p, c = params[i], cols[i]
        n = p["rows"]
        host = DataFrame([Series("key", Column.int64(c["key"].numpy(n))), Series("value", Column.int64(c["value"].numpy(n))),
                          Series("x", Column.float64(c["x"].numpy(n), datagen.valid_from_words(c["x.valid"], n)))])
        q = batch_query(p["threshold"])
        lf = type(q)(host, q.filters, q.keys, q.aggs, q.maintain_order, True)
        res = gpu_ctx.groupby(*jobs[i])
        got = result_frame(res, jobs[i][0], jobs[i][1])
        res.release()
        exp, _ = oracle.collect(lf)
        assert_frames_close(got, exp, 1)


def _serve_tables(gpu_ctx, seed=0, count=12):
    """Device tables of 1e6 to 4e6 rows whose key is Int64, Float64 (values
    k / 8 with -0.0, NaN and nulls among them) or dictionary String."""
    rng = np.random.default_rng(seed)
    labels = datagen.dict_labels()
    tabs = []
    for t in range(count):
        n = int(rng.integers(1_000_000, 4_000_000))
        card = int(np.exp(rng.uniform(0, np.log(1e5))))
        kind = t % 3
        kint = rng.integers(0, max(card, 1), n)
        if kind == 0:
            key = Column.int64(kint)
        elif kind == 1:
            f = kint / 8.0
            f[kint == 0] = -0.0
            f[kint == 1] = np.nan
            key = Column.float64(f, rng.random(n) > 0.01)
        else:
            key = Column.dictionary(kint % len(labels), labels)
        df = DataFrame([Series("key", key), Series("value", Column.int64(rng.integers(-10**6, 10**6, n))),
                        Series("x", Column.float64(rng.normal(size=n), rng.random(n) > 0.05))])
        dev = {c: gpu_ctx.upload(df.column(c).column) for c in df.columns}
        tabs.append((df, dev))
    return tabs


def test_serve_poisson_with_float_keys_alone_equals_mix(gpu_ctx):
    from bench import batch_query

    rng = np.random.default_rng(4)
    tabs = _serve_tables(gpu_ctx)
    n_q = 300
    arrivals = np.cumsum(rng.exponential(1.0 / 2000.0, n_q))
    work = []
    for i in range(n_q):
        df, dev = tabs[int(rng.integers(0, len(tabs)))]
        n = df.height()
        length = int(np.exp(rng.uniform(np.log(1e3), np.log(n))))
        off = int(rng.integers(0, n - length + 1))
        q = batch_query(int(rng.integers(-10**6, 10**6)))
        plan = compile_plan(q)
        sl = {c: dev[c].slice(off, length) for c in ("key", "value", "x")}
        work.append((df, off, length, q, (plan, [sl[k] for k in plan.keys], [sl[c] for c in plan.columns])))
    import threading
    import time

    results = [None] * n_q
    with Server(gpu_ctx, lanes=2, max_batch=64) as srv:
        tickets = [0] * n_q
        t0 = time.perf_counter()
        for i in range(n_q):
            d = t0 + arrivals[i] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets[i] = srv.submit(work[i][4])
        for i in range(n_q):
            results[i] = srv.wait(tickets[i])
        stats = srv.stats()
    assert stats["completed"] == n_q and stats["max_batch"] > 1
    for i, (df, off, length, q, job) in enumerate(work):
        got = results[i]
        alone = gpu_ctx.groupby(*job)
        na = len(job[0].aggs)
        assert got.groups == alone.groups, i
        for (va, ma), (vb, mb) in zip(_result_arrays(got, 1, na), _result_arrays(alone, 1, na)):
            assert np.array_equal(ma, mb), i
            assert np.array_equal(va[ma].view(np.uint8), vb[mb].view(np.uint8)), i
        if i % 25 == 0:
            host = DataFrame([Series(c, Column(df.column(c).column.dtype, df.column(c).column.values[off:off + length],
                                               df.column(c).column.is_valid()[off:off + length],
                                               df.column(c).column.labels)) for c in df.columns])
            lf = type(q)(host, q.filters, q.keys, q.aggs, q.maintain_order, True)
            exp, _ = oracle.collect(lf)
            assert_frames_close(result_frame(got, job[0], job[1]), exp, 1)
        alone.release()
        got.release()
