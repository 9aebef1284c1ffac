"""The serving loop (dfr_server_*): queries submitted concurrently from
several threads, admitted by a device-memory budget and run in overlapping
batches, each give the same result as the query alone through dfr_groupby,
bit for bit (rows compared sorted by key: group order is unspecified)."""
import threading

import numpy as np
import pytest

from dataframe_mojo_amd import Column, DataFrame, Series, _lib, col, len_
from dataframe_mojo_amd.plan import compile_plan
from dataframe_mojo_amd.runtime import Server, result_frame
from oracle import oracle

from test_gpu_batch import frame, frames_equal_bitwise, AGG_SETS
from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu


def tables(gpu_ctx, count=4):
    """Device tables of 20k-200k rows and their host frames."""
    out = []
    for t in range(count):
        df = frame(20000 * (t + 1) + 977 * t, 50 + t, [5, 300, 4000, 40000][t % 4])
        dev = {name: gpu_ctx.upload(df.column(name).column) for name in df.columns}
        out.append((df, dev))
    return out


def window_query(df, dev, rng, i):
    """A query over a random row window of a table (slices of the device
    columns, the matching rows of the host frame)."""
    n = df.height()
    length = int(rng.integers(1, n))
    off = int(rng.integers(0, n - length + 1))
    host = DataFrame([Series(c, Column(df.column(c).column.dtype, df.column(c).column.values[off:off + length],
                                       df.column(c).column.is_valid()[off:off + length],
                                       df.column(c).column.labels)) for c in df.columns])
    thr = int(rng.integers(-10**6, 10**6))
    keys = ["k"] if i % 4 else ["k", "s"]
    lf = host.filter(col("v") > thr).group_by(*keys).agg(AGG_SETS[i % len(AGG_SETS)])
    plan = compile_plan(lf)
    sl = {name: dev[name].slice(off, length) for name in {*plan.keys, *plan.columns}}
    return lf, (plan, [sl[k] for k in plan.keys], [sl[c] for c in plan.columns])


@pytest.mark.parametrize("lanes,budget", [(1, 0), (2, 0), (3, 1 << 20)])
def test_served_results_equal_alone(lanes, budget, gpu_ctx):
    rng = np.random.default_rng(lanes)
    tabs = tables(gpu_ctx)
    work = [window_query(*tabs[i % len(tabs)], rng, i) for i in range(48)]
    tickets = [None] * len(work)
    results = [None] * len(work)
    with Server(gpu_ctx, lanes=lanes, max_batch=8, budget=budget) as srv:
        def client(part):
            for i in part:
                tickets[i] = srv.submit(work[i][1])
            for i in part:
                results[i] = srv.wait(tickets[i])

        threads = [threading.Thread(target=client, args=(range(k, len(work), 4),)) for k in range(4)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        stats = srv.stats()
        assert stats["submitted"] == stats["completed"] == len(work)
        assert stats["batches"] >= 1 and stats["queued"] == 0 and stats["in_flight_bytes"] == 0
        if lanes > 1 and budget == 0:
            assert stats["max_batch"] > 1  # arrivals during a batch were batched together
        for (lf, job), got in zip(work, results):
            plan, keys, cols = job
            alone = gpu_ctx.groupby(plan, keys, cols)
            frames_equal_bitwise(result_frame(got, plan, keys), result_frame(alone, plan, keys), len(plan.keys))
            assert got.latency_ms > 0
            alone.release()
            got.release()


def test_served_window_matches_oracle(gpu_ctx):
    rng = np.random.default_rng(9)
    tabs = tables(gpu_ctx, 2)
    work = [window_query(*tabs[i % 2], rng, i) for i in range(6)]
    with Server(gpu_ctx, lanes=2) as srv:
        ts = [srv.submit(job) for _, job in work]
        for (lf, (plan, keys, _)), t in zip(work, ts):
            r = srv.wait(t)
            exp, _ = oracle.collect(lf)
            assert_same(result_frame(r, plan, keys), exp, False, ("x_min",), len(plan.keys))
            r.release()


def test_server_errors_and_tickets(gpu_ctx):
    bad = DataFrame([Series("k", Column.int64([1, 1])), Series("v", Column.int64([2**63 - 1, 1]))])
    lf = bad.lazy().group_by("k").agg(col("v").sum())
    plan = compile_plan(lf)
    dev = {n: gpu_ctx.upload(bad.column(n).column) for n in bad.columns}
    with Server(gpu_ctx, lanes=1) as srv:
        t = srv.submit((plan, [dev["k"]], [dev["v"]]))
        with pytest.raises(_lib.DataFault, match="Int64 expression sum overflow"):
            srv.wait(t)
        with pytest.raises(_lib.DfrError):
            srv.wait(t)  # answered once


def test_result_released_after_server_close(gpu_ctx):
    # a served result keeps its lane's context alive: it can be read and
    # released after the server is gone (the lane goes with the last result)
    rng = np.random.default_rng(12)
    tabs = tables(gpu_ctx, 2)
    work = [window_query(*tabs[i % 2], rng, i) for i in range(5)]
    srv = Server(gpu_ctx, lanes=2, max_batch=4)
    ts = [srv.submit(job) for _, job in work]
    got = [srv.wait(t) for t in ts[:4]]  # the last ticket is never waited for
    srv.close()
    for (lf, (plan, keys, _)), r in zip(work, got):
        exp, _ = oracle.collect(lf)
        assert_same(result_frame(r, plan, keys), exp, False, ("x_min",), len(plan.keys))
        r.release()
