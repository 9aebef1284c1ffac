"""Benchmark of the filtered group-by path on one MI355X.

Workloads (columns generated on the device, resident in HBM):
  zipf_200m (default): 200,000,000 rows — key Int32 Zipf(1.1) over 1e6
    values, a Int64 uniform [-1e9, 1e9], x Float64 N(0,1) with 5% nulls,
    y Int64 uniform [0, 1e6]; filter(a > 0).group_by(key)
    .agg(sum/count/min/max(y), sum/mean/var/min(x)).
  dict_500m: 500,000,000 rows — k1 dictionary String of 1000 labels, k2
    Int64 uniform [0, 1e7), v Float64 lognormal(0, 2) with 2% nulls;
    filter(k2 % 3 != 0).group_by(k1, k2).agg(count/sum/mean(v)).
  batch_4096: 4096 independent queries, rows log-uniform in [1e3, 1e7]
    (about 4.45e9 rows in all), key Int64 of cardinality log-uniform in
    [1, 1e5], value Int64 uniform [-1e6, 1e6], x Float64 N(0,1) with 5%
    nulls; filter(value > t_q).group_by(key).agg(len, sum/max(value),
    sum/mean(x)); one dfr_groupby_batch call per step, against one
    dfr_groupby call per query.
  serve: Poisson arrivals at 2000 queries/s of that query distribution over
    row windows of 64 device-resident tables (1e6 to 1e8 rows; Int64,
    dictionary String and Float64 keys) through the serving loop (dfr_server_*); p50 and
    p99 latency and queries/s, against serial one-call-at-a-time execution
    of the same arrivals.
A step is one whole query through the C ABI (dfr_groupby: table set-up,
fused scan, partition passes, compaction, finish) with its inputs already on
the device and its results left on the device. Prints one JSON line.

The default run (zipf_200m) also measures, after its own line's figures,
dict_500m, batch_4096 (graph captured) and serve at shorter settings, and
reports them as the extra keys "dict_500m", "batch_4096" and "serve"
(--extras 0 leaves them out).

  python bench.py [--workload W] [--steps K] [--warmup W] [--rows N] [--cpu-sample S] [--extras 0|1]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


from dataframe_mojo_amd.workloads import (  # noqa: E402  (the workloads bench.py times)
    batch_query, dict_query, host_dict, host_zipf, wrap_dict, wrap_zipf, zipf_query)


def traffic_of(workload: str, kernel: str, rows: int):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    counter summary (two rocprofv3 --pmc runs, FETCH_SIZE and WRITE_SIZE,
    scaled by their readings for a copy of known size), when that summary
    covers this kernel at this size; else None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if t.get("kernel") != kernel or t.get("rows") not in (None, rows):
        return None
    return t["traffic"]


def cpu_threads() -> int:
    """The host threads a CPU baseline uses: the machine's, at most 16 (a
    one-GPU box's share of its host)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def cpu_baseline(spec, cols, sample: int, threads: int):
    """The oracle's C restatement of the reference CPU algorithm on the first
    `sample` rows of the same device columns, copied to the host untimed:
    with threads > 1, morsel-parallel workers with per-worker hash
    partitions and partition-parallel aggregation and finish
    (dfo_groupby_parallel). Returns rows/s and the best time."""
    from oracle import oracle

    n = sample
    frame = spec["host"](cols, n)
    q = spec["query"]()
    lf = type(q)(frame, q.filters, q.keys, q.aggs, q.maintain_order, True)
    times = []
    for _ in range(2):
        t0 = time.perf_counter()
        out, _ = oracle.collect(lf, threads=threads)
        times.append(time.perf_counter() - t0)
    return n / min(times), min(times), frame, lf, out


def dd_error(ctx, spec, cols, frame, lf, cpu_out):
    """Per-group error of the Float64 sum and variance against a
    double-double reference (double-double sums, two-pass variances), for
    the device result and the CPU (oracle) result on the same sample:
    max over groups of |r - dd| / max(1, |dd|)."""
    from dataframe_mojo_amd.plan import compile_plan
    from dataframe_mojo_amd.runtime import result_frame
    from oracle import oracle

    n = frame.height()
    dev = spec["wrap"](ctx, cols, n)
    plan = compile_plan(lf)
    res = ctx.groupby(plan, [dev[k] for k in plan.keys], [dev[c] for c in plan.columns])
    gpu_out = result_frame(res, plan, [dev[k] for k in plan.keys])
    res.release()
    keep = oracle.kept_rows(frame, lf.filters)
    key = frame.column(spec["dd"]["key"]).column
    x = frame.column(spec["dd"]["x"]).column
    uniq, gid = np.unique(key.values[keep], return_inverse=True)
    hi, lo, cnt, var = oracle.dd_moments(gid, x.values[keep], x.is_valid()[keep], len(uniq), ddof=1)
    dd_sum = hi + lo
    out = {"sample_rows": n, "groups": int(len(uniq))}
    for who, fr in (("gpu", gpu_out), ("cpu", cpu_out)):
        kc = fr.column(spec["dd"]["key"]).column
        order = np.argsort(kc.values, kind="stable")
        assert np.array_equal(kc.values[order], uniq), who
        for agg, ref, mask in (("sum", dd_sum, np.ones(len(uniq), bool)), ("var", var, cnt > 1)):
            c = fr.column(spec["dd"][agg]).column
            vals = np.where(c.is_valid(), c.values, np.nan)[order]
            err = np.abs(vals[mask] - ref[mask]) / np.maximum(1.0, np.abs(ref[mask]))
            out[f"{who}_{agg}_max"] = float(err.max()) if err.size else 0.0
    return out


def bench_batch(args) -> dict:
    """batch_4096: 4096 independent filtered group-by queries (rows
    log-uniform in [1e3, 1e7], key cardinality log-uniform in [1, 1e5],
    threshold uniform) as one dfr_groupby_batch call per step, against the
    same queries one dfr_groupby call each."""
    from dataframe_mojo_amd import datagen, runtime
    from dataframe_mojo_amd.plan import compile_plan
    from dataframe_mojo_amd.runtime import PreparedBatch

    params = datagen.batch_4096_params(args.queries, args.seed)
    total_rows = sum(p["rows"] for p in params)
    log(f"generating {len(params)} queries' columns on the device: {total_rows} rows")
    cols = datagen.batch_4096_device(params, args.seed)
    ctx = runtime.context(0)
    ctx.tune("batch_graph", 1 if args.batch_graph else 0)
    jobs, prepared = [], []
    for p, c in zip(params, cols):
        n = p["rows"]
        dev = {"key": ctx.wrap_array(c["key"], length=n), "value": ctx.wrap_array(c["value"], length=n),
               "x": ctx.wrap_array(c["x"], c["x.valid"], length=n)}
        plan = compile_plan(batch_query(p["threshold"]))
        job = (plan, [dev[k] for k in plan.keys], [dev[k] for k in plan.columns])
        jobs.append(job)
        prepared.append(ctx.prepare(*job))
    batch = PreparedBatch(jobs)

    def release(rs):
        for r in rs:
            if not isinstance(r, Exception):
                r.release()

    log("warm-up")
    for _ in range(args.warmup):
        release(batch.run(ctx))
    log(f"timing {args.steps} batched steps")
    times, dev_ms, phases = [], [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        out = batch.run(ctx)
        times.append(time.perf_counter() - t0)
        first = next(r for r in out if not isinstance(r, Exception))
        dev_ms.append(first.device_ms)
        phases.append(first.phase_ms)
        groups = sum(r.groups for r in out if not isinstance(r, Exception))
        release(out)
    batch_ms = statistics.mean(times) * 1e3
    log(f"batched: {batch_ms:.2f} ms/step (device {statistics.mean(dev_ms):.2f} ms), {groups} groups")
    stream_ms = None
    if args.batch_graph:
        # the same batch launched kernel by kernel on the stream, for comparison
        ctx.tune("batch_graph", 0)
        release(batch.run(ctx))
        st = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            out = batch.run(ctx)
            st.append(time.perf_counter() - t0)
            release(out)
        stream_ms = statistics.mean(st) * 1e3
        ctx.tune("batch_graph", 1)
        log(f"batched without the graph: {stream_ms:.2f} ms/step")

    # plain: one launch sequence per query, the same queries, warm
    plain, plain_ms = [], None
    if args.plain:
        for q in prepared[:64]:
            q.run().release()
        log("timing the plain one-launch-per-query pass")
        t0 = time.perf_counter()
        for q in prepared:
            plain.append(q.run())
        plain_ms = (time.perf_counter() - t0) * 1e3
        log(f"plain: {plain_ms:.1f} ms for {len(prepared)} queries")

    # identical results: every 16th query, batched vs plain, bit for bit
    out = batch.run(ctx)
    checked = 0
    for i in range(0, len(jobs) if plain else 0, 16):
        a, b = out[i], plain[i]
        if isinstance(a, Exception):
            continue
        assert a.groups == b.groups, (i, a.groups, b.groups)
        for k in range(len(jobs[i][1])):
            ka, kb = a.key(k)[1], b.key(k)[1]
            oa, ob = np.argsort(ka, kind="stable"), np.argsort(kb, kind="stable")
            assert np.array_equal(ka[oa], kb[ob]), i
        for j in range(len(jobs[i][0].aggs)):
            _, va, ma = a.agg(j)
            _, vb, mb = b.agg(j)
            assert np.array_equal(ma[oa], mb[ob]) and np.array_equal(va[oa].view(np.uint8), vb[ob].view(np.uint8)), (i, j)
        checked += 1
    release(out)
    release(plain)
    log(f"batched == plain on {checked} queries")
    ctx.tune("batch_graph", 0)

    bytes_per_row = 8 + 8 + 8 + 1 / 8  # key, value, x, x validity
    out_bytes = groups * (8 + 5 * 8 + 3 / 8)
    agg_ms = statistics.mean(p[1] for p in phases) if not args.batch_graph else None
    kernel_ms = agg_ms if agg_ms else statistics.mean(dev_ms)
    achieved = total_rows * bytes_per_row / (kernel_ms / 1e3)
    line = {
        "metric": "filtered group-by rows/s (batched)",
        "value": total_rows / (batch_ms / 1e3),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": batch_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64+f64",
        "data": "synthetic (seeded counter-based generator, device resident)",
        "config": {"workload": "batch_4096", "queries": len(params), "rows": total_rows, "groups": groups,
                   "rows_per_query": "log-uniform [1e3, 1e7]", "key_cardinality": "log-uniform [1, 1e5]",
                   "filter": "value > t, t uniform in [-1e6, 1e6]", "aggs": "len, sum/max(value), sum/mean(x)",
                   "graph": bool(args.batch_graph)},
        "batched_vs_plain": plain_ms / batch_ms if plain_ms else None,
        "plain_ms": plain_ms,
        "stream_ms": stream_ms,
        "identical_results_checked": checked,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None,
                     "kernel": "k_batch_aggregate" if agg_ms else "whole batch", "kernel_ms": kernel_ms,
                     "bytes_per_row": bytes_per_row},
        "query_hbm": {"alg_bytes": total_rows * bytes_per_row + out_bytes,
                      "GB/s": (total_rows * bytes_per_row + out_bytes) / (batch_ms / 1e3) / 1e9},
        "phases_ms": dict(zip(["init", "aggregate", "compact", "finish"],
                              [statistics.mean(p[i] for p in phases) for i in range(4)])),
        "cpu_baseline": None,
    }
    return line


def bench_serve(args) -> dict:
    """serve: Poisson arrivals (args.rate queries/s for args.duration s) of
    the batched workload's query distribution over 64 device-resident
    tables, through the serving loop, against the same arrivals served one
    dfr_groupby call at a time."""
    import threading

    from dataframe_mojo_amd import datagen, runtime
    from dataframe_mojo_amd.plan import compile_plan
    from dataframe_mojo_amd.runtime import Server

    rng = np.random.default_rng(args.seed)
    n_tables = args.tables
    rows = np.exp(rng.uniform(np.log(1e6), np.log(args.table_rows_max), n_tables)).astype(np.int64)
    card = np.exp(rng.uniform(0.0, np.log(1e5), n_tables)).astype(np.int64).clip(1)
    log(f"generating {n_tables} tables on the device: {int(rows.sum())} rows")
    ctx = runtime.context(0)
    ctx.tune("batch_graph", args.serve_graph)  # the lanes take the context's tuning
    labels = datagen.dict_labels()
    kinds = [{"int": 0, "dict": 1, "float": 2}[k] for k in args.serve_keys.split(",")]
    tabs = []
    for t in range(n_tables):
        n, base = int(rows[t]), 5000 + 8 * t
        kind = kinds[t % len(kinds)]  # 0 Int64, 1 dictionary String, 2 Float64 keys
        dict_key = kind == 1
        hi = min(int(card[t]), 1000) if dict_key else int(card[t])
        if kind == 2:
            key = datagen.device_float_keys(n, args.seed, base, hi)
        else:
            key = datagen.device_uniform(n, args.seed, base, 0, hi - 1, dtype=np.uint32 if dict_key else np.int64)
        value = datagen.device_uniform(n, args.seed, base + 1, -1_000_000, 1_000_000)
        x = datagen.device_normal(n, args.seed, base + 2, 0.0, 1.0)
        xv = datagen.device_validity(n, args.seed, base + 3, 0.05)
        dev = {"key": ctx.wrap_array(key, labels=labels if dict_key else None, dict_codes=dict_key),
               "value": ctx.wrap_array(value), "x": ctx.wrap_array(x, xv)}
        tabs.append((n, dev, (key, value, x, xv)))
    # the query stream: arrival times and (table, window, threshold)
    n_q = int(args.rate * args.duration)
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, n_q))
    jobs = []
    total_rows = 0
    for i in range(n_q):
        t = int(rng.integers(0, n_tables))
        n, dev, _ = tabs[t]
        length = int(np.exp(rng.uniform(np.log(1e3), np.log(min(1e7, n)))))
        off = int(rng.integers(0, n - length + 1))
        thr = int(rng.integers(-1_000_000, 1_000_001))
        plan = compile_plan(batch_query(thr))
        sl = {c: dev[c].slice(off, length) for c in ("key", "value", "x")}
        jobs.append(ctx.prepare(plan, [sl[k] for k in plan.keys], [sl[c] for c in plan.columns]))
        total_rows += length
    log(f"{n_q} queries over {total_rows} rows, {args.rate:.0f}/s for {args.duration:.1f} s")

    # warm both paths
    for j in jobs[:32]:
        j.run().release()

    def percentiles(lat):
        return float(np.percentile(lat, 50)), float(np.percentile(lat, 99))

    def served(arr, js):
        """The serving loop under the arrival times `arr`: a submitter thread
        keeps the schedule, waiter threads collect the server's latencies."""
        nq = len(js)
        lat = np.zeros(nq)
        with Server(ctx, lanes=args.lanes, max_batch=args.max_batch, budget=args.serve_budget) as srv:
            tickets = [0] * nq
            ready = [threading.Event() for _ in range(nq)]
            t0 = time.perf_counter()

            def submitter():
                for i in range(nq):
                    d = t0 + arr[i] - time.perf_counter()
                    if d > 0:
                        time.sleep(d)
                    tickets[i] = srv.submit(js[i])
                    ready[i].set()

            def waiter(k, m):
                for i in range(k, nq, m):
                    ready[i].wait()
                    r = srv.wait(tickets[i])
                    lat[i] = r.latency_ms
                    r.release()

            th = [threading.Thread(target=submitter)] + [threading.Thread(target=waiter, args=(k, 4)) for k in range(4)]
            for x in th:
                x.start()
            for x in th:
                x.join()
            secs = time.perf_counter() - t0
            stats = srv.stats()
        p = percentiles(lat)
        return {"p50_ms": p[0], "p99_ms": p[1], "qps": nq / secs, "batches": stats["batches"],
                "max_batch": stats["max_batch"], "concurrent_lanes_max": stats["concurrent_max"]}

    def serial(arr, js):
        """One dfr_groupby at a time under the same arrivals (queueing
        included)."""
        nq = len(js)
        lat = np.zeros(nq)
        t0 = time.perf_counter()
        for i in range(nq):
            now = time.perf_counter() - t0
            if now < arr[i]:
                time.sleep(arr[i] - now)
            js[i].run().release()
            lat[i] = (time.perf_counter() - t0 - arr[i]) * 1e3
        p = percentiles(lat)
        return {"p50_ms": p[0], "p99_ms": p[1], "qps": nq / (time.perf_counter() - t0)}

    srv_res = served(arrivals, jobs)
    s50, s99 = srv_res["p50_ms"], srv_res["p99_ms"]
    served_s = n_q / srv_res["qps"]
    log(f"served: {srv_res}")
    plain_res = serial(arrivals, jobs)
    log(f"serial plain: {plain_res}")
    # overload: the same queries arriving `overload` times faster
    over = {}
    if args.overload > 1:
        arr2 = arrivals / args.overload
        over = {"rate": args.rate * args.overload, "served": served(arr2, jobs), "serial_plain": serial(arr2, jobs)}
        log(f"overload x{args.overload}: {over}")

    # the same result alone and in the mix: a sample, served again in a mix
    sample = list(range(0, n_q, max(1, n_q // 64)))
    checked = 0
    with Server(ctx, lanes=args.lanes, max_batch=args.max_batch) as srv:
        ts = [srv.submit(jobs[i]) for i in sample]
        for i, t in zip(sample, ts):
            a = srv.wait(t)
            b = jobs[i].run()
            ka, kb = a.key(0)[1], b.key(0)[1]
            oa, ob = np.argsort(ka, kind="stable"), np.argsort(kb, kind="stable")
            assert a.groups == b.groups and np.array_equal(ka[oa], kb[ob]), i
            for j in range(len(jobs[i].plan.aggs)):
                _, va, ma = a.agg(j)
                _, vb, mb = b.agg(j)
                assert np.array_equal(ma[oa], mb[ob]) and np.array_equal(va[oa].view(np.uint8), vb[ob].view(np.uint8))
            a.release()
            b.release()
            checked += 1
    log(f"served == alone on {checked} queries")
    line = {
        "metric": "served filtered group-by p99 latency",
        "value": s99,
        "unit": "ms",
        "n_gpus": 1,
        "steps": n_q,
        "warmup": 32,
        "ms_per_step": served_s * 1e3 / n_q,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64+f64",
        "data": "synthetic (seeded counter-based generator, device resident)",
        "config": {"workload": "serve", "tables": n_tables, "table_rows": "log-uniform [1e6, %.0e]" % args.table_rows_max,
                   "keys": "Int64, dictionary String and Float64 (cardinality log-uniform [1, 1e5]), in turn",
                   "arrivals": f"Poisson {args.rate:.0f}/s for {args.duration} s", "queries": n_q,
                   "query": "window rows log-uniform [1e3, 1e7], value > t, len/sum/max(value), sum/mean(x)",
                   "lanes": args.lanes, "max_batch": args.max_batch, "budget_bytes": args.serve_budget,
                   "graph": bool(args.serve_graph)},
        "served": {**srv_res, "rows_per_s": total_rows / served_s},
        "serial_plain": plain_res,
        "overload": over,
        "identical_results_checked": checked,
        "roofline": None,
        "cpu_baseline": None,
    }
    return line


WORKLOADS = {
    "zipf_200m": {
        "rows": 200_000_000, "query": zipf_query, "wrap": wrap_zipf, "host": host_zipf,
        "bytes_per_row": 4 + 8 + 8 + 8 + 1 / 8,  # key, a, x, y, x validity
        "out_bytes_per_group": 4 + 8 * 8 + 7 / 8,  # key; 8 aggregates; 7 nullable
        "config": {"key": "Int32 Zipf(1.1) over 1e6", "filter": "a > 0",
                   "aggs": "sum/count/min/max(y), sum/mean/var/min(x)"},
        "dd": {"key": "key", "x": "x", "sum": "x_sum", "var": "x_var"},
    },
    "dict_500m": {
        "rows": 500_000_000, "query": dict_query, "wrap": wrap_dict, "host": host_dict,
        "bytes_per_row": 4 + 8 + 8 + 1 / 8,  # k1 codes, k2, v, v validity
        "out_bytes_per_group": 4 + 8 + 8 * 3 + 2 / 8,  # k1 code, k2; count, sum, mean (mean nullable)
        "config": {"key": "(dictionary String of 1000 labels, Int64 uniform [0, 1e7))", "filter": "k2 % 3 != 0",
                   "aggs": "count/sum/mean(v), v Float64 lognormal(0, 2), 2% nulls"},
    },
}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="zipf_200m", choices=sorted([*WORKLOADS, "batch_4096", "serve"]))
    ap.add_argument("--rate", type=float, default=2000.0, help="serve: arrivals per second")
    ap.add_argument("--duration", type=float, default=5.0, help="serve: seconds of arrivals")
    ap.add_argument("--tables", type=int, default=64, help="serve: device-resident tables")
    ap.add_argument("--table-rows-max", type=float, default=1e8, help="serve: largest table")
    ap.add_argument("--lanes", type=int, default=2, help="serve: server lanes")
    ap.add_argument("--max-batch", type=int, default=256, help="serve: queries per batch at most")
    ap.add_argument("--serve-budget", type=int, default=0, help="serve: admission budget in bytes (0: none)")
    ap.add_argument("--overload", type=float, default=4.0, help="serve: also run the arrivals this many times faster")
    ap.add_argument("--serve-keys", default="int,dict,float", help="serve: key types of the tables, in turn")
    ap.add_argument("--serve-graph", type=int, default=0, help="serve: lanes capture each batch as a HIP graph")
    ap.add_argument("--queries", type=int, default=4096, help="batch_4096: queries per batch")
    ap.add_argument("--batch-graph", type=int, default=0, help="batch_4096: run the batch as a captured HIP graph")
    ap.add_argument("--plain", type=int, default=1,
                    help="batch_4096: also time one dfr_groupby call per query and compare results (0: batch only)")
    ap.add_argument("--rows", type=int, default=0, help="rows (default: the workload's own)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=50_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--extras", type=int, default=1,
                    help="zipf_200m: also measure dict_500m, batch_4096 (graph) and serve, as extra keys")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="first run one streaming copy of this many bytes (a known-size kernel for byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        return 2

    if args.workload == "batch_4096":
        print(json.dumps(bench_batch(args)), flush=True)
        return 0
    if args.workload == "serve":
        print(json.dumps(bench_serve(args)), flush=True)
        return 0

    from dataframe_mojo_amd import datagen

    if args.copy_calibration:
        rate = datagen.stream_rate(args.copy_calibration, copy=True, reps=1)
        log(f"calibration copy of {args.copy_calibration} bytes: {rate / 1e9:.0f} GB/s (read + write)")
    line, ctx = bench_query(args, args.workload, args.steps, args.warmup, args.rows, not args.no_cpu)
    if args.extras and args.workload == "zipf_200m":
        line.update(extras(args, ctx))
    print(json.dumps(line), flush=True)
    return 0


def extras(args, ctx) -> dict:
    """The other configurations at shorter settings, each summarised: the
    500M-row dictionary query, the 4096-query batch with its launch sequence
    captured as a graph, and the serving loop."""
    import gc

    out = {}
    d, _ = bench_query(args, "dict_500m", 5, 2, 0, False)
    out["dict_500m"] = {"ms_per_query": d["ms_per_step"], "rows_per_s": d["value"], "groups": d["config"]["groups"],
                        "query_hbm_frac": d["query_hbm"]["frac"], "scan_frac": d["roofline"]["frac"],
                        "phases_ms": d["phases_ms"], "ordered_ms_per_query": d["ordered_ms_per_query"]}
    gc.collect()
    ctx.trim()
    b_args = argparse.Namespace(**{**vars(args), "steps": 3, "warmup": 1, "batch_graph": 1})
    b = bench_batch(b_args)
    out["batch_4096"] = {"batched_ms": b["ms_per_step"], "plain_ms": b["plain_ms"],
                         "batched_vs_plain": b["batched_vs_plain"], "graph": True, "rows": b["config"]["rows"],
                         "stream_ms": b["stream_ms"],
                         "identical_results_checked": b["identical_results_checked"],
                         "rows_per_s": b["value"], "query_hbm_GBps": b["query_hbm"]["GB/s"]}
    gc.collect()
    ctx.trim()
    s_args = argparse.Namespace(**{**vars(args), "duration": 2.0, "overload": 4.0})
    sv = bench_serve(s_args)
    out["serve"] = {"rate": s_args.rate, "queries": sv["config"]["queries"], "served": sv["served"],
                    "serial_plain": sv["serial_plain"], "overload": sv["overload"],
                    "identical_results_checked": sv["identical_results_checked"]}
    gc.collect()
    ctx.trim()
    return out


def bench_query(args, workload: str, steps: int, warmup: int, rows: int, with_cpu: bool):
    """One-query workloads (zipf_200m, dict_500m): the JSON line's fields."""
    import gc

    from dataframe_mojo_amd import datagen, runtime
    from dataframe_mojo_amd.plan import compile_plan

    spec = WORKLOADS[workload]
    n = rows or spec["rows"]
    log(f"generating {workload} columns on the device: {n} rows")
    cols = getattr(datagen, f"{workload}_device")(n, args.seed)
    ctx = runtime.context(0)
    dev = spec["wrap"](ctx, cols, n)
    dplan = compile_plan(spec["query"]())
    keys = [dev[k] for k in dplan.keys]
    inputs = [dev[c] for c in dplan.columns]
    log("warm-up")
    groups = 0
    for _ in range(warmup):
        r = ctx.groupby(dplan, keys, inputs)
        groups = r.groups
        r.release()
    log(f"timing {steps} steps ({groups} groups)")
    phases = []
    jit_before = ctx.observe()["jit_launches"]
    t0 = time.perf_counter()
    for _ in range(steps):
        r = ctx.groupby(dplan, keys, inputs)
        phases.append(r.phase_ms)
        r.release()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / steps
    # pass A ran as the kernel compiled for the query's shape, or precompiled
    kernel = "dfr_scan_jit" if ctx.observe()["jit_launches"] > jit_before else "k_scan_spill"
    scan_ms = statistics.mean(p[1] for p in phases)
    bytes_per_row = spec["bytes_per_row"]
    scan_bytes = bytes_per_row * n
    achieved = scan_bytes / (scan_ms / 1e3)
    alg_bytes = scan_bytes + spec["out_bytes_per_group"] * groups
    log(f"{ms_per_step:.3f} ms/step, scan {scan_ms:.3f} ms, {achieved / 1e9:.1f} GB/s")

    # the same query under maintain_order (ordered output) for reference
    dplan_o = compile_plan(spec["query"](True))
    ctx.groupby(dplan_o, keys, inputs).release()  # warm-up: its shape's kernels are compiled here
    ordered = []
    for _ in range(max(3, steps // 4)):
        r = ctx.groupby(dplan_o, keys, inputs)
        ordered.append(r.device_ms)
        r.release()

    cpu = None
    dd = None
    if with_cpu:
        sample = min(args.cpu_sample, n)
        threads = cpu_threads()
        log(f"CPU baseline on {sample} rows, {threads} threads")
        rate, secs, frame, lf, cpu_out = cpu_baseline(spec, cols, sample, threads)
        cpu = {"value": rate, "unit": "rows/s", "cores": threads, "kind": "port",
               "sample": f"first {sample} rows of the same workload; restated reference CPU algorithm in C "
                         f"(oracle dfo_groupby_parallel: {threads} workers over contiguous rows, per-worker hash "
                         f"partitions, partition-parallel aggregation and finish), host filter included, "
                         f"best of 2 ({secs:.2f} s)"}
        if spec.get("dd"):
            log("per-group error against double-double on the same sample")
            dd = dd_error(ctx, spec, cols, frame, lf, cpu_out)
            log(f"double-double error: {dd}")
        del frame, lf, cpu_out

    line = {
        "metric": "filtered group-by rows/s",
        "value": n / (ms_per_step / 1e3),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64+f64",
        "data": "synthetic (seeded counter-based generator, device resident)",
        "config": {"workload": workload, "rows": n, "groups": groups, **spec["config"], "maintain_order": False},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic_of(workload, kernel, n),
                     "kernel": kernel, "kernel_ms": scan_ms, "bytes_per_row": bytes_per_row,
                     "alg_bytes": scan_bytes},
        "query_hbm": {"alg_bytes": alg_bytes, "GB/s": alg_bytes / (ms_per_step / 1e3) / 1e9,
                      "frac": alg_bytes / (ms_per_step / 1e3) / HBM_PEAK},
        "phases_ms": {k: statistics.mean(p[i] for p in phases)
                      for i, k in enumerate(["init", "scan", "partition_and_order", "finish"])},
        "ordered_ms_per_query": statistics.mean(ordered),
        "cpu_baseline": cpu,
    }
    if dd is not None:
        line["dd_error"] = dd
    del dev, keys, inputs, cols
    gc.collect()
    return line, ctx


if __name__ == "__main__":
    sys.exit(main())
