"""Benchmark of the filtered group-by path on one MI355X.

Workloads (columns generated on the device, resident in HBM):
  zipf_200m (default): 200,000,000 rows — key Int32 Zipf(1.1) over 1e6
    values, a Int64 uniform [-1e9, 1e9], x Float64 N(0,1) with 5% nulls,
    y Int64 uniform [0, 1e6]; filter(a > 0).group_by(key)
    .agg(sum/count/min/max(y), sum/mean/var/min(x)).
  dict_500m: 500,000,000 rows — k1 dictionary String of 1000 labels, k2
    Int64 uniform [0, 1e7), v Float64 lognormal(0, 2) with 2% nulls;
    filter(k2 % 3 != 0).group_by(k1, k2).agg(count/sum/mean(v)).
A step is one whole query through the C ABI (dfr_groupby: table set-up,
fused scan, partition passes, compaction, finish) with its inputs already on
the device and its results left on the device. Prints one JSON line.

  python bench.py [--workload W] [--steps K] [--warmup W] [--rows N] [--cpu-sample S]
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


def zipf_query(maintain_order: bool = False):
    from dataframe_mojo_amd import Column, DataFrame, Series, col

    schema = DataFrame([
        Series("key", Column.int32([])),
        Series("a", Column.int64([])),
        Series("x", Column.float64([])),
        Series("y", Column.int64([])),
    ])
    return (schema.filter(col("a") > 0).group_by("key", maintain_order=maintain_order).agg(
        col("y").sum().alias("y_sum"), col("y").count().alias("y_count"), col("y").min().alias("y_min"),
        col("y").max().alias("y_max"), col("x").sum().alias("x_sum"), col("x").mean().alias("x_mean"),
        col("x").var().alias("x_var"), col("x").min().alias("x_min")))


def dict_query(maintain_order: bool = False):
    from dataframe_mojo_amd import Column, DataFrame, Series, col

    schema = DataFrame([
        Series("k1", Column.dictionary([], [])),
        Series("k2", Column.int64([])),
        Series("v", Column.float64([])),
    ])
    return (schema.filter(col("k2") % 3 != 0).group_by("k1", "k2", maintain_order=maintain_order).agg(
        col("v").count().alias("v_count"), col("v").sum().alias("v_sum"), col("v").mean().alias("v_mean")))


def wrap_dict(ctx, cols, n):
    from dataframe_mojo_amd import datagen

    return {
        "k1": ctx.wrap_array(cols["k1"], labels=datagen.dict_labels(), dict_codes=True, length=n),
        "k2": ctx.wrap_array(cols["k2"], length=n),
        "v": ctx.wrap_array(cols["v"], cols["v.valid"], length=n),
    }


def host_dict(cols, n):
    from dataframe_mojo_amd import Column, DataFrame, Series, datagen

    return DataFrame([
        Series("k1", Column.dictionary(cols["k1"].numpy(n).astype(np.int64), datagen.dict_labels())),
        Series("k2", Column.int64(cols["k2"].numpy(n))),
        Series("v", Column.float64(cols["v"].numpy(n), datagen.valid_from_words(cols["v.valid"], n))),
    ])


def host_zipf(cols, n):
    from dataframe_mojo_amd import Column, DataFrame, Series, datagen

    return DataFrame([
        Series("key", Column.int32(cols["key"].numpy(n))),
        Series("a", Column.int64(cols["a"].numpy(n))),
        Series("x", Column.float64(cols["x"].numpy(n), datagen.valid_from_words(cols["x.valid"], n))),
        Series("y", Column.int64(cols["y"].numpy(n))),
    ])


def wrap_zipf(ctx, cols, n):
    return {
        "key": ctx.wrap_array(cols["key"], length=n),
        "a": ctx.wrap_array(cols["a"], length=n),
        "x": ctx.wrap_array(cols["x"], cols["x.valid"], length=n),
        "y": ctx.wrap_array(cols["y"], length=n),
    }


def cpu_baseline(spec, cols, sample: int):
    """The oracle (a sequential C restatement of the reference CPU sink) on
    the first `sample` rows of the same device columns, copied to the host
    untimed. Returns rows/s."""
    from oracle import oracle

    n = sample
    frame = spec["host"](cols, n)
    q = spec["query"]()
    lf = type(q)(frame, q.filters, q.keys, q.aggs, q.maintain_order, True)
    times = []
    for _ in range(2):
        t0 = time.perf_counter()
        oracle.collect(lf)
        times.append(time.perf_counter() - t0)
    return n / min(times), min(times)


WORKLOADS = {
    "zipf_200m": {
        "rows": 200_000_000, "query": zipf_query, "wrap": wrap_zipf, "host": host_zipf,
        "bytes_per_row": 4 + 8 + 8 + 8 + 1 / 8,  # key, a, x, y, x validity
        "out_bytes_per_group": 4 + 8 * 8 + 7 / 8,  # key; 8 aggregates; 7 nullable
        "config": {"key": "Int32 Zipf(1.1) over 1e6", "filter": "a > 0",
                   "aggs": "sum/count/min/max(y), sum/mean/var/min(x)"},
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
    ap.add_argument("--workload", default="zipf_200m", choices=sorted(WORKLOADS))
    ap.add_argument("--rows", type=int, default=0, help="rows (default: the workload's own)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=20_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="first run one streaming copy of this many bytes (a known-size kernel for byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        return 2

    from dataframe_mojo_amd import datagen, runtime
    from dataframe_mojo_amd.plan import compile_plan

    if args.copy_calibration:
        rate = datagen.stream_rate(args.copy_calibration, copy=True, reps=1)
        log(f"calibration copy of {args.copy_calibration} bytes: {rate / 1e9:.0f} GB/s (read + write)")
    spec = WORKLOADS[args.workload]
    n = args.rows or spec["rows"]
    log(f"generating {args.workload} columns on the device: {n} rows")
    cols = getattr(datagen, f"{args.workload}_device")(n, args.seed)
    ctx = runtime.context(0)
    dev = spec["wrap"](ctx, cols, n)
    dplan = compile_plan(spec["query"]())
    keys = [dev[k] for k in dplan.keys]
    inputs = [dev[c] for c in dplan.columns]
    log("warm-up")
    groups = 0
    for _ in range(args.warmup):
        r = ctx.groupby(dplan, keys, inputs)
        groups = r.groups
        r.release()
    log(f"timing {args.steps} steps ({groups} groups)")
    phases = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = ctx.groupby(dplan, keys, inputs)
        phases.append(r.phase_ms)
        r.release()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    scan_ms = statistics.mean(p[1] for p in phases)
    bytes_per_row = spec["bytes_per_row"]
    scan_bytes = bytes_per_row * n
    achieved = scan_bytes / (scan_ms / 1e3)
    alg_bytes = scan_bytes + spec["out_bytes_per_group"] * groups
    log(f"{ms_per_step:.3f} ms/step, scan {scan_ms:.3f} ms, {achieved / 1e9:.1f} GB/s")

    # the same query under maintain_order (ordered output) for reference
    dplan_o = compile_plan(spec["query"](True))
    ordered = []
    for _ in range(max(3, args.steps // 4)):
        r = ctx.groupby(dplan_o, keys, inputs)
        ordered.append(r.device_ms)
        r.release()

    cpu = None
    if not args.no_cpu:
        sample = min(args.cpu_sample, n)
        log(f"CPU baseline on {sample} rows")
        rate, secs = cpu_baseline(spec, cols, sample)
        cpu = {"value": rate, "unit": "rows/s", "cores": 1, "kind": "port",
               "sample": f"first {sample} rows of the same workload, sequential C oracle (restated reference CPU sink), best of 2 ({secs:.2f} s)"}

    line = {
        "metric": "filtered group-by rows/s",
        "value": n / (ms_per_step / 1e3),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64+f64",
        "data": "synthetic (seeded counter-based generator, device resident)",
        "config": {"workload": args.workload, "rows": n, "groups": groups, **spec["config"], "maintain_order": False},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None, "kernel": "k_aggregate",
                     "kernel_ms": scan_ms, "bytes_per_row": bytes_per_row},
        "query_hbm": {"alg_bytes": alg_bytes, "GB/s": alg_bytes / (ms_per_step / 1e3) / 1e9,
                      "frac": alg_bytes / (ms_per_step / 1e3) / HBM_PEAK},
        "phases_ms": {k: statistics.mean(p[i] for p in phases) for i, k in enumerate(["init", "scan", "order", "finish"])},
        "ordered_ms_per_query": statistics.mean(ordered),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
