"""Benchmark of the filtered group-by path on one MI355X.

Default workload (zipf_200m): 200,000,000 rows resident in HBM —
  key Int32 Zipf(1.1) over 1e6 values, a Int64 uniform [-1e9, 1e9],
  x Float64 N(0,1) with 5% nulls, y Int64 uniform [0, 1e6];
  filter(a > 0).group_by(key).agg(sum/count/min/max(y), sum/mean/var/min(x)).
A step is one whole query through the C ABI (dfr_groupby: table set-up,
fused scan, compaction, finish) with its inputs already on the device and
its results left on the device. Prints one JSON line.

  python bench.py [--steps K] [--warmup W] [--rows N] [--cpu-sample S]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

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


def wrap_zipf(ctx, cols, n):
    return {
        "key": ctx.wrap_array(cols["key"], length=n),
        "a": ctx.wrap_array(cols["a"], length=n),
        "x": ctx.wrap_array(cols["x"], cols["x.valid"], length=n),
        "y": ctx.wrap_array(cols["y"], length=n),
    }


def cpu_baseline(cols, sample: int):
    """The oracle (a sequential C restatement of the reference CPU sink) on
    the first `sample` rows of the same device columns, copied to the host
    untimed. Returns rows/s."""
    from dataframe_mojo_amd import Column, DataFrame, Series, datagen
    from oracle import oracle

    n = sample
    frame = DataFrame([
        Series("key", Column.int32(cols["key"].numpy(n))),
        Series("a", Column.int64(cols["a"].numpy(n))),
        Series("x", Column.float64(cols["x"].numpy(n), datagen.valid_from_words(cols["x.valid"], n))),
        Series("y", Column.int64(cols["y"].numpy(n))),
    ])
    q = zipf_query()
    lf = type(q)(frame, q.filters, q.keys, q.aggs, q.maintain_order, True)
    times = []
    for _ in range(2):
        t0 = time.perf_counter()
        oracle.collect(lf)
        times.append(time.perf_counter() - t0)
    return n / min(times), min(times)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=200_000_000)
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
    n = args.rows
    log(f"generating zipf_200m columns on the device: {n} rows")
    cols = datagen.zipf_200m_device(n, args.seed)
    ctx = runtime.context(0)
    dev = wrap_zipf(ctx, cols, n)
    dplan = compile_plan(zipf_query())
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
    bytes_per_row = 4 + 8 + 8 + 8 + 1 / 8  # key, a, x, y, x validity
    scan_bytes = bytes_per_row * n
    achieved = scan_bytes / (scan_ms / 1e3)
    log(f"{ms_per_step:.3f} ms/step, scan {scan_ms:.3f} ms, {achieved / 1e9:.1f} GB/s")

    # the same query under maintain_order (ordered output) for reference
    dplan_o = compile_plan(zipf_query(True))
    ordered = []
    for _ in range(max(3, args.steps // 4)):
        r = ctx.groupby(dplan_o, keys, inputs)
        ordered.append(r.device_ms)
        r.release()

    cpu = None
    if not args.no_cpu:
        sample = min(args.cpu_sample, n)
        log(f"CPU baseline on {sample} rows")
        rate, secs = cpu_baseline(cols, sample)
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
        "config": {"workload": "zipf_200m", "rows": n, "groups": groups, "key": "Int32 Zipf(1.1) over 1e6",
                   "filter": "a > 0", "aggs": "sum/count/min/max(y), sum/mean/var/min(x)", "maintain_order": False},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None, "kernel": "k_aggregate",
                     "kernel_ms": scan_ms, "bytes_per_row": bytes_per_row},
        "phases_ms": {k: statistics.mean(p[i] for p in phases) for i, k in enumerate(["init", "scan", "order", "finish"])},
        "ordered_ms_per_query": statistics.mean(ordered),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
