"""Summaries of rocprofv3 output for the kernels of the group-by path.

Two uses, both from the repository root after a profiled run of bench.py:

  kernel statistics (rocprofv3 --kernel-trace --stats ... -o run):
    python -m dataframe_mojo_amd.profiling stats DIR/run_kernel_stats.csv \\
        --title "..." --command "..."  > profiles/kernels_<workload>.md

  HBM traffic of one kernel per launch, from two counter runs of the same
  command (rocprofv3 --pmc FETCH_SIZE, then --pmc WRITE_SIZE) that each
  start with a streaming copy of a known size (bench.py --copy-calibration
  BYTES): each counter is scaled by what it reads for that copy, so the
  figures are bytes, whatever the counter's unit and request granularity.
    python -m dataframe_mojo_amd.profiling traffic FETCH_DIR/run_counter_collection.csv \\
        WRITE_DIR/run_counter_collection.csv --kernel dfr_scan_jit --copy-bytes BYTES \\
        --workload zipf_200m --rows 200000000 --command "..." > profiles/traffic_<workload>.json

  every counter of a counter run, as the mean per dispatch of each kernel
  whose name contains one of the given names:
    python -m dataframe_mojo_amd.profiling counters DIR/run_counter_collection.csv \
        --kernels k_batch_aggregate,dfr_scan_jit

bench.py reports a workload's ``roofline.traffic`` from
profiles/traffic_<workload>.json when its kernel is the one the bench ran.
"""
from __future__ import annotations

import argparse
import csv
import json
import statistics
import sys

COPY_KERNEL = "k_stream_copy"


def kernel_stats_md(path: str, title: str, command: str, top: int = 25) -> str:
    rows = list(csv.DictReader(open(path)))
    out = [f"# {title}", "", "Command (from the repository root, on one MI355X):", "", "```", command, "```", "",
           "rocprofv3 kernel statistics (every dispatch of the run, column generation included):", "",
           "| Kernel | Calls | Avg µs | Min µs | Max µs | % of GPU time |", "|---|---:|---:|---:|---:|---:|"]
    for r in rows[:top]:
        name = r["Name"].replace("|", "\\|")
        if len(name) > 90:
            name = name[:87] + "..."
        out.append(f"| `{name}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | {float(r['MinNs']) / 1e3:.1f} | "
                   f"{float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.2f} |")
    return "\n".join(out) + "\n"


def _counter(path: str, counter: str) -> dict[str, list[float]]:
    """Per kernel name, the counter's value for each dispatch."""
    per: dict[str, list[float]] = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        per.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return per


def _find(per: dict[str, list[float]], name: str) -> list[float]:
    hits = [v for k, v in per.items() if name in k]
    if not hits:
        raise KeyError(f"no dispatch of a kernel named like {name!r}")
    return [x for v in hits for x in v]


def traffic(fetch_csv: str, write_csv: str, kernel: str, copy_bytes: int) -> dict:
    """Bytes a launch of `kernel` reads from and writes to HBM (mean over its
    dispatches), with each counter scaled by its reading for the copy."""
    fetch = _counter(fetch_csv, "FETCH_SIZE")
    write = _counter(write_csv, "WRITE_SIZE")
    cf = statistics.mean(_find(fetch, COPY_KERNEL))
    cw = statistics.mean(_find(write, COPY_KERNEL))
    f_scale = copy_bytes / cf  # bytes per counter unit, reads
    w_scale = copy_bytes / cw  # writes
    kf = _find(fetch, kernel)
    kw = _find(write, kernel)
    return {
        "kernel": kernel,
        "dispatches": len(kf),
        "fetch_bytes": statistics.mean(kf) * f_scale,
        "write_bytes": statistics.mean(kw) * w_scale,
        "traffic": statistics.mean(kf) * f_scale + statistics.mean(kw) * w_scale,
        "copy_bytes": copy_bytes,
        "fetch_per_unit": f_scale,
        "write_per_unit": w_scale,
        "copy_fetch_reading": cf,
        "copy_write_reading": cw,
    }


def counter_means(path: str, names: list[str]) -> dict[str, dict]:
    """Per name: the dispatches of the kernels whose name contains it, and
    each counter's mean per dispatch over them."""
    sums: dict[str, dict[str, float]] = {n: {} for n in names}
    disp: dict[str, set] = {n: set() for n in names}
    for r in csv.DictReader(open(path)):
        hit = next((n for n in names if n in r["Kernel_Name"]), None)
        if hit is None:
            continue
        c = r["Counter_Name"]
        sums[hit][c] = sums[hit].get(c, 0.0) + float(r["Counter_Value"])
        disp[hit].add(r["Dispatch_Id"])
    return {n: {"dispatches": len(disp[n]),
                "mean": {c: v / max(1, len(disp[n])) for c, v in sorted(sums[n].items())}} for n in names}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m dataframe_mojo_amd.profiling")
    sub = ap.add_subparsers(dest="what", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--title", required=True)
    s.add_argument("--command", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_csv")
    t.add_argument("write_csv")
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--workload", required=True)
    t.add_argument("--rows", type=int, required=True)
    t.add_argument("--command", required=True)
    c = sub.add_parser("counters")
    c.add_argument("csv")
    c.add_argument("--kernels", required=True)
    a = ap.parse_args(argv)
    if a.what == "stats":
        sys.stdout.write(kernel_stats_md(a.csv, a.title, a.command))
    elif a.what == "counters":
        json.dump(counter_means(a.csv, a.kernels.split(",")), sys.stdout, indent=1)
        sys.stdout.write("\n")
    else:
        out = traffic(a.fetch_csv, a.write_csv, a.kernel, a.copy_bytes)
        out["workload"] = a.workload
        out["rows"] = a.rows
        out["command"] = a.command
        json.dump(out, sys.stdout, indent=1)
        sys.stdout.write("\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
