"""The autotuner and placement calibration of libdfrocm on one device.

    python -m dataframe_mojo_amd.autotune [--out PATH] [--quick]

Writes a calibration profile (JSON, version 2) that the library applies to
every context it creates on a matching device (see csrc/calib.hip) and that
the provider reads for placement costs (calibration.py):

  identity   device name, arch, compute units and the library's build id:
             a profile is used only by the build and hardware it was
             measured on (the reference's identity-matched profiles,
             dataframe_metal/calibration.mojo:1-41)
  tuning     launch settings picked by measurement: workgroup sizes, LDS
             table bytes and partition radix bits, per class of query
  transfer   host <-> device ByteCost fits and the context start-up time
             (dataframe/devices.mojo:43-80)
  groupby    the device's cost model of a filtered group-by:
             fixed_ns + ns_per_row * rows + ns_per_group * groups
  tuned      what each tuned setting was measured on and what it gained

Tuning is coordinate descent: each knob of a class in turn takes the
candidate value with the lowest time (median of the timed runs) on that
class's workload at full size, given the values chosen so far; a candidate
must beat the current value by `min_gain` to replace it. Every candidate's
result is compared with the built-in settings' result on the same inputs
(group count, keys and every aggregate bit for bit, after sorting by key:
all state words merge exactly, so any launch shape gives identical bits); a
candidate that differs, or that the library refuses, is never chosen.

The classes and their workloads (workloads.py, synthetic device-resident
columns of datagen.py):
  one_level  zipf_200m (200M rows, ~1e6 groups): pass A geometry and LDS
             table, first-level partitions (radix bits), pass C geometry,
             the scatter's chunk, the hot-key share
  two_level  dict_500m (500M rows, ~3.3e8 groups): first-level partitions
             of a two-level query, the second-level table and its geometry
  batch      4096 queries of the batched distribution: the batch kernel's
             workgroup size, workgroups per CU, LDS table bytes, row tile,
             and the row count from which a query takes the partitioned
             passes
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import statistics
import subprocess
import sys
import time

import numpy as np

from . import _lib, datagen, runtime, workloads
from .plan import compile_plan

VERSION = 2

# class -> (knob, candidates), tried in this order
KNOBS = {
    "one_level": [
        ("scan_block", [512, 1024]),
        ("scan_lds_bytes", [98304, 131072, 159744]),
        ("scan_unroll", [1, 2]),
        ("max_parts", [512, 1024, 2048, 4096]),
        ("part_block", [512, 1024]),
        ("part_lds_bytes", [65536, 98304, 131072]),
        ("scatter_chunk", [8192, 16384]),
        ("hot_share", [15, 30, 50]),
    ],
    "two_level": [
        ("two_level_parts", [256, 512, 1024]),
        ("sub_block", [256, 512]),
        ("sub_lds_bytes", [32768, 49152, 65536]),
        ("sub_fill", [30, 50, 70]),
        ("sub_wg_per_cu", [-1, 2]),
    ],
    "batch": [
        ("batch_tile", [32768, 65536, 131072, 262144]),
        ("block", [256, 512, 1024]),
        ("wg_per_cu", [1, 2, 4]),
        ("lds_bytes", [32768, 65536, 98304]),
        ("batch_part_rows", [1 << 21, 1 << 22, 1 << 23]),
    ],
}


def log(msg: str) -> None:
    print(f"[autotune] {msg}", file=sys.stderr, flush=True)


# ------------------------------------------------------------ identity


def identity(device: int = 0) -> dict:
    """The device's identity as a profile records it; read without a HIP
    context (KFD topology) plus the loaded library's build id."""
    import ctypes

    lib = _lib.load()
    info = _lib.DeviceInfo()
    e = ctypes.c_void_p()
    _lib.check(lib.dfr_device_info_get(device, ctypes.byref(info), ctypes.byref(e)), e)
    return {"device": info.name.decode(), "arch": info.arch.decode(), "compute_units": int(info.compute_units),
            "build_id": lib.dfr_build_id().decode()}


def profile_name(ident: dict) -> str:
    return f"{ident['arch']}-{ident['compute_units']}cu.json"


# ------------------------------------------------------------ workloads


class Workload:
    """A class's workload: run(ctx) executes it once and returns
    (seconds, digest) where digest identifies the result bits."""

    def run(self, ctx) -> tuple[float, str]:
        raise NotImplementedError


_MIX = np.uint64(0x9E3779B97F4A7C15)


def _mix(x: np.ndarray) -> np.ndarray:
    """SplitMix64's finaliser over uint64 (wrapping)."""
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def _bits(vals: np.ndarray, valid: np.ndarray) -> np.ndarray:
    b = vals.view(np.uint32 if vals.itemsize == 4 else np.uint64).astype(np.uint64)
    return np.where(valid, b, np.uint64(0xFFFFFFFFFFFFFFFF))


def _digest_result(res, n_keys: int, n_aggs: int) -> str:
    """Group count, keys and every aggregate's bits, independent of row
    order and coupled by row: each row's hash of its key words is mixed
    with each of its values, and the wrapping sums over rows are hashed
    (a value moved to another group's row changes the digest)."""
    with np.errstate(over="ignore"):
        row = np.zeros(res.groups, dtype=np.uint64)
        for k in range(n_keys):
            _, kv, km = res.key(k)
            row = _mix(row * _MIX + _bits(kv, km))
        sums = [res.groups, int(row.sum(dtype=np.uint64))]
        for a in range(n_aggs):
            _, v, m = res.agg(a)
            sums.append(int(_mix(row + _bits(v, m) * _MIX + np.uint64(a)).sum(dtype=np.uint64)))
    return hashlib.sha256(repr(sums).encode()).hexdigest()


class OneQuery(Workload):
    def __init__(self, ctx, name: str, rows: int):
        self.name = name
        if name == "zipf_200m":
            self.cols = datagen.zipf_200m_device(rows)
            self.dev = workloads.wrap_zipf(ctx, self.cols, rows)
            self.plan = compile_plan(workloads.zipf_query())
        else:
            self.cols = datagen.dict_500m_device(rows)
            self.dev = workloads.wrap_dict(ctx, self.cols, rows)
            self.plan = compile_plan(workloads.dict_query())
        self.keys = [self.dev[k] for k in self.plan.keys]
        self.columns = [self.dev[c] for c in self.plan.columns]

    def run(self, ctx, digest: bool = False):
        t0 = time.perf_counter()
        res = ctx.groupby(self.plan, self.keys, self.columns)
        t = time.perf_counter() - t0
        d = _digest_result(res, len(self.plan.keys), len(self.plan.aggs)) if digest else ""
        res.release()
        return t, d


class Batch(Workload):
    def __init__(self, ctx, queries: int):
        from .runtime import PreparedBatch

        self.params = datagen.batch_4096_params(queries, 0)
        self.cols = datagen.batch_4096_device(self.params, 0)
        jobs = []
        for p, c in zip(self.params, self.cols):
            n = p["rows"]
            dev = {"key": ctx.wrap_array(c["key"], length=n), "value": ctx.wrap_array(c["value"], length=n),
                   "x": ctx.wrap_array(c["x"], c["x.valid"], length=n)}
            plan = compile_plan(workloads.batch_query(p["threshold"]))
            jobs.append((plan, [dev[k] for k in plan.keys], [dev[k] for k in plan.columns]))
        self.jobs = jobs
        self.batch = PreparedBatch(jobs)

    def run(self, ctx, digest: bool = False):
        t0 = time.perf_counter()
        out = self.batch.run(ctx)
        t = time.perf_counter() - t0
        h = hashlib.sha256()
        for i, r in enumerate(out):
            if isinstance(r, Exception):
                raise r
            h.update(str(r.groups).encode())
            if digest and i % 64 == 0:
                h.update(_digest_result(r, 1, len(self.jobs[i][0].aggs)).encode())
        for r in out:
            r.release()
        return t, h.hexdigest() if digest else ""


# ------------------------------------------------------------ tuning


def measure(ctx, work: Workload, reps: int) -> tuple[float, str]:
    """Median seconds over `reps` runs after one untimed run (which also
    compiles the shape's kernels), and the result digest."""
    _, d = work.run(ctx, digest=True)
    times = [work.run(ctx)[0] for _ in range(reps)]
    return statistics.median(times), d


def tune_class(ctx, cls: str, work: Workload, reps: int, min_gain: float, knobs=None) -> tuple[dict, dict]:
    """Coordinate descent over the class's knobs; returns (chosen values,
    provenance)."""
    knobs = knobs or KNOBS[cls]
    base_t, base_d = measure(ctx, work, reps)
    log(f"{cls}: built-in settings {base_t * 1e3:.3f} ms")
    chosen, notes = {}, {}
    best_t = base_t
    for key, cands in knobs:
        results = {}
        for v in cands:
            try:
                ctx.tune(key, v)
                t, d = measure(ctx, work, reps)
            except Exception as exc:  # refused by the library, or a failed run
                log(f"{cls}: {key}={v} refused: {exc}")
                continue
            finally:
                ctx.tune(key, chosen.get(key, default_value(key)))
            if d != base_d:
                log(f"{cls}: {key}={v} changed the result: never chosen")
                continue
            results[v] = t
            log(f"{cls}: {key}={v} {t * 1e3:.3f} ms")
        if not results:
            continue
        v_best = min(results, key=results.get)
        cur = chosen.get(key, default_value(key))
        t_cur = results.get(cur, best_t)
        if v_best != cur and results[v_best] < t_cur * (1 - min_gain):
            chosen[key] = v_best
            best_t = results[v_best]
        else:
            best_t = min(best_t, t_cur)
        ctx.tune(key, chosen.get(key, cur))
        notes[key] = {"ms": {str(k): round(t * 1e3, 4) for k, t in results.items()},
                      "chosen": chosen.get(key, cur)}
    confirm = None
    if chosen:
        # the chosen settings against the built-in ones, measured back to
        # back: a descent step can win on noise, and steps can add up to
        # nothing
        t_new, _ = measure(ctx, work, reps)
        for k in chosen:
            ctx.tune(k, default_value(k))
        t_old, _ = measure(ctx, work, reps)
        confirm = {"chosen_ms": round(t_new * 1e3, 4), "builtin_ms": round(t_old * 1e3, 4)}
        if t_new < t_old * (1 - min_gain):
            for k, v in chosen.items():
                ctx.tune(k, v)
            best_t = t_new
        else:
            log(f"{cls}: {chosen} not confirmed ({t_new * 1e3:.3f} against {t_old * 1e3:.3f} ms): built-in kept")
            chosen = {}
            best_t = t_old
    log(f"{cls}: tuned {best_t * 1e3:.3f} ms (built-in {base_t * 1e3:.3f} ms): {chosen}")
    return chosen, {"workload": work.__class__.__name__, "builtin_ms": round(base_t * 1e3, 4),
                    "tuned_ms": round(best_t * 1e3, 4), "knobs": notes, "confirm": confirm}


_DEFAULTS: dict = {}


def default_value(key: str) -> int:
    """A knob's built-in value (read from a context created without a
    profile, before any knob is varied)."""
    return _DEFAULTS[key]


def read_defaults(ctx, classes) -> None:
    for cls in classes:
        for key, _ in KNOBS[cls]:
            _DEFAULTS[key] = ctx.tune_get(key)


# ------------------------------------------------------------ placement costs


def fit_linear(xs: list[list[float]], ys: list[float]) -> list[float]:
    """Least squares y ~ c0 + sum c_i x_i with non-negative coefficients
    (coefficients that come out negative are dropped and the rest refitted)."""
    X = np.array([[1.0, *x] for x in xs])
    y = np.array(ys)
    keep = list(range(X.shape[1]))
    while True:
        c, *_ = np.linalg.lstsq(X[:, keep], y, rcond=None)
        neg = [k for k, v in zip(keep, c) if v < 0]
        if not neg:
            out = [0.0] * X.shape[1]
            for k, v in zip(keep, c):
                out[k] = float(v)
            return out
        keep = [k for k in keep if k not in neg]
        if not keep:
            return [0.0] * X.shape[1]


def transfer_costs(reps: int = 3) -> dict:
    """ByteCost fits of pageable host <-> device copies (the provider's
    uploads and result copies) from 1 MiB to 256 MiB."""
    sizes = [1 << 20, 4 << 20, 16 << 20, 64 << 20, 256 << 20]
    up, down = [], []
    for n in sizes:
        host = np.ones(n // 8, dtype=np.uint64)
        tu, td = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            d = datagen.DeviceArray.from_numpy(host)
            tu.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            d.numpy()
            td.append(time.perf_counter() - t0)
            d.free()
        up.append(min(tu) * 1e9)
        down.append(min(td) * 1e9)
    cu = fit_linear([[s] for s in sizes], up)
    cd = fit_linear([[s] for s in sizes], down)
    return {"to_device": [cu[0], cu[1]], "to_host": [cd[0], cd[1]],
            "samples": {"bytes": sizes, "to_device_ns": up, "to_host_ns": down}}


def startup_ns(device: int) -> float:
    """A new process's first context: HIP runtime start-up and context set-up
    (timed inside a child process around dfr_context_create)."""
    code = (
        "import ctypes, time, sys\n"
        f"sys.path.insert(0, {os.path.dirname(os.path.dirname(os.path.abspath(__file__)))!r})\n"
        "from dataframe_mojo_amd import _lib\n"
        "lib = _lib.load()\n"
        "h = ctypes.c_void_p(); e = ctypes.c_void_p()\n"
        "t0 = time.perf_counter_ns()\n"
        f"rc = lib.dfr_context_create({device}, 0, ctypes.byref(h), ctypes.byref(e))\n"
        "t = time.perf_counter_ns() - t0\n"
        "lib.dfr_context_destroy(h)\n"
        "print(t if rc == 0 else -1)\n"
    )
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    try:
        return float(out.stdout.strip().splitlines()[-1])
    except (ValueError, IndexError):
        return -1.0


def groupby_costs(ctx, reps: int = 3, rows_list=(100_000, 1_000_000, 10_000_000, 100_000_000),
                  domains=(100, 10_000, 1_000_000)) -> dict:
    """The device cost model of a filtered group-by over resident columns:
    wall time of dfr_groupby (inputs on the device, results left there),
    fitted as fixed_ns + ns_per_row * rows + ns_per_group * groups."""
    samples = []
    n_max = max(rows_list)
    value = datagen.device_uniform(n_max, 7, 200, -1_000_000, 1_000_000)
    x = datagen.device_normal(n_max, 7, 201)
    xv = datagen.device_validity(n_max, 7, 202, 0.05)
    plan = compile_plan(workloads.batch_query(0))
    for d in domains:
        key = datagen.device_uniform(n_max, 7, 203 + d % 97, 0, d - 1)
        for n in rows_list:
            dev = {"key": ctx.wrap_array(key, length=n), "value": ctx.wrap_array(value, length=n),
                   "x": ctx.wrap_array(x, xv, length=n)}
            keys, cols = [dev[k] for k in plan.keys], [dev[c] for c in plan.columns]
            ctx.groupby(plan, keys, cols).release()  # warm (compiles the shape)
            ts, g = [], 0
            for _ in range(reps):
                t0 = time.perf_counter()
                r = ctx.groupby(plan, keys, cols)
                ts.append(time.perf_counter() - t0)
                g = r.groups
                r.release()
            samples.append({"rows": n, "groups": g, "ns": min(ts) * 1e9})
            log(f"groupby cost: {n} rows, {g} groups: {min(ts) * 1e3:.3f} ms")
    c = fit_linear([[s["rows"], s["groups"]] for s in samples], [s["ns"] for s in samples])
    return {"fixed_ns": c[0], "ns_per_row": c[1], "ns_per_group": c[2], "samples": samples,
            "query": "filter(value > t).group_by(key).agg(len, sum/max(value), sum/mean(x)), Int64 key"}


# ------------------------------------------------------------ main


def calibrate(device: int = 0, classes=("one_level", "two_level", "batch"), reps: int = 3,
              min_gain: float = 0.02, quick: bool = False, costs: bool = True) -> dict:
    ident = identity(device)
    log(f"identity {ident}")
    os.environ["DATAFRAME_ROCM_CALIBRATION"] = "off"  # tune from the built-in settings
    ctx = runtime.Context(device)
    read_defaults(ctx, classes)
    tuning, tuned = {}, {}
    for cls in classes:
        if cls == "one_level":
            work = OneQuery(ctx, "zipf_200m", 20_000_000 if quick else 200_000_000)
        elif cls == "two_level":
            work = OneQuery(ctx, "dict_500m", 50_000_000 if quick else 500_000_000)
        else:
            work = Batch(ctx, 256 if quick else 4096)
        chosen, prov = tune_class(ctx, cls, work, reps, min_gain)
        # the class's choices stay set while the next class is tuned
        tuning.update(chosen)
        tuned[cls] = prov
        del work
        ctx.trim()
    prof = {
        "version": VERSION,
        "identity": ident,
        "tuning": tuning,
        "tuned": tuned,
        "measured": time.strftime("%Y-%m-%dT%H:%M:%SZ", time.gmtime()),
    }
    if costs:
        log("placement costs")
        prof["transfer"] = dict(transfer_costs(), startup_ns=startup_ns(device))
        prof["groupby"] = groupby_costs(ctx, rows_list=(100_000, 1_000_000, 10_000_000) if quick else
                                        (100_000, 1_000_000, 10_000_000, 100_000_000))
    ctx.close()
    return prof


def merge_profiles(base: dict, new: dict) -> dict:
    """`new` over `base`, both measured with the same identity: tuning and
    provenance joined (new wins), costs taken from whichever has them."""
    if base.get("identity") != new.get("identity") or base.get("version") != new.get("version"):
        raise ValueError("profiles of different identities do not merge")
    out = dict(base)
    out.update({k: v for k, v in new.items() if k not in ("tuning", "tuned")})
    out["tuning"] = {**base.get("tuning", {}), **new.get("tuning", {})}
    out["tuned"] = {**base.get("tuned", {}), **new.get("tuned", {})}
    for k in ("transfer", "groupby"):
        if k not in new and k in base:
            out[k] = base[k]
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--out", default="", help="profile path (default: calibration/<arch>-<CUs>cu.json in the package)")
    ap.add_argument("--classes", default="one_level,two_level,batch")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--min-gain", type=float, default=0.02)
    ap.add_argument("--quick", action="store_true", help="smaller workloads (a smoke run, not a calibration)")
    ap.add_argument("--no-costs", action="store_true", help="tune only (no placement cost capture)")
    ap.add_argument("--merge", default="", help="profiles (comma separated) whose tuning and costs this run's extend")
    args = ap.parse_args(argv)
    prof = calibrate(args.device, tuple(c for c in args.classes.split(",") if c), args.reps, args.min_gain,
                     args.quick, not args.no_costs)
    for path in [p for p in args.merge.split(",") if p]:
        with open(path) as f:
            prof = merge_profiles(json.load(f), prof)
    out = args.out or os.path.join(os.path.dirname(os.path.abspath(__file__)), "calibration",
                                   profile_name(prof["identity"]))
    os.makedirs(os.path.dirname(out) or ".", exist_ok=True)
    with open(out, "w") as f:
        json.dump(prof, f, indent=1, sort_keys=True)
    print(json.dumps({"profile": out, "tuning": prof["tuning"]}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
