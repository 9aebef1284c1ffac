"""Seeded synthetic columns for the filtered group-by workloads.

The generator is counter based: value r of a column is a pure function of
(seed, stream, r), so a column can be produced on the GPU (libdfrgen.so,
csrc/datagen.hip) or, for integer columns and validity bits, on the host by
the numpy functions below, bit for bit the same.

    base = mix64(seed ^ (stream * 0xD1B54A32D192ED03))
    u(r) = mix64(base + (r + 1) * 0x9E3779B97F4A7C15)      (mod 2^64)

uniform [lo, hi]  lo + u(r) % (hi - lo + 1)
validity          valid unless (u(r) >> 11) < round(p_null * 2^53)
zipf(s, m)        1 + first i with (u(r) >> 11) < cdf[i], cdf the Zipf(s)
                  distribution over 1..m scaled to 2^53
normal(mu, sd)    Box-Muller on u(2r), u(2r+1) (device only: libm differs)

Workloads (the column sets the benchmark and the tests run):
  groupby_1m        1,000,000 rows, seed 0: key Int64 uniform [0, 64),
                    amount Int64 uniform [-1000, 1000], no nulls;
                    filter(amount > 0).group_by(key, maintain_order=True)
                    .agg(sum(amount), count())
  zipf_200m         200,000,000 rows: key Int32 Zipf(1.1) over 1e6 values,
                    a Int64 uniform [-1e9, 1e9] (filter a > 0), x Float64
                    N(0, 1) with 5% nulls, y Int64 uniform [0, 1e6];
                    sum/count/min/max(y), sum/mean/var/min(x)
  dict_500m         500,000,000 rows: k1 dictionary String of 1000 labels
                    (uniform codes), k2 Int64 uniform [0, 1e7), v Float64
                    lognormal(0, 2) with 2% nulls; filter k2 % 3 != 0;
                    count/sum/mean(v)
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
GOLDEN = np.uint64(0x9E3779B97F4A7C15)
STREAM_MUL = np.uint64(0xD1B54A32D192ED03)

# stream ids of the workload columns
STREAMS = {
    "key": 1,
    "amount": 2,
    "a": 3,
    "x": 4,
    "x.valid": 5,
    "y": 6,
    "k1": 7,
    "k2": 8,
    "v": 9,
    "v.valid": 10,
}


def mix64(z: np.ndarray) -> np.ndarray:
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def stream_base(seed: int, stream: int) -> np.uint64:
    with np.errstate(over="ignore"):
        s = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) ^ (np.uint64(stream) * STREAM_MUL)
    return mix64(np.array([s], dtype=np.uint64))[0]


def draws(seed: int, stream: int, start: int, n: int) -> np.ndarray:
    base = stream_base(seed, stream)
    r = np.arange(start + 1, start + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return mix64(base + r * GOLDEN)


def uniform(seed: int, stream: int, n: int, lo: int, hi: int, dtype=np.int64, start: int = 0) -> np.ndarray:
    """Uniform integers in [lo, hi], as the device generator makes them."""
    u = draws(seed, stream, start, n)
    span = np.uint64((hi - lo + 1) & 0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        v = np.uint64(lo & 0xFFFFFFFFFFFFFFFF) + (u % span if span else u)
    return v.astype(np.uint64).view(np.int64).astype(dtype)


def null_threshold(p_null: float) -> int:
    return int(round(p_null * (1 << 53)))


def validity_bits(seed: int, stream: int, n: int, p_null: float, start: int = 0) -> np.ndarray:
    """Boolean validity per row (True = valid)."""
    u = draws(seed, stream, start, n)
    return (u >> np.uint64(11)) >= np.uint64(null_threshold(p_null))


def pack_bits(valid: np.ndarray) -> np.ndarray:
    """LSB-first packed bytes (Arrow order)."""
    return np.packbits(np.asarray(valid, dtype=bool), bitorder="little")


def zipf_cdf(s: float, m: int) -> np.ndarray:
    """Zipf(s) over 1..m as an increasing uint64 table scaled to 2^53."""
    k = np.arange(1, m + 1, dtype=np.float64)
    p = k ** (-s)
    c = np.cumsum(p)
    c /= c[-1]
    t = np.round(c * float(1 << 53)).astype(np.uint64)
    t[-1] = np.uint64(1 << 53)
    return t


def zipf(seed: int, stream: int, n: int, cdf: np.ndarray, start: int = 0) -> np.ndarray:
    x = draws(seed, stream, start, n) >> np.uint64(11)
    return (np.searchsorted(cdf, x, side="right") + 1).astype(np.int32)


def groupby_1m(n: int = 1_000_000, seed: int = 0) -> dict:
    """The 1M-row parity workload: key Int64 uniform [0, 64), amount Int64
    uniform [-1000, 1000], no nulls."""
    return {
        "key": uniform(seed, STREAMS["key"], n, 0, 63),
        "amount": uniform(seed, STREAMS["amount"], n, -1000, 1000),
    }


# ---------------------------------------------------------------- device side

_gen = None


def _libgen():
    global _gen
    if _gen is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdfrgen.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run python -m dataframe_mojo_amd.build")
        lib = ctypes.CDLL(path)
        u64, i64, vp = ctypes.c_uint64, ctypes.c_int64, ctypes.c_void_p
        lib.dfrg_alloc.argtypes = [ctypes.POINTER(vp), u64]
        lib.dfrg_free.argtypes = [vp]
        lib.dfrg_copy_to_host.argtypes = [vp, vp, u64]
        lib.dfrg_copy_to_device.argtypes = [vp, vp, u64]
        lib.dfrg_fill_zero.argtypes = [vp, u64]
        lib.dfrg_uniform.argtypes = [vp, ctypes.c_int, u64, u64, u64, i64, i64]
        lib.dfrg_f64_keys.argtypes = [vp, vp, u64, ctypes.c_double]
        lib.dfrg_zipf_i32.argtypes = [vp, u64, u64, u64, vp, u64, i64]
        lib.dfrg_normal_f64.argtypes = [vp, u64, u64, u64, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        lib.dfrg_validity.argtypes = [vp, u64, u64, u64, u64]
        lib.dfrg_stream.argtypes = [vp, vp, u64, ctypes.c_int, ctypes.c_int]
        lib.dfrg_stream.restype = ctypes.c_double
        _gen = lib
    return _gen


def _check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceArray:
    """A device buffer of `n` elements of a numpy dtype (HIP memory owned by
    libdfrgen; the same HIP runtime as libdfrocm)."""

    def __init__(self, n: int, dtype):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(_libgen().dfrg_alloc(ctypes.byref(p), max(self.nbytes, 16)), "dfrg_alloc")
        self.ptr = p.value

    def data_ptr(self) -> int:
        return self.ptr

    def numpy(self, count: int | None = None, start: int = 0) -> np.ndarray:
        count = self.n - start if count is None else count
        out = np.empty(count, dtype=self.dtype)
        if count:
            _check(_libgen().dfrg_copy_to_host(out.ctypes.data, self.ptr + start * self.dtype.itemsize,
                                               count * self.dtype.itemsize), "dfrg_copy_to_host")
        return out

    @staticmethod
    def from_numpy(a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(a.shape[0], a.dtype)
        if a.nbytes:
            _check(_libgen().dfrg_copy_to_device(d.ptr, a.ctypes.data, a.nbytes), "dfrg_copy_to_device")
        return d

    def free(self):
        if self.ptr:
            _libgen().dfrg_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_uniform(n, seed, stream, lo, hi, dtype=np.int64) -> DeviceArray:
    t = DeviceArray(n, dtype)
    _check(_libgen().dfrg_uniform(t.ptr, t.dtype.itemsize, n, seed, stream, lo, hi), "dfrg_uniform")
    return t


def device_float_keys(n, seed, stream, card, scale=0.125) -> DeviceArray:
    """Float64 keys k * scale for k uniform in [0, card) (k = 0 as -0.0);
    the host twin is float_keys."""
    t = device_uniform(n, seed, stream, 0, max(card, 1) - 1)
    _check(_libgen().dfrg_f64_keys(t.ptr, t.ptr, n, scale), "dfrg_f64_keys")
    t.dtype = np.dtype(np.float64)
    return t


def float_keys(seed, stream, n, card, scale=0.125) -> np.ndarray:
    k = uniform(seed, stream, n, 0, max(card, 1) - 1)
    out = k.astype(np.float64) * scale
    out[k == 0] = -0.0
    return out


def device_zipf(n, seed, stream, s, m) -> DeviceArray:
    cdf = DeviceArray.from_numpy(zipf_cdf(s, m))
    t = DeviceArray(n, np.int32)
    _check(_libgen().dfrg_zipf_i32(t.ptr, n, seed, stream, cdf.ptr, m, 1), "dfrg_zipf_i32")
    cdf.free()
    return t


def device_normal(n, seed, stream, mu=0.0, sigma=1.0, lognormal=False) -> DeviceArray:
    t = DeviceArray(n, np.float64)
    _check(_libgen().dfrg_normal_f64(t.ptr, n, seed, stream, mu, sigma, int(lognormal)), "dfrg_normal_f64")
    return t


def device_validity(n, seed, stream, p_null) -> DeviceArray:
    """Validity as uint64 words (LSB-first, bit r of word r // 64), with one
    spare zero word at the end."""
    t = DeviceArray((n + 63) // 64 + 1, np.uint64)
    _check(_libgen().dfrg_fill_zero(t.ptr, t.nbytes), "dfrg_fill_zero")
    _check(_libgen().dfrg_validity(t.ptr, n, seed, stream, null_threshold(p_null)), "dfrg_validity")
    return t


def stream_rate(nbytes: int, copy: bool = False, blocks: int = 0, reps: int = 5) -> float:
    """Achievable HBM rate in bytes/s: a streaming read (or copy) of nbytes,
    best of `reps`. A copy counts read + write bytes."""
    nbytes = nbytes // 16 * 16
    src = DeviceArray(nbytes // 8, np.uint64)
    _check(_libgen().dfrg_fill_zero(src.ptr, src.nbytes), "dfrg_fill_zero")
    dst = DeviceArray(nbytes // 8, np.uint64) if copy else None
    blocks = blocks or 256 * 64
    ms = _libgen().dfrg_stream(src.ptr, dst.ptr if dst else None, nbytes, blocks, reps)
    if ms <= 0:
        raise RuntimeError("dfrg_stream failed")
    return (2 if copy else 1) * nbytes / (ms / 1e3)


def valid_from_words(words: DeviceArray, n: int) -> np.ndarray:
    """Host boolean validity of the first n rows of device validity words."""
    w = words.numpy((n + 63) // 64)
    return np.unpackbits(w.view(np.uint8), bitorder="little")[:n].astype(bool)


def zipf_200m_device(n: int = 200_000_000, seed: int = 0) -> dict:
    """Device columns of the 200M-row Zipf workload (see module doc)."""
    return {
        "key": device_zipf(n, seed, STREAMS["key"], 1.1, 1_000_000),
        "a": device_uniform(n, seed, STREAMS["a"], -1_000_000_000, 1_000_000_000),
        "x": device_normal(n, seed, STREAMS["x"], 0.0, 1.0),
        "x.valid": device_validity(n, seed, STREAMS["x.valid"], 0.05),
        "y": device_uniform(n, seed, STREAMS["y"], 0, 1_000_000),
    }


def dict_500m_device(n: int = 500_000_000, seed: int = 0) -> dict:
    """Device columns of the 500M-row dictionary-key workload."""
    return {
        "k1": device_uniform(n, seed, STREAMS["k1"], 0, 999, dtype=np.uint32),
        "k2": device_uniform(n, seed, STREAMS["k2"], 0, 9_999_999),
        "v": device_normal(n, seed, STREAMS["v"], 0.0, 2.0, lognormal=True),
        "v.valid": device_validity(n, seed, STREAMS["v.valid"], 0.02),
    }


def batch_4096_params(n_queries: int = 4096, seed: int = 0, rows_lo: float = 1e3, rows_hi: float = 1e7) -> list:
    """Per-query parameters of the batched workload: rows log-uniform in
    [rows_lo, rows_hi], key cardinality log-uniform in [1, 1e5], filter
    threshold uniform in [-1e6, 1e6] (from a seeded host generator)."""
    rng = np.random.default_rng(seed)
    rows = np.exp(rng.uniform(np.log(rows_lo), np.log(rows_hi), n_queries)).astype(np.int64)
    card = np.exp(rng.uniform(0.0, np.log(1e5), n_queries)).astype(np.int64).clip(1)
    thr = rng.integers(-1_000_000, 1_000_001, n_queries)
    return [{"rows": int(r), "card": int(c), "threshold": int(t)} for r, c, t in zip(rows, card, thr)]


def batch_4096_device(params: list, seed: int = 0) -> list:
    """Device columns of every batched query (query q draws from streams
    100 + 4q ..): key Int64 uniform in [0, card), value Int64 uniform in
    [-1e6, 1e6], x Float64 N(0, 1) with 5% nulls."""
    out = []
    for q, p in enumerate(params):
        n, base = p["rows"], 100 + 4 * q
        out.append({
            "key": device_uniform(n, seed, base, 0, p["card"] - 1),
            "value": device_uniform(n, seed, base + 1, -1_000_000, 1_000_000),
            "x": device_normal(n, seed, base + 2, 0.0, 1.0),
            "x.valid": device_validity(n, seed, base + 3, 0.05),
        })
    return out


def dict_labels(m: int = 1000) -> list[str]:
    """The 1000 category labels of the dictionary workload (code i -> label)."""
    return [f"cat-{i:04d}" for i in range(m)]
