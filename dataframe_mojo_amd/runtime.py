"""Device contexts, device columns and query execution over libdfrocm.

A `Context` owns one device, one HIP stream and one HBM budget
(DATAFRAME_ACCEL_MEMORY_LIMIT, as dataframe_metal/metal.mojo:36-46 reads it).
Columns reach the device by upload (host arrays) or by wrapping device
memory that already holds them (zero copy), which is the
device column hand-off of storage/residence.mojo.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

from . import _lib
from .frame import FLOAT64, INT32, INT64, STRING, ChunkedColumn, Column, DataFrame, Series
from .plan import DTYPE_CODE, DevicePlan

MEMORY_LIMIT_ENV = "DATAFRAME_ACCEL_MEMORY_LIMIT"
DEVICE_ENV = "DATAFRAME_ACCEL_DEVICE"

_NP_OF = {_lib.DFR_INT32: np.int32, _lib.DFR_INT64: np.int64, _lib.DFR_FLOAT64: np.float64, _lib.DFR_DICT: np.uint32}
_NAME_OF = {_lib.DFR_INT32: INT32, _lib.DFR_INT64: INT64, _lib.DFR_FLOAT64: FLOAT64, _lib.DFR_DICT: STRING}


def memory_limit() -> int:
    """DATAFRAME_ACCEL_MEMORY_LIMIT in bytes (0 = the device's free memory);
    a bad value raises, as the reference's runtime configuration does."""
    raw = os.environ.get(MEMORY_LIMIT_ENV, "").strip()
    if not raw:
        return 0
    try:
        v = int(raw)
    except ValueError:
        raise ValueError(f"{MEMORY_LIMIT_ENV}={raw!r} is not a byte count") from None
    if v <= 0:
        raise ValueError(f"{MEMORY_LIMIT_ENV} must be positive")
    return v


def device_ordinal() -> int:
    raw = os.environ.get(DEVICE_ENV, "").strip()
    if not raw:
        return 0
    try:
        v = int(raw)
    except ValueError:
        raise ValueError(f"{DEVICE_ENV}={raw!r} is not a device ordinal") from None
    if v < 0:
        raise ValueError(f"{DEVICE_ENV} must be non-negative")
    return v


def _err():
    return ctypes.c_void_p()


class DeviceColumn:
    """A column on the device: a libdfrocm handle plus what keeps its memory
    alive (for wrapped device memory) and its labels (dictionary Strings)."""

    def __init__(self, ctx: "Context", handle: ctypes.c_void_p, dtype: int, length: int, labels=None, keepalive=None):
        self.ctx = ctx
        self.handle = handle
        self.dtype = dtype
        self.length = length
        self.labels = labels
        self.keepalive = keepalive

    def null_count(self) -> int:
        return int(self.ctx.lib.dfr_column_null_count(self.handle))

    def distinct(self) -> int:
        """Upper estimate of the distinct non-null values (made at hand-off)."""
        return int(self.ctx.lib.dfr_column_distinct(self.handle))

    def slice(self, offset: int, length: int) -> "DeviceColumn":
        """Rows [offset, offset + length) as a column of their own, borrowing
        this one's buffers and statistics (no device work)."""
        h = ctypes.c_void_p()
        e = _err()
        _lib.check(self.ctx.lib.dfr_column_slice(self.handle, offset, length, ctypes.byref(h), ctypes.byref(e)), e)
        return DeviceColumn(self.ctx, h, self.dtype, length, self.labels, keepalive=self)

    def release(self):
        if self.handle:
            self.ctx.lib.dfr_column_release(self.handle)
            self.handle = None
            self.keepalive = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Result:
    """A group-by result resident on the device until copied out."""

    def __init__(self, ctx: "Context", handle: ctypes.c_void_p, n_keys: int, n_aggs: int):
        self.ctx = ctx
        self.handle = handle
        self.n_keys = n_keys
        self.n_aggs = n_aggs
        self.groups = int(ctx.lib.dfr_result_groups(handle))
        self._timing = None  # read on first use: a batch hands out thousands of results

    def _times(self):
        if self._timing is None:
            if not self.handle:
                raise ValueError("result released")
            lib, h = self.ctx.lib, self.handle
            self._timing = (float(lib.dfr_result_device_ms(h)), int(lib.dfr_result_device_bytes(h)),
                            [float(lib.dfr_result_phase_ms(h, p)) for p in range(4)])
        return self._timing

    @property
    def device_ms(self) -> float:
        """Device time of the launch sequence that made this result (a
        batch's whole sequence for a batched query)."""
        return self._times()[0]

    @property
    def device_bytes(self) -> int:
        return self._times()[1]

    @property
    def phase_ms(self) -> list:
        return self._times()[2]

    def _copy(self, fn, i: int, dtype: int):
        n = self.groups
        vals = np.empty(n, dtype=_NP_OF[dtype])
        bits = np.empty((n + 7) // 8 + 1, dtype=np.uint8)
        e = _err()
        _lib.check(fn(self.handle, i, vals.ctypes.data_as(ctypes.c_void_p), bits.ctypes.data_as(ctypes.c_void_p), ctypes.byref(e)), e)
        valid = np.unpackbits(bits, bitorder="little")[:n].astype(bool)
        return vals, valid

    def key(self, k: int):
        dt = self.ctx.lib.dfr_result_key_dtype(self.handle, k)
        return dt, *self._copy(self.ctx.lib.dfr_result_copy_key, k, dt)

    def agg(self, a: int):
        dt = self.ctx.lib.dfr_result_agg_dtype(self.handle, a)
        return dt, *self._copy(self.ctx.lib.dfr_result_copy_agg, a, dt)

    def first_rows(self) -> np.ndarray:
        out = np.empty(self.groups, dtype=np.int64)
        e = _err()
        _lib.check(self.ctx.lib.dfr_result_copy_first_rows(self.handle, out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(e)), e)
        return out

    def residence(self, kind: int, index: int, part: int) -> "DeviceResidence":
        """One output buffer as a device residence (storage/residence.mojo:
        context, bytes, and the _CopyToHost/_ReleaseDevice callbacks by
        address); it keeps this result's device memory alive."""
        h = ctypes.c_void_p()
        n = ctypes.c_int64()
        e = _err()
        _lib.check(self.ctx.lib.dfr_result_buffer(self.handle, kind, index, part, ctypes.byref(h), ctypes.byref(n),
                                                  ctypes.byref(e)), e)
        lib = self.ctx.lib
        return DeviceResidence(
            backend="amd", device=self.ctx.device, context=h.value, bytes=n.value,
            to_host=ctypes.cast(lib.dfr_buffer_copy_to_host, ctypes.c_void_p).value,
            release=ctypes.cast(lib.dfr_buffer_release, ctypes.c_void_p).value)

    def release(self):
        if self.handle:
            self.ctx.lib.dfr_result_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


# (context, host destination address, bytes) -> 0 on success; (context) -> None
# (storage/residence.mojo:20-22)
_COPY_TO_HOST = ctypes.CFUNCTYPE(ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64)
_RELEASE_DEVICE = ctypes.CFUNCTYPE(None, ctypes.c_int64)


class DeviceResidence:
    """A buffer left on the device, as the reference's _DeviceResidence
    holds it (storage/residence.mojo:33-80): the first host read copies the
    bytes once through the `to_host` callback; `release` (or the object's
    end) calls the `release` callback. Both are called by address."""

    def __init__(self, backend: str, device: int, context: int, bytes: int, to_host: int, release: int):
        self.backend, self.device, self.context, self.bytes = backend, device, context, bytes
        self.to_host, self._release = to_host, release
        self._host = None

    def host(self) -> np.ndarray:
        if self._host is None:
            buf = np.empty(max(self.bytes, 1), dtype=np.uint8)
            status = _COPY_TO_HOST(self.to_host)(self.context, buf.ctypes.data, self.bytes)
            if status != 0:
                raise RuntimeError(f"copying a {self.backend} device buffer to the host failed (status {status})")
            self._host = buf[: self.bytes]
        return self._host

    def release(self):
        if self.context:
            _RELEASE_DEVICE(self._release)(self.context)
            self.context = 0

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def build_query(plan: DevicePlan, keys: list, columns: list):
    """The dfr_query struct (and the arrays it points at) of a compiled plan
    over device columns."""
    q = _lib.Query()
    key_arr = (ctypes.c_void_p * max(1, len(keys)))(*[k.handle for k in keys])
    col_arr = (ctypes.c_void_p * max(1, len(columns)))(*[c.handle for c in columns])
    q.n_keys = len(keys)
    q.n_columns = len(columns)
    q.keys = ctypes.cast(key_arr, ctypes.POINTER(ctypes.c_void_p))
    q.columns = ctypes.cast(col_arr, ctypes.POINTER(ctypes.c_void_p))
    prog = plan.program
    q.predicate.n_leaves = len(prog.leaves)
    for i in range(8):
        q.predicate.truth[i] = prog.truth[i] & 0xFFFFFFFF
    for i, leaf in enumerate(prog.leaves):
        L = q.predicate.leaves[i]
        L.column = plan.columns.index(leaf.column)
        L.op = leaf.op
        L.flags = leaf.flags
        L.literal = plan.columns.index(leaf.column2) if leaf.column2 is not None else leaf.literal
        L.modulus = leaf.modulus
    aggs = (_lib.Agg * max(1, len(plan.aggs)))()
    for i, (op, c, ddof, min_count) in enumerate(plan.aggs):
        aggs[i].op = op
        aggs[i].column = max(c, 0)
        aggs[i].ddof = ddof
        aggs[i].min_count = min_count
    q.n_aggs = len(plan.aggs)
    q.aggs = ctypes.cast(aggs, ctypes.POINTER(_lib.Agg))
    q.maintain_order = 1 if plan.maintain_order else 0
    return q, (key_arr, col_arr, aggs)


def _fill_predicate(pred, prog, names: list):
    """A dfr_predicate from a compiled Program whose leaves name columns of
    `names` (their index is the leaf's column)."""
    pred.n_leaves = len(prog.leaves)
    for i in range(8):
        pred.truth[i] = prog.truth[i] & 0xFFFFFFFF
    for i, leaf in enumerate(prog.leaves):
        L = pred.leaves[i]
        L.column = names.index(leaf.column)
        L.op = leaf.op
        L.flags = leaf.flags
        L.literal = names.index(leaf.column2) if leaf.column2 is not None else leaf.literal
        L.modulus = leaf.modulus


def build_value_program(spec, keep) -> "_lib.ValueProgram":
    """The dfr_value_program of a ValueSpec; `keep` is the query's filter
    Program (rows it drops are neither evaluated nor faulted)."""
    vp = _lib.ValueProgram()
    vp.n_ops = len(spec.ops)
    for i, (op, arg, dtype, literal) in enumerate(spec.ops):
        o = vp.ops[i]
        o.op, o.arg, o.dtype = op, arg, dtype
        o.literal = literal
    vp.n_conditions = len(spec.conditions)
    for c, prog in enumerate(spec.conditions):
        _fill_predicate(vp.conditions[c], prog, spec.inputs)
    _fill_predicate(vp.keep, keep, spec.inputs)
    return vp


class Context:
    def __init__(self, device: int = 0, budget: int = 0):
        self.lib = _lib.load()
        self.device = device
        h = ctypes.c_void_p()
        e = _err()
        _lib.check(self.lib.dfr_context_create(device, budget, ctypes.byref(h), ctypes.byref(e)), e)
        self.handle = h
        self.lock = threading.Lock()

    def close(self):
        if self.handle:
            self.lib.dfr_context_destroy(self.handle)
            self.handle = None

    def trim(self):
        """Returns the device memory the context caches for reuse."""
        self.lib.dfr_context_trim(self.handle)

    def tune(self, key: str, value: int):
        e = _err()
        _lib.check(self.lib.dfr_context_tune(self.handle, key.encode(), int(value), ctypes.byref(e)), e)

    def tune_get(self, key: str) -> int:
        v = ctypes.c_int64()
        e = _err()
        _lib.check(self.lib.dfr_context_tune_get(self.handle, key.encode(), ctypes.byref(v), ctypes.byref(e)), e)
        return int(v.value)

    def calibration(self) -> str:
        """What became of the device's calibration profile at creation:
        "applied: <path> (...)" or "none: <reason>"."""
        t = ctypes.c_void_p()
        e = _err()
        _lib.check(self.lib.dfr_context_calibration(self.handle, ctypes.byref(t), ctypes.byref(e)), e)
        try:
            return ctypes.string_at(t.value).decode()
        finally:
            self.lib.dfr_free(t)

    def observe(self) -> dict:
        o = _lib.Observation()
        self.lib.dfr_context_observe(self.handle, ctypes.byref(o))
        return {f: getattr(o, f) for f, _ in o._fields_}

    # columns -------------------------------------------------------------
    def upload(self, column: Column) -> DeviceColumn:
        dt = DTYPE_CODE[column.dtype]
        if isinstance(column, ChunkedColumn):
            # every chunk in one transfer, joined on the device
            from .selection import upload_pieces

            h = upload_pieces(self, dt, [(c, 0, None, len(c)) for c in column.chunks], len(column.labels or []))
            return DeviceColumn(self, h, dt, len(column), column.labels)
        vals = np.ascontiguousarray(column.values)
        bits = column.validity_bytes()
        h = ctypes.c_void_p()
        e = _err()
        _lib.check(
            self.lib.dfr_column_upload(
                self.handle,
                dt,
                vals.ctypes.data_as(ctypes.c_void_p),
                0,
                len(vals),
                None if bits is None else bits.ctypes.data_as(ctypes.c_void_p),
                0,
                len(column.labels or []),
                ctypes.byref(h),
                ctypes.byref(e),
            ),
            e,
        )
        return DeviceColumn(self, h, dt, len(vals), column.labels)

    def wrap(self, dtype: int, values_ptr: int, length: int, valid_ptr: int = 0, bit_offset: int = 0,
             offset: int = 0, labels=None, keepalive=None) -> DeviceColumn:
        h = ctypes.c_void_p()
        e = _err()
        _lib.check(
            self.lib.dfr_column_wrap(
                self.handle, dtype, ctypes.c_void_p(values_ptr), offset, length,
                ctypes.c_void_p(valid_ptr) if valid_ptr else None, bit_offset,
                len(labels or []), ctypes.byref(h), ctypes.byref(e),
            ),
            e,
        )
        return DeviceColumn(self, h, dtype, length, labels, keepalive)

    def wrap_array(self, values, valid_words=None, labels=None, dict_codes=False, length=None) -> DeviceColumn:
        """Wraps device arrays in place (anything with data_ptr(), dtype and
        n: datagen.DeviceArray); `valid_words` are uint64 LSB-first words."""
        dt = {np.dtype(np.int32): _lib.DFR_INT32, np.dtype(np.uint32): _lib.DFR_DICT,
              np.dtype(np.int64): _lib.DFR_INT64, np.dtype(np.float64): _lib.DFR_FLOAT64}[np.dtype(values.dtype)]
        if dict_codes:
            dt = _lib.DFR_DICT
        n = values.n if length is None else length
        return self.wrap(dt, values.data_ptr(), n, valid_words.data_ptr() if valid_words is not None else 0,
                         labels=labels, keepalive=(values, valid_words))

    # queries ---------------------------------------------------------------
    def groupby(self, plan: DevicePlan, keys: list, columns: list) -> Result:
        q, keep = build_query(plan, keys, columns)
        h = ctypes.c_void_p()
        e = _err()
        with self.lock:
            _lib.check(self.lib.dfr_groupby(self.handle, ctypes.byref(q), ctypes.byref(h), ctypes.byref(e)), e)
        del keep
        return Result(self, h, len(keys), len(plan.aggs))


    def compute(self, spec, keep, inputs: list) -> DeviceColumn:
        """dfr_column_compute: a ValueSpec's column over device `inputs`
        (one per spec.inputs), for the rows `keep` (a Program) keeps."""
        vp = build_value_program(spec, keep)
        arr = (ctypes.c_void_p * len(inputs))(*[c.handle for c in inputs])
        h = ctypes.c_void_p()
        e = _err()
        with self.lock:
            _lib.check(self.lib.dfr_column_compute(self.handle, ctypes.byref(vp), arr, len(inputs), ctypes.byref(h),
                                                   ctypes.byref(e)), e)
        return DeviceColumn(self, h, DTYPE_CODE[spec.dtype], inputs[0].length if inputs else 0)

    def groupby_into(self, plan: DevicePlan, keys: list, columns: list, out) -> int:
        """dfr_groupby_into: the query into caller-owned device buffers,
        allocating nothing. `out` is (capacity, [(values_ptr, validity_ptr)]
        per key, [(values_ptr, validity_ptr)] per aggregate, first_rows_ptr
        or 0). Returns the group count."""
        capacity, kbufs, abufs, first = out
        q, keep = build_query(plan, keys, columns)
        ko = (_lib.OutputColumn * max(1, len(kbufs)))(*[_lib.OutputColumn(v, m) for v, m in kbufs])
        ao = (_lib.OutputColumn * max(1, len(abufs)))(*[_lib.OutputColumn(v, m) for v, m in abufs])
        o = _lib.Output(capacity, ctypes.cast(ko, ctypes.POINTER(_lib.OutputColumn)),
                        ctypes.cast(ao, ctypes.POINTER(_lib.OutputColumn)), first or None)
        g = ctypes.c_int64()
        e = _err()
        with self.lock:
            _lib.check(self.lib.dfr_groupby_into(self.handle, ctypes.byref(q), ctypes.byref(o), ctypes.byref(g),
                                                 ctypes.byref(e)), e)
        del keep
        return g.value

    def groupby_batch(self, jobs) -> list:
        """Runs [(plan, keys, columns), ...] (or a PreparedBatch) as one
        dfr_groupby_batch call. Returns one entry per job: a Result, or the
        DfrError that query raised (the batch itself raises only when it
        could not run)."""
        prepared = jobs if isinstance(jobs, PreparedBatch) else PreparedBatch(jobs)
        return prepared.run(self)

    def prepare(self, plan: DevicePlan, keys: list, columns: list) -> "PreparedQuery":
        """A query whose C descriptor is built once, for repeated runs."""
        return PreparedQuery(self, plan, keys, columns)


class PreparedQuery:
    """One query's dfr_query, built once."""

    def __init__(self, ctx: "Context", plan: DevicePlan, keys: list, columns: list):
        self.ctx, self.plan, self.keys = ctx, plan, keys
        self.query, self._keep = build_query(plan, keys, columns)
        self._keep = (self._keep, keys, columns)

    def run(self) -> Result:
        h = ctypes.c_void_p()
        e = _err()
        with self.ctx.lock:
            _lib.check(self.ctx.lib.dfr_groupby(self.ctx.handle, ctypes.byref(self.query), ctypes.byref(h),
                                                ctypes.byref(e)), e)
        return Result(self.ctx, h, len(self.keys), len(self.plan.aggs))


class PreparedBatch:
    """The dfr_query array of a batch, built once."""

    def __init__(self, jobs: list):
        self.jobs = list(jobs)
        n = len(self.jobs)
        self.queries = (_lib.Query * max(n, 1))()
        self._keep = []
        for i, (plan, keys, columns) in enumerate(self.jobs):
            q, k = build_query(plan, keys, columns)
            self.queries[i] = q
            self._keep.append((k, keys, columns))

    def run(self, ctx: "Context") -> list:
        n = len(self.jobs)
        if n == 0:
            return []
        handles = (ctypes.c_void_p * n)()
        statuses = (ctypes.c_int32 * n)()
        errors = (ctypes.c_void_p * n)()
        e = _err()
        with ctx.lock:
            _lib.check(ctx.lib.dfr_groupby_batch(ctx.handle, self.queries, n, handles, statuses, errors,
                                                 ctypes.byref(e)), e)
        out = []
        for i, (plan, keys, _) in enumerate(self.jobs):
            if statuses[i] != _lib.DFR_OK:
                out.append(_lib.error_of(statuses[i], ctypes.c_void_p(errors[i])))
            else:
                out.append(Result(ctx, ctypes.c_void_p(handles[i]), len(keys), len(plan.aggs)))
        return out


class Server:
    """The serving loop (dfr_server_*): `lanes` worker threads with their own
    streams take waiting queries as batches, admitted by `budget` device
    bytes (0: no limit); submit() returns a ticket, wait() its Result."""

    def __init__(self, ctx: "Context", lanes: int = 2, max_batch: int = 256, budget: int = 0):
        self.ctx = ctx
        self.lib = ctx.lib
        h = ctypes.c_void_p()
        e = _err()
        _lib.check(self.lib.dfr_server_create(ctx.handle, lanes, max_batch, budget, ctypes.byref(h), ctypes.byref(e)), e)
        self.handle = h
        self._jobs: dict = {}

    def submit(self, job) -> int:
        """Queues a PreparedQuery (or (plan, keys, columns)); returns its ticket."""
        pq = job if isinstance(job, PreparedQuery) else PreparedQuery(self.ctx, *job)
        t = ctypes.c_int64()
        e = _err()
        _lib.check(self.lib.dfr_server_submit(self.handle, ctypes.byref(pq.query), ctypes.byref(t), ctypes.byref(e)), e)
        self._jobs[t.value] = pq
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1) -> Result:
        """The ticket's Result; its `latency_ms` is the server's time from
        submission to the end of the query's batch."""
        h = ctypes.c_void_p()
        lat = ctypes.c_int64()
        e = _err()
        st = self.lib.dfr_server_wait(self.handle, ticket, timeout_ms, ctypes.byref(h), ctypes.byref(lat),
                                      ctypes.byref(e))
        if st == _lib.DFR_PENDING:
            if e.value:
                self.lib.dfr_free(e)
            raise TimeoutError(f"ticket {ticket} is not done")
        pq = self._jobs.pop(ticket, None)
        _lib.check(st, e)
        r = Result(self.ctx, h, len(pq.keys), len(pq.plan.aggs))
        r.latency_ms = lat.value / 1e6
        return r

    def stats(self) -> dict:
        out = _lib.ServerStats()
        self.lib.dfr_server_stats_get(self.handle, ctypes.byref(out))
        return {f: getattr(out, f) for f, _ in out._fields_}

    def close(self):
        if self.handle:
            self.lib.dfr_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


_contexts: dict = {}
_ctx_lock = threading.Lock()


def context(device: int | None = None) -> Context:
    """The process's context for a device (created on first use)."""
    if device is None:
        device = device_ordinal()
    with _ctx_lock:
        ctx = _contexts.get(device)
        if ctx is None:
            ctx = Context(device, memory_limit())
            _contexts[device] = ctx
        return ctx


def result_frame(res: Result, plan: DevicePlan, key_cols: list) -> DataFrame:
    """Copies a device result into a host DataFrame: keys (dictionary keys
    decoded through their labels), then the aggregates."""
    series = []
    for k, name in enumerate(plan.keys):
        dt, vals, valid = res.key(k)
        if dt == _lib.DFR_DICT:
            col = Column(STRING, vals, valid, key_cols[k].labels)
        else:
            col = Column(_NAME_OF[dt], vals, valid)
        series.append(Series(name, col))
    for a, name in enumerate(plan.agg_names):
        dt, vals, valid = res.agg(a)
        series.append(Series(name, Column(_NAME_OF[dt], vals, valid)))
    return DataFrame(series)


def run_plan(ctx: Context, frame: DataFrame, plan: DevicePlan) -> tuple[DataFrame, Result]:
    """Uploads the columns a plan reads, runs it, and returns the host frame."""
    uploaded: dict = {}

    def dev(name):
        if name not in uploaded:
            spec = next((v for v in plan.values if v.name == name), None)
            if spec is None:
                uploaded[name] = ctx.upload(frame.column(name).column)
            else:
                # an aggregate's expression, computed for the kept rows
                uploaded[name] = ctx.compute(spec, plan.program, [dev(i) for i in spec.inputs])
        return uploaded[name]

    keys = [dev(k) for k in plan.keys]
    cols = [dev(c) for c in plan.columns]
    try:
        res = ctx.groupby(plan, keys, cols)
        out = result_frame(res, plan, keys)
    finally:
        for d in uploaded.values():
            d.release()
    return out, res
