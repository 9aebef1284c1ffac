"""Selection hand-off from CPU producers: the batches a CPU pipeline hands
to a selected consumer without materialising them (the reference's
SelectionBatch / SelectionBatches, vector/selection_batches.mojo:13-160,
consumed through QueryContext._selection_execute,
query_execution/placement.mojo:285-330).

A batch keeps, per output column, its source column, a start, and either
a row list (shared by the columns one filter selected) or none (a
contiguous slice of `count` rows from start). The device path uploads each
plan column's batches with dfr_column_upload_selection: only the source
rows each batch selects are copied, in one transfer, and the device
gathers them into one column; no host copy of the selected rows is made.
`materialize` is the reference's explicit adapter for consumers that need
ordinary columns (host).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .frame import STRING, ChunkedColumn, Column, DataFrame, LazyFrame, Series


@dataclass
class SelectionBatch:
    """vector/selection_batches.mojo:13-75."""

    morsel: int
    columns: list  # source Column per output column
    starts: list  # int per column
    lists: list  # index into `rows` per column, or -1 for a contiguous slice
    rows: list  # uint32 row lists
    count: int

    def validate(self):
        """Every mapping checked before a consumer submits work
        (selection_batches.mojo:34-58)."""
        width = len(self.columns)
        if self.count < 0 or len(self.starts) != width or len(self.lists) != width:
            raise ValueError("selection batch has inconsistent dimensions")
        largest = []
        for r in self.rows:
            if len(r) != self.count:
                raise ValueError("selection batch row list has the wrong length")
            largest.append(int(np.max(r)) if len(r) else -1)
        for c in range(width):
            start, lst = self.starts[c], self.lists[c]
            if start < 0 or lst < -1 or lst >= len(self.rows):
                raise ValueError("selection batch has an invalid column mapping")
            last = self.count - 1 if lst < 0 else largest[lst]
            if start > len(self.columns[c]) or (last >= 0 and last >= len(self.columns[c]) - start):
                raise ValueError("selection batch row is outside its source column")

    def piece(self, c: int):
        """Column c of this batch as (source, start, rows or None, count)."""
        lst = self.lists[c]
        rows = None if lst < 0 else np.ascontiguousarray(self.rows[lst], dtype=np.uint32)
        return self.columns[c], self.starts[c], rows, self.count

    def materialize(self, names: list) -> DataFrame:
        """The explicit host adapter (selection_batches.mojo:60-75)."""
        self.validate()
        out = []
        for c, name in enumerate(names):
            src, start, rows, count = self.piece(c)
            idx = start + (np.arange(count) if rows is None else rows.astype(np.int64))
            out.append(Series(name, Column(src.dtype, src.values[idx], src.is_valid()[idx], src.labels)))
        return DataFrame(out)


@dataclass
class SelectionBatches:
    """selection_batches.mojo:78-160: batches in morsel order over one
    schema (`names` and a dtype per column)."""

    names: list
    dtypes: list
    batches: list = field(default_factory=list)
    rows: int = 0

    def append(self, batch: SelectionBatch):
        if batch.count < 0:
            raise ValueError("selection batch cardinality overflow")
        self.rows += batch.count
        self.batches.append(batch)

    def validate(self):
        previous = -1
        for b in self.batches:
            b.validate()
            if b.morsel < previous or len(b.columns) != len(self.names):
                raise ValueError("selection batches have inconsistent order or schema")
            previous = b.morsel
            for c, src in enumerate(b.columns):
                if src.dtype != self.dtypes[c]:
                    raise ValueError("selection batches have inconsistent column types")
                if src.dtype == STRING and src.labels != b.columns[c].labels:
                    raise ValueError("String sources share one dictionary")

    def labels(self, c: int):
        for b in self.batches:
            return b.columns[c].labels
        return [] if self.dtypes[c] == STRING else None

    def materialize(self) -> DataFrame:
        """One host frame of every batch's rows, in order."""
        self.validate()
        parts = [b.materialize(self.names) for b in self.batches]
        cols = []
        for c, name in enumerate(self.names):
            chunks = [p.column(name).column for p in parts] or [Column(self.dtypes[c], [], None, self.labels(c))]
            cols.append(Series(name, ChunkedColumn(chunks)))
        return DataFrame(cols)


def upload_pieces(ctx, dtype: int, pieces: list, dict_size: int):
    """dfr_column_upload_selection over (source Column, start, rows or None,
    count) pieces; returns the column handle (ctypes.c_void_p). The sources'
    buffers stay alive for the call (the library copies what it reads)."""
    arr = (_lib.Selection * max(1, len(pieces)))()
    keep = []
    for i, (src, start, rows, count) in enumerate(pieces):
        vals = np.ascontiguousarray(src.values)
        bits = src.validity_bytes()
        keep.extend([vals, bits, rows])
        s = arr[i]
        s.source.values = vals.ctypes.data if len(vals) else None
        s.source.offset = 0
        s.source.length = len(vals)
        s.source.validity = bits.ctypes.data if bits is not None else None
        s.source.bit_offset = 0
        s.start = int(start)
        s.rows = rows.ctypes.data if rows is not None else None
        s.count = int(count)
    h = ctypes.c_void_p()
    e = ctypes.c_void_p()
    _lib.check(ctx.lib.dfr_column_upload_selection(ctx.handle, dtype, arr, len(pieces), dict_size,
                                                    ctypes.byref(h), ctypes.byref(e)), e)
    del keep
    return h


def selection_execute(plan: LazyFrame, batches: SelectionBatches, device: int = 0) -> DataFrame:
    """Runs `plan` (filter -> group_by -> agg over the batches' schema) on
    the device over the batches, as the reference's selected consumer
    (QueryContext._selection_execute, planning/context.mojo:266,367)."""
    from . import runtime
    from .plan import DTYPE_CODE, compile_plan

    batches.validate()
    schema = DataFrame([Series(n, Column(d, [], None, batches.labels(c)))
                        for c, (n, d) in enumerate(zip(batches.names, batches.dtypes))])
    lf = LazyFrame(schema, plan.filters, plan.keys, plan.aggs, plan.maintain_order, plan.grouped)
    dplan = compile_plan(lf)
    ctx = runtime.context(device)
    uploaded = {}

    def dev(name):
        if name not in uploaded:
            c = batches.names.index(name)
            pieces = [b.piece(c) for b in batches.batches]
            labels = batches.labels(c)
            h = upload_pieces(ctx, DTYPE_CODE[batches.dtypes[c]], pieces, len(labels or []))
            uploaded[name] = runtime.DeviceColumn(ctx, h, DTYPE_CODE[batches.dtypes[c]], batches.rows, labels)
        return uploaded[name]

    keys = [dev(k) for k in dplan.keys]
    cols = [dev(c) for c in dplan.columns]
    try:
        res = ctx.groupby(dplan, keys, cols)
        out = runtime.result_frame(res, dplan, keys)
        res.release()
    finally:
        for d in uploaded.values():
            d.release()
    return out
