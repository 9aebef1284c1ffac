"""Typed, nullable columns and frames: the host-side shape of the reference's
storage (dataframe/storage/column.mojo:18-34 Column[T] with LSB-first
validity; storage/string.mojo:35-104 StringColumn with dictionary
StringCodes) and its eager/lazy API surface for this path
(frame.mojo:2244 filter, :3329-3395 group_by; lazy_api.mojo:418 filter,
:439-470 group_by, :3223 LazyGroupBy.agg, :1737 collect).

Only the filter -> group_by -> agg path is provided; every other frame
operation of the reference stays in the reference.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Iterable, Sequence

import numpy as np

from .expr import Expr, col

INT32 = "int32"
INT64 = "int64"
FLOAT64 = "float64"
STRING = "string"

_NP = {INT32: np.int32, INT64: np.int64, FLOAT64: np.float64}


class Column:
    """Values plus an optional validity mask (True = valid).

    A STRING column is held dictionary encoded: `codes` (uint32) index
    `labels`, and equal labels share one code (the canonical form the
    reference groups by, kernels/canonical_labels.mojo:13-47)."""

    def __init__(self, dtype: str, values, valid=None, labels: Sequence[str] | None = None):
        self.dtype = dtype
        if dtype == STRING:
            self.values = np.ascontiguousarray(values, dtype=np.uint32)
            self.labels = list(labels or [])
        else:
            self.values = np.ascontiguousarray(values, dtype=_NP[dtype])
            self.labels = None
        if valid is not None:
            valid = np.ascontiguousarray(valid, dtype=bool)
            if valid.shape != self.values.shape:
                raise ValueError("validity length differs from the values")
            if valid.all():
                valid = None
        self.valid = valid

    # constructors -------------------------------------------------------
    @staticmethod
    def int64(values, valid=None) -> "Column":
        return Column(INT64, values, valid)

    @staticmethod
    def int32(values, valid=None) -> "Column":
        return Column(INT32, values, valid)

    @staticmethod
    def float64(values, valid=None) -> "Column":
        return Column(FLOAT64, values, valid)

    @staticmethod
    def string(items: Iterable[str | None]) -> "Column":
        """Dictionary-encodes Python strings (None is null), labels in order
        of first appearance."""
        index: dict[str, int] = {}
        codes, valid = [], []
        for s in items:
            if s is None:
                codes.append(0)
                valid.append(False)
                continue
            codes.append(index.setdefault(s, len(index)))
            valid.append(True)
        return Column(STRING, np.array(codes, dtype=np.uint32), np.array(valid, dtype=bool), list(index))

    @staticmethod
    def dictionary(codes, labels: Sequence[str], valid=None) -> "Column":
        """A dictionary String column whose labels may repeat; codes are
        made canonical (one code per distinct label)."""
        labels = list(labels)
        first: dict[str, int] = {}
        remap = np.empty(len(labels), dtype=np.uint32)
        canon: list[str] = []
        for i, s in enumerate(labels):
            if s not in first:
                first[s] = len(canon)
                canon.append(s)
            remap[i] = first[s]
        codes = np.asarray(codes, dtype=np.uint32)
        if len(labels):
            codes = remap[np.minimum(codes, len(labels) - 1)]
        return Column(STRING, codes, valid, canon)

    def __len__(self) -> int:
        return int(self.values.shape[0])

    def null_count(self) -> int:
        return 0 if self.valid is None else int((~self.valid).sum())

    def is_valid(self) -> np.ndarray:
        return np.ones(len(self), dtype=bool) if self.valid is None else self.valid

    def to_list(self) -> list:
        ok = self.is_valid()
        if self.dtype == STRING:
            return [self.labels[int(c)] if v else None for c, v in zip(self.values, ok)]
        return [v.item() if o else None for v, o in zip(self.values, ok)]

    def validity_bytes(self) -> np.ndarray | None:
        if self.valid is None:
            return None
        return np.packbits(self.valid, bitorder="little")


class ChunkedColumn(Column):
    """A column held as several chunks of one dtype (a chunked Series,
    series.mojo chunks(); String chunks share one dictionary, as the
    reference groups chunked String keys only when every chunk carries
    codes into one dictionary, series.mojo:337-370). The device path
    uploads the chunks as they are (dfr_column_upload_selection, one
    transfer, joined on the device); `values` and `valid` join them on the
    host only when host code reads them."""

    def __init__(self, chunks: Sequence[Column]):
        chunks = list(chunks)
        if not chunks:
            raise ValueError("a chunked column needs at least one chunk")
        self.dtype = chunks[0].dtype
        if any(c.dtype != self.dtype for c in chunks):
            raise ValueError("chunks of one column share a dtype")
        if self.dtype == STRING and any(c.labels != chunks[0].labels for c in chunks):
            raise ValueError("String chunks share one dictionary")
        self.chunks = chunks
        self.labels = chunks[0].labels
        self._joined = None

    def _join(self):
        if self._joined is None:
            vals = np.concatenate([c.values for c in self.chunks])
            valid = np.concatenate([c.is_valid() for c in self.chunks])
            self._joined = (vals, None if valid.all() else valid)
        return self._joined

    @property
    def values(self):
        return self._join()[0]

    @property
    def valid(self):
        return self._join()[1]

    def __len__(self) -> int:
        return sum(len(c) for c in self.chunks)


@dataclass
class Series:
    name: str
    column: Column

    def __len__(self) -> int:
        return len(self.column)

    @property
    def dtype(self) -> str:
        return self.column.dtype

    def to_list(self) -> list:
        return self.column.to_list()


class DataFrame:
    def __init__(self, series: Sequence[Series]):
        self._series = list(series)
        names = [s.name for s in self._series]
        if len(set(names)) != len(names):
            raise ValueError("duplicate column names")
        lengths = {len(s) for s in self._series}
        if len(lengths) > 1:
            raise ValueError("columns differ in length")

    @property
    def columns(self) -> list[str]:
        return [s.name for s in self._series]

    def height(self) -> int:
        return len(self._series[0]) if self._series else 0

    def width(self) -> int:
        return len(self._series)

    def column(self, name: str) -> Series:
        for s in self._series:
            if s.name == name:
                return s
        raise KeyError(f"no column named {name!r}")

    def __getitem__(self, name: str) -> Series:
        return self.column(name)

    def to_dict(self) -> dict:
        return {s.name: s.to_list() for s in self._series}

    def rows(self) -> list[tuple]:
        cols = [s.to_list() for s in self._series]
        return list(zip(*cols))

    def lazy(self) -> "LazyFrame":
        return LazyFrame(self)

    # eager API: one-node lazy plans, as frame.mojo:2244-2290 does
    def filter(self, predicate: Expr) -> "LazyFrame":
        return self.lazy().filter(predicate)

    def group_by(self, *keys, maintain_order: bool = False) -> "LazyGroupBy":
        return self.lazy().group_by(*keys, maintain_order=maintain_order)

    def __repr__(self) -> str:
        head = " | ".join(self.columns)
        body = "\n".join(" | ".join("null" if v is None else str(v) for v in row) for row in self.rows()[:20])
        return f"{head}\n{body}"


@dataclass
class LazyFrame:
    """A source frame, its filters, and optionally one grouped aggregation."""

    source: DataFrame
    filters: list = field(default_factory=list)
    keys: list = field(default_factory=list)
    aggs: list = field(default_factory=list)
    maintain_order: bool = False
    grouped: bool = False

    def filter(self, predicate: Expr) -> "LazyFrame":
        if self.grouped:
            raise ValueError("filter after agg is not part of this path")
        return LazyFrame(self.source, self.filters + [predicate], [], [], False, False)

    def group_by(self, *keys, maintain_order: bool = False) -> "LazyGroupBy":
        flat: list = []
        for k in keys:
            if isinstance(k, (list, tuple)):
                flat.extend(k)
            else:
                flat.append(k)
        names = []
        for k in flat:
            if isinstance(k, Expr):
                if k.kind != "col":
                    raise ValueError("group keys are column names")
                names.append(k.args[0])
            else:
                names.append(str(k))
        return LazyGroupBy(self, names, maintain_order)

    def collect(self, engine: str = "accel") -> DataFrame:
        from . import provider

        if engine not in ("accel", "auto"):
            raise ValueError(f"engine {engine!r}: this package provides the accelerator engine only")
        if engine == "auto" and not provider.installed():
            raise RuntimeError("no device provider installed")
        return provider.execute(self)[0]

    def profile(self):
        from . import provider

        return provider.execute_profiled(self)

    def explain(self) -> str:
        from . import provider

        return provider.describe(self)


@dataclass
class LazyGroupBy:
    frame: LazyFrame
    keys: list
    maintain_order: bool

    def agg(self, *exprs) -> LazyFrame:
        flat: list = []
        for e in exprs:
            if isinstance(e, (list, tuple)):
                flat.extend(e)
            else:
                flat.append(e)
        return LazyFrame(self.frame.source, list(self.frame.filters), list(self.keys), flat, self.maintain_order, True)

    def len(self, name: str = "len") -> LazyFrame:
        from .expr import len_

        return self.agg(len_().alias(name))

    def collect(self, engine: str = "accel") -> DataFrame:
        raise ValueError("call agg(...) before collect()")


def frame_from_dict(data: dict, dtypes: dict | None = None) -> DataFrame:
    """Convenience: {name: list} -> DataFrame, inferring int64/float64/string."""
    out = []
    for name, vals in data.items():
        dt = (dtypes or {}).get(name)
        vals = list(vals)
        sample = next((v for v in vals if v is not None), 0)
        if dt is None:
            dt = STRING if isinstance(sample, str) else FLOAT64 if isinstance(sample, float) else INT64
        if dt == STRING:
            c = Column.string(vals)
        else:
            valid = np.array([v is not None for v in vals], dtype=bool)
            arr = np.array([0 if v is None else v for v in vals], dtype=_NP[dt])
            c = Column(dt, arr, valid)
        out.append(Series(name, c))
    return DataFrame(out)


__all__ = [
    "Column",
    "Series",
    "DataFrame",
    "LazyFrame",
    "LazyGroupBy",
    "INT32",
    "INT64",
    "FLOAT64",
    "STRING",
    "col",
    "frame_from_dict",
]
