"""On-disk inputs: Parquet columns read into this package's frames with
String columns kept as dictionary codes, never decoded to strings, so they
feed the device's dictionary-code keys (DFR_DICT) directly.

This follows the reference's dictionary stream
(native/dfparquet/dfparquet.cc:1-17, `dfq_read_parquet_stream_dict`: a
string column stays dictionary-encoded, and each batch's codes index one
running dictionary per column that only grows) and its StringCodes, which a
String column keeps beside its strings when its source stored it
dictionary-encoded (storage/string.mojo:35-49). The reference's reader is
Arrow C++ behind a C ABI; here the same Arrow library's Parquet reader
(pyarrow) decodes the pages.

Each read batch becomes one chunk of a ChunkedColumn, so the chunks reach
the device in one transfer and are joined there
(dfr_column_upload_selection). Every chunk's codes index the column's final
running dictionary, in which equal labels share one code (the canonical
form the device groups by, kernels/canonical_labels.mojo:13-47).
"""
from __future__ import annotations

import numpy as np

from .frame import FLOAT64, INT32, INT64, STRING, ChunkedColumn, Column, DataFrame, Series
from .plan import Unsupported


class _RunningDictionary:
    """Labels in first-seen order; codes of later batches extend it."""

    def __init__(self):
        self.labels: list[str] = []
        self.index: dict[str, int] = {}

    def remap(self, batch_labels) -> np.ndarray:
        """The running code of each of a batch dictionary's entries."""
        out = np.empty(len(batch_labels), dtype=np.uint32)
        for i, s in enumerate(batch_labels):
            code = self.index.get(s)
            if code is None:
                code = self.index[s] = len(self.labels)
                self.labels.append(s)
            out[i] = code
        return out


def _numeric(arr, dtype: str) -> Column:
    import pyarrow as pa
    import pyarrow.compute as pc

    np_type = {INT32: np.int32, INT64: np.int64, FLOAT64: np.float64}[dtype]
    valid = None
    if arr.null_count:
        valid = np.asarray(arr.is_valid())
        arr = pc.fill_null(arr, pa.scalar(0, type=arr.type))
    return Column(dtype, np.asarray(arr.to_numpy(zero_copy_only=False), dtype=np_type), valid)


def _dtype_of(field) -> str:
    import pyarrow as pa

    t = field.type
    if pa.types.is_dictionary(t):
        t = t.value_type
    if pa.types.is_int32(t):
        return INT32
    if pa.types.is_int64(t):
        return INT64
    if pa.types.is_float64(t):
        return FLOAT64
    if pa.types.is_string(t) or pa.types.is_large_string(t):
        return STRING
    raise Unsupported(f"Parquet column {field.name!r} of type {field.type} is not on the device path")


def read_parquet(path: str, columns: list | None = None, batch_rows: int = 1 << 20) -> DataFrame:
    """The file's columns (all, or `columns`) as a frame of chunked columns,
    one chunk per batch of at most `batch_rows` rows. Int32, Int64 and
    Float64 columns keep their values and nulls; String columns become
    dictionary codes into one running dictionary per column."""
    import pyarrow.parquet as pq

    schema = pq.read_schema(path)
    names = list(columns) if columns is not None else list(schema.names)
    dtypes = {n: _dtype_of(schema.field(n)) for n in names}
    strings = [n for n in names if dtypes[n] == STRING]
    pf = pq.ParquetFile(path, read_dictionary=strings)
    chunks: dict[str, list] = {n: [] for n in names}
    dicts = {n: _RunningDictionary() for n in strings}
    for batch in pf.iter_batches(batch_size=batch_rows, columns=names, use_threads=True):
        for n in names:
            arr = batch.column(batch.schema.get_field_index(n))
            if dtypes[n] != STRING:
                chunks[n].append(_numeric(arr, dtypes[n]))
                continue
            remap = dicts[n].remap(arr.dictionary.to_pylist())
            idx = arr.indices
            valid = np.asarray(idx.is_valid()) if idx.null_count else None
            raw = np.asarray(idx.fill_null(0).to_numpy(zero_copy_only=False), dtype=np.int64)
            codes = remap[raw] if len(remap) else np.zeros(len(raw), dtype=np.uint32)
            chunks[n].append((codes, valid))
    out = []
    for n in names:
        if dtypes[n] == STRING:
            labels = dicts[n].labels
            parts = [Column(STRING, c, v, labels) for c, v in chunks[n]]
        else:
            parts = chunks[n]
        if not parts:
            parts = [Column(dtypes[n], [], None, dicts[n].labels if n in dicts else None)]
        out.append(Series(n, ChunkedColumn(parts)))
    return DataFrame(out)
