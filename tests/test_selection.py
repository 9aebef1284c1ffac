"""Chunked columns and selection hand-off (the reference's SelectionBatch,
vector/selection_batches.mojo:13-160): host validation and the host
adapter on the CPU; on the device, a query over chunked or selected inputs
equals the oracle's answer over the materialised frame (keys, validity and
integer aggregates exact, Float64 within 1e-12 relative), including odd
chunk lengths, row lists, nulls, dictionary keys, empty batches, and
maintain_order (first rows counted across batches)."""
import math

import numpy as np
import pytest

from dataframe_mojo_amd import ChunkedColumn, Column, DataFrame, Series, col, len_
from dataframe_mojo_amd.selection import SelectionBatch, SelectionBatches, selection_execute
from oracle import oracle

LABELS = [f"k{i}" for i in range(37)]


def _source(rng, dtype, n, p_null=0.1):
    valid = rng.random(n) > p_null
    if dtype == "int64":
        return Column.int64(rng.integers(-1000, 1000, n), valid)
    if dtype == "float64":
        return Column.float64(rng.normal(size=n), valid)
    if dtype == "dict":
        return Column.dictionary(rng.integers(0, len(LABELS), n), LABELS, valid)
    return Column.int32(rng.integers(0, 50, n), valid)


def _batches(rng, n_batches, with_lists=True, empty_every=0):
    names, dtypes = ["k", "d", "v", "x"], ["int32", "string", "int64", "float64"]
    sb = SelectionBatches(names, dtypes)
    for b in range(n_batches):
        n = int(rng.integers(1, 3000))
        srcs = [_source(rng, "int32", n), _source(rng, "dict", n), _source(rng, "int64", n),
                _source(rng, "float64", n)]
        count = 0 if (empty_every and b % empty_every == 0) else int(rng.integers(0, n + 1))
        start = int(rng.integers(0, n - count + 1)) if not with_lists else int(rng.integers(0, max(1, n // 4)))
        rows = []
        lists = [-1] * 4
        if with_lists and count:
            span = n - start
            rows.append(np.sort(rng.choice(span, size=min(count, span), replace=False)).astype(np.uint32))
            count = len(rows[0])
            lists = [0, 0, 0, -1] if b % 2 else [0, 0, 0, 0]
            if lists[3] < 0:
                # a column read as a contiguous slice of the same count
                start3 = int(rng.integers(0, n - count + 1))
                sb.append(SelectionBatch(b, srcs, [start, start, start, start3], lists, rows, count))
                continue
        sb.append(SelectionBatch(b, srcs, [start] * 4, lists, rows, count))
    return sb


def _plan(frame, ordered):
    return (frame.filter(col("v") > -500).group_by("k", "d", maintain_order=ordered)
            .agg(len_(), col("v").sum().alias("v_sum"), col("v").min().alias("v_min"),
                 col("x").mean().alias("x_mean"), col("x").var().alias("x_var"), col("x").count().alias("x_n")))


def _rows_by_key(d, n_keys=2):
    names = list(d)
    rows = list(zip(*d.values()))
    rows.sort(key=lambda r: tuple((v is None, v if v is not None else 0) for v in r[:n_keys]))
    return {name: [r[i] for r in rows] for i, name in enumerate(names)}


def _assert_same(got, exp, sort=False):
    g, e = got.to_dict(), exp.to_dict()
    assert list(g) == list(e)
    if sort:
        g, e = _rows_by_key(g), _rows_by_key(e)
    for name in e:
        assert len(g[name]) == len(e[name]), name
        for a, b in zip(g[name], e[name]):
            if isinstance(b, float) and a is not None:
                assert (math.isnan(a) and math.isnan(b)) or abs(a - b) <= 1e-12 * max(1.0, abs(b)), (name, a, b)
            else:
                assert a == b, (name, a, b)


def test_validate_rejects_rows_outside_the_source():
    src = Column.int64(np.arange(10))
    SelectionBatch(0, [src], [8], [-1], [], 2).validate()
    with pytest.raises(ValueError):
        SelectionBatch(0, [src], [9], [-1], [], 2).validate()
    with pytest.raises(ValueError):
        SelectionBatch(0, [src], [3], [0], [np.array([1, 7], dtype=np.uint32)], 2).validate()
    with pytest.raises(ValueError):
        SelectionBatch(0, [src], [0], [0], [np.array([1], dtype=np.uint32)], 2).validate()
    sb = SelectionBatches(["a"], ["int64"])
    sb.append(SelectionBatch(2, [src], [0], [-1], [], 1))
    sb.append(SelectionBatch(1, [src], [0], [-1], [], 1))
    with pytest.raises(ValueError):
        sb.validate()


def test_materialize_takes_the_selected_rows():
    src = Column.int64(np.arange(100, 110), np.arange(10) % 3 != 0)
    b = SelectionBatch(0, [src, src], [2, 5], [0, -1], [np.array([0, 3, 4], dtype=np.uint32)], 3)
    f = b.materialize(["a", "b"])
    assert f.column("a").column.to_list() == [102, 105, None]
    assert f.column("b").column.to_list() == [105, None, 107]


def test_chunked_column_joins_on_the_host_only_when_read():
    c = ChunkedColumn([Column.int64([1, 2]), Column.int64([3], [False])])
    assert len(c) == 3 and c.to_list() == [1, 2, None]
    with pytest.raises(ValueError):
        ChunkedColumn([Column.int64([1]), Column.int32([1])])


@pytest.mark.gpu
@pytest.mark.parametrize("seed,ordered", [(1, False), (2, True), (3, False)])
def test_chunked_columns_match_the_oracle(seed, ordered):
    rng = np.random.default_rng(seed)
    chunks = {"k": [], "d": [], "v": [], "x": []}
    for _ in range(int(rng.integers(1, 9))):
        n = int(rng.integers(0, 5000))
        chunks["k"].append(_source(rng, "int32", n))
        chunks["d"].append(_source(rng, "dict", n))
        chunks["v"].append(_source(rng, "int64", n))
        chunks["x"].append(_source(rng, "float64", n))
    frame = DataFrame([Series(k, ChunkedColumn(v)) for k, v in chunks.items()])
    joined = DataFrame([Series(k, Column(c.dtype, c.values, c.valid, c.labels))
                        for k, c in ((k, ChunkedColumn(v)) for k, v in chunks.items())])
    got = _plan(frame, ordered).collect(engine="accel")
    exp = oracle.collect(_plan(joined, ordered))[0]
    _assert_same(got, exp, sort=not ordered)


@pytest.mark.gpu
@pytest.mark.parametrize("seed,ordered,lists,empty", [(4, False, True, 0), (5, True, True, 3), (6, False, False, 2)])
def test_selection_batches_match_the_oracle(seed, ordered, lists, empty):
    rng = np.random.default_rng(seed)
    sb = _batches(rng, int(rng.integers(1, 12)), with_lists=lists, empty_every=empty)
    host = sb.materialize()
    got = selection_execute(_plan(host, ordered), sb)
    exp = oracle.collect(_plan(host, ordered))[0]
    _assert_same(got, exp, sort=not ordered)


@pytest.mark.gpu
def test_selection_row_outside_the_source_is_invalid_before_any_transfer():
    from dataframe_mojo_amd import _lib, runtime
    from dataframe_mojo_amd.selection import upload_pieces

    ctx = runtime.context(0)
    src = Column.int64(np.arange(10))
    with pytest.raises(_lib.DfrError) as e:
        upload_pieces(ctx, _lib.DFR_INT64, [(src, 4, np.array([0, 6], dtype=np.uint32), 2)], 0)
    assert e.value.status == _lib.DFR_INVALID
