"""On-disk inputs: Parquet String columns read as dictionary codes into one
running dictionary per column (native/dfparquet/dfparquet.cc:1-17), so they
feed the device's dictionary-code keys without decoding. Host: the codes
decode to exactly what pyarrow reads, across row groups whose dictionaries
differ, with nulls. Device: a group-by over the file equals the oracle's
over the same frame."""
import math

import numpy as np
import pyarrow as pa
import pyarrow.parquet as pq
import pytest

from dataframe_mojo_amd import col, len_
from dataframe_mojo_amd.parquet import read_parquet
from oracle import oracle


def _write(path, seed=0, groups=5, rows=3000):
    rng = np.random.default_rng(seed)
    writer = None
    expected = {"cat": [], "k": [], "v": [], "x": []}
    for g in range(groups):
        # each row group draws from its own, partly overlapping label set
        labels = [f"lab{(g * 7 + i) % 40:02d}" for i in range(12)]
        cat = [None if rng.random() < 0.05 else labels[int(rng.integers(0, 12))] for _ in range(rows)]
        k = rng.integers(0, 30, rows).astype(np.int64)
        v = pa.array(rng.integers(-100, 100, rows).astype(np.int32), mask=rng.random(rows) < 0.1)
        x = pa.array(rng.normal(size=rows), mask=rng.random(rows) < 0.05)
        t = pa.table({"cat": pa.array(cat, pa.string()), "k": k, "v": v, "x": x})
        if writer is None:
            writer = pq.ParquetWriter(path, t.schema, use_dictionary=True)
        writer.write_table(t)
        expected["cat"] += cat
        expected["k"] += k.tolist()
        expected["v"] += v.to_pylist()
        expected["x"] += x.to_pylist()
    writer.close()
    return expected


def test_dictionary_codes_decode_to_the_file_strings(tmp_path):
    path = str(tmp_path / "d.parquet")
    exp = _write(path)
    frame = read_parquet(path, batch_rows=1000)
    cat = frame.column("cat").column
    assert cat.dtype == "string" and len(cat.chunks) == 15
    assert cat.to_list() == exp["cat"]
    # one canonical code per label, shared by every chunk
    assert len(set(cat.labels)) == len(cat.labels)
    assert all(c.labels == cat.labels for c in cat.chunks)
    assert frame.column("k").column.to_list() == exp["k"]
    assert frame.column("v").column.to_list() == exp["v"]
    got_x = frame.column("x").column.to_list()
    assert all((a is None and b is None) or a == b for a, b in zip(got_x, exp["x"]))


def test_unsupported_parquet_types_raise(tmp_path):
    from dataframe_mojo_amd.plan import Unsupported

    path = str(tmp_path / "b.parquet")
    pq.write_table(pa.table({"b": pa.array([True, False])}), path)
    with pytest.raises(Unsupported):
        read_parquet(path)


@pytest.mark.gpu
@pytest.mark.parametrize("ordered", [False, True])
def test_groupby_over_parquet_dictionary_keys_matches_the_oracle(tmp_path, ordered):
    path = str(tmp_path / "g.parquet")
    _write(path, seed=3, groups=6, rows=20000)
    frame = read_parquet(path, batch_rows=7000)
    lf = (frame.filter(col("v") > -50).group_by("cat", "k", maintain_order=ordered)
          .agg(len_(), col("v").sum().alias("v_sum"), col("x").mean().alias("x_mean"), col("x").max().alias("x_max")))
    got = lf.collect(engine="accel").to_dict()
    exp = oracle.collect(lf)[0].to_dict()
    if not ordered:
        def order(d):
            rows = sorted(zip(*d.values()), key=lambda r: ((r[0] is None, r[0] or ""), r[1]))
            return {n: [r[i] for r in rows] for i, n in enumerate(d)}
        got, exp = order(got), order(exp)
    assert list(got) == list(exp)
    for name in exp:
        for a, b in zip(got[name], exp[name]):
            if isinstance(b, float):
                assert (math.isnan(a) and math.isnan(b)) or abs(a - b) <= 1e-12 * max(1.0, abs(b)), (name, a, b)
            else:
                assert a == b, (name, a, b)
