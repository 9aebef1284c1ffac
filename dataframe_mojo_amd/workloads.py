"""The workloads the benchmark times and the autotuner measures: their
queries over typed columns and the device-resident column sets of
datagen.py.

  zipf_200m   filter(a > 0).group_by(key).agg(sum/count/min/max(y),
              sum/mean/var/min(x)); key Int32 Zipf(1.1) over 1e6
  dict_500m   filter(k2 % 3 != 0).group_by(k1, k2).agg(count/sum/mean(v));
              k1 a dictionary String of 1000 labels, k2 Int64 in [0, 1e7)
  batch       filter(value > t).group_by(key).agg(len, sum/max(value),
              sum/mean(x)), one query per threshold t
"""
import numpy as np


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


def dict_query(maintain_order: bool = False):
    from dataframe_mojo_amd import Column, DataFrame, Series, col

    schema = DataFrame([
        Series("k1", Column.dictionary([], [])),
        Series("k2", Column.int64([])),
        Series("v", Column.float64([])),
    ])
    return (schema.filter(col("k2") % 3 != 0).group_by("k1", "k2", maintain_order=maintain_order).agg(
        col("v").count().alias("v_count"), col("v").sum().alias("v_sum"), col("v").mean().alias("v_mean")))


def wrap_dict(ctx, cols, n):
    from dataframe_mojo_amd import datagen

    return {
        "k1": ctx.wrap_array(cols["k1"], labels=datagen.dict_labels(), dict_codes=True, length=n),
        "k2": ctx.wrap_array(cols["k2"], length=n),
        "v": ctx.wrap_array(cols["v"], cols["v.valid"], length=n),
    }


def host_dict(cols, n):
    from dataframe_mojo_amd import Column, DataFrame, Series, datagen

    return DataFrame([
        Series("k1", Column.dictionary(cols["k1"].numpy(n).astype(np.int64), datagen.dict_labels())),
        Series("k2", Column.int64(cols["k2"].numpy(n))),
        Series("v", Column.float64(cols["v"].numpy(n), datagen.valid_from_words(cols["v.valid"], n))),
    ])


def host_zipf(cols, n):
    from dataframe_mojo_amd import Column, DataFrame, Series, datagen

    return DataFrame([
        Series("key", Column.int32(cols["key"].numpy(n))),
        Series("a", Column.int64(cols["a"].numpy(n))),
        Series("x", Column.float64(cols["x"].numpy(n), datagen.valid_from_words(cols["x.valid"], n))),
        Series("y", Column.int64(cols["y"].numpy(n))),
    ])


def wrap_zipf(ctx, cols, n):
    return {
        "key": ctx.wrap_array(cols["key"], length=n),
        "a": ctx.wrap_array(cols["a"], length=n),
        "x": ctx.wrap_array(cols["x"], cols["x.valid"], length=n),
        "y": ctx.wrap_array(cols["y"], length=n),
    }


def batch_query(threshold: int):
    from dataframe_mojo_amd import Column, DataFrame, Series, col, len_

    schema = DataFrame([
        Series("key", Column.int64([])),
        Series("value", Column.int64([])),
        Series("x", Column.float64([])),
    ])
    return schema.filter(col("value") > threshold).group_by("key").agg(
        len_(), col("value").sum().alias("v_sum"), col("value").max().alias("v_max"),
        col("x").sum().alias("x_sum"), col("x").mean().alias("x_mean"))
