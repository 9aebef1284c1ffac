"""dataframe_mojo_amd — an MI355X-native accelerator provider for
dataframe_mojo's filter -> group_by -> agg path.

The device work runs in libdfrocm.so (hand-written HIP for gfx950, C ABI in
include/dfrocm.h). This package is the host side: the provider's thirteen
boundary functions (`provider`), the frame/expression surface the path
needs (`frame`, `expr`), the plan compiler (`plan`), device contexts and
columns, batched execution and the serving loop (`runtime`), chunked and
selected inputs (`selection`), the autotuner and placement calibration
(`autotune`, `calibration`).
"""
from .expr import col, len_, lit, when
from .frame import FLOAT64, INT32, INT64, STRING, ChunkedColumn, Column, DataFrame, LazyFrame, Series, frame_from_dict
from . import provider

__all__ = [
    "col",
    "lit",
    "len_",
    "when",
    "Column",
    "ChunkedColumn",
    "Series",
    "DataFrame",
    "LazyFrame",
    "frame_from_dict",
    "INT32",
    "INT64",
    "FLOAT64",
    "STRING",
    "provider",
]
