"""ctypes binding of libdfrocm (include/dfrocm.h).

This is the Python twin of the binding a Mojo provider makes with
``OwnedDLHandle`` (dataframe/io/native_library.mojo:22-110): the library is
found next to this file or through ``DATAFRAME_ROCM_LIBRARY``; every call
returns a status and a malloc'd error string freed with ``dfr_free``.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_ENV = "DATAFRAME_ROCM_LIBRARY"

DFR_OK = 0
DFR_UNSUPPORTED = 1
DFR_DATA_FAULT = 2
DFR_CAPACITY = 3
DFR_DEVICE = 4
DFR_INVALID = 5
DFR_PENDING = 6

DFR_INT32 = 1
DFR_INT64 = 2
DFR_FLOAT64 = 3
DFR_DICT = 4

DFR_BUFFER_KEY = 0
DFR_BUFFER_AGG = 1
DFR_BUFFER_FIRST_ROWS = 2
DFR_BUFFER_VALUES = 0
DFR_BUFFER_VALIDITY = 1

EQ, NE, LT, LE, GT, GE, IS_NULL, IS_NOT_NULL = range(8)
LEAF_MOD = 1
LEAF_NOT = 2
LEAF_COLUMN = 4
MAX_LEAVES = 8

AGG_LEN, AGG_COUNT, AGG_SUM, AGG_MIN, AGG_MAX, AGG_MEAN, AGG_VAR, AGG_STD = range(8)

# every symbol include/dfrocm.h declares
EXPORTS = [
    "dfr_abi_version",
    "dfr_free",
    "dfr_device_count",
    "dfr_device_info_get",
    "dfr_context_create",
    "dfr_context_destroy",
    "dfr_context_observe",
    "dfr_context_tune",
    "dfr_context_tune_get",
    "dfr_context_trim",
    "dfr_build_id",
    "dfr_context_calibration",
    "dfr_column_upload",
    "dfr_column_wrap",
    "dfr_column_release",
    "dfr_column_length",
    "dfr_column_null_count",
    "dfr_column_distinct",
    "dfr_groupby",
    "dfr_groupby_batch",
    "dfr_groupby_into",
    "dfr_column_slice",
    "dfr_column_upload_selection",
    "dfr_column_compute",
    "dfr_server_create",
    "dfr_server_submit",
    "dfr_server_wait",
    "dfr_server_stats_get",
    "dfr_server_destroy",
    "dfr_result_groups",
    "dfr_result_key_dtype",
    "dfr_result_agg_dtype",
    "dfr_result_copy_key",
    "dfr_result_copy_agg",
    "dfr_result_copy_first_rows",
    "dfr_result_device_bytes",
    "dfr_result_device_ms",
    "dfr_result_phase_ms",
    "dfr_result_release",
    "dfr_result_buffer",
    "dfr_buffer_copy_to_host",
    "dfr_buffer_release",
]


class DeviceInfo(ctypes.Structure):
    _fields_ = [
        ("ordinal", ctypes.c_int32),
        ("compute_units", ctypes.c_int32),
        ("memory_bytes", ctypes.c_int64),
        ("name", ctypes.c_char * 64),
        ("arch", ctypes.c_char * 32),
    ]


class Source(ctypes.Structure):
    """dfr_source: a host source column window."""

    _fields_ = [
        ("values", ctypes.c_void_p),
        ("offset", ctypes.c_int64),
        ("length", ctypes.c_int64),
        ("validity", ctypes.c_void_p),
        ("bit_offset", ctypes.c_int64),
    ]


class Selection(ctypes.Structure):
    """dfr_selection: one batch of a chunked or selected column."""

    _fields_ = [
        ("source", Source),
        ("start", ctypes.c_int64),
        ("rows", ctypes.c_void_p),
        ("count", ctypes.c_int64),
    ]


class Observation(ctypes.Structure):
    _fields_ = [
        (n, ctypes.c_int64)
        for n in (
            "launches",
            "completed",
            "upload_bytes",
            "download_bytes",
            "reserved_bytes",
            "peak_reserved_bytes",
            "waits",
            "queries",
            "jit_compiles",
            "jit_launches",
            "hot_keys",
        )
    ]


class Leaf(ctypes.Structure):
    _fields_ = [
        ("column", ctypes.c_int32),
        ("op", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("literal", ctypes.c_int64),
        ("modulus", ctypes.c_int64),
    ]


class Predicate(ctypes.Structure):
    _fields_ = [
        ("n_leaves", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("truth", ctypes.c_uint32 * 8),
        ("leaves", Leaf * MAX_LEAVES),
    ]


MAX_VALUE_OPS = 32
MAX_CONDITIONS = 4
VAL_COLUMN, VAL_LITERAL, VAL_NULL, VAL_ADD, VAL_SUB, VAL_MUL, VAL_DIV, VAL_NEG, VAL_ABS, VAL_CASE = range(10)


class ValueOp(ctypes.Structure):
    _fields_ = [
        ("op", ctypes.c_int32),
        ("arg", ctypes.c_int32),
        ("dtype", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("literal", ctypes.c_int64),
    ]


class ValueProgram(ctypes.Structure):
    _fields_ = [
        ("n_ops", ctypes.c_int32),
        ("n_conditions", ctypes.c_int32),
        ("ops", ValueOp * MAX_VALUE_OPS),
        ("conditions", Predicate * MAX_CONDITIONS),
        ("keep", Predicate),
    ]


class Agg(ctypes.Structure):
    _fields_ = [
        ("op", ctypes.c_int32),
        ("column", ctypes.c_int32),
        ("ddof", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("min_count", ctypes.c_int64),
    ]


class Query(ctypes.Structure):
    _fields_ = [
        ("n_keys", ctypes.c_int32),
        ("n_columns", ctypes.c_int32),
        ("keys", ctypes.POINTER(ctypes.c_void_p)),
        ("columns", ctypes.POINTER(ctypes.c_void_p)),
        ("predicate", Predicate),
        ("n_aggs", ctypes.c_int32),
        ("maintain_order", ctypes.c_int32),
        ("aggs", ctypes.POINTER(Agg)),
    ]


class OutputColumn(ctypes.Structure):
    _fields_ = [("values", ctypes.c_void_p), ("validity", ctypes.c_void_p)]


class Output(ctypes.Structure):
    _fields_ = [
        ("capacity", ctypes.c_int64),
        ("keys", ctypes.POINTER(OutputColumn)),
        ("aggs", ctypes.POINTER(OutputColumn)),
        ("first_rows", ctypes.c_void_p),
    ]


def library_path() -> str:
    return os.environ.get(LIBRARY_ENV) or os.path.join(HERE, "libdfrocm.so")


_lib = None


def load(required: bool = True):
    """Loads libdfrocm; raises when it is missing and `required`."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        if required:
            raise RuntimeError(
                f"libdfrocm is not built ({path} missing); run python -m dataframe_mojo_amd.build"
            )
        return None
    lib = ctypes.CDLL(path)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    err = ctypes.POINTER(ctypes.c_char_p)
    errp = ctypes.POINTER(ctypes.c_void_p)
    lib.dfr_abi_version.restype = ctypes.c_int
    lib.dfr_free.argtypes = [vp]
    lib.dfr_device_count.argtypes = [ctypes.POINTER(i32), errp]
    lib.dfr_device_info_get.argtypes = [i32, ctypes.POINTER(DeviceInfo), errp]
    lib.dfr_context_create.argtypes = [i32, i64, ctypes.POINTER(vp), errp]
    lib.dfr_context_destroy.argtypes = [vp]
    lib.dfr_context_observe.argtypes = [vp, ctypes.POINTER(Observation)]
    lib.dfr_context_tune.argtypes = [vp, ctypes.c_char_p, i64, errp]
    lib.dfr_context_tune_get.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(i64), errp]
    lib.dfr_context_trim.argtypes = [vp]
    lib.dfr_build_id.argtypes = []
    lib.dfr_build_id.restype = ctypes.c_char_p
    lib.dfr_context_calibration.argtypes = [vp, ctypes.POINTER(vp), errp]
    lib.dfr_column_upload.argtypes = [vp, i32, vp, i64, i64, vp, i64, i64, ctypes.POINTER(vp), errp]
    lib.dfr_column_wrap.argtypes = [vp, i32, vp, i64, i64, vp, i64, i64, ctypes.POINTER(vp), errp]
    lib.dfr_column_release.argtypes = [vp]
    lib.dfr_column_length.argtypes = [vp]
    lib.dfr_column_length.restype = i64
    lib.dfr_column_null_count.argtypes = [vp]
    lib.dfr_column_null_count.restype = i64
    lib.dfr_column_distinct.argtypes = [vp]
    lib.dfr_column_distinct.restype = i64
    lib.dfr_groupby.argtypes = [vp, ctypes.POINTER(Query), ctypes.POINTER(vp), errp]
    lib.dfr_groupby_into.argtypes = [vp, ctypes.POINTER(Query), ctypes.POINTER(Output), ctypes.POINTER(i64), errp]
    lib.dfr_column_slice.argtypes = [vp, i64, i64, ctypes.POINTER(vp), errp]
    lib.dfr_column_upload_selection.argtypes = [vp, i32, ctypes.POINTER(Selection), i32, i64, ctypes.POINTER(vp), errp]
    lib.dfr_column_compute.argtypes = [vp, ctypes.POINTER(ValueProgram), ctypes.POINTER(vp), i32, ctypes.POINTER(vp), errp]
    lib.dfr_server_create.argtypes = [vp, i32, i32, i64, ctypes.POINTER(vp), errp]
    lib.dfr_server_submit.argtypes = [vp, ctypes.POINTER(Query), ctypes.POINTER(i64), errp]
    lib.dfr_server_wait.argtypes = [vp, i64, i64, ctypes.POINTER(vp), ctypes.POINTER(i64), errp]
    lib.dfr_server_stats_get.argtypes = [vp, ctypes.POINTER(ServerStats)]
    lib.dfr_server_destroy.argtypes = [vp]
    lib.dfr_groupby_batch.argtypes = [vp, ctypes.POINTER(Query), i32, ctypes.POINTER(vp), ctypes.POINTER(i32),
                                      ctypes.POINTER(vp), errp]
    lib.dfr_result_groups.argtypes = [vp]
    lib.dfr_result_groups.restype = i64
    lib.dfr_result_key_dtype.argtypes = [vp, i32]
    lib.dfr_result_agg_dtype.argtypes = [vp, i32]
    lib.dfr_result_copy_key.argtypes = [vp, i32, vp, vp, errp]
    lib.dfr_result_copy_agg.argtypes = [vp, i32, vp, vp, errp]
    lib.dfr_result_copy_first_rows.argtypes = [vp, vp, errp]
    lib.dfr_result_device_bytes.argtypes = [vp]
    lib.dfr_result_device_bytes.restype = i64
    lib.dfr_result_device_ms.argtypes = [vp]
    lib.dfr_result_device_ms.restype = ctypes.c_double
    lib.dfr_result_phase_ms.argtypes = [vp, i32]
    lib.dfr_result_phase_ms.restype = ctypes.c_double
    lib.dfr_result_release.argtypes = [vp]
    lib.dfr_result_buffer.argtypes = [vp, i32, i32, i32, ctypes.POINTER(vp), ctypes.POINTER(i64), errp]
    lib.dfr_buffer_copy_to_host.argtypes = [vp, vp, i64]
    lib.dfr_buffer_copy_to_host.restype = i64
    lib.dfr_buffer_release.argtypes = [vp]
    _lib = lib
    return lib


class ServerStats(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int64) for f in
                ("submitted", "completed", "batches", "max_batch", "concurrent_max", "queued", "in_flight_bytes")]


class DfrError(RuntimeError):
    """A libdfrocm failure: `status` is the DFR_* code."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


class DataFault(DfrError):
    """An aggregate out of its result type's range (the reference's
    "... expression sum overflow")."""


class Unsupported(DfrError):
    """Work the device path does not take; raised before submission."""


def error_of(status: int, err: ctypes.c_void_p) -> DfrError:
    """The exception for a failed status; frees the library's message."""
    lib = load()
    msg = ctypes.string_at(err.value).decode() if err.value else f"libdfrocm status {status}"
    if err.value:
        lib.dfr_free(err)
    cls = {DFR_DATA_FAULT: DataFault, DFR_UNSUPPORTED: Unsupported}.get(status, DfrError)
    return cls(status, msg)


def check(status: int, err: ctypes.c_void_p) -> None:
    if status == DFR_OK:
        return
    raise error_of(status, err)
