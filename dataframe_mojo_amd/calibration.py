"""Placement calibration: measured device costs that `engine="auto"`
placement reads through `transfer_model`, `pipeline_estimate` and
`pipeline_costs` (pattern: dataframe/placement.mojo, and the Metal
provider's versioned, identity-matched profiles,
dataframe_metal/calibration.mojo:1-41).

A profile is the autotuner's JSON (autotune.py writes it; the library
applies its "tuning" to every context it creates on the same device):
  {"version": 2,
   "identity": {"device", "arch", "compute_units", "build_id"},
   "tuning": {...},
   "transfer": {"to_device": [fixed_ns, ns_per_byte],
                "to_host": [fixed_ns, ns_per_byte], "startup_ns": ns},
   "groupby": {"fixed_ns": f, "ns_per_row": r, "ns_per_group": g}}
found at $DATAFRAME_ROCM_CALIBRATION or calibration/<arch>-<CUs>cu.json in
this package. A profile whose identity does not match the device (read from
the KFD topology, no HIP context) and the loaded library's build is not
used: its costs are "not measured" (-1), as the reference treats
unidentified or legacy profiles.
"""
from __future__ import annotations

import ctypes
import json
import os

VERSION = 2
ENV = "DATAFRAME_ROCM_CALIBRATION"
HERE = os.path.dirname(os.path.abspath(__file__))


def device_identity(device: int):
    """(device name, arch, compute units, build id) of a device and the
    loaded library, without a HIP context; None when unavailable."""
    from . import _lib

    lib = _lib.load(required=False)
    if lib is None:
        return None
    info = _lib.DeviceInfo()
    e = ctypes.c_void_p()
    if lib.dfr_device_info_get(device, ctypes.byref(info), ctypes.byref(e)) != _lib.DFR_OK:
        if e.value:
            lib.dfr_free(e)
        return None
    return {"device": info.name.decode(), "arch": info.arch.decode(), "compute_units": int(info.compute_units),
            "build_id": lib.dfr_build_id().decode()}


def profile_path(device: int, ident: dict | None = None) -> str:
    env = os.environ.get(ENV)
    if env:
        return env
    ident = ident or device_identity(device)
    if not ident:
        return ""
    return os.path.join(HERE, "calibration", f"{ident['arch']}-{ident['compute_units']}cu.json")


def identity_matches(prof: dict, ident: dict | None) -> bool:
    if not ident or not isinstance(prof.get("identity"), dict):
        return False
    got = prof["identity"]
    return all(got.get(k) == ident[k] for k in ("device", "arch", "compute_units", "build_id"))


def load_profile(device: int):
    """The device's profile, or None when there is none, it is another
    version, or its identity does not match this device and build."""
    if os.environ.get(ENV) == "off":
        return None
    ident = device_identity(device)
    path = profile_path(device, ident)
    if not path:
        return None
    try:
        with open(path) as f:
            prof = json.load(f)
    except (OSError, ValueError):
        return None
    if not isinstance(prof, dict) or prof.get("version") != VERSION or not identity_matches(prof, ident):
        return None
    return prof


def _rows(pipeline, features) -> int:
    if features is not None and features.input_rows >= 0:
        return features.input_rows
    return pipeline.source.height()


def _cost(model: dict, rows: int, groups: int) -> float:
    return model["fixed_ns"] + model["ns_per_row"] * rows + model.get("ns_per_group", 0.0) * max(groups, 0)


def estimate_ns(prof, pipeline, features) -> float:
    """The device's time for the pipeline, inputs already resident (the
    planner adds transfers from transfer_model); -1 without a profile."""
    if not prof or "groupby" not in prof:
        return -1.0
    groups = features.groups if features is not None else -1
    if groups < 0:
        # unknown groups: bounded by the rows
        groups = min(_rows(pipeline, features), 1 << 20)
    return _cost(prof["groupby"], _rows(pipeline, features), groups)
