"""Placement calibration: measured device costs that `engine="auto"`
placement reads through `transfer_model`, `pipeline_estimate` and
`pipeline_costs` (pattern: dataframe/placement.mojo, the Metal provider's
versioned, identity-matched profiles in dataframe_metal/calibration.mojo).

A profile is JSON:
  {"version": 1, "device": "<name (arch)>",
   "transfer": {"to_device": [fixed_ns, ns_per_byte],
                "to_host": [fixed_ns, ns_per_byte], "startup_ns": ns},
   "groupby": {"fixed_ns": f, "ns_per_row": r, "ns_per_group": g},
   "cpu": {"fixed_ns": f, "ns_per_row": r, "ns_per_group": g}}
found at $DATAFRAME_ROCM_CALIBRATION or calibration_<ordinal>.json next to
this module. Missing or mismatched profiles mean "not measured" (-1).
"""
from __future__ import annotations

import json
import os

VERSION = 1
ENV = "DATAFRAME_ROCM_CALIBRATION"


def profile_path(device: int) -> str:
    return os.environ.get(ENV) or os.path.join(os.path.dirname(os.path.abspath(__file__)), f"calibration_{device}.json")


def load_profile(device: int):
    path = profile_path(device)
    try:
        with open(path) as f:
            prof = json.load(f)
    except (OSError, ValueError):
        return None
    if prof.get("version") != VERSION:
        return None
    return prof


def save_profile(device: int, prof: dict, path: str | None = None) -> str:
    prof = dict(prof, version=VERSION)
    path = path or profile_path(device)
    with open(path, "w") as f:
        json.dump(prof, f, indent=1, sort_keys=True)
    return path


def _rows(pipeline, features) -> int:
    if features is not None and features.input_rows >= 0:
        return features.input_rows
    return pipeline.source.height()


def _cost(model: dict, rows: int, groups: int) -> float:
    return model["fixed_ns"] + model["ns_per_row"] * rows + model.get("ns_per_group", 0.0) * max(groups, 0)


def estimate_ns(prof, pipeline, features) -> float:
    if not prof or "groupby" not in prof:
        return -1.0
    groups = features.groups if features is not None else -1
    return _cost(prof["groupby"], _rows(pipeline, features), groups)


def cpu_estimate_ns(prof, pipeline, features) -> float:
    if not prof or "cpu" not in prof:
        return -1.0
    groups = features.groups if features is not None else -1
    return _cost(prof["cpu"], _rows(pipeline, features), groups)
