"""The ROCm accelerator provider: the thirteen functions of the reference's
provider boundary (providers/cpu/dataframe_provider/__init__.mojo:15-70;
contract: docs/design/accelerator-contract.md:20-47), backed by libdfrocm.

  installed, execute, describe, execute_profiled      whole plans
  handoff_supports, handoff_execute, handoff_estimate  hand-off (a discrete
                                                       device declines it)
  devices, transfer_model, pipeline_supports,          device placement
  pipeline_estimate, pipeline_costs, run_pipeline

`installed`, `describe`, `devices`, `transfer_model`, `pipeline_supports`,
`pipeline_estimate` and `pipeline_costs` never create a HIP context: they
read sysfs and the plan only.
"""
from __future__ import annotations

import ctypes
import os
import time
from dataclasses import dataclass

from . import _lib
from . import calibration
from .frame import Column, DataFrame, LazyFrame, Series
from .plan import Unsupported, compile_plan

BACKEND = "amd"
MEMORY_DISCRETE = 0
MEMORY_UNIFIED = 1


@dataclass
class DeviceInfo:
    """devices.mojo:24-40."""

    backend: str
    ordinal: int
    name: str
    memory: int
    budget_bytes: int


@dataclass
class ByteCost:
    """`fixed_ns + ns_per_byte * bytes`; negative fields mean not measured."""

    fixed_ns: float
    ns_per_byte: float

    def measured(self) -> bool:
        return self.fixed_ns >= 0 and self.ns_per_byte >= 0

    def cost(self, nbytes: int) -> float:
        return self.fixed_ns + self.ns_per_byte * nbytes


@dataclass
class TransferModel:
    to_device: ByteCost
    to_host: ByteCost
    startup_ns: float

    @staticmethod
    def unmeasured() -> "TransferModel":
        return TransferModel(ByteCost(-1, -1), ByteCost(-1, -1), -1)

    def measured(self) -> bool:
        return self.to_device.measured() and self.to_host.measured()


@dataclass
class PipelineFeatures:
    """devices.mojo PipelineFeatures: -1 means unknown."""

    input_rows: int = -1
    input_bytes: int = -1
    output_rows: int = -1
    groups: int = -1
    build_rows: int = -1
    collect_request: bool = False

    @staticmethod
    def rows(n: int) -> "PipelineFeatures":
        return PipelineFeatures(input_rows=n)


# ------------------------------------------------------------ whole plans


def installed() -> bool:
    """The library is present and a KFD device node exists; no HIP call."""
    return os.path.exists(_lib.library_path()) and os.path.exists("/dev/kfd")


def execute(plan: LazyFrame) -> tuple[DataFrame, DataFrame]:
    """Runs the whole plan on the device; raises `Unsupported` before any
    device work when the plan is not filter -> group_by -> agg over
    supported types."""
    from . import runtime

    dplan = compile_plan(plan)
    ctx = runtime.context()
    t0 = time.perf_counter()
    out, res = runtime.run_plan(ctx, plan.source, dplan)
    wall = (time.perf_counter() - t0) * 1e3
    report = _report(ctx, res, wall)
    res.release()
    return out, report


def execute_profiled(plan: LazyFrame) -> tuple[DataFrame, DataFrame]:
    return execute(plan)


def describe(plan: LazyFrame) -> str:
    try:
        dplan = compile_plan(plan)
    except Exception as exc:  # errors become text
        return f"ENGINE accel: not on the ROCm device path: {exc}\n"
    prog = dplan.program
    lines = [
        "ENGINE accel: ROCm (libdfrocm, gfx950)",
        f"  GROUP BY {', '.join(dplan.keys)}{' (maintain_order)' if dplan.maintain_order else ''}",
        f"  FILTER {len(prog.leaves)} leaves, truth table {[hex(t) for t in prog.truth[: max(1, (1 << len(prog.leaves)) // 32)]]}",
        f"  AGG {', '.join(dplan.agg_names)}",
    ]
    prof = calibration.load_profile(0)
    if prof is not None:
        rows = plan.source.height()
        est = calibration.estimate_ns(prof, plan, PipelineFeatures.rows(rows))
        lines.append(f"  PLACE device 0: {est / 1e6:.3f} ms for {rows} resident rows (calibrated "
                     f"{prof['identity']['device']}, build {prof['identity']['build_id']})")
    return "\n".join(lines) + "\n"


def _report(ctx, res, wall_ms: float) -> DataFrame:
    obs = ctx.observe()
    names = ["device_ms", "wall_ms", "groups", "launches", "completed", "upload_bytes", "download_bytes", "queries"]
    vals = [res.device_ms, wall_ms, float(res.groups)] + [float(obs[k]) for k in names[3:]]
    return DataFrame([Series("metric", Column.string(names)), Series("value", Column.float64(vals))])


# ------------------------------------------------------------ hand-off


def handoff_supports(region: LazyFrame) -> bool:
    return False  # a discrete device runs whole pipelines, not CPU regions


def handoff_execute(region: LazyFrame) -> DataFrame:
    raise RuntimeError("the ROCm provider does not accept hand-off regions (discrete device)")


def handoff_estimate(region: LazyFrame, rows: int) -> tuple[str, float]:
    return ("", -1.0)


# ------------------------------------------------------------ devices


def devices() -> list[DeviceInfo]:
    lib = _lib.load(required=False)
    if lib is None or not os.path.exists("/dev/kfd"):
        return []
    n = ctypes.c_int32(0)
    e = ctypes.c_void_p()
    if lib.dfr_device_count(ctypes.byref(n), ctypes.byref(e)) != _lib.DFR_OK:
        if e.value:
            lib.dfr_free(e)
        return []
    from .runtime import memory_limit

    out = []
    for i in range(n.value):
        info = _lib.DeviceInfo()
        if lib.dfr_device_info_get(i, ctypes.byref(info), ctypes.byref(e)) != _lib.DFR_OK:
            if e.value:
                lib.dfr_free(e)
            continue
        budget = memory_limit() or int(info.memory_bytes)
        out.append(DeviceInfo(BACKEND, i, f"{info.name.decode()} ({info.arch.decode()})", MEMORY_DISCRETE, budget))
    return out


def transfer_model(device: int) -> TransferModel:
    prof = calibration.load_profile(device)
    if prof is None or "transfer" not in prof:
        return TransferModel.unmeasured()
    t = prof["transfer"]
    return TransferModel(ByteCost(*t["to_device"]), ByteCost(*t["to_host"]), float(t.get("startup_ns", -1)))


def pipeline_supports(pipeline: LazyFrame, device: int) -> bool:
    try:
        compile_plan(pipeline)
        return device >= 0
    except Exception:
        return False


def pipeline_estimate(pipeline: LazyFrame, device: int, features: PipelineFeatures) -> float:
    if not pipeline_supports(pipeline, device):
        return -1.0
    prof = calibration.load_profile(device)
    return calibration.estimate_ns(prof, pipeline, features) if prof else -1.0


def pipeline_costs(pipeline: LazyFrame, device: int, features: PipelineFeatures) -> tuple[float, float]:
    """(CPU, device) nanoseconds. The device side is the calibrated model;
    the CPU side is -1: this provider measures no CPU executor, so the
    planner keeps its generic CPU calibration (accelerator-contract.md:47,
    "A provider without matched CPU measurements can return
    (-1, pipeline_estimate(...))")."""
    return (-1.0, pipeline_estimate(pipeline, device, features))


def run_pipeline(pipeline: LazyFrame, device: int) -> DataFrame:
    from . import runtime

    dplan = compile_plan(pipeline)
    out, res = runtime.run_plan(runtime.context(device), pipeline.source, dplan)
    res.release()
    return out


__all__ = [
    "installed",
    "execute",
    "describe",
    "execute_profiled",
    "handoff_supports",
    "handoff_execute",
    "handoff_estimate",
    "devices",
    "transfer_model",
    "pipeline_supports",
    "pipeline_estimate",
    "pipeline_costs",
    "run_pipeline",
    "Unsupported",
    "DeviceInfo",
    "TransferModel",
    "ByteCost",
    "PipelineFeatures",
]
