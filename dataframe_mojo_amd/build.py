"""Builds the package's native libraries in-tree for gfx950.

libdfrocm.so  the C-ABI provider library (csrc/dfrocm.hip, csrc/groupby.hip,
              csrc/server.hip, csrc/jit.hip; links hiprtc, which compiles pass A
              for a query's shape at run time from the headers in csrc/)
libdfrgen.so  seeded synthetic column generators (csrc/datagen.hip), used by
              the benchmark and the tests to create device-resident inputs
dfr_jitc      the runtime compiler libdfrocm runs as a child process for its
              shape-specialised kernels (csrc/jitc.cpp, host code over hiprtc)

Run as ``python -m dataframe_mojo_amd.build`` or through
``__graft_entry__.build()``.
"""
from __future__ import annotations

import concurrent.futures
import hashlib
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ARCH = os.environ.get("DFR_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wno-unused-value",
    "-Wno-unused-result",
    "-Wno-pass-failed",
    # packed FP32 instructions stay out of every kernel
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]

LIBS = {
    "libdfrocm.so": {
        "sources": ["dfrocm.hip", "groupby.hip", "server.hip", "jit.hip", "calib.hip", "selection.hip"],
        "extra": [f"-Wl,--version-script={os.path.join(CSRC, 'libdfrocm.map')}", "-lhiprtc", "-ldl"],
    },
    "libdfrgen.so": {"sources": ["datagen.hip"], "extra": []},
}


def _run(cmd: list[str], verbose: bool, what: str) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stderr)
        raise RuntimeError(f"{what} failed (exit {proc.returncode})")


def build_id() -> str:
    """The identity of this source tree (kernels, runtime, headers): a
    calibration profile measured with other code is not applied."""
    h = hashlib.sha256()
    root = os.path.dirname(HERE)
    paths = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC))
    paths.append(os.path.join(root, "include", "dfrocm.h"))
    for path in paths:
        if os.path.isfile(path):
            h.update(os.path.basename(path).encode())
            with open(path, "rb") as f:
                h.update(f.read())
    return h.hexdigest()[:16]


def build(verbose: bool = False) -> list[str]:
    """Compiles every library; returns their paths. Raises on failure.

    Each source compiles to its own object (in parallel; every kernel is
    launched from the source that defines it, so the objects need no
    device-code linking), then the objects link into the library."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    jobs = []
    bid = build_id()
    for name, spec in LIBS.items():
        for src in spec["sources"]:
            path = os.path.join(CSRC, src)
            obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
            flags = [f for f in COMMON if f != "-shared"]
            if src == "calib.hip":
                flags.append(f'-DDFR_BUILD_ID="{bid}"')
            jobs.append(([HIPCC, *flags, "-c", "-o", obj, path], f"compiling {src}"))
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        for f in [pool.submit(_run, cmd, verbose, what) for cmd, what in jobs]:
            f.result()
    # the runtime compiler program (host code only)
    jitc = os.path.join(HERE, "dfr_jitc")
    rocm = os.path.dirname(os.path.dirname(os.path.realpath(HIPCC)))
    _run([os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", f"-I{rocm}/include",
          "-o", jitc + ".tmp", os.path.join(CSRC, "jitc.cpp"), f"-L{rocm}/lib", "-lhiprtc", f"-Wl,-rpath,{rocm}/lib"],
         verbose, "building dfr_jitc")
    os.replace(jitc + ".tmp", jitc)
    out = [jitc]
    for name, spec in LIBS.items():
        target = os.path.join(HERE, name)
        objs = [os.path.join(objdir, os.path.splitext(s)[0] + ".o") for s in spec["sources"]]
        tmp = target + ".tmp"
        _run([HIPCC, *COMMON, *spec["extra"], "-o", tmp, *objs], verbose, f"linking {name}")
        os.replace(tmp, target)
        out.append(target)
    return out


if __name__ == "__main__":
    for p in build(verbose=True):
        print(p)
