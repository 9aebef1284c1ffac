"""Builds the package's native libraries in-tree for gfx950.

libdfrocm.so  the C-ABI provider library (csrc/dfrocm.hip, csrc/groupby.hip,
              csrc/server.hip, csrc/jit.hip; links hiprtc, which compiles pass A
              for a query's shape at run time from the headers in csrc/)
libdfrgen.so  seeded synthetic column generators (csrc/datagen.hip), used by
              the benchmark and the tests to create device-resident inputs

Run as ``python -m dataframe_mojo_amd.build`` or through
``__graft_entry__.build()``.
"""
from __future__ import annotations

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
        "sources": ["dfrocm.hip", "groupby.hip", "server.hip", "jit.hip"],
        "extra": [f"-Wl,--version-script={os.path.join(CSRC, 'libdfrocm.map')}", "-lhiprtc", "-ldl"],
    },
    "libdfrgen.so": {"sources": ["datagen.hip"], "extra": []},
}


def build(verbose: bool = False) -> list[str]:
    """Compiles every library; returns their paths. Raises on failure."""
    out = []
    for name, spec in LIBS.items():
        target = os.path.join(HERE, name)
        sources = [os.path.join(CSRC, s) for s in spec["sources"]]
        sources = [s for s in sources if os.path.exists(s)]
        tmp = target + ".tmp"
        cmd = [HIPCC, *COMMON, *spec["extra"], "-o", tmp, *sources]
        if verbose:
            print(" ".join(cmd), flush=True)
        proc = subprocess.run(cmd, capture_output=True, text=True)
        if proc.returncode != 0:
            sys.stderr.write(proc.stderr)
            raise RuntimeError(f"building {name} failed (exit {proc.returncode})")
        os.replace(tmp, target)
        out.append(target)
    return out


if __name__ == "__main__":
    for p in build(verbose=True):
        print(p)
