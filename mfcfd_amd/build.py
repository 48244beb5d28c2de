"""Builds the HIP library libmfcfd_hip.so in-tree (hipcc, gfx950).

Run ``python -m mfcfd_amd.build`` (or ``__graft_entry__.build()``).  hipcc
cross-compiles for gfx950 without a GPU, so this runs anywhere the ROCm toolchain is.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmfcfd_hip.so")
ARCH = os.environ.get("MFCFD_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: the FP64 path rounds in the reference's operation order (no FMA
# contraction), so its results can be compared with the FP64 oracle to ~1e-13.
# Packed FP32 instructions are disabled for the whole device build.
HIPCC_FLAGS = [
    "-O3", "-std=c++20", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the MI355X library cannot be built")


def build(verbose: bool = False) -> str:
    """Compile csrc/mfcfd.hip into mfcfd_amd/libmfcfd_hip.so; returns the path."""
    src = os.path.join(CSRC, "mfcfd.hip")
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
