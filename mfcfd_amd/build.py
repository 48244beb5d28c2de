"""Builds the HIP library libmfcfd_hip.so in-tree (hipcc, gfx950) and the Fortran host
program fortran/meshfree_host (amdflang, linked to the library and to the HDF5 C
library it reads point.h5 with).

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

# The library's translation units and their floating-point flags: mfcfd.hip (host side,
# FP64 kernels, adjoint) without contraction; mfcfd_f32.hip (the mixed path's FP32
# kernels, held to a tolerance against FP64, not to rounding parity) with FMA
# contraction; mfcfd_lsq32.hip (the mixed path's LSQ kernel) also without the IEEE mode
# bit and with NaN operands assumed absent, so that its stencil extremes are single
# max/min instructions.
UNITS = [("mfcfd.hip", ["-ffp-contract=off"]), ("mfcfd_f32.hip", ["-ffp-contract=fast"]),
         ("mfcfd_lsq32.hip", ["-ffp-contract=fast", "-fno-honor-nans", "-mno-amdgpu-ieee"]),
         ("mfcfd_noalloc.hip", ["-ffp-contract=off"])]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the MI355X library cannot be built")


FORTRAN = os.path.join(HERE, "fortran")
HOST = os.path.join(FORTRAN, "meshfree_host")
HOST_SOURCES = ["mfcfd_api.f90", "host_io.f90", "meshfree_host.f90"]


def amdflang() -> str | None:
    for cand in (os.environ.get("FC_AMD"), "/opt/rocm/bin/amdflang", "/opt/rocm/lib/llvm/bin/amdflang",
                 shutil.which("amdflang")):
        if cand and os.path.exists(cand):
            return cand
    return None


def hdf5_libdir() -> str | None:
    for d in (os.environ.get("MFCFD_HDF5_LIBDIR"), "/opt/conda/lib"):
        if d and os.path.exists(os.path.join(d, "libhdf5.so")):
            return d
    return None


def build_host(verbose: bool = False) -> str | None:
    """Compile the Fortran host program (None when amdflang or libhdf5 is missing)."""
    fc, h5 = amdflang(), hdf5_libdir()
    if fc is None or h5 is None:
        print("meshfree_host not built: amdflang or libhdf5 missing", file=sys.stderr)
        return None
    import tempfile
    with tempfile.TemporaryDirectory() as mod:
        cmd = [fc, "-O2", "-J", mod] + [os.path.join(FORTRAN, f) for f in HOST_SOURCES] + [
            "-o", HOST + ".tmp", f"-L{HERE}", "-lmfcfd_hip", "-Wl,-rpath,$ORIGIN/..",
            f"-L{h5}", "-lhdf5", f"-Wl,-rpath,{h5}"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
    os.replace(HOST + ".tmp", HOST)
    return HOST


def build(verbose: bool = False) -> str:
    """Compile the units of UNITS and link them into mfcfd_amd/libmfcfd_hip.so;
    returns the path."""
    import tempfile
    common = [f for f in HIPCC_FLAGS if f not in ("-shared", "-ffp-contract=off")]
    with tempfile.TemporaryDirectory() as tmpd:
        objs = []
        for src, flags in UNITS:
            obj = os.path.join(tmpd, src.replace(".hip", ".o"))
            cmd = [hipcc()] + common + flags + ["-c", "-o", obj, os.path.join(CSRC, src)]
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.check_call(cmd)
            objs.append(obj)
        tmp = LIB + ".tmp"
        cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
