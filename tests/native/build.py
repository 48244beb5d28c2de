"""Builds tests/native/libadjcheck.so (host-only wrappers of the device header's
forward and reverse functions, for CPU finite-difference checks).  Test
infrastructure: the shipped library does not use it."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
LIB = os.path.join(HERE, "libadjcheck.so")


def build() -> str:
    from mfcfd_amd.build import HIPCC_FLAGS, hipcc
    src = os.path.join(HERE, "adjcheck.hip")
    subprocess.check_call([hipcc()] + HIPCC_FLAGS + ["-o", LIB + ".tmp", src])
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build())
