"""Builds tests/cpp/bin/test_hip_bfe (g++, host code only) against
include/pagmo_hip/hip_bfe.hpp, the pagmo stand-in of tests/cpp/pagmo_compat and
libpagmo_hip.so.  Run after `python -m pagmo2_amd.build`."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(HERE, "bin", "test_hip_bfe")


def build() -> str:
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-Wall",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "include", "pagmo_hip"),
           "-I", os.path.join(HERE, "pagmo_compat"), "-I", os.path.join(ROOT, "oracle"),
           os.path.join(HERE, "test_hip_bfe.cpp"), "-L", os.path.join(ROOT, "pagmo2_amd"), "-lpagmo_hip",
           "-Wl,-rpath,$ORIGIN/../../../pagmo2_amd", "-pthread", "-o", BIN + ".tmp"]
    subprocess.check_call(cmd)
    os.replace(BIN + ".tmp", BIN)
    return BIN


if __name__ == "__main__":
    print(build())
