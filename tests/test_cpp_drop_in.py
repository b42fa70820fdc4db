"""hip_bfe (include/pagmo_hip/hip_bfe.hpp) used from C++ exactly as pagmo's
UDBFE plugin contract prescribes; see tests/cpp/test_hip_bfe.cpp."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "cpp"))
import build as cpp_build  # noqa: E402


def _binary():
    if not os.path.exists(cpp_build.BIN):
        cpp_build.build()
    return cpp_build.BIN


def _run(mode):
    r = subprocess.run([_binary(), mode], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr


def test_contract_cpu():
    _run("cpu")


@pytest.mark.gpu
def test_drop_in_gpu():
    _run("gpu")
