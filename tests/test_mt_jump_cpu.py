"""The jump-ahead polynomials of pagmo2_amd/csrc/mt_jump_host.hpp against
std::mt19937::discard, on the CPU (tests/cpp/test_mt_jump.cpp)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_jump_polynomials_match_discard():
    out = os.path.join(HERE, "cpp", "bin", "test_mt_jump")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "pagmo2_amd", "csrc"),
                           os.path.join(HERE, "cpp", "test_mt_jump.cpp"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    r = subprocess.run([out], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
