"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer, on the CPU.

* The oracle (oracle/liboracle_san.so, `make -C oracle san`, g++) runs the
  oracle's and the migration restatement's test suites.
* The library's host code (pagmo2_amd/libpagmo_hip_san.so,
  pagmo2_amd.build.build_sanitized: hipcc with the sanitizers on the host
  side only) runs the C ABI's validation suite: argument checks, messages,
  the per-thread context and the no-device paths.
Each suite runs in a child process with the matching sanitizer runtime
preloaded (and libstdc++, so that ASan intercepts C++ throws); any report
fails the child, and so the test.
"""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SAN = os.path.join(ROOT, "oracle", "liboracle_san.so")


def _gcc_lib(name):
    return subprocess.check_output(["gcc", f"-print-file-name={name}"], text=True).strip()


def _clang_asan():
    hits = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return hits[-1] if hits else None


def _run(env_extra, tests):
    env = dict(os.environ)
    env.update(env_extra)
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0:exitcode=23"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1:exitcode=23"
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu", *tests],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, (r.stdout[-4000:], r.stderr[-4000:])
    assert " passed" in r.stdout


def test_oracle_sanitized():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"])
    asan, stdcxx = _gcc_lib("libasan.so"), _gcc_lib("libstdc++.so")
    if not os.path.isabs(asan) or not os.path.exists(asan):
        pytest.skip("gcc's ASan runtime is not installed")
    _run({"LD_PRELOAD": f"{asan} {stdcxx}", "PGH_ORACLE_LIB": ORACLE_SAN},
         ["tests/test_oracle.py", "tests/test_migration_cpu.py"])


def test_capi_sanitized():
    sys.path.insert(0, ROOT)
    import pagmo2_amd.build as B
    asan = _clang_asan()
    if asan is None:
        pytest.skip("clang's ASan runtime is not installed")
    if not os.path.exists(B.SAN_LIB):
        B.build_sanitized()
    _run({"LD_PRELOAD": f"{asan} {_gcc_lib('libstdc++.so')}", "PGH_LIB": B.SAN_LIB}, ["tests/test_capi_cpu.py"])
