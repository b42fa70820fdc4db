"""Build recipe of libpagmo_hip.so (gfx950) -- hipcc directly, in-tree.

    python -m pagmo2_amd.build        # or pagmo2_amd.build.build()

Each csrc/*.hip is compiled to an object with hipcc and the objects are linked
into pagmo2_amd/libpagmo_hip.so.  Device code is built with
-ffp-contract=off (parity-critical arithmetic must not be fused) and without
packed FP32 instructions.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpagmo_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-command-line-argument",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *HIPFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    # hipcc's host pass warns that it does not know the packed-fp32 feature: harmless
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


# Host code with AddressSanitizer and UndefinedBehaviorSanitizer (device code
# as in the product build): libpagmo_hip_san.so, for the CPU-only validation
# runs of tests/test_sanitize_cpu.py, which preload clang's ASan runtime.
SAN_OBJ = os.path.join(HERE, "_build_san")
SAN_LIB = os.path.join(HERE, "libpagmo_hip_san.so")
SAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]
CLANGXX = os.environ.get("PGH_CLANGXX", "/opt/rocm/llvm/bin/clang++")


def _compile_san(src: str) -> str:
    obj = os.path.join(SAN_OBJ, os.path.basename(src) + ".o")
    flags = [f if f != "-O3" else "-O1" for f in HIPFLAGS]
    cmd = [HIPCC, *flags, "-g", *SAN_HOST, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc (sanitized host) failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build_sanitized(verbose: bool = False) -> str:
    os.makedirs(SAN_OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile_san, srcs))
    tmp = SAN_LIB + ".tmp"
    # the host link: clang++ with the sanitizers' shared runtime (preloaded
    # by the process that loads the library)
    cmd = [CLANGXX, "-shared", "-fPIC", "-fsanitize=address,undefined", "-shared-libsan", "-o", tmp, *objs,
           "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"sanitized link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, SAN_LIB)
    if verbose:
        print(f"built {SAN_LIB}")
    return SAN_LIB


def clean() -> None:
    shutil.rmtree(OBJ, ignore_errors=True)
    shutil.rmtree(SAN_OBJ, ignore_errors=True)
    for lib in (LIB, SAN_LIB):
        if os.path.exists(lib):
            os.remove(lib)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
