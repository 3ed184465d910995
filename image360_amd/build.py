"""Builds the gfx950 HIP library ``image360_amd/lib/libi360hip.so`` in tree.

    python -m image360_amd.build

Every ``csrc/*.hip`` file is compiled with hipcc for gfx950 only, then
linked into one shared library exporting the C ABI of
``include/i360hip.h``.  The C++ host side (``host/*.cpp``: the
IComputeBackend implementation and its factory glue) is plain C++17 over that
ABI, built with g++ into ``lib/libi360hip_backend.so``, and
``tests/cpp/hip_backend_test.cpp`` into ``lib/hip_backend_test``.  All device code is compiled with
``-ffp-contract=off`` (the bit-exact kernels depend on it; the fast kernels
request fused multiply-adds explicitly with fmaf) and without packed-FP32
instruction selection.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB_DIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIB_DIR, "libi360hip.so")
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    build_host()
    if verbose:
        print("built", LIB)
    return LIB


HOST = os.path.join(PKG, "host")
HOST_LIB = os.path.join(LIB_DIR, "libi360hip_backend.so")
HOST_TEST = os.path.join(LIB_DIR, "hip_backend_test")
CXX = os.environ.get("CXX", shutil.which("g++") or "g++")
HOST_FLAGS = ["-std=c++17", "-O2", "-Wall", "-Wextra", "-Wno-unused-parameter",
              f"-I{HOST}", f"-I{os.path.join(HOST, 'contract')}", f"-I{os.path.join(ROOT, 'include')}"]


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed:\n{r.stderr[-6000:]}")


def build_host() -> None:
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))
    _run([CXX, *HOST_FLAGS, "-fPIC", "-shared", *srcs, f"-L{LIB_DIR}", "-li360hip", "-Wl,-rpath,$ORIGIN",
          "-o", HOST_LIB + ".tmp"], "host library build")
    os.replace(HOST_LIB + ".tmp", HOST_LIB)
    test_src = os.path.join(ROOT, "tests", "cpp", "hip_backend_test.cpp")
    _run([CXX, *HOST_FLAGS, test_src, f"-L{LIB_DIR}", "-li360hip_backend", "-li360hip", "-lpthread",
          "-Wl,-rpath,$ORIGIN", "-o", HOST_TEST + ".tmp"], "host test build")
    os.replace(HOST_TEST + ".tmp", HOST_TEST)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
