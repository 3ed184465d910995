"""The C++ HipComputeBackend (image360_amd/host), the IComputeBackend
implementation a reference build links, driven by tests/cpp/
hip_backend_test.cpp.  Without a GPU: its contract (uninitialised errors,
NO_GPU, burst ops NOT_SUPPORTED, factory fallback).  On the GPU: every
operation on seeded inputs through the C++ interface, outputs compared with
the oracle bit for bit, plus delegation, validation messages and concurrent
calls on one instance."""
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.gpu_util import note

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "image360_amd", "lib",
                   "hip_backend_test")


def _run(*args, timeout=120):
    assert os.path.exists(EXE), "hip_backend_test is not built (python -m image360_amd.build)"
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=timeout)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return r.stdout


def test_contract_without_gpu():
    out = _run("cpu")
    assert "ok Auto falls back to the supplied CPU backend" in out
    assert "ok Initialize initialises the host delegate" in out


ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("sanitizer", ["thread", "address,undefined"])
def test_host_layer_concurrency_under_sanitizers(tmp_path, sanitizer):
    """HipComputeBackend and the factory, built with a sanitizer and linked
    against the CPU stand-in of the C ABI (tests/cpp/abi_stub.cpp): eight
    threads on one instance mixing succeeding calls, failing C ABI calls with
    thread-specific messages, validation failures and delegated operations.
    The sanitizer's report (a data race, a memory error, undefined
    behaviour) fails the run."""
    exe = str(tmp_path / "host_concurrency_test")
    host = os.path.join(ROOT, "image360_amd", "host")
    cmd = ["g++", "-std=c++17", "-O1", "-g", f"-fsanitize={sanitizer}", "-fno-omit-frame-pointer",
           "-fno-sanitize-recover=all", "-w", f"-I{host}", f"-I{os.path.join(host, 'contract')}",
           f"-I{os.path.join(ROOT, 'include')}", os.path.join(host, "HipComputeBackend.cpp"),
           os.path.join(host, "HipBackendFactory.cpp"), os.path.join(ROOT, "tests", "cpp", "abi_stub.cpp"),
           os.path.join(ROOT, "tests", "cpp", "host_concurrency_test.cpp"), "-lpthread", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66", ASAN_OPTIONS="detect_leaks=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "ok concurrent results" in r.stdout and "ok error texts whole" in r.stdout


def _load(d):
    arrs = {}
    with open(os.path.join(d, "manifest.txt")) as f:
        for line in f:
            name, dtype, *shape = line.split()
            dt = {"u16": np.uint16, "f32": np.float32, "f64": np.float64}[dtype]
            arrs[name] = np.fromfile(os.path.join(d, name + ".bin"), dtype=dt).reshape([int(s) for s in shape])
    return arrs


@pytest.mark.gpu
def test_operations_through_cpp_interface(tmp_path):
    note("cpp backend gpu")
    out = _run("gpu", str(tmp_path))
    assert "ok concurrent calls on one instance match sequential results" in out
    a = _load(str(tmp_path))
    want = po.warp_perspective(a["warp_src"], a["warp_out"].shape[1], a["warp_out"].shape[0], a["warp_h"], 5, -3)
    assert np.array_equal(a["warp_out"], want)
    assert np.array_equal(a["median_out"], po.median_stack(list(a["median_in"]), 2.0))
    assert np.array_equal(a["gain_out"], po.apply_gain(a["gain_in"], 1.37))
    cam_mul = np.array([2.0, 1.0, 1.5, 1.0], np.float32)
    rgb_cam = np.array([[1.6, -0.4, -0.2, 0.0], [-0.2, 1.5, -0.3, 0.0], [0.05, -0.5, 1.45, 0.0]], np.float32)
    assert np.array_equal(a["demosaic_out"], po.demosaic(a["demosaic_in"], 512, cam_mul, rgb_cam, 0x94949494))
    assert np.array_equal(a["blend_out"], po.blend_chunk(list(a["blend_in"]), 5))
    # colour transfer: within the tolerance the C library's powf / cbrtf leave open
    rm, rs = po.lab_stats(a["lab_ref"])
    tm, ts = po.lab_stats(a["lab_tgt"])
    assert np.abs(a["lab_stats"] - np.array([rm, rs, tm, ts])).max() <= 1e-4
    d = np.abs(a["reinhard_out"].astype(np.int32) - po.reinhard(a["lab_tgt"], tm, ts, rm, rs).astype(np.int32))
    assert d.max() <= 2 and (d == 0).mean() >= 0.99
