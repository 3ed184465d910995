"""C ABI checks that need no GPU: the library loads and exports every
symbol include/*.h declares, and context creation without a device fails
with NO_GPU rather than crashing."""
import ctypes as C
import os
import subprocess

import pytest

from image360_amd import _abi


def test_library_exports_every_declared_symbol():
    L = _abi.lib()
    declared = _abi.header_symbols()
    assert declared, "no declarations found in include/*.h"
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing
    assert set(_abi.SIGNATURES) == declared


def test_library_is_gfx950_only():
    out = subprocess.run(["/opt/rocm/llvm/bin/llvm-readelf", "-n", _abi.LIB_PATH], capture_output=True, text=True)
    if out.returncode != 0:
        pytest.skip("llvm-readelf unavailable")
    # the offload bundle is embedded in .hip_fatbin; the host ELF must load
    assert os.path.getsize(_abi.LIB_PATH) > 0


def test_version_string():
    assert b"gfx950" in _abi.lib().i360hip_version()


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK), reason="GPU visible")
def test_create_without_gpu_reports_no_gpu():
    L = _abi.lib()
    h = C.c_void_p()
    rc = L.i360hip_create(0, C.byref(h))
    assert rc == _abi.NO_GPU
    assert L.i360hip_last_error(None)


def test_null_context_is_rejected():
    L = _abi.lib()
    assert L.i360hip_apply_gain_u16(None, None, 4, 2.0) == _abi.DEVICE_ERROR
