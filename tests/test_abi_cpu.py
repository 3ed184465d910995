"""C ABI checks that need no GPU: the library loads and exports every
symbol include/*.h declares, and context creation without a device fails
with NO_GPU rather than crashing."""
import ctypes as C
import os

import pytest

from image360_amd import _abi


def test_library_exports_every_declared_symbol():
    L = _abi.lib()
    declared = _abi.header_symbols()
    assert declared, "no declarations found in include/*.h"
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing
    assert set(_abi.SIGNATURES) == declared


def _offload_bundles(blob):
    """Entries (triple, code object bytes) of every clang offload bundle in
    `blob`: magic, uint64 count, then per entry uint64 offset, size and
    triple length and the triple (offsets from the bundle's start)."""
    import struct

    magic = b"__CLANG_OFFLOAD_BUNDLE__"
    out, pos = [], blob.find(magic)
    while pos >= 0:
        (n,) = struct.unpack_from("<Q", blob, pos + len(magic))
        q = pos + len(magic) + 8
        for _ in range(n):
            off, size, tlen = struct.unpack_from("<QQQ", blob, q)
            q += 24
            triple = blob[q:q + tlen].decode()
            q += tlen
            out.append((triple, blob[pos + off: pos + off + size]))
        pos = blob.find(magic, pos + 1)
    return out


def test_library_is_gfx950_only():
    # every device code object embedded in the library (one bundle per
    # translation unit) targets gfx950 and nothing else
    with open(_abi.LIB_PATH, "rb") as f:
        blob = f.read()
    entries = _offload_bundles(blob)
    device = [(t, co) for t, co in entries if not t.startswith("host-")]
    assert len(device) >= 5, [t for t, _ in entries]
    for triple, co in device:
        assert triple == "hipv4-amdgcn-amd-amdhsa--gfx950", triple
        # an AMDGPU ELF code object (e_machine 0xE0) whose note names gfx950
        assert co[:4] == b"\x7fELF" and int.from_bytes(co[18:20], "little") == 0xE0, triple
        assert b"amdgcn-amd-amdhsa--gfx950" in co


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


def test_python_constants_match_the_header():
    import re

    from image360_amd import _abi

    text = open(os.path.join(_abi.HEADER_DIR, "i360hip.h")).read()
    vals = {m.group(1): int(m.group(2), 0) for m in re.finditer(r"\b(I360HIP_[A-Z0-9_]+)\s*=\s*(0x[0-9a-fA-F]+|\d+)", text)}
    vals.update({m.group(1): int(m.group(2), 0)
                 for m in re.finditer(r"#define\s+(I360HIP_[A-Z0-9_]+)\s+(0x[0-9a-fA-F]+|\d+)", text)})
    want = {"I360HIP_PIX_F16": _abi.PIX_F16, "I360HIP_PIX_F32": _abi.PIX_F32, "I360HIP_PIX_U16": _abi.PIX_U16,
            "I360HIP_RENDER_FP64": _abi.RENDER_FP64, "I360HIP_RENDER_PLAIN": _abi.RENDER_PLAIN,
            "I360HIP_RENDER_PRECISION": _abi.RENDER_PRECISION, "I360HIP_RENDER_BATCHED": _abi.RENDER_BATCHED,
            "I360HIP_RENDER_SURROGATE": _abi.RENDER_SURROGATE, "I360HIP_RENDER_OUT_RGB48": _abi.RENDER_OUT_RGB48}
    for name, v in want.items():
        assert vals.get(name) == v, (name, vals.get(name), v)
