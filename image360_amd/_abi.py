"""ctypes binding of the C ABI declared in ``include/i360hip.h``.

The library is loaded from ``image360_amd/lib/libi360hip.so`` (built in tree
by ``python -m image360_amd.build``).  There is no fallback: a missing
library raises immediately, and so does a missing gfx950 device when a
context is created.
"""
from __future__ import annotations

import ctypes as C
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "lib", "libi360hip.so")
HEADER_DIR = os.path.join(ROOT, "include")

SUCCESS, NO_GPU, INVALID_PARAM, OUT_OF_MEMORY, KERNEL_LAUNCH_FAILED, DEVICE_ERROR, NOT_SUPPORTED = range(7)
RESULT_NAMES = {0: "SUCCESS", 1: "NO_GPU", 2: "INVALID_PARAM", 3: "OUT_OF_MEMORY", 4: "KERNEL_LAUNCH_FAILED",
                5: "DEVICE_ERROR", 6: "NOT_SUPPORTED"}

_LIB = None

vp = C.c_void_p
i32 = C.c_int
u16p = C.POINTER(C.c_uint16)
f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)
sz = C.c_size_t


class DeviceInfo(C.Structure):
    _fields_ = [("device_id", C.c_int), ("name", C.c_char * 256), ("total_memory", C.c_size_t),
                ("free_memory", C.c_size_t), ("compute_major", C.c_int), ("compute_minor", C.c_int),
                ("max_threads_per_block", C.c_int), ("multiprocessor_count", C.c_int),
                ("has_matrix_cores", C.c_int), ("arch", C.c_char * 64)]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("launches", C.c_int), ("total_ms", C.c_double), ("alg_bytes", C.c_double)]


class Frame(C.Structure):
    _fields_ = [("pixels", C.c_void_p), ("width", C.c_int), ("height", C.c_int), ("format", C.c_int),
                ("gain", C.c_float), ("h", C.c_double * 9)]


class ChunkDesc(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("width", C.c_int), ("height", C.c_int), ("first", C.c_int),
                ("count", C.c_int), ("out", C.c_void_p), ("owner", C.c_void_p)]


class ExportChunk(C.Structure):
    _fields_ = [("x_offset", C.c_int), ("y_offset", C.c_int), ("width", C.c_int), ("height", C.c_int),
                ("pixels", C.c_void_p), ("buf_width", C.c_int), ("buf_height", C.c_int)]


PIX_F16, PIX_F32, PIX_U16 = 0, 1, 2
RENDER_FP64, RENDER_PLAIN, RENDER_PRECISION, RENDER_BATCHED, RENDER_SURROGATE = 0, 1, 2, 3, 4
RENDER_NAMES = {0: "fp64", 1: "plain", 2: "precision", 3: "batched", 4: "surrogate"}
RENDER_OUT_RGB48 = 0x100  # OR-ed into a render mode: outputs are RGB48 (uint16) blob payloads


# name -> (restype, argtypes)
class Tuning(C.Structure):
    _fields_ = [("chunk_groups", C.c_int), ("chamfer_strip", C.c_int), ("grid_share", C.c_int),
                ("ms", C.c_double)]


SIGNATURES = {
    "i360hip_create": (i32, [i32, C.POINTER(vp)]),
    "i360hip_destroy": (None, [vp]),
    "i360hip_get_device_info": (i32, [vp, C.POINTER(DeviceInfo)]),
    "i360hip_last_error": (C.c_char_p, [vp]),
    "i360hip_thread_error": (C.c_char_p, []),
    "i360hip_version": (C.c_char_p, []),
    "i360hip_profile_enable": (i32, [vp, i32]),
    "i360hip_profile_read": (i32, [vp, C.POINTER(KernelStat), i32, C.POINTER(i32)]),
    "i360hip_warp_perspective_u16": (i32, [vp, vp, i32, i32, vp, i32, i32, vp, i32, i32]),
    "i360hip_median_stack_u16": (i32, [vp, vp, i32, vp, i32, i32, C.c_float]),
    "i360hip_render_chunks": (i32, [vp, C.POINTER(Frame), i32, C.POINTER(ChunkDesc), i32, C.POINTER(i32), i32, i32,
                                    vp]),
    "i360hip_render_map_error": (i32, [vp, C.POINTER(Frame), i32, C.POINTER(ChunkDesc), i32, C.POINTER(i32),
                                       C.POINTER(C.c_double)]),
    "i360hip_apply_gain_u16": (i32, [vp, vp, i32, C.c_float]),
    "i360hip_demosaic_bayer_u16": (i32, [vp, vp, i32, i32, C.c_uint16, vp, vp, C.c_uint32, vp]),
    "i360hip_blend_chunk_u16": (i32, [vp, vp, i32, i32, i32, i32, vp, vp]),
    "i360hip_median_stack_u16_dev": (i32, [vp, vp, i32, vp, sz, C.c_float, vp]),
    "i360hip_demosaic_bayer_u16_dev": (i32, [vp, vp, i32, i32, C.c_uint16, vp, vp, C.c_uint32, vp, vp]),
    "i360hip_demosaic_bayer_f32_dev": (i32, [vp, vp, i32, i32, C.c_uint16, vp, vp, C.c_uint32, vp, vp]),
    "i360hip_warp_perspective_u16_dev": (i32, [vp, vp, i32, i32, vp, i32, i32, vp, i32, i32, vp]),
    "i360hip_apply_gain_u16_dev": (i32, [vp, vp, sz, C.c_float, vp]),
    "i360hip_blend_chunk_u16_dev": (i32, [vp, vp, i32, i32, i32, i32, vp, vp, vp]),
    "i360hip_surrogate_forward": (i32, [vp, vp, vp, vp, i32, i32, i32, vp, vp]),
    "i360hip_surrogate_gradients": (i32, [vp, vp, vp, vp, vp, i32, i32, i32, vp, vp, vp, vp]),
    "i360hip_surrogate_train_step": (i32, [vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, C.c_float, i32, vp, vp]),
    "i360hip_surrogate_round_bf16": (i32, [vp, vp, vp, vp]),
    "i360hip_surrogate_set": (i32, [vp, vp, vp, C.c_float]),
    "i360hip_surrogate_set_max_error": (i32, [vp, C.c_float]),
    "i360hip_surrogate_chunk_level": (i32, [i32, i32, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
    "i360hip_surrogate_features": (i32, [vp, C.POINTER(Frame), i32, C.POINTER(ChunkDesc), i32, C.POINTER(i32), i32,
                                         vp, vp, vp]),
    "i360hip_surrogate_stats": (i32, [vp, C.POINTER(i32), C.POINTER(i32)]),
    "i360hip_render_autotune": (i32, [vp, C.POINTER(Frame), i32, C.POINTER(ChunkDesc), i32, C.POINTER(i32), i32,
                                      C.POINTER(Tuning)]),
    "i360hip_render_set_tuning": (i32, [vp, C.POINTER(Tuning)]),
    "i360hip_export_scale": (C.c_float, [i32, i32, i32]),
    "i360hip_export_placement": (None, [i32, i32, i32, i32, C.c_float, i32p]),
    "i360hip_export_size": (i32, [i32, i32, i32, i32p, i32p]),
    "i360hip_export_composite_dev": (i32, [vp, C.POINTER(ExportChunk), i32, i32, i32, i32, i32, vp, vp]),
    "i360hip_export_composite": (i32, [vp, C.POINTER(ExportChunk), i32, i32, i32, i32, i32, vp]),
    "i360hip_pixel_buffer_to_blob": (i32, [vp, vp, i32, i32, vp, sz]),
    "i360hip_pixel_buffer_from_blob": (i32, [vp, vp, sz, vp, sz, C.POINTER(i32), C.POINTER(i32), C.POINTER(sz)]),
    "i360hip_rgb48_from_render_dev": (i32, [vp, vp, sz, vp, vp]),
    "i360hip_compute_lab_stats": (i32, [vp, vp, i32, i32, vp, vp]),
    "i360hip_apply_reinhard_u16": (i32, [vp, vp, i32, i32, vp, vp, vp, vp]),
    "i360hip_compute_lab_stats_dev": (i32, [vp, vp, i32, i32, vp, vp, vp]),
    "i360hip_compute_lab_stats_dev_noalloc": (i32, [vp, vp, i32, i32, vp, vp, vp]),
    "i360hip_apply_reinhard_u16_dev": (i32, [vp, vp, i32, i32, vp, vp, vp, vp, vp]),
}


def header_symbols():
    """Every function name declared in include/*.h."""
    names = set()
    for f in sorted(os.listdir(HEADER_DIR)):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(HEADER_DIR, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(i360hip_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


class HipLibraryMissing(RuntimeError):
    pass


class HipError(RuntimeError):
    def __init__(self, code, where, msg):
        super().__init__(f"{where}: {RESULT_NAMES.get(code, code)} ({code}): {msg}")
        self.code = code


def lib():
    global _LIB
    if _LIB is None:
        # One HIP runtime per process: torch ships its own libamdhip64.so.7;
        # when torch is importable, load it first so this library binds to
        # the same runtime (same soname) instead of a second copy, which
        # would leave torch without devices.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise HipLibraryMissing(f"{LIB_PATH} is not built; run `python -m image360_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB
