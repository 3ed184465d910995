"""Python mirror of ``WindowsApp::Compute::IComputeBackend`` over the C ABI.

Method names, argument meaning, result codes and error behaviour follow the
reference interface (WindowsApp.Compute/HeaderFiles/IComputeBackend.h:16-169):
every operation returns a ComputeResult code, writes into caller-allocated
output arrays and leaves the error text in ``GetLastError()``.  Device-array
variants (``*_dev``) take raw device pointers (ints, e.g. from
``torch.Tensor.data_ptr()``) and an optional stream handle.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import (DEVICE_ERROR, INVALID_PARAM, NOT_SUPPORTED, NO_GPU, SUCCESS, HipError, lib)

DEFAULT_BANDS = 5


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _c_contig(a, dtype):
    if not (isinstance(a, np.ndarray) and a.dtype == dtype and a.flags.c_contiguous):
        raise TypeError(f"expected a C-contiguous {np.dtype(dtype).name} array")
    return a


def _ordered(fn):
    """Device-pointer calls without a stream run on the context's own
    stream; the buffers they take are usually torch tensors written and read
    on torch's streams, so such a call is fenced by device synchronisations
    on both sides.  Calls given a stream are asynchronous on it."""
    import functools
    import inspect

    pos = list(inspect.signature(fn).parameters).index("stream")

    @functools.wraps(fn)
    def wrapper(*args, **kwargs):
        stream = kwargs.get("stream", args[pos] if len(args) > pos else None)
        if stream is not None:
            return fn(*args, **kwargs)
        import torch

        torch.cuda.synchronize()
        rc = fn(*args, **kwargs)
        torch.cuda.synchronize()
        return rc

    return wrapper


class HipComputeBackend:
    """IComputeBackend over the HIP library (one device context)."""

    def __init__(self, device: int = 0):
        self.device = device
        self._ctx = None
        self._error = ""

    # ---- lifetime (IComputeBackend.h:21-25)
    def Initialize(self) -> int:
        if self._ctx is not None:
            return SUCCESS
        L = lib()
        h = C.c_void_p()
        rc = L.i360hip_create(self.device, C.byref(h))
        if rc != SUCCESS:
            self._error = (L.i360hip_last_error(None) or b"").decode()
            return rc
        self._ctx = h
        return SUCCESS

    def Shutdown(self) -> None:
        if self._ctx is not None:
            lib().i360hip_destroy(self._ctx)
            self._ctx = None

    def IsInitialized(self) -> bool:
        return self._ctx is not None

    def GetGpuInfo(self) -> dict:
        info = _abi.DeviceInfo()
        if self._ctx is None:
            return {"deviceId": -1}
        lib().i360hip_get_device_info(self._ctx, C.byref(info))
        return {"deviceId": info.device_id, "name": info.name.decode(), "totalMemory": info.total_memory,
                "freeMemory": info.free_memory, "computeMajor": info.compute_major,
                "computeMinor": info.compute_minor, "maxThreadsPerBlock": info.max_threads_per_block,
                "multiProcessorCount": info.multiprocessor_count, "hasTensorCores": bool(info.has_matrix_cores),
                "arch": info.arch.decode()}

    def GetLastError(self) -> str:
        if self._ctx is None:
            return self._error or "Not initialized."
        return (lib().i360hip_last_error(self._ctx) or b"").decode()

    def _ready(self) -> int:
        if self._ctx is None:
            self._error = "Not initialized."
            return DEVICE_ERROR
        return SUCCESS

    # ---- host-array operations (synchronous)
    def WarpPerspective(self, src, srcW, srcH, dst, dstW, dstH, homography, offsetX, offsetY) -> int:
        if (rc := self._ready()) != SUCCESS:
            return rc
        src = _c_contig(src, np.uint16)
        dst = _c_contig(dst, np.uint16)
        m = np.ascontiguousarray(homography, dtype=np.float32).reshape(9)
        return lib().i360hip_warp_perspective_u16(self._ctx, _ptr(src), srcW, srcH, _ptr(dst), dstW, dstH, _ptr(m),
                                                 offsetX, offsetY)

    def MedianStack(self, inputs, numInputs, output, width, height, sigmaThreshold=2.0) -> int:
        if (rc := self._ready()) != SUCCESS:
            return rc
        planes = [_c_contig(p, np.uint16) for p in inputs[:numInputs]]
        arr = (C.c_void_p * max(1, len(planes)))(*[_ptr(p) for p in planes])
        output = _c_contig(output, np.uint16)
        return lib().i360hip_median_stack_u16(self._ctx, arr, numInputs, _ptr(output), width, height,
                                              float(sigmaThreshold))

    def ApplyGain(self, data, numPixels, gain) -> int:
        if (rc := self._ready()) != SUCCESS:
            return rc
        data = _c_contig(data, np.uint16)
        return lib().i360hip_apply_gain_u16(self._ctx, _ptr(data), numPixels, float(gain))

    def DemosaicBayer(self, cfaData, width, height, blackLevel, camMul, rgbCam, filters, rgbOut) -> int:
        if (rc := self._ready()) != SUCCESS:
            return rc
        cfa = _c_contig(cfaData, np.uint16)
        cm = np.ascontiguousarray(camMul, dtype=np.float32).reshape(4)
        rc_ = np.ascontiguousarray(rgbCam, dtype=np.float32).reshape(12)
        out = _c_contig(rgbOut, np.uint16)
        return lib().i360hip_demosaic_bayer_u16(self._ctx, _ptr(cfa), width, height, blackLevel, _ptr(cm), _ptr(rc_),
                                                filters, _ptr(out))

    def BlendChunkContributors(self, warpedBuffers, width, height, outResult, numBands=DEFAULT_BANDS,
                               ownerOut=None) -> int:
        """Kernels::SeamBlend::BlendChunkContributors (SeamBlendKernels.h:91-92)."""
        if (rc := self._ready()) != SUCCESS:
            return rc
        bufs = [_c_contig(b, np.uint16) for b in warpedBuffers]
        arr = (C.c_void_p * max(1, len(bufs)))(*[_ptr(b) for b in bufs])
        out = _c_contig(outResult, np.uint16)
        own = None if ownerOut is None else _ptr(_c_contig(ownerOut, np.int32))
        return lib().i360hip_blend_chunk_u16(self._ctx, arr, len(bufs), width, height, numBands, _ptr(out), own)

    # ---- optimize-stage colour operations (GainColorOps.cpp:103-157)
    def ComputeLabStats(self, rgb, width, height, outMean, outStd) -> int:
        """outMean / outStd: float64 arrays of 3, written in place."""
        if (rc := self._ready()) != SUCCESS:
            return rc
        rgb = _c_contig(rgb, np.uint16)
        m = _c_contig(outMean, np.float64)
        sd = _c_contig(outStd, np.float64)
        return lib().i360hip_compute_lab_stats(self._ctx, _ptr(rgb), width, height, _ptr(m), _ptr(sd))

    def ApplyReinhardColorTransfer(self, rgbInOut, width, height, srcMean, srcStd, refMean, refStd) -> int:
        if (rc := self._ready()) != SUCCESS:
            return rc
        rgb = _c_contig(rgbInOut, np.uint16)
        v = [np.ascontiguousarray(x, dtype=np.float64).reshape(3) for x in (srcMean, srcStd, refMean, refStd)]
        return lib().i360hip_apply_reinhard_u16(self._ctx, _ptr(rgb), width, height, *[_ptr(x) for x in v])

    def lab_stats_dev(self, rgb_ptr, width, height, stream=None):
        """(rc, mean[3], std[3]) of a device RGB48 image; returns once the result is on the host."""
        m, sd = np.zeros(3), np.zeros(3)
        rc = self._lab_stats_dev(rgb_ptr, width, height, m, sd, stream=stream)
        return rc, m, sd

    @_ordered
    def _lab_stats_dev(self, rgb_ptr, width, height, m, sd, stream=None) -> int:
        return lib().i360hip_compute_lab_stats_dev(self._ctx, rgb_ptr, width, height, _ptr(m), _ptr(sd), stream)

    @_ordered
    def reinhard_dev(self, rgb_ptr, width, height, src_mean, src_std, ref_mean, ref_std, stream=None) -> int:
        v = [np.ascontiguousarray(x, dtype=np.float64).reshape(3) for x in (src_mean, src_std, ref_mean, ref_std)]
        return lib().i360hip_apply_reinhard_u16_dev(self._ctx, rgb_ptr, width, height, *[_ptr(x) for x in v], stream)

    # Align / burst operations are not on this backend's path: the reference
    # GPU backends answer NOT_SUPPORTED for the burst family
    # (CudaPipeline.cpp:794-839); hosts delegate the rest to their CPU backend.
    def BlockMatchAlign(self, *a, **k) -> int:
        self._error = "BlockMatchAlign is not supported by the HIP backend."
        return NOT_SUPPORTED

    RobustMergeAccumulate = TileFftMerge = StructureTensorKernelRegression = BlockMatchAlign

    # ---- device-pointer operations (asynchronous on `stream`)
    @_ordered
    def median_stack_dev(self, input_ptrs, output_ptr, num_scalars, sigma=2.0, stream=None) -> int:
        arr = (C.c_void_p * max(1, len(input_ptrs)))(*input_ptrs)
        return lib().i360hip_median_stack_u16_dev(self._ctx, arr, len(input_ptrs), output_ptr, num_scalars,
                                                  float(sigma), stream)

    @_ordered
    def demosaic_dev(self, cfa_ptr, width, height, black, cam_mul, rgb_cam, filters, out_ptr, f32=False,
                     stream=None) -> int:
        cm = np.ascontiguousarray(cam_mul, dtype=np.float32).reshape(4)
        rc_ = np.ascontiguousarray(rgb_cam, dtype=np.float32).reshape(12)
        fn = lib().i360hip_demosaic_bayer_f32_dev if f32 else lib().i360hip_demosaic_bayer_u16_dev
        return fn(self._ctx, cfa_ptr, width, height, black, _ptr(cm), _ptr(rc_), filters, out_ptr, stream)

    @_ordered
    def warp_dev(self, src_ptr, srcW, srcH, dst_ptr, dstW, dstH, inv_h, offX, offY, stream=None) -> int:
        m = np.ascontiguousarray(inv_h, dtype=np.float32).reshape(9)
        return lib().i360hip_warp_perspective_u16_dev(self._ctx, src_ptr, srcW, srcH, dst_ptr, dstW, dstH, _ptr(m),
                                                      offX, offY, stream)

    @_ordered
    def gain_dev(self, ptr, n, gain, stream=None) -> int:
        return lib().i360hip_apply_gain_u16_dev(self._ctx, ptr, n, float(gain), stream)

    @_ordered
    def blend_dev(self, warped_ptrs, width, height, out_ptr, num_bands=DEFAULT_BANDS, owner_ptr=None,
                  stream=None) -> int:
        arr = (C.c_void_p * max(1, len(warped_ptrs)))(*warped_ptrs)
        return lib().i360hip_blend_chunk_u16_dev(self._ctx, arr, len(warped_ptrs), width, height, num_bands,
                                                 out_ptr, owner_ptr, stream)

    @staticmethod
    def _render_args(frames, chunks, contributors):
        fa = (_abi.Frame * max(1, len(frames)))()
        for i, f in enumerate(frames):
            fa[i].pixels = f["ptr"]
            fa[i].width = f["w"]
            fa[i].height = f["h"]
            fa[i].format = f.get("fmt", _abi.PIX_F16)
            fa[i].gain = f.get("gain", 1.0)
            for k, v in enumerate(np.asarray(f["H"], dtype=np.float64).reshape(9)):
                fa[i].h[k] = float(v)
        ca = (_abi.ChunkDesc * max(1, len(chunks)))()
        for i, c in enumerate(chunks):
            ca[i].x, ca[i].y, ca[i].width, ca[i].height = c["x"], c["y"], c["w"], c["h"]
            ca[i].first, ca[i].count = c["first"], c["count"]
            ca[i].out = c["out"]
            ca[i].owner = c.get("owner") or None
        cont = (C.c_int * max(1, len(contributors)))(*[int(v) for v in contributors])
        return fa, ca, cont

    @_ordered
    def render_chunks(self, frames, chunks, contributors, num_bands, mode, stream=None) -> int:
        """i360hip_render_chunks: frames = [{ptr, w, h, fmt, gain, H}], chunks =
        [{x, y, w, h, first, count, out, owner}] (device pointers)."""
        fa, ca, cont = self._render_args(frames, chunks, contributors)
        return lib().i360hip_render_chunks(self._ctx, fa, len(frames), ca, len(chunks), cont, num_bands, mode, stream)

    def render_autotune(self, frames, chunks, contributors, num_bands) -> dict:
        """i360hip_render_autotune: picks and keeps the fastest launch
        configuration of the batched form for these chunks (synchronous)."""
        fa, ca, cont = self._render_args(frames, chunks, contributors)
        t = _abi.Tuning()
        self.check(lib().i360hip_render_autotune(self._ctx, fa, len(frames), ca, len(chunks), cont, num_bands,
                                                 C.byref(t)), "render_autotune")
        return {"chunk_groups": t.chunk_groups, "chamfer_strip": t.chamfer_strip, "grid_share": t.grid_share,
                "ms": t.ms}

    def render_set_tuning(self, chunk_groups=0, chamfer_strip=0, grid_share=0) -> None:
        t = _abi.Tuning(chunk_groups, chamfer_strip, grid_share, 0.0)
        self.check(lib().i360hip_render_set_tuning(self._ctx, C.byref(t)), "render_set_tuning")

    def render_map_error(self, frames, chunks, contributors) -> float:
        fa, ca, cont = self._render_args(frames, chunks, contributors)
        out = C.c_double(0)
        self.check(lib().i360hip_render_map_error(self._ctx, fa, len(frames), ca, len(chunks), cont, C.byref(out)),
                   "render_map_error")
        return out.value

    # ---- learned surrogate (device pointers; see include/i360hip.h) ----------
    @_ordered
    def surrogate_forward(self, weights, biases, feat, n, h, w, out, stream=None) -> int:
        return lib().i360hip_surrogate_forward(self._ctx, weights, biases, feat, n, h, w, out, stream)

    @_ordered
    def surrogate_gradients(self, weights, biases, feat, target, n, h, w, grad_w, grad_b, loss, stream=None) -> int:
        return lib().i360hip_surrogate_gradients(self._ctx, weights, biases, feat, target, n, h, w, grad_w, grad_b,
                                                 loss, stream)

    @_ordered
    def surrogate_train_step(self, master, biases, adam, weights, feat, target, n, h, w, lr, step, loss,
                             stream=None) -> int:
        return lib().i360hip_surrogate_train_step(self._ctx, master, biases, adam, weights, feat, target, n, h, w,
                                                  lr, step, loss, stream)

    @_ordered
    def surrogate_round_bf16(self, master, weights, stream=None) -> int:
        return lib().i360hip_surrogate_round_bf16(self._ctx, master, weights, stream)

    def surrogate_set(self, weights, biases, threshold_mse=1e-5) -> int:
        return lib().i360hip_surrogate_set(self._ctx, weights, biases, threshold_mse)

    def surrogate_set_max_error(self, max_sq_err) -> int:
        return lib().i360hip_surrogate_set_max_error(self._ctx, max_sq_err)

    @staticmethod
    def surrogate_chunk_level(w, h, num_bands):
        """(K, level width, level height); K = 0 when the chunk has < 4 levels."""
        k, lw, lh = C.c_int(0), C.c_int(0), C.c_int(0)
        rc = lib().i360hip_surrogate_chunk_level(w, h, num_bands, C.byref(k), C.byref(lw), C.byref(lh))
        if rc != SUCCESS:
            raise HipError(rc, "surrogate_chunk_level", "invalid arguments")
        return k.value, lw.value, lh.value

    @_ordered
    def surrogate_features(self, frames, chunks, contributors, num_bands, feat, target, stream=None) -> int:
        fa, ca, cont = self._render_args(frames, chunks, contributors)
        return lib().i360hip_surrogate_features(self._ctx, fa, len(frames), ca, len(chunks), cont, num_bands, feat,
                                                target, stream)

    def surrogate_stats(self):
        t, f = C.c_int(0), C.c_int(0)
        self.check(lib().i360hip_surrogate_stats(self._ctx, C.byref(t), C.byref(f)), "surrogate_stats")
        return t.value, f.value

    # ---- preview export (PanoramaExporter.cpp:9-24, 34-98) ----
    @staticmethod
    def export_scale(total_w, total_h, max_dim) -> float:
        """ComputeExportScale (PanoramaExporter.cpp:9-14)."""
        return float(lib().i360hip_export_scale(total_w, total_h, max_dim))

    @staticmethod
    def export_placement(x_offset, y_offset, width, height, scale):
        """ComputeChunkPlacement (PanoramaExporter.cpp:16-24): (destX, destY, destW, destH)."""
        out = (C.c_int32 * 4)()
        lib().i360hip_export_placement(x_offset, y_offset, width, height, scale, out)
        return tuple(out)

    @staticmethod
    def export_size(total_w, total_h, max_dim):
        """(out_w, out_h) of the composite, or None where ExportPreviewJpeg returns false."""
        w, h = C.c_int32(0), C.c_int32(0)
        if lib().i360hip_export_size(total_w, total_h, max_dim, C.byref(w), C.byref(h)) != SUCCESS:
            return None
        return w.value, h.value

    @_ordered
    def export_composite_dev(self, chunks, total_w, total_h, max_dim, fmt, out_ptr, stream=None) -> int:
        """chunks: sequence of (x_offset, y_offset, width, height, pixels_ptr or None, buf_w, buf_h);
        fmt: _abi.PIX_U16 (RGB48) or _abi.PIX_F32 (render output)."""
        arr = (_abi.ExportChunk * max(1, len(chunks)))()
        for i, (x, y, w, h, p, bw, bh) in enumerate(chunks):
            arr[i] = _abi.ExportChunk(x, y, w, h, p, bw, bh)
        return lib().i360hip_export_composite_dev(self._ctx, arr, len(chunks), total_w, total_h, max_dim, fmt,
                                                  out_ptr, stream)

    def ExportComposite(self, chunks, total_w, total_h, max_dim, out):
        """Host-buffer composite: chunks = [(x_offset, y_offset, width, height, buf or None)] with buf an
        (h, w, 3) uint16 (RGB48) or float32 array, all of one dtype; out: (out_h, out_w, 3) uint8."""
        if (rc := self._ready()) != SUCCESS:
            return rc
        bufs = [None if c[4] is None else np.ascontiguousarray(c[4]) for c in chunks]
        kinds = {b.dtype for b in bufs if b is not None}
        if len(kinds) > 1:
            raise TypeError("all chunk buffers of one call must share a dtype")
        fmt = _abi.PIX_F32 if kinds == {np.dtype(np.float32)} else _abi.PIX_U16
        arr = (_abi.ExportChunk * max(1, len(chunks)))()
        for i, (c, b) in enumerate(zip(chunks, bufs)):
            arr[i] = _abi.ExportChunk(c[0], c[1], c[2], c[3], None if b is None else b.ctypes.data,
                                      0 if b is None else b.shape[1], 0 if b is None else b.shape[0])
        out = _c_contig(out, np.uint8)
        return lib().i360hip_export_composite(self._ctx, arr, len(chunks), total_w, total_h, max_dim, fmt, _ptr(out))

    def pixel_buffer_to_blob(self, rgb48_ptr, width, height):
        """(rc, bytes) of StorageEngine::WritePixelBuffer's framing of a device RGB48 buffer."""
        import torch

        torch.cuda.synchronize()  # the buffer is usually written on torch's streams
        blob = np.zeros(8 + width * height * 6, np.uint8)
        rc = lib().i360hip_pixel_buffer_to_blob(self._ctx, rgb48_ptr, width, height, _ptr(blob), blob.size)
        return rc, blob.tobytes()

    def pixel_buffer_from_blob(self, blob: bytes, rgb48_ptr, capacity_values):
        """(rc, width, height, values) after uploading a blob's payload to a device buffer."""
        import torch

        torch.cuda.synchronize()
        b = np.frombuffer(blob, np.uint8)
        w, h, n = C.c_int32(0), C.c_int32(0), C.c_size_t(0)
        rc = lib().i360hip_pixel_buffer_from_blob(self._ctx, b.ctypes.data if b.size else None, b.size, rgb48_ptr,
                                                  capacity_values, C.byref(w), C.byref(h), C.byref(n))
        return rc, w.value, h.value, n.value

    @_ordered
    def rgb48_from_render_dev(self, rgb_ptr, num_values, out_ptr, stream=None) -> int:
        return lib().i360hip_rgb48_from_render_dev(self._ctx, rgb_ptr, num_values, out_ptr, stream)

    def profile_enable(self, on=True) -> None:
        lib().i360hip_profile_enable(self._ctx, 1 if on else 0)

    def profile_read(self) -> dict:
        """{kernel: {"launches", "total_ms", "alg_bytes"}} from event timing."""
        arr = (_abi.KernelStat * 64)()
        n = C.c_int(0)
        lib().i360hip_profile_read(self._ctx, arr, 64, C.byref(n))
        return {arr[i].name.decode(): {"launches": arr[i].launches, "total_ms": arr[i].total_ms,
                                       "alg_bytes": arr[i].alg_bytes} for i in range(n.value)}

    def check(self, rc: int, where: str) -> None:
        if rc != SUCCESS:
            raise HipError(rc, where, self.GetLastError())

    def __enter__(self):
        rc = self.Initialize()
        if rc != SUCCESS:
            raise HipError(rc, "Initialize", self.GetLastError())
        return self

    def __exit__(self, *exc):
        self.Shutdown()

    def __del__(self):
        try:
            self.Shutdown()
        except Exception:
            pass
