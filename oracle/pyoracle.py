"""ctypes binding of the oracle (TEST INFRASTRUCTURE ONLY).

Only the test suite, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module, and only as the checker.  It binds
``oracle/liboracle.so`` (the scalar restatement of the reference CPU backend)
and, when present, ``oracle/_ref/libi360ref.so`` (the reference's own
hot-path sources compiled by ``make -C oracle ref``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ORC = None
_REF = None

u16p = C.POINTER(C.c_uint16)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int32)
u8p = C.POINTER(C.c_uint8)


def _p(a, t):
    return a.ctypes.data_as(t)


def lib():
    global _ORC
    if _ORC is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run: make -C oracle)")
        _ORC = C.CDLL(path)
    return _ORC


def ref():
    """The compiled reference sources, or None where they were not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libi360ref.so")
        if not os.path.exists(path):
            return None
        _REF = C.CDLL(path)
    return _REF


_REF_NC = None


def ref_nc():
    """The compiled reference sources built with floating-point contraction
    off (no fused multiply-adds in the AVX units), or None."""
    global _REF_NC
    if _REF_NC is None:
        path = os.path.join(_HERE, "_ref", "libi360ref_nc.so")
        if not os.path.exists(path):
            return None
        _REF_NC = C.CDLL(path)
    return _REF_NC


def _ptr_array(planes, t=u16p):
    arr = (t * len(planes))()
    for i, p in enumerate(planes):
        arr[i] = p.ctypes.data_as(t)
    return arr


# ----------------------------------------------------------------- oracle --
def median_stack(planes, sigma=2.0):
    planes = [np.ascontiguousarray(p, dtype=np.uint16) for p in planes]
    out = np.zeros_like(planes[0])
    lib().orc_median_stack_u16(_ptr_array(planes), len(planes), _p(out, u16p), C.c_longlong(out.size),
                               C.c_float(sigma))
    return out


def warp_perspective(src, dst_w, dst_h, invH, off_x=0, off_y=0):
    src = np.ascontiguousarray(src, dtype=np.uint16)
    h, w = src.shape[:2]
    m = np.ascontiguousarray(invH, dtype=np.float32).reshape(9)
    out = np.zeros((dst_h, dst_w, 3), np.uint16)
    lib().orc_warp_perspective_u16(_p(src, u16p), w, h, _p(out, u16p), dst_w, dst_h, _p(m, f32p), off_x, off_y)
    return out


def apply_gain(data, gain):
    d = np.ascontiguousarray(data, dtype=np.uint16).copy()
    lib().orc_apply_gain_u16(_p(d, u16p), C.c_longlong(d.size), C.c_float(gain))
    return d


def demosaic(cfa, black, cam_mul, rgb_cam, filters, f32=False):
    cfa = np.ascontiguousarray(cfa, dtype=np.uint16)
    h, w = cfa.shape
    cm = np.ascontiguousarray(cam_mul, dtype=np.float32)
    rc = np.ascontiguousarray(rgb_cam, dtype=np.float32).reshape(12)
    if f32:
        out = np.zeros((h, w, 3), np.float32)
        lib().orc_demosaic_bayer_f32(_p(cfa, u16p), w, h, C.c_uint16(black), _p(cm, f32p), _p(rc, f32p),
                                     C.c_uint32(filters), _p(out, f32p))
    else:
        out = np.zeros((h, w, 3), np.uint16)
        lib().orc_demosaic_bayer_u16(_p(cfa, u16p), w, h, C.c_uint16(black), _p(cm, f32p), _p(rc, f32p),
                                     C.c_uint32(filters), _p(out, u16p))
    return out


def invert_homography(hm):
    m = np.ascontiguousarray(hm, dtype=np.float32).reshape(9)
    out = np.zeros(9, np.float32)
    ok = lib().orc_invert_homography(_p(m, f32p), _p(out, f32p))
    return out if ok else None


def invert_homography_f64(hm):
    m = np.ascontiguousarray(hm, dtype=np.float64).reshape(9)
    out = np.zeros(9, np.float64)
    ok = lib().orc_invert_homography_f64(_p(m, f64p), _p(out, f64p))
    return out if ok else None


def coverage_distance(cov):
    cov = np.ascontiguousarray(cov, dtype=np.uint8)
    h, w = cov.shape
    out = np.zeros((h, w), np.float32)
    lib().orc_coverage_distance(_p(cov, u8p), w, h, _p(out, f32p))
    return out


def ownership(dists):
    """ComputeOwnership: argmax distance > 0, ties to the lowest index."""
    dists = [np.ascontiguousarray(d, dtype=np.float32) for d in dists]
    h, w = dists[0].shape
    owner = np.zeros((h, w), np.int32)
    lib().orc_ownership(_ptr_array(dists, f32p), len(dists), w, h, _p(owner, i32p))
    return owner


def cover_box(cov):
    cov = np.ascontiguousarray(cov, dtype=np.uint8)
    h, w = cov.shape
    box = np.zeros(4, np.int32)
    ok = lib().orc_cover_box(_p(cov, u8p), w, h, _p(box, i32p))
    return box if ok else None


def refine_seam_dp(imgI, imgJ, distI, distJ, ci, cj, boxI, boxJ, ov, owner):
    """RefineSeamDP of pair (ci, cj) on RGB48 images, updating `owner` in place."""
    imgI = np.ascontiguousarray(imgI, dtype=np.uint16)
    imgJ = np.ascontiguousarray(imgJ, dtype=np.uint16)
    distI = np.ascontiguousarray(distI, dtype=np.float32)
    distJ = np.ascontiguousarray(distJ, dtype=np.float32)
    assert owner.dtype == np.int32 and owner.flags.c_contiguous
    w = owner.shape[1]
    b = [np.ascontiguousarray(x, dtype=np.int32) for x in (boxI, boxJ, ov)]
    lib().orc_refine_seam_dp(_p(imgI, u16p), _p(imgJ, u16p), _p(distI, f32p), _p(distJ, f32p), w, ci, cj,
                             _p(b[0], i32p), _p(b[1], i32p), _p(b[2], i32p), _p(owner, i32p))


def blend_chunk(warped, num_bands=5, want_owner=False):
    warped = [np.ascontiguousarray(x, dtype=np.uint16) for x in warped]
    h, w = warped[0].shape[:2]
    out = np.zeros((h, w, 3), np.uint16)
    owner = np.zeros((h, w), np.int32)
    lib().orc_blend_chunk_u16(_ptr_array(warped), len(warped), w, h, num_bands, _p(out, u16p), _p(owner, i32p))
    return (out, owner) if want_owner else out


def multiband_compose(warped, owner, num_bands=5):
    warped = [np.ascontiguousarray(x, dtype=np.uint16) for x in warped]
    h, w = warped[0].shape[:2]
    covs = [np.ascontiguousarray((x != 0).any(axis=2).astype(np.uint8)) for x in warped]
    owner = np.ascontiguousarray(owner, dtype=np.int32)
    out = np.zeros((h, w, 3), np.uint16)
    lib().orc_multiband_compose_u16(_ptr_array(warped), _ptr_array(covs, u8p), _p(owner, i32p), len(warped), w, h,
                                    num_bands, _p(out, u16p))
    return out


def clamp_bands(w, h, n):
    return lib().orc_clamp_bands(w, h, n)


def render_chunk(srcs, hinv, gains, x0, y0, w, h, num_bands, fp64=True, owner_in=None, want_owner=False):
    """Float chunk render; srcs are float32 RGB frames, hinv canvas->source."""
    srcs = [np.ascontiguousarray(s, dtype=np.float32) for s in srcs]
    c = len(srcs)
    sw = np.array([s.shape[1] for s in srcs], np.int32)
    sh = np.array([s.shape[0] for s in srcs], np.int32)
    g = np.ascontiguousarray(gains, dtype=np.float32)
    out = np.zeros((h, w, 3), np.float32)
    owner = np.zeros((h, w), np.int32)
    oin = None if owner_in is None else np.ascontiguousarray(owner_in, dtype=np.int32)
    if fp64:
        m = np.ascontiguousarray(hinv, dtype=np.float64).reshape(c * 9)
        lib().orc_render_chunk_f64(_ptr_array(srcs, f32p), _p(sw, i32p), _p(sh, i32p), _p(m, f64p), _p(g, f32p), c,
                                   x0, y0, w, h, num_bands, _p(out, f32p), _p(owner, i32p),
                                   None if oin is None else _p(oin, i32p))
    else:
        m = np.ascontiguousarray(hinv, dtype=np.float32).reshape(c * 9)
        lib().orc_render_chunk_f32(_ptr_array(srcs, f32p), _p(sw, i32p), _p(sh, i32p), _p(m, f32p), _p(g, f32p), c,
                                   x0, y0, w, h, num_bands, _p(out, f32p), _p(owner, i32p),
                                   None if oin is None else _p(oin, i32p))
    return (out, owner) if want_owner else out


def warp_float(src, hinv, gain, x0, y0, w, h, fp64=True):
    """One contributor's float warp into a w x h chunk at canvas (x0, y0):
    (rgb (h, w, 3), coverage (h, w) uint8).  src: float32 RGB frame."""
    src = np.ascontiguousarray(src, dtype=np.float32)
    sh, sw = src.shape[:2]
    cov = np.zeros((h, w), np.uint8)
    if fp64:
        m = np.ascontiguousarray(hinv, dtype=np.float64).reshape(9)
        rgb = np.zeros((h, w, 3), np.float64)
        lib().orc_warp_float_f64(_p(src, f32p), sw, sh, _p(m, f64p), C.c_float(gain), x0, y0, w, h, _p(rgb, f64p),
                                 _p(cov, u8p))
    else:
        m = np.ascontiguousarray(hinv, dtype=np.float32).reshape(9)
        rgb = np.zeros((h, w, 3), np.float32)
        lib().orc_warp_float_f32(_p(src, f32p), sw, sh, _p(m, f32p), C.c_float(gain), x0, y0, w, h, _p(rgb, f32p),
                                 _p(cov, u8p))
    return rgb, cov


class planted:
    """Context manager planting a mistake in the float warp (1: w10 and w01
    swapped; 2: gain applied after the clamp), for tests that show a pin
    fails on it."""

    def __init__(self, which):
        self.which = which

    def __enter__(self):
        lib().orc_render_plant(self.which)
        return self

    def __exit__(self, *exc):
        lib().orc_render_plant(0)
        return False


# ---------------------------------------------------------------- export --
def export_scale(total_w, total_h, max_dim):
    f = lib().orc_export_scale
    f.restype = C.c_float
    f.argtypes = [C.c_int, C.c_int, C.c_int]
    return f(total_w, total_h, max_dim)


def export_placement(x, y, w, h, scale):
    out = (C.c_int * 4)()
    f = lib().orc_export_placement
    f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.POINTER(C.c_int)]
    f(x, y, w, h, scale, out)
    return tuple(out)


def export_size(total_w, total_h, max_dim):
    w, h = C.c_int(0), C.c_int(0)
    ok = lib().orc_export_size(total_w, total_h, max_dim, C.byref(w), C.byref(h))
    return (w.value, h.value) if ok else None


def rgb48_from_render(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, np.uint16)
    lib().orc_rgb48_from_render(_p(x, f32p), C.c_longlong(x.size), _p(out, u16p))
    return out


def export_composite(chunks, total_w, total_h, max_dim):
    """chunks: list of (x_offset, y_offset, width, height, buf) with buf an
    (h, w, 3) uint16 or float32 array, or None.  All buffers of one call
    share one dtype.  Returns the (out_h, out_w, 3) uint8 composite, or None
    where the reference returns false."""
    size = export_size(total_w, total_h, max_dim)
    if size is None:
        return None
    n = len(chunks)
    bufs = [None if c[4] is None else np.ascontiguousarray(c[4]) for c in chunks]
    kinds = {b.dtype for b in bufs if b is not None}
    assert len(kinds) <= 1
    is_f32 = int(kinds == {np.dtype(np.float32)})
    rect = np.array([c[:4] for c in chunks] or [[0, 0, 0, 0]], np.int32).reshape(-1)
    bw = np.array([0 if b is None else b.shape[1] for b in bufs] or [0], np.int32)
    bh = np.array([0 if b is None else b.shape[0] for b in bufs] or [0], np.int32)
    ptrs = (C.c_void_p * max(1, n))()
    for i, b in enumerate(bufs):
        ptrs[i] = None if b is None else b.ctypes.data
    out = np.empty((size[1], size[0], 3), np.uint8)
    lib().orc_export_composite(_p(rect, i32p), ptrs, _p(bw, i32p), _p(bh, i32p), n, total_w, total_h, max_dim,
                               is_f32, _p(out, u8p))
    return out


# ----------------------------------------------------------------- colour --
def rgb_to_lab(rgb):
    rgb = np.ascontiguousarray(rgb, dtype=np.uint16)
    out = np.empty(rgb.shape, np.float32)
    lib().orc_rgb_to_lab(_p(rgb, u16p), C.c_longlong(rgb.size // 3), _p(out, f32p))
    return out


def lab_to_rgb(lab):
    lab = np.ascontiguousarray(lab, dtype=np.float32)
    out = np.empty(lab.shape, np.uint16)
    lib().orc_lab_to_rgb(_p(lab, f32p), C.c_longlong(lab.size // 3), _p(out, u16p))
    return out


def _stats(fn, rgb):
    rgb = np.ascontiguousarray(rgb, dtype=np.uint16)
    h, w = rgb.shape[0], rgb.size // (3 * rgb.shape[0])
    m, s = np.zeros(3), np.zeros(3)
    fn(_p(rgb, u16p), w, h, _p(m, f64p), _p(s, f64p))
    return m, s


def _transfer(fn, rgb, sm, ss, rm, rs):
    out = np.ascontiguousarray(rgb, dtype=np.uint16).copy()
    h, w = out.shape[0], out.size // (3 * out.shape[0])
    args = [np.ascontiguousarray(v, dtype=np.float64).reshape(3) for v in (sm, ss, rm, rs)]
    fn(_p(out, u16p), w, h, *[_p(v, f64p) for v in args])
    return out


def lab_stats(rgb):
    """(mean[3], std[3]) of an (h, w, 3) RGB48 image (GainColorOps.cpp:103-130)."""
    return _stats(lib().orc_lab_stats, rgb)


def reinhard(rgb, src_mean, src_std, ref_mean, ref_std):
    """Reinhard transfer of an (h, w, 3) RGB48 image (GainColorOps.cpp:132-157)."""
    return _transfer(lib().orc_reinhard, rgb, src_mean, src_std, ref_mean, ref_std)


# -------------------------------------------------------------- reference --
def ref_tier():
    r = ref()
    return -1 if r is None else r.ref_simd_tier()


def ref_median_stack(planes, sigma=2.0, tier=0, lib=None):
    planes = [np.ascontiguousarray(p, dtype=np.uint16) for p in planes]
    out = np.zeros_like(planes[0])
    n = out.size // 3
    (lib or ref()).ref_median_stack(tier, _ptr_array(planes), len(planes), _p(out, u16p), n, 1, C.c_float(sigma))
    return out


def ref_median_stack_ptrs(plane_ptrs, out_ptr, num_scalars, sigma=2.0, tier=0, lib=None):
    """The reference MedianStack over `num_scalars` (a multiple of 3) values
    at raw host addresses (one band of the planes)."""
    arr = (u16p * len(plane_ptrs))()
    for i, p in enumerate(plane_ptrs):
        arr[i] = C.cast(C.c_void_p(p), u16p)
    (lib or ref()).ref_median_stack(tier, arr, len(plane_ptrs), C.cast(C.c_void_p(out_ptr), u16p), num_scalars // 3, 1,
                           C.c_float(sigma))


def ref_demosaic_ptrs(cfa_ptr, w, h, black, cam_mul, rgb_cam, filters, out_ptr, tier=0):
    """The reference DemosaicBayer on a w x h CFA block at a raw host address."""
    cm = np.ascontiguousarray(cam_mul, dtype=np.float32)
    rc = np.ascontiguousarray(rgb_cam, dtype=np.float32).reshape(12)
    ref().ref_demosaic_bayer(tier, C.cast(C.c_void_p(cfa_ptr), u16p), w, h, C.c_uint16(black), _p(cm, f32p),
                             _p(rc, f32p), C.c_uint32(filters), C.cast(C.c_void_p(out_ptr), u16p))


def ref_warp_perspective(src, dst_w, dst_h, invH, off_x=0, off_y=0, tier=0):
    src = np.ascontiguousarray(src, dtype=np.uint16)
    h, w = src.shape[:2]
    m = np.ascontiguousarray(invH, dtype=np.float32).reshape(9)
    out = np.zeros((dst_h, dst_w, 3), np.uint16)
    ref().ref_warp_perspective(tier, _p(src, u16p), w, h, _p(out, u16p), dst_w, dst_h, _p(m, f32p), off_x, off_y)
    return out


def ref_demosaic(cfa, black, cam_mul, rgb_cam, filters, tier=0, lib=None):
    cfa = np.ascontiguousarray(cfa, dtype=np.uint16)
    h, w = cfa.shape
    cm = np.ascontiguousarray(cam_mul, dtype=np.float32)
    rc = np.ascontiguousarray(rgb_cam, dtype=np.float32).reshape(12)
    out = np.zeros((h, w, 3), np.uint16)
    (lib or ref()).ref_demosaic_bayer(tier, _p(cfa, u16p), w, h, C.c_uint16(black), _p(cm, f32p), _p(rc, f32p),
                             C.c_uint32(filters), _p(out, u16p))
    return out


def ref_apply_gain(data, gain):
    d = np.ascontiguousarray(data, dtype=np.uint16).copy()
    ref().ref_apply_gain(_p(d, u16p), d.size, C.c_float(gain))
    return d


def ref_lab_stats(rgb):
    return _stats(ref().ref_lab_stats, rgb)


def ref_reinhard(rgb, src_mean, src_std, ref_mean, ref_std):
    return _transfer(ref().ref_reinhard, rgb, src_mean, src_std, ref_mean, ref_std)


def ref_invert_homography(hm):
    m = np.ascontiguousarray(hm, dtype=np.float32).reshape(9)
    out = np.zeros(9, np.float32)
    ok = ref().ref_invert_homography(_p(m, f32p), _p(out, f32p))
    return out if ok else None


def ref_blend_chunk(warped):
    warped = [np.ascontiguousarray(x, dtype=np.uint16) for x in warped]
    h, w = warped[0].shape[:2]
    out = np.zeros((h, w, 3), np.uint16)
    ref().ref_blend_chunk(_ptr_array(warped), len(warped), w, h, _p(out, u16p))
    return out


def ref_blend_chunk_bands(warped, num_bands):
    warped = [np.ascontiguousarray(x, dtype=np.uint16) for x in warped]
    h, w = warped[0].shape[:2]
    out = np.zeros((h, w, 3), np.uint16)
    ref().ref_blend_chunk_bands(_ptr_array(warped), len(warped), w, h, num_bands, _p(out, u16p))
    return out


def ref_blend_owner(warped):
    warped = [np.ascontiguousarray(x, dtype=np.uint16) for x in warped]
    h, w = warped[0].shape[:2]
    owner = np.zeros((h, w), np.int32)
    ref().ref_blend_owner(_ptr_array(warped), len(warped), w, h, _p(owner, i32p))
    return owner


def ref_coverage_distance(cov):
    cov = np.ascontiguousarray(cov, dtype=np.uint8)
    h, w = cov.shape
    out = np.zeros((h, w), np.float32)
    ref().ref_coverage_distance(_p(cov, u8p), w, h, _p(out, f32p))
    return out


def ref_multiband_compose(warped, owner, num_bands):
    warped = [np.ascontiguousarray(x, dtype=np.uint16) for x in warped]
    h, w = warped[0].shape[:2]
    dists = [ref_coverage_distance((x != 0).any(axis=2).astype(np.uint8)) for x in warped]
    owner = np.ascontiguousarray(owner, dtype=np.int32)
    out = np.zeros((h, w, 3), np.uint16)
    ref().ref_multiband_compose(_ptr_array(warped), _ptr_array(dists, f32p), _p(owner, i32p), len(warped), w, h,
                                num_bands, _p(out, u16p))
    return out


# ------------------------------------------------------------- surrogate --
SURR_WEIGHTS = 42624
SURR_BIASES = 164


def surr_forward(w, b, feat):
    """Network output (n, h, w, 4) in binary64, rounded to float."""
    w = np.ascontiguousarray(w, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    feat = np.ascontiguousarray(feat, np.float32)
    n, h, wd, _ = feat.shape
    out = np.zeros((n, h, wd, 4), np.float32)
    lib().orc_surr_forward(_p(w, f32p), _p(b, f32p), _p(feat, f32p), n, h, wd, _p(out, f32p))
    return out


def surr_loss_grad(w, b, feat, target, le_in=None, want_le=False):
    """(loss, grad_w, grad_b[, le]) in binary64; le_in freezes the error
    term's target per pixel."""
    w = np.ascontiguousarray(w, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    feat = np.ascontiguousarray(feat, np.float32)
    target = np.ascontiguousarray(target, np.float32)
    n, h, wd, _ = feat.shape
    gw = np.zeros(SURR_WEIGHTS, np.float64)
    gb = np.zeros(SURR_BIASES, np.float64)
    loss = np.zeros(1, np.float64)
    le = np.zeros(n * h * wd, np.float64)
    lin = None if le_in is None else np.ascontiguousarray(le_in, np.float64)
    lib().orc_surr_loss_grad(_p(w, f32p), _p(b, f32p), _p(feat, f32p), _p(target, f32p), n, h, wd,
                             _p(gw, f64p), _p(gb, f64p), _p(loss, f64p),
                             None if lin is None else _p(lin, f64p), _p(le, f64p))
    return (float(loss[0]), gw, gb, le) if want_le else (float(loss[0]), gw, gb)
