"""Pins of the float chunk-render oracle (oracle/oracle_render.cpp), on the
CPU.  The float render is the reference's RGB48 render carried out on
[0,1] floats, so it is pinned to the RGB48 oracle (itself pinned to the
compiled reference by the golden vectors): on integer-translation warps
(exact samples) of sources with no zero pixel, the two renders must make
the same ownership decisions and agree to within the RGB48 rounding."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.render_cases import psnr, small_case


def _translation_case(seed, n=3, cw=150, ch=90):
    rng = np.random.default_rng(seed)
    srcs, hs = [], []
    for k in range(n):
        fw, fh = int(rng.integers(60, 110)), int(rng.integers(50, 80))
        img = rng.integers(1, 65536, size=(fh, fw, 3)).astype(np.uint16)
        # smooth it a little so seams are not pure noise
        img = np.clip(img.astype(np.float64) * 0.3 + 30000, 1, 65535).astype(np.uint16)
        tx, ty = int(rng.integers(-20, cw - 40)), int(rng.integers(-15, ch - 30))
        srcs.append(img)
        hs.append(np.array([[1, 0, tx], [0, 1, ty], [0, 0, 1.0]]))
    return srcs, hs, cw, ch


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_float_render_agrees_with_rgb48_render(seed):
    srcs, hs, cw, ch = _translation_case(seed)
    # RGB48 path: reference warp (binary32 inverse) + BlendChunkContributors
    warped = [po.warp_perspective(s, cw, ch, po.invert_homography(h.astype(np.float32))) for s, h in zip(srcs, hs)]
    out16, own16 = po.blend_chunk(warped, 5, want_owner=True)
    # float path on the same samples scaled to [0,1]
    fsrc = [s.astype(np.float32) / np.float32(65535) for s in srcs]
    hinv = np.stack([po.invert_homography_f64(h) for h in hs])
    outf, ownf = po.render_chunk(fsrc, hinv, np.ones(len(srcs), np.float32), 0, 0, cw, ch, 5, fp64=True,
                                 want_owner=True)
    assert np.array_equal(ownf, own16)
    owned = own16 >= 0
    diff = np.abs(outf.astype(np.float64) * 65535 - out16.astype(np.float64))
    # RGB48 output rounds to nearest (<= 0.5) and is computed in binary32
    assert diff[owned].max() <= 0.55
    assert not outf[~owned].any()


def test_single_full_contributor_reconstructs_input():
    # the Laplacian pyramid telescopes: one contributor owning every pixel
    # comes back unchanged (up to binary64 rounding)
    rng = np.random.default_rng(0)
    src = rng.random((70, 90, 3)).astype(np.float32)
    hinv = po.invert_homography_f64(np.array([[1, 0, -5.0], [0, 1, -3.0], [0, 0, 1]]))[None]
    out = po.render_chunk([src], hinv, np.ones(1, np.float32), 0, 0, 80, 60, 5, fp64=True)
    assert np.max(np.abs(out - src[3:63, 5:85])) < 1e-12


def test_constant_scene_stays_constant():
    # two sources that together own every pixel: the weights sum to 1
    # everywhere and the [1,4,6,4,1]/16 blur preserves constants exactly
    src = np.full((64, 64, 3), 0.4375, np.float32)
    hs = [np.array([[1, 0, 0.0], [0, 1, 0.0], [0, 0, 1]]), np.array([[1, 0, 40.0], [0, 1, 0.0], [0, 0, 1]])]
    hinv = np.stack([po.invert_homography_f64(h) for h in hs])
    out, own = po.render_chunk([src, src], hinv, np.ones(2, np.float32), 0, 0, 100, 60, 5, fp64=True,
                               want_owner=True)
    assert (own >= 0).all()
    assert np.max(np.abs(out - 0.4375)) < 1e-12


def test_binary32_oracle_close_to_binary64():
    g, frames = small_case(seed=3)
    c = g.chunks[0]
    srcs = [frames[i].astype(np.float32) for i in c.contributors]
    h64 = np.stack([po.invert_homography_f64(g.H[i]) for i in c.contributors])
    h32 = np.stack([po.invert_homography(g.H[i].astype(np.float32)) for i in c.contributors])
    o64, w64 = po.render_chunk(srcs, h64, g.gains[c.contributors], c.x, c.y, c.w, c.h, 5, True, want_owner=True)
    o32, w32 = po.render_chunk(srcs, h32, g.gains[c.contributors], c.x, c.y, c.w, c.h, 5, False, owner_in=w64,
                               want_owner=True)
    assert psnr(o32, o64) > 80
    assert np.max(np.abs(o32 - o64)) < 1e-3


# ---- the float warp's fractional path (oracle_render.cpp WarpF) ---------
# The render tests compare the GPU with this oracle, so its bilinear
# weights, perspective divide at non-integer positions, validity rule and
# gain order are pinned here: against the reference's own RGB48 warp
# (golden vectors written by the compiled reference), a closed form and the
# RGB48 oracle's gain, each shown to fail on a planted mistake.

def _warp_pin_residual(golden, fp64):
    """Largest |trunc(float warp * 65535) - reference RGB48 warp| over every
    golden warp case, with the float warp run on the sources scaled to
    [0,1]; also the number of pixels the two disagree on covering."""
    g = golden("warp")
    cases = [(g["smoke_src"], g["smoke_h"], (0, 0), g["smoke_out"])]
    cases += [(g["rand_src"], g["rand_h"][k], tuple(g["rand_off"][k]), g["rand_out"][k]) for k in range(6)]
    worst, cover_diff, covered = 0, 0, 0
    for src16, hm, (ox, oy), ref in cases:
        src = src16.astype(np.float32) / np.float32(65535)
        hinv = hm.astype(np.float64) if fp64 else hm
        h, w = ref.shape[:2]
        rgb, cov = po.warp_float(src, hinv, 1.0, int(ox), int(oy), w, h, fp64=fp64)
        got = np.floor(rgb.astype(np.float64) * 65535.0)
        on = cov.astype(bool)
        covered += int(on.sum())
        # the reference writes (0,0,0) where it has no sample
        cover_diff += int((~on & ref.any(axis=2)).sum())
        if on.any():
            worst = max(worst, int(np.abs(got[on] - ref[on].astype(np.float64)).max()))
    return worst, cover_diff, covered


@pytest.mark.parametrize("fp64", [False, True])
def test_float_warp_matches_reference_rgb48_warp(golden, fp64):
    # rotated, scaled and perspective homographies (one with its singular
    # line across the output), offsets, fractional samples everywhere
    worst, cover_diff, covered = _warp_pin_residual(golden, fp64)
    assert covered > 20000
    assert cover_diff == 0
    # binary32: the reference's own coordinate arithmetic, so only the value
    # scaling and its truncation differ; binary64 coordinates move a sample
    # by ~1e-5 px, worth < 2 LSB at full-range texel differences
    assert worst <= (2 if fp64 else 1), worst


def test_float_warp_reference_pin_fails_on_swapped_weights(golden):
    with po.planted(1):
        worst, _, _ = _warp_pin_residual(golden, fp64=False)
    assert worst > 1000


def _ramp_case(gain=1.0):
    sw, sh = 97, 61
    y, x = np.mgrid[0:sh, 0:sw].astype(np.float64)
    a = np.array([0.05, 0.2, 0.1])
    b = np.array([0.8, 0.1, -0.3]) / sw
    d = np.array([0.07, 0.6, 0.5]) / sh
    src = (a + x[..., None] * b + y[..., None] * d).astype(np.float32)
    fwd = np.array([[0.93, -0.21, 31.4], [0.17, 1.08, 12.7], [4.1e-4, -2.7e-4, 1.0]])
    hinv = np.linalg.inv(fwd)
    x0, y0, w, h = 7, -3, 150, 110
    X, Y = np.meshgrid(np.arange(x0, x0 + w, dtype=np.float64), np.arange(y0, y0 + h, dtype=np.float64))
    m = hinv.reshape(9)
    q = (m[6] * X + m[7] * Y) + m[8]
    u = ((m[0] * X + m[1] * Y) + m[2]) / q
    v = ((m[3] * X + m[4] * Y) + m[5]) / q
    fu, fv = np.floor(u), np.floor(v)
    valid = (np.abs(q) >= 1e-10) & (fu >= 0) & (fv >= 0) & (fu <= sw - 2) & (fv <= sh - 2)
    want = np.clip((a + u[..., None] * b + v[..., None] * d) * gain, 0.0, 1.0)
    return src, hinv, (x0, y0, w, h), valid, want


def _ramp_residual(gain=1.0):
    src, hinv, (x0, y0, w, h), valid, want = _ramp_case(gain)
    rgb, cov = po.warp_float(src, hinv, gain, x0, y0, w, h, fp64=True)
    assert np.array_equal(cov.astype(bool), valid)
    assert valid.mean() > 0.3
    return float(np.abs(rgb - want)[valid].max())


def test_float_warp_samples_an_affine_ramp_exactly():
    # bilinear interpolation reproduces an affine function exactly, so the
    # warp of a ramp image is the ramp at the projective source position:
    # this pins the weights, the perspective divide at non-integer
    # positions and the validity rule (up to the float32 storage of the
    # ramp, ~3e-8); with a gain that saturates, also that the gain comes
    # before the clamp
    assert _ramp_residual() < 1e-7
    assert _ramp_residual(gain=1.7) < 2e-7


def test_float_warp_ramp_pin_fails_on_planted_mistakes():
    with po.planted(1):
        assert _ramp_residual() > 1e-3
    with po.planted(2):
        assert _ramp_residual(gain=1.7) > 0.1


def _gain_residual(gain):
    rng = np.random.default_rng(17)
    src16 = rng.integers(0, 65536, size=(53, 71, 3)).astype(np.uint16)
    hm = np.array([0.97, -0.12, 9.25, 0.11, 1.02, -4.5, 3e-4, -2e-4, 1.0], np.float32)
    w, h = 80, 60
    ref = po.apply_gain(po.warp_perspective(src16, w, h, hm, 0, 0), gain)
    rgb, cov = po.warp_float(src16.astype(np.float32) / np.float32(65535), hm, gain, 0, 0, w, h, fp64=False)
    on = cov.astype(bool)
    got = np.floor(rgb.astype(np.float64) * 65535.0)
    return int(np.abs(got[on] - ref[on].astype(np.float64)).max()), float(rgb.max())


@pytest.mark.parametrize("gain", [0.61, 1.3, 1.9])
def test_float_warp_gain_matches_rgb48_apply_gain(gain):
    # the RGB48 chain truncates twice (warp, then gain): at most
    # ceil(gain) + 1 LSB apart; gain 1.9 saturates most pixels, which both
    # chains clamp to full scale
    worst, top = _gain_residual(gain)
    assert worst <= int(np.ceil(gain)) + 1, worst
    assert top <= 1.0


def test_float_warp_gain_pin_fails_on_gain_after_clamp():
    with po.planted(2):
        worst, top = _gain_residual(1.9)
    assert top > 1.0 and worst > 1000
