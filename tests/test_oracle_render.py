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
