"""The oracle (and so the GPU, bit-exact against it) against every SIMD tier
of the reference's CPU kernels, compiled from the reference's sources
(oracle/_ref).

The reference defines each kernel by its Scalar tier: the AVX2 and AVX-512
tiers vectorise the same binary32 operations in the same order.  Built by
MSVC (the reference's Windows build, /fp:precise) they give the same bits.
Built by GCC with -mfma, GCC's default -ffp-contract=fast fuses a*b+c into
one FMA in those units (MedianStack's sum of squares and clip bounds, the
demosaic's colour matrix), so a few values differ there.  The same sources
built with -ffp-contract=off (libi360ref_nc.so) match the oracle on every
tier: the median and demosaic parity is with the reference's arithmetic,
not with one compiler's contraction choices."""
import numpy as np
import pytest

from image360_amd import synth
from oracle import pyoracle as po

pytestmark = pytest.mark.skipif(po.ref() is None or po.ref_nc() is None,
                                reason="the reference sources were not compiled (make -C oracle ref)")


@pytest.fixture(scope="module")
def burst_planes():
    frames, shifts = synth.burst_frames(w=1200, h=900, device="cpu")
    w, h = 1200, 900
    n = (h - 9) * w
    hf = [f.numpy().view(np.uint16).reshape(-1) for f in frames]
    return [f[(8 - dy) * w + (8 - dx):(8 - dy) * w + (8 - dx) + n].copy() for f, (dx, dy) in zip(hf, shifts)]


def test_median_every_tier_without_contraction(burst_planes):
    want = po.median_stack(burst_planes, 2.0)
    for tier in (0, 1, 2):
        got = po.ref_median_stack(burst_planes, 2.0, tier=tier, lib=po.ref_nc())
        assert np.array_equal(got, want), tier
    assert np.array_equal(po.ref_median_stack(burst_planes, 2.0, tier=0), want)
    fused = po.ref_median_stack(burst_planes, 2.0, tier=2)
    print(f"GCC-contracted AVX-512 median: {int((fused != want).sum())} of {want.size} values differ")


def test_demosaic_every_tier_without_contraction():
    g, cfas = synth.small_bayer()
    args = (synth.BAYER_BLACK, synth.BAYER_CAM_MUL, synth.BAYER_RGB_CAM, synth.BAYER_RGGB)
    for cfa in cfas[:2]:
        want = po.demosaic(cfa, *args)
        for tier in (0, 1, 2):
            assert np.array_equal(po.ref_demosaic(cfa, *args, tier=tier, lib=po.ref_nc()), want), tier
        assert np.array_equal(po.ref_demosaic(cfa, *args, tier=0), want)
        fused = po.ref_demosaic(cfa, *args, tier=2)
        print(f"GCC-contracted AVX-512 demosaic: {int((fused != want).sum())} of {want.size} values differ")
