"""The oracle restatement against the golden vectors written by the
reference's own compiled sources (tests/golden/gen_golden.py)."""
import numpy as np
import pytest

from oracle import pyoracle as po


def test_median_smoke_shape(golden):
    g = golden("median")
    assert np.array_equal(po.median_stack(list(g["smoke_in"])), g["smoke_out"])


def test_median_burst(golden):
    g = golden("median")
    assert np.array_equal(po.median_stack(list(g["burst_in"])), g["burst_out"])


def test_median_every_count(golden):
    g = golden("median")
    allin, allout = g["all_in"], g["all_out"]
    for n in range(1, 33):
        for k, sigma in enumerate((2.0, 1.0, 0.5)):
            assert np.array_equal(po.median_stack(list(allin[:n]), sigma), allout[n - 1, k]), (n, sigma)


def test_warp(golden):
    g = golden("warp")
    assert np.array_equal(po.warp_perspective(g["smoke_src"], 40, 30, g["smoke_h"]), g["smoke_out"])
    for k in range(len(g["rand_h"])):
        ox, oy = (int(v) for v in g["rand_off"][k])
        out = po.warp_perspective(g["rand_src"], 101, 67, g["rand_h"][k], ox, oy)
        assert np.array_equal(out, g["rand_out"][k]), k


@pytest.mark.parametrize("layout,filters", [("RGGB", 0x94949494), ("BGGR", 0x16161616),
                                            ("GRBG", 0x61616161), ("GBRG", 0x49494949)])
def test_demosaic_layouts(golden, layout, filters):
    g = golden("demosaic")
    cam = np.array([1.8, 1.0, 1.5, 1.0], np.float32)
    rc = np.array([[1.9, -0.7, -0.2, 0.0], [-0.3, 1.5, -0.2, 0.0], [0.0, -0.4, 1.6, 0.0]], np.float32)
    assert np.array_equal(po.demosaic(g["smoke_cfa"], 64, cam, rc, filters), g[f"smoke_out_{layout}"])


def test_demosaic_random_and_tiny(golden):
    g = golden("demosaic")
    assert np.array_equal(po.demosaic(g["rand_cfa"], 512, g["rand_cam"], g["rand_rgbcam"], 0x94949494), g["rand_out"])
    cam = np.array([1.8, 1.0, 1.5, 1.0], np.float32)
    rc = np.array([[1.9, -0.7, -0.2, 0.0], [-0.3, 1.5, -0.2, 0.0], [0.0, -0.4, 1.6, 0.0]], np.float32)
    for key in g.files:
        if key.startswith("tiny_cfa_"):
            sz = key[len("tiny_cfa_"):]
            assert np.array_equal(po.demosaic(g[key], 64, cam, rc, 0x94949494), g["tiny_out_" + sz]), sz


def test_demosaic_f32_truncates_to_u16(golden):
    g = golden("demosaic")
    f = po.demosaic(g["rand_cfa"], 512, g["rand_cam"], g["rand_rgbcam"], 0x94949494, f32=True)
    assert np.array_equal(f.astype(np.uint16), g["rand_out"])


def test_gain(golden):
    g = golden("gain")
    for k, gain in enumerate(g["gains"]):
        assert np.array_equal(po.apply_gain(g["in"], float(gain)), g["out"][k]), gain


def test_invert_homography(golden):
    g = golden("homography")
    for k in range(len(g["h"])):
        r = po.invert_homography(g["h"][k])
        if g["ok"][k]:
            assert r is not None and np.array_equal(r, g["inv"][k])
        else:
            assert r is None


@pytest.mark.parametrize("scene", ["ghost", "gap", "route", "c0", "c1", "c2", "c3"])
def test_blend_chunk(golden, scene):
    g = golden("blend")
    bufs = list(g[f"{scene}_in"])
    out, owner = po.blend_chunk(bufs, 5, want_owner=True)
    assert np.array_equal(owner, g[f"{scene}_owner"])
    assert np.array_equal(out, g[f"{scene}_out"])
    for i, b in enumerate(bufs):
        d = po.coverage_distance((b != 0).any(axis=2).astype(np.uint8))
        assert np.array_equal(d, g[f"{scene}_dist"][i])


def test_compose_seven_bands(golden):
    g = golden("compose7")
    assert np.array_equal(po.multiband_compose(list(g["inp"]), g["owner"], 7), g["out"])
