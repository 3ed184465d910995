"""Pins of the export oracle (oracle/oracle_export.cpp) and of the host-side
arithmetic the C ABI exports for it.

The reference's own unit tests (PanoramaExporterTests.cpp:16-66) fix the
scale and the placement.  The composite has no golden vectors in the
reference, so it is pinned by closed forms (scale 1 copies the chunks >> 8,
scale 1/2 takes every second pixel, other scales are checked against exact
rational division rounded to binary32) and by the overwrite rules of the
reference's in-order loop (PanoramaExporter.cpp:53-98)."""
from fractions import Fraction

import numpy as np
import pytest

from oracle import pyoracle as po


# -------------------------------------------------- reference unit tests --
def test_scale_shrinks_to_fit_longest_edge():
    # PanoramaExporterTests.cpp:16-23
    assert abs(po.export_scale(8192, 4096, 2048) - 0.25) < 1e-5


def test_scale_handles_taller_than_wide():
    # PanoramaExporterTests.cpp:25-32
    assert abs(po.export_scale(2048, 4096, 1024) - 0.25) < 1e-5


def test_placement_scales_offset_and_size():
    # PanoramaExporterTests.cpp:34-50
    assert po.export_placement(4096, 8192, 4096, 4096, 0.25) == (1024, 2048, 1024, 1024)


def test_placement_never_zero_sized():
    # PanoramaExporterTests.cpp:52-66
    x, y, w, h = po.export_placement(0, 0, 100, 100, 0.001)
    assert w >= 1 and h >= 1


def test_size_rejects_what_the_reference_rejects():
    # PanoramaExporter.cpp:36, 47
    assert po.export_size(100, 50, 0) is None
    assert po.export_size(0, 50, 10) is None
    assert po.export_size(100, -1, 10) is None
    assert po.export_size(100, 50, 10) == (10, 5)
    assert po.export_size(1000, 1, 10) == (10, 1)  # clamped to 1


# ------------------------------------------------------ helpers (pins) --
def _f32_nearest(q: Fraction) -> np.float32:
    """The binary32 value nearest to the rational q (ties to even), found
    exactly from the candidates around float(q)."""
    f = np.float32(float(q))
    cands = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
    best = min(cands, key=lambda c: (abs(Fraction(float(c)) - q), int(np.float32(c).view(np.uint32)) & 1))
    return np.float32(best)


def _index_chunk(w, h, cid=0):
    """RGB48 buffer whose pixel (x, y) reads back, after >> 8, as (x, y, cid)."""
    buf = np.zeros((h, w, 3), np.uint16)
    buf[..., 0] = (np.arange(w, dtype=np.uint16) << 8)[None, :] | 0x55
    buf[..., 1] = (np.arange(h, dtype=np.uint16) << 8)[:, None] | 0xAA
    buf[..., 2] = cid << 8
    return buf


# ---------------------------------------------------------- closed forms --
def test_scale_one_is_a_copy_shifted_by_8():
    rng = np.random.default_rng(1)
    tw, th = 96, 64
    tiles = [(x, y, 32, 32) for y in (0, 32) for x in (0, 32, 64)]
    bufs = [rng.integers(0, 65536, (32, 32, 3), dtype=np.uint16) for _ in tiles]
    out = po.export_composite([t + (b,) for t, b in zip(tiles, bufs)], tw, th, max(tw, th))
    want = np.zeros((th, tw, 3), np.uint8)
    for (x, y, w, h), b in zip(tiles, bufs):
        want[y:y + h, x:x + w] = (b >> 8).astype(np.uint8)
    assert np.array_equal(out, want)


def test_scale_half_takes_every_second_pixel():
    rng = np.random.default_rng(2)
    tw, th = 128, 90  # longest 128 -> max 64 gives scale 0.5 exactly
    tiles = [(0, 0, 64, 90), (64, 0, 64, 90)]
    bufs = [rng.integers(0, 65536, (90, 64, 3), dtype=np.uint16) for _ in tiles]
    out = po.export_composite([t + (b,) for t, b in zip(tiles, bufs)], tw, th, 64)
    assert out.shape == (45, 64, 3)
    want = np.zeros_like(out)
    for (x, y, w, h), b in zip(tiles, bufs):
        want[y // 2:y // 2 + h // 2, x // 2:x // 2 + w // 2] = (b[::2, ::2] >> 8).astype(np.uint8)
    assert np.array_equal(out, want)


@pytest.mark.parametrize("total,max_dim", [(300, 100), (999, 333), (250, 77), (255, 254), (7, 3), (169, 91),
                                           (225, 135), (120, 36), (50, 36)])
def test_sampling_is_truncated_binary32_division(total, max_dim):
    """srcX = (int)(dx / scale) with a correctly rounded binary32 quotient:
    derived here from exact rational arithmetic, independent of the C
    code.  A reciprocal multiply or a rounding instead of a truncation
    fails some of these scales."""
    scale = np.float32(max_dim) / np.float32(total)  # binary32 division, as the reference
    assert np.float32(po.export_scale(total, total, max_dim)) == scale
    buf = _index_chunk(min(total, 255), min(total, 255))
    out = po.export_composite([(0, 0, total, total, buf)], total, total, max_dim)
    for d in range(out.shape[1]):
        sx = _exact_src(d, scale)
        if sx < buf.shape[1]:
            assert out[0, d, 0] == sx, (d, sx)   # column d samples source column sx
            assert out[d, 0, 1] == sx, (d, sx)   # row d samples source row sx
        else:
            assert out[0, d].tolist() == [0, 0, 0], d


def _exact_src(d, scale):
    return int(_f32_nearest(Fraction(float(np.float32(d))) / Fraction(float(scale))))


def test_sampling_check_has_teeth():
    """A reciprocal multiply, the usual shortcut for the division, picks
    other source pixels than the exact derivation above at these scales."""
    for total, max_dim in [(169, 91), (225, 135), (120, 36), (50, 36)]:
        scale = np.float32(max_dim) / np.float32(total)
        recip = np.float32(1) / scale
        assert any(int(np.float32(d) * recip) != _exact_src(d, scale) for d in range(max_dim)), total


# ----------------------------------------------------- overwrite rules --
def test_later_chunk_wins_where_placements_overlap():
    a = _index_chunk(40, 40, cid=1)
    b = _index_chunk(40, 40, cid=2)
    out = po.export_composite([(0, 0, 40, 40, a), (20, 10, 40, 40, b)], 60, 50, 60)
    assert (out[10:, 20:, 2] == 2).all()
    assert (out[:10, :40, 2] == 1).all() and (out[:40, :20, 2] == 1).all()
    out2 = po.export_composite([(20, 10, 40, 40, b), (0, 0, 40, 40, a)], 60, 50, 60)
    assert (out2[:40, :40, 2] == 1).all()


def test_source_outside_buffer_leaves_earlier_value():
    a = _index_chunk(40, 40, cid=1)
    small = _index_chunk(16, 24, cid=2)  # buffer smaller than its 40x40 chunk
    out = po.export_composite([(0, 0, 40, 40, a), (0, 0, 40, 40, small)], 40, 40, 40)
    assert (out[:24, :16, 2] == 2).all()
    assert (out[24:, :, 2] == 1).all() and (out[:, 16:, 2] == 1).all()


def test_missing_outputs_and_off_canvas_chunks_write_nothing():
    a = _index_chunk(20, 20, cid=3)
    out = po.export_composite([(0, 0, 20, 20, None), (100, 100, 20, 20, a), (-30, -30, 20, 20, a)], 40, 40, 40)
    assert not out.any()


def test_negative_offsets_clip():
    a = _index_chunk(20, 20, cid=3)
    out = po.export_composite([(-5, -7, 20, 20, a)], 40, 40, 40)
    # destination (0, 0) is dx = 5, dy = 7 of the chunk
    assert out[0, 0].tolist() == [5, 7, 3]
    assert (out[13:, :] == 0).all() and (out[:, 15:] == 0).all()


# ------------------------------------------------------- RGB48 rounding --
def test_rgb48_rounding_hand_values():
    v = np.array([0.0, 1.0, 0.5, -0.1, 1.5, 1.0 / 65535, 0.49 / 65535, 0.51 / 65535, np.nan, np.inf, -np.inf],
                 np.float32)
    want = [0, 65535, 32768, 0, 65535, 1, 0, 1, 65535, 65535, 0]
    assert po.rgb48_from_render(v).tolist() == want


def test_float_composite_rounds_then_shifts():
    rng = np.random.default_rng(3)
    f = rng.random((24, 32, 3), dtype=np.float32)
    out = po.export_composite([(0, 0, 32, 24, f)], 32, 24, 32)
    assert np.array_equal(out, (po.rgb48_from_render(f) >> 8).astype(np.uint8))


# ------------------------------------- boundary host arithmetic (no GPU) --
def test_boundary_host_arithmetic_matches_oracle():
    from image360_amd.backend import HipComputeBackend as B

    rng = np.random.default_rng(4)
    for _ in range(2000):
        tw, th = (int(v) for v in rng.integers(1, 200000, 2))
        md = int(rng.integers(1, 20000))
        assert B.export_scale(tw, th, md) == po.export_scale(tw, th, md)
        assert B.export_size(tw, th, md) == po.export_size(tw, th, md)
        s = np.float32(po.export_scale(tw, th, md))
        x, y = (int(v) for v in rng.integers(-5000, 200000, 2))
        w, h = (int(v) for v in rng.integers(1, 9000, 2))
        assert B.export_placement(x, y, w, h, float(s)) == po.export_placement(x, y, w, h, float(s))
    assert B.export_size(10, 10, 0) is None and B.export_size(0, 10, 5) is None
