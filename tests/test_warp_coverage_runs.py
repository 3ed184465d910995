"""The sheared-row numbering of a 64x16 tile's four-pixel runs
(shear_tile_px in image360_amd/csrc/render_float.hip), used by the warp's
coverage store for interior tiles and by the ownership kernel's distance
loads, emulated on the CPU: the 256 runs are the tile's pixels, each once,
for every strip width S the chamfer uses (pixel (x, y) at sheared row
y + (x >> log2 S), pitch a multiple of 16); and a wavefront's 64 runs touch
few 128-byte lines of a byte plane, where a per-pixel store touches one line
per column group."""
import math

import pytest

TW, TH = 64, 16


def shear_tile_px(t, shift):
    S, G = 1 << shift, TW >> shift
    q, k = t >> (shift - 2), t & ((S >> 2) - 1)

    def tri(v):
        d = int((math.sqrt(8.0 * v + 1.0) - 1.0) * 0.5)
        if (d + 1) * (d + 2) // 2 <= v:
            d += 1
        if d * (d + 1) // 2 > v:
            d -= 1
        return d

    T1, T3 = G * (G + 1) // 2, G * (G - 1) // 2
    if q < T1:
        d = tri(q)
        gl = q - d * (d + 1) // 2
    elif q < TH * G - T3:
        q2 = q - T1
        d = G + q2 // G
        gl = q2 - (d - G) * G
    else:
        q3 = TH * G - 1 - q
        dd = tri(q3)
        d = TH + G - 2 - dd
        gl = G - 1 - (q3 - dd * (dd + 1) // 2)
    return gl * S + 4 * k, d - gl


def shear_at(x, y, shift, pitch):
    return (y + (x >> shift)) * pitch + x


def shear_tile_dword(t, tx0, ty0, shift, pitch):
    dx, dy = shear_tile_px(t, shift)
    assert dx % 4 == 0 and 0 <= dx < TW and 0 <= dy < TH
    return shear_at(tx0 + dx, ty0 + dy, shift, pitch)


@pytest.mark.parametrize("shift", [2, 3, 4])
@pytest.mark.parametrize("tx0,ty0,w", [(0, 0, 64), (128, 32, 4096), (4032, 4080, 4096), (640, 16, 1000)])
def test_dwords_cover_the_tile_bytes_once(shift, tx0, ty0, w):
    pitch = (w + 15) & ~15
    want = sorted((y + (x >> shift)) * pitch + x for y in range(ty0, ty0 + TH) for x in range(tx0, tx0 + TW))
    got = []
    for t in range(256):
        o = shear_tile_dword(t, tx0, ty0, shift, pitch)
        assert o % 4 == 0
        got.extend(range(o, o + 4))
    assert sorted(got) == want


@pytest.mark.parametrize("shift", [2, 3, 4])
def test_wavefront_lines(shift):
    pitch, tx0, ty0 = 4096, 1024, 512
    for wave in range(4):
        lines = {shear_tile_dword(t, tx0, ty0, shift, pitch) // 128 for t in range(64 * wave, 64 * wave + 64)}
        per_pixel = {((y + (x >> shift)) * pitch + x) // 128
                     for y in range(ty0 + 4 * wave, ty0 + 4 * wave + 4) for x in range(tx0, tx0 + TW)}
        assert len(lines) <= 12
        assert len(lines) < len(per_pixel) or shift == 4
