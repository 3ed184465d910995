"""GPU SeamBlend::BlendChunkContributors against the reference's own outputs
(golden vectors) and the oracle: bit-exact owner maps and RGB48 output."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.gpu_util import note

pytestmark = pytest.mark.gpu


def _blend(hip, bufs, bands=5):
    h, w = bufs[0].shape[:2]
    out = np.zeros((h, w, 3), np.uint16)
    owner = np.zeros((h, w), np.int32)
    hip.check(hip.BlendChunkContributors([np.ascontiguousarray(b) for b in bufs], w, h, out, bands, owner), "blend")
    return out, owner


@pytest.mark.parametrize("scene", ["ghost", "gap", "route", "c0", "c1", "c2", "c3"])
def test_blend_golden(hip, golden, scene):
    g = golden("blend")
    note(f"blend golden {scene}")
    out, owner = _blend(hip, list(g[f"{scene}_in"]))
    assert np.array_equal(owner, g[f"{scene}_owner"])
    assert np.array_equal(out, g[f"{scene}_out"])


def test_blend_seven_bands(hip, golden):
    g = golden("compose7")
    note("blend 7 bands")
    out, owner = _blend(hip, list(g["inp"]), 7)
    assert np.array_equal(owner, g["owner"])
    assert np.array_equal(out, g["out"])


def test_blend_empty_and_single(hip):
    note("blend c=0 / c=1")
    out = np.ones((8, 9, 3), np.uint16)
    assert hip.BlendChunkContributors([], 9, 8, out) == 0
    assert not out.any()
    img = np.random.Generator(np.random.PCG64(2)).integers(1, 60000, size=(33, 47, 3)).astype(np.uint16)
    o, own = _blend(hip, [img])
    assert np.array_equal(o, po.blend_chunk([img]))


def test_blend_render_behaviour(hip, golden):
    # tests/render_blend/main.cpp:113-115, 145-153, 193 on the GPU output
    g = golden("blend")
    out, _ = _blend(hip, list(g["ghost_in"]))
    obj = np.array([60000, 2000, 2000])
    assert np.abs(out[20, 165].astype(int) - obj).sum() < 3000
    assert np.abs(out[20, 235].astype(int) - obj).sum() < 3000
    out, _ = _blend(hip, list(g["gap_in"]))
    assert not (out == 0).all(axis=2).any()


@pytest.mark.parametrize("seed,cw,ch,c", [(1, 257, 190, 3), (2, 190, 300, 4), (3, 400, 64, 2), (4, 128, 128, 6)])
def test_blend_random_vs_oracle(hip, seed, cw, ch, c):
    rng = np.random.Generator(np.random.PCG64(seed))
    bufs = []
    for i in range(c):
        sw, sh = int(rng.integers(cw // 2, cw)), int(rng.integers(ch // 2, ch))
        img = np.clip(rng.normal(30000, 6000, size=(sh, sw, 3)), 1, 65535).astype(np.uint16)
        th = rng.uniform(-0.1, 0.1)
        fwd = np.array([[np.cos(th), -np.sin(th), rng.uniform(-sw / 3, cw - sw / 2)],
                        [np.sin(th), np.cos(th), rng.uniform(-sh / 3, ch - sh / 2)], [0, 0, 1]])
        inv = np.linalg.inv(fwd).astype(np.float32)
        bufs.append(po.warp_perspective(img, cw, ch, inv))
    note(f"blend random seed={seed} {cw}x{ch} c={c}")
    want, want_owner = po.blend_chunk(bufs, 5, want_owner=True)
    out, owner = _blend(hip, bufs)
    assert np.array_equal(owner, want_owner)
    assert np.array_equal(out, want)


@pytest.mark.parametrize("strip", ["4", "8", "16"])
@pytest.mark.parametrize("cw", [260, 268, 276])
def test_blend_ragged_width_full_rows(hip, monkeypatch, strip, cw):
    """A contributor covering whole rows of a chunk whose width is not a
    multiple of the chamfer strip: its last strip ends past the last column
    and must not touch the next row, and the backward pass must read that
    strip's own columns.  A rival covering the left part only loses every
    pixel the full contributor covers (its distance is smaller)."""
    monkeypatch.setenv("I360HIP_CHAMFER_STRIP", strip)
    ch = 37
    rng = np.random.Generator(np.random.PCG64(cw))
    full = np.clip(rng.normal(30000, 5000, size=(ch, cw, 3)), 1, 65535).astype(np.uint16)
    left = np.clip(rng.normal(20000, 5000, size=(ch, cw, 3)), 1, 65535).astype(np.uint16)
    left[:, 200:] = 0
    bufs = [left, full]
    note(f"blend ragged strip={strip} {cw}x{ch}")
    want, want_owner = po.blend_chunk(bufs, 4, want_owner=True)
    out, owner = _blend(hip, bufs, bands=4)
    assert np.array_equal(owner, want_owner)
    assert np.array_equal(out, want)


@pytest.mark.parametrize("hole", [(0, 0), (100, 120), (64, 64), (190, 70), (130, 200)])
def test_blend_seam_chain_with_uncontested_gap(hip, hole):
    """A 400-line seam (overlap 500 x 400, scanned row by row) whose chain of
    contested lines skips a run of uncontested rows: the second contributor
    has a hole there, so those rows hold no crossover.  The DP's backtrack
    walks 64-line batches and reloads when the chain jumps over a whole
    batch (a hole of 64 rows or more); bands are clipped near the hole's
    rims.  Owners and output equal the oracle's bit for bit."""
    cw, ch = 800, 400
    rng = np.random.Generator(np.random.PCG64(7 + hole[0]))
    a = np.clip(rng.normal(30000, 6000, size=(ch, cw, 3)), 1, 65535).astype(np.uint16)
    b = np.clip(rng.normal(26000, 6000, size=(ch, cw, 3)), 1, 65535).astype(np.uint16)
    a[:, 600:] = 0
    b[:, :100] = 0
    y0, n = hole
    if n:
        b[y0:y0 + n, :] = 0
    bufs = [a, b]
    note(f"blend seam gap hole={hole} {cw}x{ch}")
    want, want_owner = po.blend_chunk(bufs, 5, want_owner=True)
    out, owner = _blend(hip, bufs)
    assert np.array_equal(owner, want_owner)
    assert np.array_equal(out, want)
