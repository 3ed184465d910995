"""The render workloads at their full sizes on the GPU, against the oracle.

* Small plain method: 4 seeded 14-bit RGGB raw frames of 1024x768 ->
  demosaic (bit-exact) -> RGB48 frames resident on the GPU -> the float
  render of one 2400x1100 chunk at 5 bands (FP64 form within 1e-9 of the FP64
  oracle; plain form within 2e-6 of the binary32 oracle; precision and
  batched forms within 1e-3 and >= 50 dB of the FP64 oracle, same owner
  maps), and the reference's RGB48 chain on the same frames
  (RenderExecutor.cpp:70-140: binary32 inverse, WarpPerspective, ApplyGain,
  BlendChunkContributors) bit-exact.
* Gigapixel panorama: 48 FP16 frames of 6000x4000, 50 chunks of <= 4096^2,
  7 bands.  The 9-contributor chunk in the precision and batched forms
  against the CPU FP64 oracle; the whole canvas, batched (FP32 and RGB48
  outputs), against the GPU's FP64 form: identical owner maps, <= 1e-3
  everywhere, >= 50 dB.
* 512 uneven chunks (edges log-uniform in 256-4096, 1-6 contributors, 6
  bands): batched equal to the per-chunk precision loop bit for bit on every
  chunk, and sampled chunks against the CPU FP64 oracle.
"""
import math

import numpy as np
import pytest

from image360_amd import _abi, render, synth
from oracle import pyoracle as po
from tests.gpu_util import note
from tests.render_cases import psnr

pytestmark = pytest.mark.gpu


def _descs(geom, frames, fmt):
    return [{"ptr": f.data_ptr(), "w": int(geom.frame_w[i]), "h": int(geom.frame_h[i]), "fmt": fmt,
             "gain": float(geom.gains[i]), "H": geom.H[i]} for i, f in enumerate(frames)]


def _render(hip, geom, fdesc, chunks, bands, mode, owners=True):
    """Renders `chunks` in `mode`; returns per-chunk device outputs and owner maps."""
    import torch

    rgb48 = bool(mode & _abi.RENDER_OUT_RGB48)
    outs, owns, descs, contrib = [], [], [], []
    for c in chunks:
        o = torch.empty((c.h, c.w, 3), dtype=torch.int16 if rgb48 else torch.float32, device="cuda")
        w = torch.full((c.h, c.w), -7, dtype=torch.int32, device="cuda") if owners else None
        descs.append({"x": c.x, "y": c.y, "w": c.w, "h": c.h, "first": len(contrib), "count": len(c.contributors),
                      "out": o.data_ptr(), "owner": w.data_ptr() if owners else None})
        contrib += c.contributors
        outs.append(o)
        owns.append(w)
    hip.check(hip.render_chunks(fdesc, descs, contrib, bands, mode), f"render_chunks mode {mode}")
    torch.cuda.synchronize()
    return outs, owns


def _oracle(geom, host_frames, chunk, bands, fp64=True):
    srcs = [host_frames[i] for i in chunk.contributors]
    if fp64:
        hinv = np.stack([po.invert_homography_f64(geom.H[i]) for i in chunk.contributors])
    else:
        hinv = np.stack([po.invert_homography(geom.H[i].astype(np.float32)) for i in chunk.contributors])
    return po.render_chunk(srcs, hinv, geom.gains[chunk.contributors], chunk.x, chunk.y, chunk.w, chunk.h, bands,
                           fp64=fp64, want_owner=True)


# ------------------------------------------------------- small plain method --
@pytest.fixture(scope="module")
def small(hip):
    import torch

    geom, cfas = synth.small_bayer()
    note("small: demosaic of 4 raw frames 1024x768")
    rgb_dev, rgb_host = [], []
    for cfa in cfas:
        h, w = cfa.shape
        d_cfa = torch.from_numpy(cfa.view(np.int16)).cuda()
        out = torch.empty((h, w, 3), dtype=torch.int16, device="cuda")
        hip.check(hip.demosaic_dev(d_cfa.data_ptr(), w, h, synth.BAYER_BLACK, synth.BAYER_CAM_MUL,
                                   synth.BAYER_RGB_CAM, synth.BAYER_RGGB, out.data_ptr()), "demosaic_dev")
        torch.cuda.synchronize()
        got = out.cpu().numpy().view(np.uint16)
        want = po.demosaic(cfa, synth.BAYER_BLACK, synth.BAYER_CAM_MUL, synth.BAYER_RGB_CAM, synth.BAYER_RGGB)
        assert np.array_equal(got, want), "demosaic differs from the oracle"
        rgb_dev.append(out)
        rgb_host.append(got)
    # the float render reads an RGB48 value v as v * (1/65535) in binary32
    host_f = [r.astype(np.float32) * np.float32(1.0 / 65535.0) for r in rgb_host]
    return geom, rgb_dev, rgb_host, host_f


def test_small_demosaic_is_bit_exact(small):
    geom, rgb_dev, rgb_host, _ = small
    assert len(rgb_host) == 4 and all(r.shape == (768, 1024, 3) for r in rgb_host)
    assert max(int(r.max()) for r in rgb_host) > 20000  # the white balance lifts the 14-bit range


def test_small_render_fp64_and_plain_forms(hip, small):
    geom, rgb_dev, _, host_f = small
    fdesc = _descs(geom, rgb_dev, _abi.PIX_U16)
    c = geom.chunks[0]
    assert (c.w, c.h) == (2400, 1100) and len(c.contributors) == 4
    note("small: fp64 form")
    outs, owns = _render(hip, geom, fdesc, [c], 5, _abi.RENDER_FP64)
    want, want_own = _oracle(geom, host_f, c, 5, fp64=True)
    assert np.array_equal(owns[0].cpu().numpy(), want_own)
    assert np.abs(outs[0].cpu().numpy() - want).max() <= 1e-9
    note("small: plain form")
    outs, owns = _render(hip, geom, fdesc, [c], 5, _abi.RENDER_PLAIN)
    want32, own32 = _oracle(geom, host_f, c, 5, fp64=False)
    assert np.array_equal(owns[0].cpu().numpy(), own32)
    assert np.abs(outs[0].cpu().numpy() - want32).max() <= 2e-6


@pytest.mark.parametrize("mode", [_abi.RENDER_PRECISION, _abi.RENDER_BATCHED])
def test_small_render_precision_forms_against_fp64(hip, small, mode):
    geom, rgb_dev, _, host_f = small
    fdesc = _descs(geom, rgb_dev, _abi.PIX_U16)
    c = geom.chunks[0]
    note(f"small: mode {mode}")
    outs, owns = _render(hip, geom, fdesc, [c], 5, mode)
    want, want_own = _oracle(geom, host_f, c, 5, fp64=True)
    got = outs[0].cpu().numpy()
    assert np.array_equal(owns[0].cpu().numpy(), want_own)
    err = float(np.abs(got - want).max())
    db = psnr(got, want)
    note(f"small mode {mode}: max abs {err:.2e}, {db:.1f} dB")
    assert err <= 1e-3 and db >= 50.0


def test_small_rgb48_render_chain_is_bit_exact(hip, small):
    # RenderExecutor's own chain on the device-resident demosaiced frames:
    # binary32 inverse, warp, gain where |gain - 1| > 1e-6, blend
    import torch

    geom, rgb_dev, rgb_host, _ = small
    geom = synth.Geometry(geom.frame_w, geom.frame_h, geom.H, geom.canvas_w, geom.canvas_h, geom.chunks,
                          np.array([1.0, 1.06, 0.95, 1.0], np.float32))
    c = geom.chunks[0]
    warped = [torch.zeros((c.h, c.w, 3), dtype=torch.int16, device="cuda") for _ in c.contributors]
    out = torch.empty((c.h, c.w, 3), dtype=torch.int16, device="cuda")
    owner = torch.empty((c.h, c.w), dtype=torch.int32, device="cuda")
    note("small: RGB48 chain")
    n = render.render_chunk_u16(hip, geom, c, [f.data_ptr() for f in rgb_dev], [w.data_ptr() for w in warped],
                                out.data_ptr(), 5, owner_ptr=owner.data_ptr())
    torch.cuda.synchronize()
    assert n == 4
    want_w = []
    for i in c.contributors:
        inv = po.invert_homography(geom.H[i].astype(np.float32))
        w = po.warp_perspective(rgb_host[i], c.w, c.h, inv, c.x, c.y)
        if abs(float(geom.gains[i]) - 1.0) > 1e-6:
            w = po.apply_gain(w, float(geom.gains[i]))
        want_w.append(w)
    for k, w in enumerate(warped):
        assert np.array_equal(w.cpu().numpy().view(np.uint16), want_w[k]), k
    want, want_own = po.blend_chunk(want_w, 5, want_owner=True)
    assert np.array_equal(owner.cpu().numpy(), want_own)
    assert np.array_equal(out.cpu().numpy().view(np.uint16), want)


# ------------------------------------------------------------- gigapixel --
@pytest.fixture(scope="module")
def giga():
    import torch

    geom = synth.gigapixel()
    scene = synth.make_scene(20261203, geom.canvas_w, geom.canvas_h)
    note("gigapixel: generating 48 FP16 frames of 6000x4000")
    frames = []
    for i in range(geom.num_frames):
        f32 = synth.render_frame_torch(scene, geom.H[i], int(geom.frame_w[i]), int(geom.frame_h[i]))
        frames.append(f32.half().contiguous())
        del f32
    torch.cuda.synchronize()
    yield geom, frames
    del frames
    torch.cuda.empty_cache()


@pytest.mark.timeout(600)
def test_gigapixel_nine_contributor_chunk_against_cpu_fp64_oracle(hip, giga):
    geom, frames = giga
    fdesc = _descs(geom, frames, _abi.PIX_F16)
    k = max(range(len(geom.chunks)), key=lambda i: (len(geom.chunks[i].contributors), geom.chunks[i].area))
    c = geom.chunks[k]
    assert len(c.contributors) == 9 and (c.w, c.h) == (4096, 4096)
    host = {i: frames[i].float().cpu().numpy() for i in c.contributors}
    note(f"gigapixel: chunk {k} ({c.x},{c.y}) on the CPU FP64 oracle")
    want, want_own = _oracle(geom, host, c, 7, fp64=True)
    for mode in (_abi.RENDER_PRECISION, _abi.RENDER_BATCHED):
        note(f"gigapixel: chunk {k} mode {mode}")
        outs, owns = _render(hip, geom, fdesc, [c], 7, mode)
        got = outs[0].cpu().numpy()
        assert np.array_equal(owns[0].cpu().numpy(), want_own), mode
        err = float(np.abs(got - want).max())
        db = psnr(got, want)
        note(f"gigapixel chunk {k} mode {mode}: max abs {err:.2e}, {db:.1f} dB")
        assert err <= 1e-3 and db >= 50.0, (mode, err, db)


@pytest.mark.timeout(600)
def test_gigapixel_canvas_batched_against_fp64_form(hip, giga):
    import torch

    geom, frames = giga
    fdesc = _descs(geom, frames, _abi.PIX_F16)
    chunks = geom.chunks
    assert len(chunks) == 50 and sum(c.area for c in chunks) == 37977 * 19405
    note("gigapixel: whole canvas, FP64 form")
    ref, ref_own = _render(hip, geom, fdesc, chunks, 7, _abi.RENDER_FP64)
    note("gigapixel: whole canvas, batched form")
    got, got_own = _render(hip, geom, fdesc, chunks, 7, _abi.RENDER_BATCHED)
    sse, n, worst = 0.0, 0, 0.0
    for k in range(len(chunks)):
        assert torch.equal(got_own[k], ref_own[k]), k
        d = (got[k].double() - ref[k].double())
        worst = max(worst, float(d.abs().max()))
        sse += float((d * d).sum())
        n += d.numel()
    del got
    db = 10 * math.log10(1.0 / (sse / n))
    note(f"gigapixel canvas batched: max abs {worst:.2e}, {db:.1f} dB")
    assert worst <= 1e-3 and db >= 50.0
    note("gigapixel: whole canvas, batched form with RGB48 output")
    g16, _ = _render(hip, geom, fdesc, chunks, 7, _abi.RENDER_BATCHED | _abi.RENDER_OUT_RGB48, owners=False)
    worst16 = 0.0
    for k in range(len(chunks)):
        v = (g16[k].to(torch.int32) & 0xFFFF).to(torch.float64) / 65535.0
        worst16 = max(worst16, float((v - ref[k].double()).abs().max()))
    note(f"gigapixel canvas RGB48: max abs {worst16:.2e} after /65535")
    assert worst16 <= 1e-3 + 0.5 / 65535


@pytest.mark.timeout(600)
def test_gigapixel_surrogate_on_8192_tiles(hip, giga):
    # the learned surrogate at its stated inference shape: the same canvas
    # cut into 8192x8192 tiles (15 chunks).  On the tile with the most
    # contributors the trained network keeps >= 50 dB against the FP64 form
    # (its max error is not bounded to 1e-3: the fallback tests a predicted
    # error), and a per-pixel bound below every prediction falls back on
    # every level-K tile and renders the batched form's bytes
    import os

    import torch

    from image360_amd import surrogate as sg

    geom, frames = giga
    g8 = synth.gigapixel(chunk=8192)
    assert len(g8.chunks) == 15 and max(max(c.w, c.h) for c in g8.chunks) == 8192
    assert np.array_equal(g8.H, geom.H)  # the same frames, other chunks
    fdesc = _descs(g8, frames, _abi.PIX_F16)
    c = max(g8.chunks, key=lambda q: (len(q.contributors), q.area))
    note(f"8192 tiles: chunk ({c.x},{c.y}) {c.w}x{c.h}, {len(c.contributors)} contributors, FP64 form")
    ref, ref_own = _render(hip, g8, fdesc, [c], 7, _abi.RENDER_FP64)
    bits, b = sg.load_params(os.path.join(os.path.dirname(sg.__file__), "surrogate_weights.npz"))
    wt = torch.from_numpy(bits.view(np.int16)).cuda()
    bt = torch.from_numpy(b).cuda()
    torch.cuda.synchronize()
    hip.check(hip.surrogate_set(wt.data_ptr(), bt.data_ptr(), 1e-5), "surrogate_set")
    try:
        got, got_own = _render(hip, g8, fdesc, [c], 7, _abi.RENDER_SURROGATE)
        tiles, fb = hip.surrogate_stats()
        d = got[0].double() - ref[0].double()
        db = 10 * math.log10(1.0 / float((d * d).mean()))
        note(f"8192 tile surrogate: {fb}/{tiles} level-K tiles fall back, {db:.1f} dB, "
             f"max abs {float(d.abs().max()):.2e}")
        assert torch.equal(got_own[0], ref_own[0])  # ownership and seams are not the network's
        assert tiles > 0 and fb < tiles and db >= 50.0
        want, _ = _render(hip, g8, fdesc, [c], 7, _abi.RENDER_BATCHED, owners=False)
        hip.check(hip.surrogate_set_max_error(1e-30), "surrogate_set_max_error")
        got, _ = _render(hip, g8, fdesc, [c], 7, _abi.RENDER_SURROGATE, owners=False)
        tiles, fb = hip.surrogate_stats()
        assert fb == tiles and torch.equal(got[0], want[0])
    finally:
        hip.check(hip.surrogate_set_max_error(0.0), "surrogate_set_max_error")


# --------------------------------------------------------- uneven chunks --
@pytest.mark.timeout(600)
def test_uneven_chunks_batched_equals_per_chunk_loop(hip, giga):
    import torch

    geom, frames = giga
    u = synth.uneven_chunks(base=geom)
    fdesc = _descs(u, frames, _abi.PIX_F16)
    assert len(u.chunks) == 512
    assert all(1 <= len(c.contributors) <= 6 and 256 <= min(c.w, c.h) and max(c.w, c.h) <= 4096 for c in u.chunks)
    note("uneven: 512 chunks, per-chunk precision loop")
    ref, ref_own = _render(hip, u, fdesc, u.chunks, 6, _abi.RENDER_PRECISION)
    note("uneven: 512 chunks, batched")
    for rep in range(2):  # capture, then replay
        got, got_own = _render(hip, u, fdesc, u.chunks, 6, _abi.RENDER_BATCHED)
        for k in range(len(u.chunks)):
            assert torch.equal(got_own[k], ref_own[k]), (rep, k)
            assert torch.equal(got[k], ref[k]), (rep, k)
        del got, got_own
    # sampled chunks against the CPU FP64 oracle: the three smallest with
    # at least two contributors, and the one with the most
    cand = sorted((c.area, k) for k, c in enumerate(u.chunks) if len(c.contributors) >= 2)
    pick = [k for _, k in cand[:3]]
    pick.append(max(range(len(u.chunks)), key=lambda k: (len(u.chunks[k].contributors), -u.chunks[k].area)))
    for k in pick:
        c = u.chunks[k]
        host = {i: frames[i].float().cpu().numpy() for i in c.contributors}
        note(f"uneven: chunk {k} {c.w}x{c.h}, {len(c.contributors)} contributors, FP64 oracle")
        want, want_own = _oracle(u, host, c, 6, fp64=True)
        got = ref[k].cpu().numpy()
        assert np.array_equal(ref_own[k].cpu().numpy(), want_own), k
        err = float(np.abs(got - want).max())
        assert err <= 1e-3 and psnr(got, want) >= 50.0, (k, err)
