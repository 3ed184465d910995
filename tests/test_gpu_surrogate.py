"""The surrogate network's GPU kernels (i360hip_surrogate_*) against the
binary64 oracle: forward outputs, loss and every weight / bias gradient on
random data with BF16-rounded weights (FP32 accumulation on the GPU; the
tolerances below are those of FP32 sums of a few thousand terms)."""
import numpy as np
import pytest

from image360_amd import surrogate as sg
from oracle import pyoracle as po
from tests.gpu_util import note

pytestmark = pytest.mark.gpu


def _case(seed, n=2, h=21, w=19):
    rng = np.random.default_rng(seed)
    w32, b = sg.init_params(seed)
    b = b + rng.normal(0, 0.05, b.shape).astype(np.float32)
    feat = rng.random((n, h, w, 16)).astype(np.float32)
    target = (feat[..., 7:10] + rng.normal(0, 0.02, (n, h, w, 3))).astype(np.float32)
    return sg.bf16_bits(w32), sg.bf16_values(w32), b, feat, target


def _t(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()



def test_forward_matches_oracle(hip):
    import torch

    bits, wv, b, feat, _ = _case(1)
    n, h, w = feat.shape[:3]
    # every device buffer stays referenced until the kernels have run
    out = torch.zeros((n, h, w, 4), dtype=torch.float32, device="cuda")
    wb, bd, fd = _t(bits.view(np.int16)), _t(b), _t(feat)
    note("surrogate forward")
    hip.check(hip.surrogate_forward(wb.data_ptr(), bd.data_ptr(), fd.data_ptr(), n, h, w, out.data_ptr()),
              "surrogate_forward")
    torch.cuda.synchronize()
    want = po.surr_forward(wv, b, feat)
    got = out.cpu().numpy()
    scale = np.abs(want).max()
    err = np.max(np.abs(got - want)) / scale
    assert err <= 1e-5, err


@pytest.mark.parametrize("shape", [(2, 21, 19), (3, 32, 32), (1, 7, 40)])
def test_gradients_match_oracle(hip, shape):
    import torch

    n, h, w = shape
    bits, wv, b, feat, target = _case(2, n, h, w)
    gw = torch.zeros(sg.WEIGHTS, dtype=torch.float32, device="cuda")
    gb = torch.zeros(sg.BIASES, dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    wb, bd, fd, td = _t(bits.view(np.int16)), _t(b), _t(feat), _t(target)
    note(f"surrogate gradients {shape}")
    hip.check(hip.surrogate_gradients(wb.data_ptr(), bd.data_ptr(), fd.data_ptr(), td.data_ptr(), n, h, w,
                                      gw.data_ptr(), gb.data_ptr(), loss.data_ptr()),
              "surrogate_gradients")
    torch.cuda.synchronize()
    want_loss, want_gw, want_gb = po.surr_loss_grad(wv, b, feat, target)
    assert abs(loss.item() - want_loss) <= 1e-4 * abs(want_loss), (loss.item(), want_loss)
    g = gw.cpu().numpy().astype(np.float64)
    for j in range(sg.LAYERS):
        a, e = g[sg.woff(j): sg.woff(j + 1)], want_gw[sg.woff(j): sg.woff(j + 1)]
        assert np.max(np.abs(a - e)) <= 2e-4 * np.abs(e).max() + 1e-12, (j, np.max(np.abs(a - e)), np.abs(e).max())
    gbv = gb.cpu().numpy().astype(np.float64)
    assert np.max(np.abs(gbv - want_gb)) <= 2e-4 * np.abs(want_gb).max() + 1e-12, \
        (np.max(np.abs(gbv - want_gb)), np.abs(want_gb).max())


def test_train_steps_reduce_loss_deterministically(hip):
    import torch

    _, _, _, feat, target = _case(3, 4, 32, 32)
    n, h, w = feat.shape[:3]
    fd, td = _t(feat), _t(target)
    runs = []
    for rep in range(2):
        w32, b = sg.init_params(5)
        master, bias = _t(w32), _t(b)
        wb = _t(sg.bf16_bits(w32).view(np.int16))
        adam = torch.zeros(2 * (sg.WEIGHTS + sg.BIASES), dtype=torch.float32, device="cuda")
        loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        hist = []
        for step in range(1, 61):
            hip.check(hip.surrogate_train_step(master.data_ptr(), bias.data_ptr(), adam.data_ptr(), wb.data_ptr(),
                                               fd.data_ptr(), td.data_ptr(), n, h, w, 1e-3, step, loss.data_ptr()),
                      "surrogate_train_step")
            hist.append(loss.item())
        runs.append((hist, master.cpu().numpy(), wb.cpu().numpy()))
    hist = runs[0][0]
    note(f"surrogate loss {hist[0]:.4g} -> {hist[-1]:.4g}")
    assert hist[-1] < 0.7 * hist[0], (hist[0], hist[-1])
    assert runs[0][0] == runs[1][0] and np.array_equal(runs[0][1], runs[1][1])
    # the BF16 copy is the rounding of the master weights
    assert np.array_equal(runs[0][2].view(np.uint16), sg.bf16_bits(runs[0][1]))


# --------------------------------------------------------------- render form
def _small_pano(seed):
    return sg.panorama(seed, rows=2, cols=3, fw=1200, fh=800, focal=1000.0, chunk=1024)


def _render(hip, g, frames, mode, bands=7):
    import torch

    fd, _, _, _ = sg.describe(g, frames)
    outs, descs, contrib = [], [], []
    for c in g.chunks:
        o = torch.zeros((c.h, c.w, 3), dtype=torch.float32, device="cuda")
        descs.append({"x": c.x, "y": c.y, "w": c.w, "h": c.h, "first": len(contrib), "count": len(c.contributors),
                      "out": o.data_ptr(), "owner": None})
        contrib += c.contributors
        outs.append(o)
    hip.check(hip.render_chunks(fd, descs, contrib, bands, mode), f"render mode {mode}")
    torch.cuda.synchronize()
    return torch.cat([o.reshape(-1) for o in outs])


def _set_params(hip, bits, b, thr):
    import torch

    wt, bt = _t(bits.view(np.int16)), _t(b)
    torch.cuda.synchronize()
    hip.check(hip.surrogate_set(wt.data_ptr(), bt.data_ptr(), thr), "surrogate_set")


def test_surrogate_form_all_fallback_equals_batched(hip):
    # threshold 0: every tile falls back, so the surrogate form must
    # reproduce the plain (batched) blend exactly, deep levels included
    from image360_amd import _abi

    g, frames = _small_pano(11)
    w32, b = sg.init_params(1)
    _set_params(hip, sg.bf16_bits(w32), b, 0.0)
    note("surrogate all-fallback")
    want = _render(hip, g, frames, _abi.RENDER_BATCHED)
    for rep in range(2):  # capture, then replay
        got = _render(hip, g, frames, _abi.RENDER_SURROGATE)
        assert bool((got == want).all()), rep
    tiles, fb = hip.surrogate_stats()
    assert tiles > 0 and fb == tiles


def test_surrogate_max_error_bound(hip):
    # the per-pixel bound on the predicted squared error: more tiles fall
    # back as it tightens (the graph is re-captured on every change), and a
    # bound below every prediction renders the batched form's bytes even
    # though the mean test alone keeps tiles on the network
    import os

    from image360_amd import _abi

    g, frames = _small_pano(13)
    bits, b = sg.load_params(os.path.join(os.path.dirname(sg.__file__), "surrogate_weights.npz"))
    _set_params(hip, bits, b, 1e-5)
    want = _render(hip, g, frames, _abi.RENDER_BATCHED)
    fbs, tiles = [], 0
    try:
        for bound in (0.0, 1e-5, 1e-6, 1e-8, 1e-30):
            hip.check(hip.surrogate_set_max_error(bound), "surrogate_set_max_error")
            got = _render(hip, g, frames, _abi.RENDER_SURROGATE)
            tiles, fb = hip.surrogate_stats()
            fbs.append(fb)
        note(f"surrogate max-error bound: fallback tiles {fbs} of {tiles}")
        assert tiles > 0 and fbs[0] < tiles  # the mean test alone keeps tiles on the network
        assert fbs[0] <= fbs[1] <= fbs[2] <= fbs[3] <= fbs[4] == tiles
        assert bool((got == want).all())
        for bad in (-1.0, float("nan"), float("inf")):
            assert hip.surrogate_set_max_error(bad) == _abi.INVALID_PARAM
    finally:
        hip.check(hip.surrogate_set_max_error(0.0), "surrogate_set_max_error")


def test_surrogate_form_composite_only(hip):
    # a network that predicts no residual and no error never falls back: the
    # low band is the single-level composite, which stays close to the blend
    from image360_amd import _abi

    g, frames = _small_pano(12)
    w = np.zeros(sg.WEIGHTS, np.float32)
    b = np.zeros(sg.BIASES, np.float32)
    b[sg.boff(sg.LAYERS - 1) + 3] = -30.0
    _set_params(hip, sg.bf16_bits(w), b, 1e-5)
    note("surrogate composite-only")
    want = _render(hip, g, frames, _abi.RENDER_BATCHED)
    got = _render(hip, g, frames, _abi.RENDER_SURROGATE)
    tiles, fb = hip.surrogate_stats()
    mse = float(((got.double() - want.double()) ** 2).mean())
    note(f"composite-only: {fb}/{tiles} tiles fall back, mse vs blend {mse:.3g}")
    assert tiles > 0 and fb == 0, (tiles, fb)
    # the deep levels spread exposure differences over wide bands; a single
    # level does not, so the result differs, but not by much
    assert 0 < mse < 1e-2, mse


def test_trained_surrogate_gains_over_composite(hip):
    # trained on the plain blend of three panoramas, the network's low band
    # is closer to the blend's than the single-level composite (the method
    # without it) on a held-out panorama
    import torch

    fs, ts = [], []
    for seed in (21, 22, 23):
        g, frames = _small_pano(seed)
        fd, cd, ct, sizes = sg.describe(g, frames)
        f, t = sg.level_patches(hip, fd, cd, ct, 7, sizes, patch=16)
        fs.append(f)
        ts.append(t)
    feat, target = torch.cat(fs), torch.cat(ts)
    g, frames = _small_pano(29)
    fd, cd, ct, sizes = sg.describe(g, frames)
    vf, vt = sg.level_patches(hip, fd, cd, ct, 7, sizes, patch=16)
    note(f"surrogate training on {feat.shape[0]} patches, {vf.shape[0]} held out")
    assert feat.shape[0] >= 50 and vf.shape[0] >= 10
    tr = sg.Trainer(hip, seed=4)
    tr.train(feat, target, 400, batch=32, lr=2e-3, lr_end=2e-4)
    base = sg.low_band_psnr(vf, vt)
    got = sg.low_band_psnr(vf, vt, tr.predict(vf))
    note(f"held-out low band: composite {base:.2f} dB, surrogate {got:.2f} dB")
    assert got > base + 1.0


def test_surrogate_form_rgb48_output(hip):
    # the RGB48 output flag on the surrogate form (chunk groups, per-group
    # network) writes the FP32 output's RGB48 rounding, with the same
    # fallback decisions
    import torch

    from image360_amd import _abi
    from oracle import pyoracle as po

    g, frames = _small_pano(13)
    w32, b = sg.init_params(2)
    _set_params(hip, sg.bf16_bits(w32), b, 1e-4)
    fd, _, _, _ = sg.describe(g, frames)
    note("surrogate form, RGB48 output")
    outs = {}
    for mode in (_abi.RENDER_SURROGATE, _abi.RENDER_SURROGATE | _abi.RENDER_OUT_RGB48):
        bufs, descs, contrib = [], [], []
        for c in g.chunks:
            o = torch.zeros((c.h, c.w, 3), dtype=torch.float32, device="cuda")
            descs.append({"x": c.x, "y": c.y, "w": c.w, "h": c.h, "first": len(contrib),
                          "count": len(c.contributors), "out": o.data_ptr(), "owner": None})
            contrib += c.contributors
            bufs.append(o)
        hip.check(hip.render_chunks(fd, descs, contrib, 7, mode), f"render mode {mode}")
        torch.cuda.synchronize()
        outs[mode] = ([o.cpu().numpy() for o in bufs], hip.surrogate_stats())
    (f32, st32), (raw, st16) = outs[_abi.RENDER_SURROGATE], outs[_abi.RENDER_SURROGATE | _abi.RENDER_OUT_RGB48]
    assert st32 == st16 and st32[0] > 0
    for k, c in enumerate(g.chunks):
        got = raw[k].view(np.uint16).reshape(-1)[: c.h * c.w * 3].reshape(c.h, c.w, 3)
        assert np.array_equal(got, po.rgb48_from_render(f32[k])), k


def test_surrogate_form_same_bytes_for_any_chunk_grouping(hip):
    # the surrogate stages run per chunk group on concurrent streams; the
    # output and the fallback decisions must not depend on the grouping
    from image360_amd import _abi

    g, frames = _small_pano(14)
    w32, b = sg.init_params(3)
    _set_params(hip, sg.bf16_bits(w32), b, 1e-4)
    res = []
    try:
        for groups in (1, 2, 4):
            hip.render_set_tuning(chunk_groups=groups)
            note(f"surrogate form, {groups} chunk group(s), {len(g.chunks)} chunks")
            res.append((_render(hip, g, frames, _abi.RENDER_SURROGATE), hip.surrogate_stats()))
    finally:
        hip.render_set_tuning()
    for out, stats in res[1:]:
        assert bool((out == res[0][0]).all()) and stats == res[0][1]
