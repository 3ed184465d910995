"""Pins of the surrogate network oracle (oracle/surrogate_oracle.cpp), on
the CPU: closed-form forward passes of hand-set networks (centre tap =
identity, off-centre tap = shift with zero padding) and central finite
differences of the loss against the oracle's analytic gradients."""
import numpy as np

from image360_amd import surrogate as sg
from oracle import pyoracle as po


def _identity_net(shift_x=0):
    """Layer 0 copies feature channels 0..15 to 0..15 at the tap (0, shift_x);
    hidden layers copy channels at the centre; the last layer maps channel 7
    to output 0 (x64) and channel 0 to output 3."""
    w = np.zeros(sg.WEIGHTS, np.float32)
    b = np.zeros(sg.BIASES, np.float32)
    for j in range(sg.LAYERS):
        cin, cout = sg.cin(j), sg.cout(j)
        wj = np.zeros((3, 3, cin, cout), np.float32)
        if j == 0:
            for c in range(16):
                wj[1, 1 + shift_x, c, c] = 1.0
        elif j < sg.LAYERS - 1:
            for c in range(32):
                wj[1, 1, c, c] = 1.0
        else:
            wj[1, 1, 7, 0] = 64.0
            wj[1, 1, 0, 3] = 1.0
        w[sg.woff(j): sg.woff(j + 1)] = wj.ravel()
    return w, b


def test_forward_identity_and_shift():
    rng = np.random.default_rng(1)
    feat = rng.random((2, 5, 7, 16)).astype(np.float32)
    w, b = _identity_net()
    out = po.surr_forward(w, b, feat)
    assert np.allclose(out[..., 0], 64 * feat[..., 7], rtol=1e-6)
    assert np.allclose(out[..., 3], feat[..., 0], rtol=1e-6)
    assert not out[..., 1:3].any()
    # tap (ky=1, kx=2) reads the right neighbour; zero past the right edge
    w, b = _identity_net(shift_x=1)
    out = po.surr_forward(w, b, feat)
    assert np.allclose(out[:, :, :-1, 0], 64 * feat[:, :, 1:, 7], rtol=1e-6)
    assert not out[:, :, -1, 0].any()


def test_relu_clips_negative():
    feat = -np.ones((1, 3, 3, 16), np.float32)
    w, b = _identity_net()
    assert not po.surr_forward(w, b, feat).any()


def _fd(f, x, idx, h):
    """Central and one-sided differences of f at x[idx] (actual float steps)."""
    xp, xm = x.copy(), x.copy()
    xp[idx] += h
    xm[idx] -= h
    dp, dm = float(xp[idx]) - float(x[idx]), float(x[idx]) - float(xm[idx])
    f0, fp, fm = f(x), f(xp), f(xm)
    return (fp - fm) / (dp + dm), (fp - f0) / dp, (f0 - fm) / dm


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(7)
    w, b = sg.init_params(seed=3)
    w = sg.bf16_values(w)
    b = b + rng.normal(0, 0.05, b.shape).astype(np.float32)
    feat = rng.random((1, 6, 5, 16)).astype(np.float32)
    target = (feat[..., 7:10] + rng.normal(0, 0.01, (1, 6, 5, 3))).astype(np.float32)
    loss, gw, gb, le = po.surr_loss_grad(w, b, feat, target, want_le=True)
    assert loss > 0
    # the loss with the error target frozen is the function the gradient
    # belongs to; a step that crosses a ReLU kink shows as one-sided
    # differences that disagree, and such samples are not compared
    checked = 0
    idxs = list(rng.choice(sg.WEIGHTS, 60, replace=False)) + [sg.woff(5) + 3, sg.woff(0) + 100]
    for idx in idxs:
        c, fwd, bwd = _fd(lambda v: po.surr_loss_grad(v, b, feat, target, le_in=le)[0], w, idx, 1e-4)
        if abs(fwd - bwd) > 1e-3 * max(abs(c), 1e-6):
            continue
        checked += 1
        assert abs(c - gw[idx]) <= 1e-4 * abs(gw[idx]) + 1e-7 * np.abs(gw).max(), (idx, c, gw[idx])
    for idx in range(0, sg.BIASES, 5):
        c, fwd, bwd = _fd(lambda v: po.surr_loss_grad(w, v, feat, target, le_in=le)[0], b, idx, 1e-4)
        if abs(fwd - bwd) > 1e-3 * max(abs(c), 1e-6):
            continue
        checked += 1
        assert abs(c - gb[idx]) <= 1e-4 * abs(gb[idx]) + 1e-7 * np.abs(gb).max(), (idx, c, gb[idx])
    assert checked >= 80


def test_gradient_check_catches_a_wrong_gradient():
    # the comparison above has teeth: a gradient taken without the ReLU mask
    # of one layer (a plausible slip) fails it
    rng = np.random.default_rng(7)
    w, b = sg.init_params(seed=3)
    w = sg.bf16_values(w)
    feat = rng.random((1, 6, 5, 16)).astype(np.float32)
    target = (feat[..., 7:10] + 0.01).astype(np.float32)
    _, gw, _, le = po.surr_loss_grad(w, b, feat, target, want_le=True)
    wrong = gw.copy()
    wrong[sg.woff(2): sg.woff(3)] *= 1.5
    bad = 0
    for idx in rng.choice(np.arange(sg.woff(2), sg.woff(3)), 20, replace=False):
        c, fwd, bwd = _fd(lambda v: po.surr_loss_grad(v, b, feat, target, le_in=le)[0], w, idx, 1e-4)
        if abs(c) > 1e-6 and abs(c - wrong[idx]) > 1e-4 * abs(wrong[idx]) + 1e-7 * np.abs(gw).max():
            bad += 1
    assert bad > 0
