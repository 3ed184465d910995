"""Learned blend surrogate: parameter helpers and the GPU training driver.

The network, its kernels and its loss are defined by the C ABI
(include/i360hip.h, i360hip_surrogate_*); this module holds the host side:
He-initialised parameters, BF16 rounding, and a training loop that calls
the GPU train-step kernels on batches of feature / target patches.
"""
from __future__ import annotations

import numpy as np

LAYERS = 6
IN = 16
HIDDEN = 32
OUT = 4
WEIGHTS = 42624
BIASES = 164


def cin(j: int) -> int:
    return IN if j == 0 else HIDDEN


def cout(j: int) -> int:
    return OUT if j == LAYERS - 1 else HIDDEN


def woff(j: int) -> int:
    return sum(9 * cin(k) * cout(k) for k in range(j))


def boff(j: int) -> int:
    return sum(cout(k) for k in range(j))


def init_params(seed: int = 0):
    """He-normal FP32 weights (HWIO per layer), zero biases except the
    error head's, which starts at log10 of a 50 dB error (-5)."""
    rng = np.random.default_rng(seed)
    w = np.zeros(WEIGHTS, np.float32)
    for j in range(LAYERS):
        n = 9 * cin(j) * cout(j)
        w[woff(j): woff(j) + n] = rng.normal(0.0, np.sqrt(2.0 / (9 * cin(j))), n).astype(np.float32)
    b = np.zeros(BIASES, np.float32)
    b[boff(LAYERS - 1) + 3] = -5.0
    return w, b


def bf16_bits(x):
    """Round-to-nearest-even BF16 bit patterns of float32 values."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def bf16_values(x):
    """float32 values of the BF16 rounding of x."""
    return (bf16_bits(x).astype(np.uint32) << 16).view(np.float32)


# ------------------------------------------------------------ training data --
def level_patches(hip, frames_desc, chunks_desc, contributors, bands, chunk_sizes, patch=32, seam_only=True,
                  stream=None, stride=None):
    """Features / targets of the plain blend at each chunk's surrogate level
    (i360hip_surrogate_features), cut into patch x patch tiles; returns two
    torch tensors (N, patch, patch, 16) and (N, patch, patch, 3) on the GPU.
    seam_only keeps tiles where two contributors share weight somewhere
    (0.02 < seam mask < 0.98)."""
    import torch

    dims = [hip.surrogate_chunk_level(w, h, bands) for (w, h) in chunk_sizes]
    npix = sum(lw * lh for k, lw, lh in dims if k > 0)
    feat = torch.empty((max(1, npix), 16), dtype=torch.float32, device="cuda")
    target = torch.empty((max(1, npix), 3), dtype=torch.float32, device="cuda")
    hip.check(hip.surrogate_features(frames_desc, chunks_desc, contributors, bands, feat.data_ptr(),
                                     target.data_ptr(), stream), "surrogate_features")
    torch.cuda.synchronize()
    fs, ts = [], []
    off = 0
    for k, lw, lh in dims:
        if k == 0:
            continue
        f = feat[off: off + lw * lh].view(lh, lw, 16)
        t = target[off: off + lw * lh].view(lh, lw, 3)
        off += lw * lh
        step = stride or patch
        for y in range(0, lh - patch + 1, step):
            for x in range(0, lw - patch + 1, step):
                fp = f[y: y + patch, x: x + patch]
                if seam_only:
                    m = fp[..., 6]
                    if not bool(((m > 0.02) & (m < 0.98)).any()):
                        continue
                fs.append(fp)
                ts.append(t[y: y + patch, x: x + patch])
    if not fs:
        return (torch.empty((0, patch, patch, 16), device="cuda"), torch.empty((0, patch, patch, 3), device="cuda"))
    return torch.stack(fs).contiguous(), torch.stack(ts).contiguous()


def panorama(seed, rows=3, cols=4, fw=3000, fh=2000, focal=2500.0, chunk=2048, gain_spread=0.15):
    """A seeded synthetic panorama for training: frames of one procedural
    scene on a rows x cols grid (25 % overlap, +-2 deg jitter) with per-frame
    exposure gains in [1-spread, 1+spread]; returns (geometry, FP16 frames
    on the GPU)."""
    from . import synth

    rng = np.random.default_rng(seed)
    H, cw, ch = synth.grid_homographies(rows, cols, fw, fh, 0.25, focal, 2.0, rng)
    n = rows * cols
    fwv, fhv = np.full(n, fw), np.full(n, fh)
    chunks = synth._attach(H, fwv, fhv, synth.chunk_grid(cw, ch, chunk))
    gains = rng.uniform(1 - gain_spread, 1 + gain_spread, n).astype(np.float32)
    g = synth.Geometry(fwv, fhv, H, cw, ch, chunks, gains)
    scene = synth.make_scene(seed, cw, ch)
    frames = [synth.render_frame_torch(scene, H[i], fw, fh).half().contiguous() for i in range(n)]
    return g, frames


def describe(g, frames, chunks=None):
    """Frame and chunk descriptors of a geometry for the feature export,
    which writes no chunk output (the descriptors' `out` is a placeholder
    that no kernel dereferences)."""
    from . import _abi

    frames_desc = [{"ptr": f.data_ptr(), "w": int(g.frame_w[i]), "h": int(g.frame_h[i]), "fmt": _abi.PIX_F16,
                    "gain": float(g.gains[i]), "H": g.H[i]} for i, f in enumerate(frames)]
    chunks = g.chunks if chunks is None else chunks
    descs, contrib = [], []
    for c in chunks:
        descs.append({"x": c.x, "y": c.y, "w": c.w, "h": c.h, "first": len(contrib), "count": len(c.contributors),
                      "out": 1, "owner": None})
        contrib += c.contributors
    return frames_desc, descs, contrib, [(c.w, c.h) for c in chunks]


# ----------------------------------------------------------------- training --
class Trainer:
    """FP32 master weights, Adam state and BF16 weights on the GPU, stepped
    by i360hip_surrogate_train_step (fenced against torch's work on the
    batches it gathers and the loss it reads)."""

    def __init__(self, hip, seed=0):
        import torch

        self.hip = hip
        w, b = init_params(seed)
        self.master = torch.from_numpy(w).cuda()
        self.bias = torch.from_numpy(b).cuda()
        self.weights = torch.from_numpy(bf16_bits(w).view(np.int16)).cuda()
        self.adam = torch.zeros(2 * (WEIGHTS + BIASES), dtype=torch.float32, device="cuda")
        self.loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        self.step = 0

    def train(self, feat, target, steps, batch=64, lr=1e-3, lr_end=1e-4, seed=0, log=None):
        import torch

        n, h, w = feat.shape[:3]
        g = torch.Generator(device="cpu").manual_seed(seed)
        losses = []
        for k in range(steps):
            idx = torch.randint(0, n, (min(batch, n),), generator=g).cuda()
            fb = feat.index_select(0, idx).contiguous()
            tb = target.index_select(0, idx).contiguous()
            self.step += 1
            cur = lr * (lr_end / lr) ** (k / max(1, steps - 1))
            self.hip.check(self.hip.surrogate_train_step(self.master.data_ptr(), self.bias.data_ptr(),
                                                         self.adam.data_ptr(), self.weights.data_ptr(),
                                                         fb.data_ptr(), tb.data_ptr(), fb.shape[0], h, w, cur,
                                                         self.step, self.loss.data_ptr()),
                           "surrogate_train_step")
            if log and (k % 100 == 0 or k == steps - 1):
                losses.append((k, self.loss.item()))
                log(f"step {k}: loss {losses[-1][1]:.5g}")
        return losses

    def predict(self, feat):
        """Network output (N, h, w, 4) for features (N, h, w, 16)."""
        import torch

        n, h, w = feat.shape[:3]
        out = torch.empty((n, h, w, 4), dtype=torch.float32, device="cuda")
        f = feat.contiguous()
        self.hip.check(self.hip.surrogate_forward(self.weights.data_ptr(), self.bias.data_ptr(), f.data_ptr(), n, h,
                                                  w, out.data_ptr()),
                       "surrogate_forward")
        return out

    def save(self, path):
        np.savez(path, weights_bf16=self.weights.cpu().numpy().view(np.uint16), biases=self.bias.cpu().numpy())


def load_params(path):
    """(BF16 weight bits as uint16, FP32 biases) from a saved file."""
    z = np.load(path, allow_pickle=False)
    return z["weights_bf16"].astype(np.uint16), z["biases"].astype(np.float32)


def low_band_psnr(feat, target, out=None):
    """PSNR (dB) of the level-K low band against the plain blend's: the
    single-level composite alone (out=None) or composite + residual."""
    import torch

    pred = feat[..., 7:10].double()
    if out is not None:
        pred = pred + out[..., 0:3].double() / 64.0
    mse = float(((pred - target.double()) ** 2).mean())
    return float("inf") if mse == 0 else 10 * np.log10(1.0 / mse)


def main():
    import argparse
    import time

    import torch

    from . import HipComputeBackend

    ap = argparse.ArgumentParser(description="train the blend surrogate on synthetic panoramas (GPU)")
    ap.add_argument("--out", default=None)
    ap.add_argument("--panoramas", type=int, default=40)
    ap.add_argument("--patches", type=int, default=20000)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--bands", type=int, default=7)
    ap.add_argument("--seed", type=int, default=20261203)
    a = ap.parse_args()
    log = lambda m: print(f"[surrogate] {m}", flush=True)  # noqa: E731
    with HipComputeBackend(0) as hip:
        fs, ts = [], []
        t0 = time.time()
        for k in range(a.panoramas):
            g, frames = panorama(a.seed + 101 * k, rows=4, cols=6)
            fd, cd, ct, sizes = describe(g, frames)
            f, t = level_patches(hip, fd, cd, ct, a.bands, sizes, stride=16)
            fs.append(f)
            ts.append(t)
            log(f"panorama {k}: {len(g.chunks)} chunks, {f.shape[0]} seam patches ({time.time() - t0:.1f} s)")
            del frames
            if sum(x.shape[0] for x in fs) >= a.patches:
                break
        feat = torch.cat(fs)[: a.patches].contiguous()
        target = torch.cat(ts)[: a.patches].contiguous()
        nval = max(1, feat.shape[0] // 10)
        tr_f, tr_t, va_f, va_t = feat[nval:], target[nval:], feat[:nval], target[:nval]
        log(f"{tr_f.shape[0]} training / {nval} held-out patches of {tuple(feat.shape[1:3])}")
        tr = Trainer(hip, seed=a.seed)
        t1 = time.time()
        tr.train(tr_f, tr_t, a.steps, a.batch, log=log)
        torch.cuda.synchronize()
        log(f"{a.steps} steps in {time.time() - t1:.1f} s")
        out = tr.predict(va_f)
        log(f"held-out low band: composite {low_band_psnr(va_f, va_t):.2f} dB, "
            f"composite + surrogate {low_band_psnr(va_f, va_t, out):.2f} dB")
        if a.out:
            tr.save(a.out)
            log(f"saved {a.out}")


if __name__ == "__main__":
    main()
