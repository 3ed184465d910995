"""Small seeded float-render cases shared by the render tests and smoke."""
import math

import numpy as np

from image360_amd import synth


def small_case(seed=3, fw=320, fh=240, n=4, chunk=None, jitter_deg=3.0, persp=1e-5, gains=None, noise=0.0):
    """n frames of fw x fh on a horizontal strip (~30 % overlap), rotation
    +-jitter_deg, perspective terms <= persp, cut from one procedural scene;
    chunks tile the canvas (chunk = edge length, None = one chunk)."""
    rng = np.random.default_rng(seed)
    Hs = []
    for k in range(n):
        th = math.radians(rng.uniform(-jitter_deg, jitter_deg))
        tx = 20 + k * 0.7 * fw
        ty = 30 + rng.uniform(-15, 15)
        Hs.append(np.array([[math.cos(th), -math.sin(th), tx], [math.sin(th), math.cos(th), ty],
                            [rng.uniform(-persp, persp), rng.uniform(-persp, persp), 1.0]]))
    Hs = np.stack(Hs)
    cw = int(20 + (n - 1) * 0.7 * fw + fw + 40)
    ch = int(fh + 90)
    fwv, fhv = np.full(n, fw), np.full(n, fh)
    rects = synth.chunk_grid(cw, ch, chunk) if chunk else [(0, 0, cw, ch)]
    chunks = synth._attach(Hs, fwv, fhv, rects)
    g = synth.Geometry(fwv, fhv, Hs, cw, ch, chunks,
                       np.asarray(gains if gains is not None else np.ones(n), np.float32))
    scene = synth.make_scene(seed, cw, ch, cell=64)
    frames = []
    for i in range(n):
        f = synth.render_frame_torch(scene, Hs[i], fw, fh, device="cpu").numpy()
        if noise:
            f = np.clip(f + rng.normal(0, noise, f.shape), 0, 1).astype(np.float32)
        frames.append(f.astype(np.float16))
    return g, frames


def grid_case(seed=7, nx=3, ny=2, fw=200, fh=150, chunk=None, jitter_deg=2.0, persp=1e-5):
    """nx x ny frames of fw x fh on a grid with ~30 % overlap both ways, so
    that a chunk holds pairs with no contributor in common (seam pairs that
    may run side by side)."""
    rng = np.random.default_rng(seed)
    Hs = []
    for r in range(ny):
        for k in range(nx):
            th = math.radians(rng.uniform(-jitter_deg, jitter_deg))
            tx = 20 + k * 0.7 * fw + rng.uniform(-6, 6)
            ty = 20 + r * 0.7 * fh + rng.uniform(-6, 6)
            Hs.append(np.array([[math.cos(th), -math.sin(th), tx], [math.sin(th), math.cos(th), ty],
                                [rng.uniform(-persp, persp), rng.uniform(-persp, persp), 1.0]]))
    Hs = np.stack(Hs)
    n = nx * ny
    cw = int(40 + (nx - 1) * 0.7 * fw + fw + 20)
    ch = int(40 + (ny - 1) * 0.7 * fh + fh + 20)
    fwv, fhv = np.full(n, fw), np.full(n, fh)
    rects = synth.chunk_grid(cw, ch, chunk) if chunk else [(0, 0, cw, ch)]
    chunks = synth._attach(Hs, fwv, fhv, rects)
    g = synth.Geometry(fwv, fhv, Hs, cw, ch, chunks, np.ones(n, np.float32))
    scene = synth.make_scene(seed, cw, ch, cell=64)
    frames = [synth.render_frame_torch(scene, Hs[i], fw, fh, device="cpu").numpy().astype(np.float16) for i in range(n)]
    return g, frames


def psnr(a, b):
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return float("inf") if mse == 0 else 10 * math.log10(1.0 / mse)
