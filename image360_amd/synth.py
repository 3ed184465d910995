"""Seeded synthetic workloads: panorama geometry, chunk grids and frames.

Geometry follows the reference's render set-up:
  * chunk grid: chunkSize tiles over the canvas, the last row/column
    clamped (ProjectManager.cpp:266-280), chunk size 4096 on a >=16 GB device
    (ProjectManager.cpp:136-142);
  * contributors: the float AABB test of each frame's projected corners
    against the chunk (OverlapCulling.cpp:10-45), capped at 32
    (RenderExecutor.cpp:59-63);
  * homographies map frame-local pixels to the canvas (HomographyMath.h:19-29).
Frames sample a procedural scene (a sum of sinusoids with periods 8-128 px
plus random axis-aligned rectangles) at H(u, v); generation runs on the GPU
through torch (plumbing only: the rendered path never calls torch).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

CHUNK = 4096


@dataclasses.dataclass
class Chunk:
    x: int
    y: int
    w: int
    h: int
    contributors: list

    @property
    def area(self):
        return self.w * self.h


@dataclasses.dataclass
class Geometry:
    frame_w: np.ndarray      # (F,) int
    frame_h: np.ndarray      # (F,) int
    H: np.ndarray            # (F, 3, 3) float64, frame-local -> canvas
    canvas_w: int
    canvas_h: int
    chunks: list
    gains: np.ndarray        # (F,) float32

    @property
    def num_frames(self):
        return len(self.frame_w)

    @property
    def output_pixels(self):
        return sum(c.area for c in self.chunks)


def _rot(yaw, pitch):
    cy, sy, cp, sp = math.cos(yaw), math.sin(yaw), math.cos(pitch), math.sin(pitch)
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rx = np.array([[1, 0, 0], [0, cp, -sp], [0, sp, cp]])
    return ry @ rx


def overlaps(Hm, fw, fh, cx, cy, cw, ch):
    """OverlapCulling.cpp:10-45 in binary32 (corner AABB vs chunk)."""
    m = Hm.astype(np.float32).reshape(9)
    minx = miny = np.float32(np.finfo(np.float32).max)
    maxx = maxy = np.float32(-np.finfo(np.float32).max)
    for u, v in ((0.0, 0.0), (float(fw), 0.0), (0.0, float(fh)), (float(fw), float(fh))):
        u, v = np.float32(u), np.float32(v)
        den = m[6] * u + m[7] * v + m[8]
        if abs(den) < np.float32(1e-10):
            den = np.float32(1e-10) if den >= 0 else np.float32(-1e-10)
        px = (m[0] * u + m[1] * v + m[2]) / den
        py = (m[3] * u + m[4] * v + m[5]) / den
        minx, miny, maxx, maxy = min(minx, px), min(miny, py), max(maxx, px), max(maxy, py)
    return bool(minx < cx + cw and maxx > cx and miny < cy + ch and maxy > cy)


def chunk_grid(canvas_w, canvas_h, size=CHUNK):
    out = []
    for y in range(0, canvas_h, size):
        for x in range(0, canvas_w, size):
            out.append((x, y, min(size, canvas_w - x), min(size, canvas_h - y)))
    return out


def _attach(H, fw, fh, rects, cap=32):
    chunks = []
    for (x, y, w, h) in rects:
        cl = [i for i in range(len(H)) if overlaps(H[i], fw[i], fh[i], x, y, w, h)]
        chunks.append(Chunk(x, y, w, h, cl[:cap]))
    return chunks


def grid_homographies(rows, cols, fw, fh, overlap, focal, jitter_deg, rng):
    """Frames on a rows x cols grid with `overlap` fractional overlap; each
    frame is rotated by yaw/pitch jitter about its own optical centre
    (H = T * K R K^-1), FP64."""
    step_x, step_y = fw * (1 - overlap), fh * (1 - overlap)
    K = np.array([[focal, 0, fw / 2], [0, focal, fh / 2], [0, 0, 1.0]])
    Ki = np.linalg.inv(K)
    Hs = []
    for r in range(rows):
        for c in range(cols):
            yaw = math.radians(rng.uniform(-jitter_deg, jitter_deg))
            pitch = math.radians(rng.uniform(-jitter_deg, jitter_deg))
            T = np.array([[1, 0, c * step_x], [0, 1, r * step_y], [0, 0, 1.0]])
            Hm = T @ K @ _rot(yaw, pitch) @ Ki
            Hs.append(Hm / Hm[2, 2])
    Hs = np.stack(Hs)
    # shift so that every projected corner lands at canvas coordinates >= 0
    pts = []
    for Hm in Hs:
        for u, v in ((0, 0), (fw, 0), (0, fh), (fw, fh)):
            p = Hm @ np.array([u, v, 1.0])
            pts.append(p[:2] / p[2])
    pts = np.array(pts)
    mn, mx = np.floor(pts.min(0)), np.ceil(pts.max(0))
    S = np.array([[1, 0, -mn[0]], [0, 1, -mn[1]], [0, 0, 1.0]])
    Hs = np.stack([S @ Hm for Hm in Hs])
    return Hs, int(mx[0] - mn[0]), int(mx[1] - mn[1])


def gigapixel(seed=20261203, rows=6, cols=8, fw=6000, fh=4000, overlap=0.25, focal=5000.0, jitter=2.0,
              chunk=CHUNK):
    """48 frames of 6000x4000 in a 6x8 grid, 25 % overlap, focal 5000 px,
    yaw/pitch jitter +-2 deg, FP64 homographies, 4096^2 chunks."""
    rng = np.random.default_rng(seed)
    H, cw, ch = grid_homographies(rows, cols, fw, fh, overlap, focal, jitter, rng)
    n = rows * cols
    fwv, fhv = np.full(n, fw), np.full(n, fh)
    chunks = _attach(H, fwv, fhv, chunk_grid(cw, ch, chunk))
    return Geometry(fwv, fhv, H, cw, ch, chunks, np.ones(n, np.float32))


def small(seed=20261203):
    """4 frames of 1024x768 cut from a 2400x1100 scene: rotation +-3 deg,
    ~30 % horizontal overlap, perspective terms <= 1e-5; one 2400x1100 chunk."""
    rng = np.random.default_rng(seed)
    fw, fh = 1024, 768
    Hs = []
    for k in range(4):
        th = math.radians(rng.uniform(-3, 3))
        tx = 40 + k * 0.7 * fw * 0.83
        ty = 150 + rng.uniform(-40, 40)
        Hm = np.array([[math.cos(th), -math.sin(th), tx], [math.sin(th), math.cos(th), ty],
                       [rng.uniform(-1e-5, 1e-5), rng.uniform(-1e-5, 1e-5), 1.0]])
        Hs.append(Hm)
    Hs = np.stack(Hs)
    fwv, fhv = np.full(4, fw), np.full(4, fh)
    chunks = _attach(Hs, fwv, fhv, [(0, 0, 2400, 1100)])
    return Geometry(fwv, fhv, Hs, 2400, 1100, chunks, np.ones(4, np.float32))


# Demosaic parameters of the small configuration's raw frames: 14-bit RGGB
# (LibRaw filters 0x94949494), black level 600 DN, daylight-like white
# balance and a mild camera-to-sRGB matrix (rows sum to 1).
BAYER_RGGB = 0x94949494
BAYER_BLACK = 600
BAYER_CAM_MUL = np.array([1.9, 1.0, 1.55, 1.0], np.float32)
BAYER_RGB_CAM = np.array([[1.45, -0.3, -0.15, 0.0], [-0.12, 1.3, -0.18, 0.0], [0.04, -0.34, 1.3, 0.0]], np.float32)


def bayer_mosaic(rgb, rng, gain=0.5, read_sigma=3.0, scale=12000.0):
    """14-bit RGGB raw of a float RGB image in [0,1] (h, w, 3): signal
    = black + scale * v, Poisson shot noise at `gain` DN per electron plus
    Gaussian read noise of `read_sigma` DN, rounded and clipped to
    0..16383.  numpy PCG64 draws, so the same seed gives the same frame on
    every machine."""
    h, w = rgb.shape[:2]
    cfa = np.empty((h, w), np.float64)
    cfa[0::2, 0::2] = rgb[0::2, 0::2, 0]
    cfa[0::2, 1::2] = rgb[0::2, 1::2, 1]
    cfa[1::2, 0::2] = rgb[1::2, 0::2, 1]
    cfa[1::2, 1::2] = rgb[1::2, 1::2, 2]
    signal = cfa * scale
    dn = rng.poisson(signal / gain) * gain + BAYER_BLACK + rng.normal(0.0, read_sigma, size=signal.shape)
    return np.clip(np.rint(dn), 0, 16383).astype(np.uint16)


def small_bayer(seed=20261203):
    """The small configuration's inputs: the small() geometry and its 4
    14-bit RGGB raw frames of 1024x768, cut from one seeded 2400x1100 scene
    (sum-of-sinusoids texture, periods 8-128 px, plus rectangles) through
    each frame's homography, with Poisson-Gaussian noise (gain 0.5, read
    sigma 3 DN).  Returns (geometry, [cfa u16 (768, 1024)])."""
    g = small(seed)
    scene = make_scene(seed, g.canvas_w, g.canvas_h, cell=256)
    rng = np.random.default_rng(seed + 11)
    cfas = []
    for i in range(g.num_frames):
        rgb = render_frame_torch(scene, g.H[i], int(g.frame_w[i]), int(g.frame_h[i]), device="cpu").numpy()
        cfas.append(bayer_mosaic(rgb.astype(np.float64), rng))
    return g, cfas


def uneven_chunks(seed=20261203, count=512, lo=256, hi=4096, base=None):
    """`count` render chunks with log-uniform edge lengths in [lo, hi], each
    placed where 1..6 frames of the gigapixel geometry overlap it."""
    rng = np.random.default_rng(seed + 1)
    g = base or gigapixel(seed)
    rects = []
    tries = 0
    while len(rects) < count and tries < 100 * count:
        tries += 1
        w = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        h = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        x = int(rng.integers(0, max(1, g.canvas_w - w)))
        y = int(rng.integers(0, max(1, g.canvas_h - h)))
        cl = [i for i in range(g.num_frames) if overlaps(g.H[i], g.frame_w[i], g.frame_h[i], x, y, w, h)]
        if 1 <= len(cl) <= 6:
            rects.append(Chunk(x, y, w, h, cl))
    return Geometry(g.frame_w, g.frame_h, g.H, g.canvas_w, g.canvas_h, rects, g.gains)


# ------------------------------------------------------------------ scene --
@dataclasses.dataclass
class Scene:
    amp: np.ndarray     # (3, K)
    kx: np.ndarray      # (3, K)  2*pi*cos(theta)/period
    ky: np.ndarray      # (3, K)
    phase: np.ndarray   # (3, K)
    base: np.ndarray    # (3,)
    cell: int
    rect_table: np.ndarray  # (ncy, ncx, 7): x0, y0, x1, y1, r, g, b


def make_scene(seed, canvas_w, canvas_h, K=6, cell=512):
    rng = np.random.default_rng(seed + 7)
    per = np.exp(rng.uniform(math.log(8), math.log(128), size=(3, K)))
    th = rng.uniform(0, math.pi, size=(3, K))
    amp = rng.uniform(0.02, 0.07, size=(3, K))
    ncx, ncy = canvas_w // cell + 1, canvas_h // cell + 1
    tab = np.zeros((ncy, ncx, 7), np.float32)
    for cy in range(ncy):
        for cx in range(ncx):
            if rng.random() < 0.6:
                x0 = cx * cell + rng.uniform(0, cell * 0.6)
                y0 = cy * cell + rng.uniform(0, cell * 0.6)
                tab[cy, cx] = [x0, y0, x0 + rng.uniform(20, cell * 0.4), y0 + rng.uniform(20, cell * 0.4),
                               *rng.uniform(0.05, 0.95, size=3)]
            else:
                tab[cy, cx] = [-1, -1, -2, -2, 0, 0, 0]
    return Scene(amp, 2 * math.pi * np.cos(th) / per, 2 * math.pi * np.sin(th) / per,
                 rng.uniform(0, 2 * math.pi, size=(3, K)), rng.uniform(0.35, 0.6, size=3), cell, tab)


def render_frame_torch(scene, Hm, fw, fh, device="cuda", rows_per_pass=500):
    """float32 RGB (fh, fw, 3) in [0,1]: the scene sampled at H(u, v)."""
    import torch

    kx = torch.tensor(scene.kx, dtype=torch.float64, device=device)
    ky = torch.tensor(scene.ky, dtype=torch.float64, device=device)
    ph = torch.tensor(scene.phase, dtype=torch.float64, device=device)
    amp = torch.tensor(scene.amp, dtype=torch.float32, device=device)
    base = torch.tensor(scene.base, dtype=torch.float32, device=device)
    tab = torch.tensor(scene.rect_table, device=device)
    Ht = torch.tensor(Hm, dtype=torch.float64, device=device)
    out = torch.empty((fh, fw, 3), dtype=torch.float32, device=device)
    u = torch.arange(fw, device=device, dtype=torch.float64)
    for y0 in range(0, fh, rows_per_pass):
        y1 = min(fh, y0 + rows_per_pass)
        v = torch.arange(y0, y1, device=device, dtype=torch.float64)
        V, U = torch.meshgrid(v, u, indexing="ij")
        q = Ht[2, 0] * U + Ht[2, 1] * V + Ht[2, 2]
        X = (Ht[0, 0] * U + Ht[0, 1] * V + Ht[0, 2]) / q
        Y = (Ht[1, 0] * U + Ht[1, 1] * V + Ht[1, 2]) / q
        blk = base.view(1, 1, 3).expand(y1 - y0, fw, 3).clone()
        for c in range(3):
            arg = X.unsqueeze(-1) * kx[c] + Y.unsqueeze(-1) * ky[c] + ph[c]
            blk[..., c] += (amp[c] * torch.sin(torch.remainder(arg, 2 * math.pi)).float()).sum(-1)
        cx = torch.clamp((X / scene.cell).long(), 0, tab.shape[1] - 1)
        cy = torch.clamp((Y / scene.cell).long(), 0, tab.shape[0] - 1)
        r = tab[cy, cx]
        inside = (X >= r[..., 0]) & (X < r[..., 2]) & (Y >= r[..., 1]) & (Y < r[..., 3])
        blk = torch.where(inside.unsqueeze(-1), r[..., 4:7], blk)
        out[y0:y1] = blk.clamp_(0.0, 1.0)
    return out


def to_u16(frame_f32):
    """[0,1] float -> RGB48 (u16 = round(v * 65535))."""
    import torch

    return torch.round(frame_f32 * 65535.0).clamp_(0, 65535).to(torch.int32).to(torch.int16)


# ------------------------------------------------------------ burst input --
def burst_frames(seed=20261203, n=16, w=4000, h=3000, max_shift=8, device="cuda"):
    """`n` 14-bit RGGB mosaics (uint16) of one seeded scene, each shifted by
    an integer (dx, dy) in [0, max_shift] with Poisson-Gaussian noise (gain
    0.5 DN per electron, read noise 3 DN); returns (frames, shifts).  The
    scene is sum-of-sinusoids texture plus rectangles."""
    import torch

    rng = np.random.default_rng(seed + 3)
    scene = make_scene(seed, w + max_shift + 1, h + max_shift + 1, cell=256)
    full = render_frame_torch(scene, np.eye(3), w + max_shift, h + max_shift, device=device)
    # RGGB mosaic of the full-size scene, 14-bit
    cfa = torch.empty((h + max_shift, w + max_shift), dtype=torch.float32, device=device)
    cfa[0::2, 0::2] = full[0::2, 0::2, 0]
    cfa[0::2, 1::2] = full[0::2, 1::2, 1]
    cfa[1::2, 0::2] = full[1::2, 0::2, 1]
    cfa[1::2, 1::2] = full[1::2, 1::2, 2]
    signal = cfa * 12000.0 + 600.0
    g = torch.Generator(device=device).manual_seed(seed)
    frames, shifts = [], []
    for k in range(n):
        dx, dy = int(rng.integers(0, max_shift + 1)), int(rng.integers(0, max_shift + 1))
        win = signal[dy: dy + h, dx: dx + w]
        e = torch.poisson(win / 0.5, generator=g) * 0.5
        dn = e + torch.randn(win.shape, generator=g, device=device) * 3.0
        frames.append(dn.round_().clamp_(0, 16383).to(torch.int32).to(torch.int16).contiguous())
        shifts.append((dx, dy))
    return frames, shifts
