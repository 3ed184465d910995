"""Writes the golden fixtures in this directory from the reference's own
hot-path sources, compiled in place by ``make -C oracle ref`` into
``oracle/_ref/libi360ref.so``.

Run in the container that holds the reference sources:

    make -C oracle ref && python tests/golden/gen_golden.py

Inputs are generated here with numpy's PCG64 from fixed seeds (the
reference's own tests draw theirs from std::mt19937 through
implementation-defined distributions, so they are not portable vectors);
the shapes and parameters follow the reference tests:
  * median: tests/engine_smoke/main.cpp:175-195 (17x13x3, 21 inputs,
    values 0..4000, three gross outliers at scalar 100), plus 14-bit
    16-input stacks and every input count 1..32;
  * warp: tests/engine_smoke/main.cpp:221-237 (40x30 ramp, 0.06 rad
    rotation + (1.7, 0.9)), plus offset / perspective / singular cases;
  * demosaic: tests/engine_smoke/main.cpp:260-275 (24x20, black 64,
    camMul {1.8,1,1.5,1}, fixed rgbCam), for all four Bayer layouts;
  * blend: tests/render_blend/main.cpp:79-193 (the three scenes), plus
    seeded multi-contributor chunks warped by the reference warp.
Each fixture is an .npz of plain arrays (no pickles).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as po  # noqa: E402

RGB_CAM = np.array([[1.9, -0.7, -0.2, 0.0], [-0.3, 1.5, -0.2, 0.0], [0.0, -0.4, 1.6, 0.0]], np.float32)
CAM_MUL = np.array([1.8, 1.0, 1.5, 1.0], np.float32)
FILTERS = {"RGGB": 0x94949494, "BGGR": 0x16161616, "GRBG": 0x61616161, "GBRG": 0x49494949}


def median_cases():
    rng = np.random.Generator(np.random.PCG64(777))
    out = {}
    # engine_smoke shape
    planes = rng.integers(0, 4001, size=(21, 13, 17, 3), dtype=np.int64).astype(np.uint16)
    planes.reshape(21, -1)[:3, 100] = 60000
    out["smoke_in"] = planes
    out["smoke_out"] = po.ref_median_stack(list(planes), 2.0, tier=0)
    # 14-bit burst-like stack: 16 noisy frames of one smooth signal + outliers
    base = rng.integers(200, 16000, size=(32, 48, 3)).astype(np.float64)
    noisy = base[None] + rng.normal(0, 40.0, size=(16, 32, 48, 3))
    noisy[rng.random(noisy.shape) < 0.02] = 16383
    planes = np.clip(noisy, 0, 16383).astype(np.uint16)
    out["burst_in"] = planes
    out["burst_out"] = po.ref_median_stack(list(planes), 2.0, tier=0)
    # every input count 1..32 with several sigma thresholds
    allin = rng.integers(0, 65536, size=(32, 7, 11, 3), dtype=np.int64).astype(np.uint16)
    allin[:, :, :3] = rng.integers(1000, 1010, size=(32, 7, 3, 3)).astype(np.uint16)  # tight clusters
    out["all_in"] = allin
    res = []
    for n in range(1, 33):
        for sigma in (2.0, 1.0, 0.5):
            res.append(po.ref_median_stack(list(allin[:n]), sigma, tier=0))
    out["all_out"] = np.stack(res).reshape(32, 3, 7, 11, 3)
    return out


def warp_cases():
    out = {}
    w, h = 40, 30
    y, x, c = np.meshgrid(np.arange(h), np.arange(w), np.arange(3), indexing="ij")
    src = ((x * 37 + y * 53 + c * 17) % 4001).astype(np.uint16)
    a = np.float32(0.06)
    ca, sa = np.cos(a, dtype=np.float32), np.sin(a, dtype=np.float32)
    hm = np.array([ca, -sa, 1.7, sa, ca, 0.9, 0, 0, 1], np.float32)
    out["smoke_src"] = src
    out["smoke_h"] = hm
    out["smoke_out"] = po.ref_warp_perspective(src, w, h, hm, 0, 0, tier=0)
    rng = np.random.Generator(np.random.PCG64(4242))
    big = rng.integers(0, 65536, size=(77, 93, 3), dtype=np.int64).astype(np.uint16)
    hs, outs, offs = [], [], []
    for k in range(6):
        th = rng.uniform(-0.2, 0.2)
        sc = rng.uniform(0.8, 1.25)
        m = np.array([[sc * np.cos(th), -sc * np.sin(th), rng.uniform(-20, 20)],
                      [sc * np.sin(th), sc * np.cos(th), rng.uniform(-20, 20)],
                      [rng.uniform(-2e-3, 2e-3), rng.uniform(-2e-3, 2e-3), 1.0]], np.float64)
        if k == 5:
            m[2] = [1e-3, 0.0, -0.05]  # singular line crosses the output
        hm = m.astype(np.float32).reshape(9)
        ox, oy = int(rng.integers(-30, 30)), int(rng.integers(-30, 30))
        hs.append(hm)
        offs.append((ox, oy))
        outs.append(po.ref_warp_perspective(big, 101, 67, hm, ox, oy, tier=0))
    out["rand_src"] = big
    out["rand_h"] = np.stack(hs)
    out["rand_off"] = np.array(offs, np.int32)
    out["rand_out"] = np.stack(outs)
    return out


def demosaic_cases():
    out = {}
    w, h = 24, 20
    yy, xx = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    cfa = (200 + (xx * 31 + yy * 19) % 3000).astype(np.uint16)
    out["smoke_cfa"] = cfa
    for name, f in FILTERS.items():
        out[f"smoke_out_{name}"] = po.ref_demosaic(cfa, 64, CAM_MUL, RGB_CAM, f, tier=0)
    rng = np.random.Generator(np.random.PCG64(99))
    cfa2 = rng.integers(0, 16384, size=(37, 53), dtype=np.int64).astype(np.uint16)
    cfa2[0, :5] = 0
    cfa2[5, 5:9] = 65535
    out["rand_cfa"] = cfa2
    cam2 = np.array([2.2, 1.0, 1.4, 1.05], np.float32)
    rc2 = np.array([[1.7, -0.5, -0.2, 10.0], [-0.25, 1.4, -0.15, -5.0], [0.05, -0.55, 1.5, 0.0]], np.float32)
    out["rand_cam"] = cam2
    out["rand_rgbcam"] = rc2
    out["rand_out"] = po.ref_demosaic(cfa2, 512, cam2, rc2, FILTERS["RGGB"], tier=0)
    for sz in [(1, 1), (1, 7), (6, 1), (2, 2), (3, 5)]:
        c3 = rng.integers(0, 16384, size=sz, dtype=np.int64).astype(np.uint16)
        out[f"tiny_cfa_{sz[0]}x{sz[1]}"] = c3
        out[f"tiny_out_{sz[0]}x{sz[1]}"] = po.ref_demosaic(c3, 64, CAM_MUL, RGB_CAM, FILTERS["RGGB"], tier=0)
    return out


def gain_cases():
    rng = np.random.Generator(np.random.PCG64(5))
    d = rng.integers(0, 65536, size=1000, dtype=np.int64).astype(np.uint16)
    out = {"in": d}
    gains = np.array([2.0, 0.5, 1.0, 1.0000005, 1.37, 0.0, -1.0, 3.3], np.float32)
    out["gains"] = gains
    out["out"] = np.stack([po.ref_apply_gain(d, float(g)) for g in gains])
    return out


def homography_cases():
    rng = np.random.Generator(np.random.PCG64(11))
    hs = []
    for _ in range(20):
        hs.append(np.concatenate([rng.uniform(-2, 2, 8), [1.0]]).astype(np.float32))
    hs.append(np.array([1, 2, 3, 2, 4, 6, 0, 0, 1], np.float32))  # singular
    hs = np.stack(hs)
    inv = np.zeros_like(hs)
    ok = np.zeros(len(hs), np.int32)
    for i, m in enumerate(hs):
        r = po.ref_invert_homography(m)
        if r is not None:
            inv[i] = r
            ok[i] = 1
    return {"h": hs, "inv": inv, "ok": ok}


def _set(buf, x0, x1, y0, y1, rgb):
    buf[y0:y1, x0:x1] = np.array(rgb, np.uint16)


def _bg(buf, x0, x1):
    h = buf.shape[0]
    for x in range(x0, x1):
        buf[:, x] = (10000 + x * 100, 20000, 30000)


def blend_scenes():
    """The three scenes of tests/render_blend/main.cpp:79-193."""
    scenes = {}
    w, h = 400, 40
    a = np.zeros((h, w, 3), np.uint16); b = np.zeros_like(a)
    _bg(a, 0, 250); _bg(b, 150, 400)
    _set(a, 160, 170, 0, h, (60000, 2000, 2000)); _set(b, 230, 240, 0, h, (60000, 2000, 2000))
    scenes["ghost"] = [a, b]
    w, h = 200, 40
    a = np.zeros((h, w, 3), np.uint16); b = np.zeros_like(a)
    _bg(a, 0, 100); _bg(b, 100, 200)
    scenes["gap"] = [a, b]
    a = np.zeros((h, w, 3), np.uint16); b = np.zeros_like(a)
    _bg(a, 0, 120); _bg(b, 80, 200)
    _set(a, 98, 102, 0, h, (60000, 2000, 2000)); _set(b, 98, 102, 0, h, (2000, 2000, 60000))
    scenes["route"] = [a, b]
    return scenes


def seeded_chunks():
    """Small chunks whose contributors are reference-warped rotated frames."""
    rng = np.random.Generator(np.random.PCG64(20261203))
    chunks = {}
    for k, (cw, ch, c) in enumerate([(96, 64, 3), (64, 96, 4), (80, 80, 2), (123, 71, 5)]):
        warped = []
        for i in range(c):
            sw, sh = int(rng.integers(50, 90)), int(rng.integers(40, 70))
            yy, xx = np.meshgrid(np.arange(sh), np.arange(sw), indexing="ij")
            img = np.zeros((sh, sw, 3))
            for ch_ in range(3):
                for _ in range(3):
                    per = rng.uniform(8, 40); th = rng.uniform(0, np.pi)
                    img[..., ch_] += rng.uniform(2000, 8000) * np.sin(
                        2 * np.pi * (xx * np.cos(th) + yy * np.sin(th)) / per + rng.uniform(0, 6.28))
            img = np.clip(img + 30000 + rng.normal(0, 200, img.shape), 1, 65535).astype(np.uint16)
            th = rng.uniform(-0.15, 0.15)
            tx, ty = rng.uniform(-40, cw - 20), rng.uniform(-30, ch - 20)
            fwd = np.array([[np.cos(th), -np.sin(th), tx], [np.sin(th), np.cos(th), ty], [0, 0, 1]])
            inv = np.linalg.inv(fwd).astype(np.float32).reshape(9)
            warped.append(po.ref_warp_perspective(img, cw, ch, inv, 0, 0, tier=0))
        chunks[f"c{k}"] = warped
    return chunks


def main():
    if po.ref() is None:
        raise SystemExit("oracle/_ref/libi360ref.so missing: run `make -C oracle ref` first")
    np.savez_compressed(os.path.join(HERE, "median.npz"), **median_cases())
    np.savez_compressed(os.path.join(HERE, "warp.npz"), **warp_cases())
    np.savez_compressed(os.path.join(HERE, "demosaic.npz"), **demosaic_cases())
    np.savez_compressed(os.path.join(HERE, "gain.npz"), **gain_cases())
    np.savez_compressed(os.path.join(HERE, "homography.npz"), **homography_cases())
    blend = {}
    for name, bufs in {**blend_scenes(), **seeded_chunks()}.items():
        blend[f"{name}_in"] = np.stack(bufs)
        blend[f"{name}_out"] = po.ref_blend_chunk(bufs)
        blend[f"{name}_owner"] = po.ref_blend_owner(bufs)
        covs = [(b != 0).any(axis=2).astype(np.uint8) for b in bufs]
        blend[f"{name}_dist"] = np.stack([po.ref_coverage_distance(cv) for cv in covs])
    np.savez_compressed(os.path.join(HERE, "blend.npz"), **blend)
    # multiband compose with 7 bands on one seeded chunk (the gigapixel level count)
    bufs = seeded_chunks()["c0"]
    owner = po.ref_blend_owner(bufs)
    np.savez_compressed(os.path.join(HERE, "compose7.npz"), inp=np.stack(bufs), owner=owner,
                        out=po.ref_multiband_compose(bufs, owner, 7))
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
