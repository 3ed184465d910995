"""GPU parity of the IComputeBackend operations against the oracle and the
golden vectors of the compiled reference (bit-exact, like the reference's
own cross-tier checks, tests/engine_smoke/main.cpp:169-296)."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.gpu_util import CAM_MUL, FILTERS, RGB_CAM, note

pytestmark = pytest.mark.gpu


def _median(hip, planes, sigma=2.0):
    planes = [np.ascontiguousarray(p, dtype=np.uint16) for p in planes]
    out = np.zeros_like(planes[0])
    h, w = planes[0].shape[:2] if planes[0].ndim == 3 else (1, planes[0].size // 3)
    if planes[0].ndim != 3:
        w = planes[0].size // 3
        h = 1
    hip.check(hip.MedianStack(planes, len(planes), out, w, h, sigma), "MedianStack")
    return out


def test_backend_identity(hip):
    info = hip.GetGpuInfo()
    note(f"device {info}")
    assert info["arch"].startswith("gfx950")


def test_median_golden(hip, golden):
    g = golden("median")
    note("median golden smoke/burst")
    assert np.array_equal(_median(hip, list(g["smoke_in"])), g["smoke_out"])
    assert np.array_equal(_median(hip, list(g["burst_in"])), g["burst_out"])
    allin, allout = g["all_in"], g["all_out"]
    for n in range(1, 33):
        for k, sigma in enumerate((2.0, 1.0, 0.5)):
            note(f"median golden n={n} sigma={sigma}")
            assert np.array_equal(_median(hip, list(allin[:n]), sigma), allout[n - 1, k]), (n, sigma)


def test_median_engine_smoke_literals(hip):
    # tests/engine_smoke/main.cpp:139-145: a gross outlier is rejected
    a, b, c = np.full(9, 100, np.uint16), np.full(9, 200, np.uint16), np.full(9, 9999, np.uint16)
    out = np.zeros(9, np.uint16)
    hip.check(hip.MedianStack([a, b, c], 3, out, 1, 3, 2.0), "MedianStack")
    assert out[0] in (100, 200)


@pytest.mark.parametrize("n", [16, 5, 32])
def test_median_random_vs_oracle(hip, n):
    rng = np.random.Generator(np.random.PCG64(1000 + n))
    base = rng.integers(0, 16384, size=(240, 321, 3))
    planes = np.clip(base[None] + rng.normal(0, 300, size=(n, 240, 321, 3)), 0, 16383).astype(np.uint16)
    planes[rng.random(planes.shape) < 0.03] = 16383
    note(f"median random n={n}")
    assert np.array_equal(_median(hip, list(planes)), po.median_stack(list(planes)))


def test_median_dev_unaligned_and_tail(hip):
    import torch

    rng = np.random.Generator(np.random.PCG64(7))
    n, count = 16, 100003  # odd length: vector tail
    host = rng.integers(0, 65536, size=(n, count + 1), dtype=np.int64).astype(np.uint16)
    dev = torch.from_numpy(host.view(np.int16)).cuda()
    out = torch.zeros(count + 1, dtype=torch.int16, device="cuda")
    base = dev.data_ptr()
    row = (count + 1) * 2
    note("median dev aligned/unaligned")
    for off in (0, 1):  # off=1 -> 2-byte aligned planes: element-load path
        ptrs = [base + i * row + off * 2 for i in range(n)]
        hip.check(hip.median_stack_dev(ptrs, out.data_ptr() + off * 2, count - off), "median_dev")
        torch.cuda.synchronize()
        got = out.cpu().numpy().view(np.uint16)[off:count]
        want = po.median_stack([host[i, off:count] for i in range(n)])
        assert np.array_equal(got, want), off


def test_median_rejects_bad_counts(hip):
    p = np.zeros(9, np.uint16)
    out = np.zeros(9, np.uint16)
    assert hip.MedianStack([p] * 33, 33, out, 1, 3) == 2
    assert hip.MedianStack([p], 0, out, 1, 3) == 2
    assert hip.MedianStack([p], 1, out, 0, 3) == 2
    assert "numInputs" in hip.GetLastError() or "Invalid" in hip.GetLastError()


@pytest.mark.parametrize("layout", list(FILTERS))
def test_demosaic_golden(hip, golden, layout):
    g = golden("demosaic")
    cfa = g["smoke_cfa"]
    out = np.zeros((20, 24, 3), np.uint16)
    note(f"demosaic golden {layout}")
    hip.check(hip.DemosaicBayer(cfa, 24, 20, 64, CAM_MUL, RGB_CAM, FILTERS[layout], out), "DemosaicBayer")
    assert np.array_equal(out, g[f"smoke_out_{layout}"])


def test_demosaic_golden_random_and_tiny(hip, golden):
    g = golden("demosaic")
    cfa = g["rand_cfa"]
    h, w = cfa.shape
    out = np.zeros((h, w, 3), np.uint16)
    hip.check(hip.DemosaicBayer(cfa, w, h, 512, g["rand_cam"], g["rand_rgbcam"], 0x94949494, out), "DemosaicBayer")
    assert np.array_equal(out, g["rand_out"])
    for key in g.files:
        if key.startswith("tiny_cfa_"):
            c = g[key]
            h, w = c.shape
            note(f"demosaic tiny {h}x{w}")
            out = np.zeros((h, w, 3), np.uint16)
            hip.check(hip.DemosaicBayer(np.ascontiguousarray(c), w, h, 64, CAM_MUL, RGB_CAM, 0x94949494, out), "tiny")
            assert np.array_equal(out, g["tiny_out_" + key[len("tiny_cfa_"):]]), key


def test_demosaic_large_and_f32(hip):
    import torch

    rng = np.random.Generator(np.random.PCG64(31))
    h, w = 301, 517
    cfa = rng.integers(0, 16384, size=(h, w), dtype=np.int64).astype(np.uint16)
    cam = np.array([2.1, 1.0, 1.6, 0.97], np.float32)
    rc = np.array([[1.62, -0.51, -0.11, 3.0], [-0.23, 1.41, -0.18, 0.0], [0.02, -0.47, 1.45, -2.5]], np.float32)
    want = po.demosaic(cfa, 200, cam, rc, 0x94949494)
    out = np.zeros((h, w, 3), np.uint16)
    note("demosaic large host")
    hip.check(hip.DemosaicBayer(cfa, w, h, 200, cam, rc, 0x94949494, out), "DemosaicBayer")
    assert np.array_equal(out, want)
    d_cfa = torch.from_numpy(cfa.view(np.int16)).cuda()
    d_out = torch.zeros((h, w, 3), dtype=torch.float32, device="cuda")
    note("demosaic f32 dev")
    hip.check(hip.demosaic_dev(d_cfa.data_ptr(), w, h, 200, cam, rc, 0x94949494, d_out.data_ptr(), f32=True), "f32")
    torch.cuda.synchronize()
    got = d_out.cpu().numpy()
    assert np.array_equal(got, po.demosaic(cfa, 200, cam, rc, 0x94949494, f32=True))
    assert np.array_equal(got.astype(np.uint16), want)


def test_warp_golden(hip, golden):
    g = golden("warp")
    out = np.zeros((30, 40, 3), np.uint16)
    note("warp golden smoke")
    hip.check(hip.WarpPerspective(g["smoke_src"], 40, 30, out, 40, 30, g["smoke_h"], 0, 0), "WarpPerspective")
    assert np.array_equal(out, g["smoke_out"])
    src = g["rand_src"]
    for k in range(len(g["rand_h"])):
        note(f"warp golden rand {k}")
        ox, oy = (int(v) for v in g["rand_off"][k])
        out = np.zeros((67, 101, 3), np.uint16)
        hip.check(hip.WarpPerspective(src, 93, 77, out, 101, 67, g["rand_h"][k], ox, oy), "WarpPerspective")
        assert np.array_equal(out, g["rand_out"][k]), k


def test_warp_identity_copies_center(hip):
    # tests/engine_smoke/main.cpp:132-138
    src = np.full((32, 32, 3), 1000, np.uint16)
    dst = np.zeros_like(src)
    hip.check(hip.WarpPerspective(src, 32, 32, dst, 32, 32, np.eye(3, dtype=np.float32), 0, 0), "identity")
    assert dst[16, 16, 0] == 1000
    assert np.all(dst[:, -1] == 0) and np.all(dst[-1] == 0)  # sx0+1 < srcW excludes the last column/row


def test_warp_large_random_vs_oracle(hip):
    rng = np.random.Generator(np.random.PCG64(77))
    src = rng.integers(0, 65536, size=(700, 1100, 3), dtype=np.int64).astype(np.uint16)
    th = 0.04
    fwd = np.array([[np.cos(th), -np.sin(th), 300.0], [np.sin(th), np.cos(th), 120.0], [3e-5, -2e-5, 1.0]])
    inv = np.linalg.inv(fwd).astype(np.float32)
    out = np.zeros((600, 900, 3), np.uint16)
    note("warp large")
    hip.check(hip.WarpPerspective(src, 1100, 700, out, 900, 600, inv, 250, 100), "WarpPerspective")
    assert np.array_equal(out, po.warp_perspective(src, 900, 600, inv, 250, 100))


def test_gain_golden(hip, golden):
    g = golden("gain")
    for k, gain in enumerate(g["gains"]):
        d = g["in"].copy()
        note(f"gain {gain}")
        hip.check(hip.ApplyGain(d, d.size, float(gain)), "ApplyGain")
        assert np.array_equal(d, g["out"][k]), gain


def test_gain_literals_and_dev(hip):
    import torch

    d = np.array([100, 200, 300, 400], np.uint16)
    hip.check(hip.ApplyGain(d, 4, 2.0), "ApplyGain")
    assert d.tolist() == [200, 400, 600, 800]  # engine_smoke/main.cpp:112-115
    rng = np.random.Generator(np.random.PCG64(3))
    host = rng.integers(0, 65536, size=100007, dtype=np.int64).astype(np.uint16)
    dev = torch.from_numpy(host.view(np.int16)).cuda()
    hip.check(hip.gain_dev(dev.data_ptr(), host.size, 1.37), "gain_dev")
    torch.cuda.synchronize()
    assert np.array_equal(dev.cpu().numpy().view(np.uint16), po.apply_gain(host, 1.37))


def test_burst_median_and_demosaic_full_size(hip):
    # the burst configuration at full size: 16 aligned 14-bit RGGB frames of
    # 4000x3000, median-of-16 bit-exact against the oracle, then the FP32
    # demosaic of the median bit-exact against the oracle
    import torch

    from image360_amd import synth

    w, h = 4000, 3000
    frames, shifts = synth.burst_frames(w=w, h=h)
    hc = h - 9
    n = hc * w
    ptrs = [f.data_ptr() + 2 * ((8 - dy) * w + (8 - dx)) for f, (dx, dy) in zip(frames, shifts)]
    med = torch.empty(n, dtype=torch.int16, device="cuda")
    note("burst median 16 x 4000x3000")
    hip.check(hip.median_stack_dev(ptrs, med.data_ptr(), n), "median_dev")
    host = [f.cpu().numpy().view(np.uint16).ravel() for f in frames]
    planes = [hf[(8 - dy) * w + (8 - dx): (8 - dy) * w + (8 - dx) + n] for hf, (dx, dy) in zip(host, shifts)]
    got = med.cpu().numpy().view(np.uint16)
    assert np.array_equal(got, po.median_stack(planes, 2.0))
    cam = [2.0, 1.0, 1.6, 1.0]
    rc = np.array([[1.7, -0.5, -0.2, 0.0], [-0.2, 1.5, -0.3, 0.0], [0.05, -0.55, 1.5, 0.0]], np.float32)
    rgb = torch.empty((hc, w, 3), dtype=torch.float32, device="cuda")
    hip.check(hip.demosaic_dev(med.data_ptr(), w, hc, 512, cam, rc, 0x94949494, rgb.data_ptr(), f32=True), "f32")
    want = po.demosaic(got.reshape(hc, w), 512, np.array(cam, np.float32), rc, 0x94949494, f32=True)
    assert np.array_equal(rgb.cpu().numpy(), want)
