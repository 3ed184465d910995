"""The chamfer kernel's systolic schedule (image360_amd/csrc/chamfer.hpp),
emulated step by step on the CPU, reproduces the serial two-pass chamfer
of the oracle bit for bit (SeamBlend.cpp:206-242): every thread reads the
neighbour strips' boundary values that were written at earlier steps,
processes its rows at step 2r + lag, and prefetches its rows P ahead."""
import numpy as np
import pytest

from oracle import pyoracle as po

f32 = np.float32
D = f32(1.41421356)


def min_strip(w):
    return 4 if w <= 4096 else (8 if w <= 8192 else 16)


def pitch(w):
    return (w + 15) & ~15


def shear_rows(w, h):
    return h + (((w - 1) // min_strip(w)) >> 1) + 1


def shear_at(x, y, S, p):
    return (y + ((x // S) >> 1)) * p + x


class Flat:
    """A flat device buffer whose every access is bounds-checked."""

    def __init__(self, n, dtype, fill):
        self.a = np.full(n, fill, dtype)

    def load(self, off, k):
        assert 0 <= off and off + k <= self.a.size, (off, k, self.a.size)
        return self.a[off:off + k].copy()

    def store(self, off, v):
        assert 0 <= off and off + len(v) <= self.a.size, (off, len(v), self.a.size)
        self.a[off:off + len(v)] = v


def shear_cov(cov, scov, box, w, h, S, box_only=False):
    """chamfer_shear_kernel: the box's rows, columns widened to 16-column
    groups, copied 8 bytes at a time to the sheared plane.  box_only: only
    the box's own pixels are written (the render's warp writes its
    footprint, which contains the box), the rest holds garbage."""
    x0, y0, x1, y1 = box
    if x1 < 0: return
    cp = pitch(w)
    if box_only:
        for y in range(y0, y1 + 1):
            for x in range(x0, x1 + 1):
                scov.store(shear_at(x, y, S, cp), np.array([cov[y, x]], np.uint8))
        return
    ga, gb = x0 & ~15, min((x1 + 16) & ~15, cp)
    for y in range(y0, y1 + 1):
        for x in range(ga, gb, 8):
            v = np.array([cov[y, x + k] if x + k < w else 0 for k in range(8)], np.uint8)
            rows = {(xx // S) >> 1 for xx in range(x, x + 8)}
            assert len(rows) == 1  # an 8-byte group never straddles two shear rows
            scov.store(shear_at(x, y, S, cp), v)


INF = np.float32(np.inf)


def kernel(scov, dist, box, w, h, S, P, FWD, nt, access_rows):
    """chamfer_kernel: neighbours beyond the processed area are constants
    (0 for an existing pixel outside the box, +inf for no pixel) taken from
    sentinel LDS slots and per-thread constants; columns past the chunk
    edge hold +inf."""
    x0, y0, x1, y1 = box
    if x1 < 0: return
    dp = pitch(w)
    xa = (x0 // S) * S
    nact = (x1 + 1 - xa + S - 1) // S
    rows = y1 - y0 + 1
    kInf = f32(w + h)
    # slot t + 1 holds strip t; slots 0 / nact + 1 the sentinels
    firstv = np.full((2, nt + 2), -7, np.float32); lastv = np.full((2, nt + 2), -7, np.float32)
    sl = f32(0) if xa > 0 else INF
    sr = f32(0) if xa + nact * S < w else INF
    lastv[:, 0] = sl; firstv[:, nact + 1] = sr
    edgePrev = (y0 > 0) if FWD else (y1 < h - 1)
    th = []
    for t in range(nt):
        lag = t if FWD else nact - 1 - t
        xs = xa + t * S
        th.append(dict(t=t, active=t < nact, xs=xs, xl=xa + min(t, nact - 1) * S, lag=lag,
                       base=lag >> 1, par=lag & 1,
                       inbox=[x0 <= xs + s <= x1 for s in range(S)], exists=[xs + s < w for s in range(S)],
                       pL0=f32(0) if (edgePrev and xs > 0) else INF,
                       pR0=f32(0) if (edgePrev and xs + S < w) else INF,
                       ring=None,
                       prev=np.array([f32(0) if (edgePrev and xs + s < w) else INF for s in range(S)], np.float32)))

    def load_row(T, r):
        # unconditional, for active and inactive threads alike (strip clamped)
        y = y0 + r if FWD else y1 - r
        if FWD:
            return scov.load(shear_at(T['xl'], y, S, dp), S)
        return dist.load(shear_at(T['xl'], y, S, dp), S)

    def fetch(T, j):
        return [load_row(T, min(max(P * j + k - T['base'], 0), rows - 1)) for k in range(P)]

    def init_of(T, raw, s):
        if not T['inbox'][s]: return f32(0)
        if not FWD: return raw[s]
        return f32(f32(raw[s]) * kInf)  # coverage byte 0 or 1

    for T in th:
        T['ring'] = fetch(T, 0)
    tmax = 2 * (rows - 1) + nact - 1
    T0 = 0
    while T0 <= tmax:
        for T in th:
            T['next'] = fetch(T, T0 // (2 * P) + 1)
        for s2 in range(2 * P):
            if T0 + s2 > tmax: break
            # all threads "in parallel": reads of LDS happen before writes of this step
            snap_first = firstv.copy(); snap_last = lastv.copy()
            for T in th:
                t = T['t']
                if not (T['active'] and (s2 & 1) == T['par']): continue
                k = s2 >> 1
                r = (T0 >> 1) + k - T['base']
                if not (0 <= r < rows): continue
                y = y0 + r if FWD else y1 - r
                access_rows.setdefault((FWD, T0 + s2, t & 1), set()).add((y + (T['xs'] // S >> 1)))
                pb, cb = (r - 1) & 1, r & 1
                lCur, rCur = snap_last[cb][t], snap_first[cb][t + 2]
                used = [lCur if FWD else rCur] + ([snap_last[pb][t], snap_first[pb][t + 2]] if r > 0 else [])
                assert all(v != -7 for v in used)  # every neighbour used was written at an earlier step
                prevL = T['pL0'] if r == 0 else snap_last[pb][t]
                prevR = T['pR0'] if r == 0 else snap_first[pb][t + 2]
                cur = np.zeros(S, np.float32)
                xs = T['xs']; prev = T['prev']; raw = T['ring'][k]
                order = range(S) if FWD else range(S - 1, -1, -1)
                for s in order:
                    b = init_of(T, raw, s)
                    if FWD:
                        b = min(b, f32((lCur if s == 0 else cur[s - 1]) + f32(1)))
                    else:
                        b = min(b, f32((rCur if s == S - 1 else cur[s + 1]) + f32(1)))
                    b = min(b, f32(prev[s] + f32(1)))
                    b = min(b, f32((prevL if s == 0 else prev[s - 1]) + D))
                    b = min(b, f32((prevR if s == S - 1 else prev[s + 1]) + D))
                    cur[s] = b if T['exists'][s] else INF
                assert xs + S <= dp  # the strip stays inside its row
                dist.store(shear_at(xs, y, S, dp), cur)  # the whole strip, as the kernel does
                T['prev'] = cur.copy()
                firstv[cb][t + 1] = cur[0]; lastv[cb][t + 1] = cur[S - 1]
        for T in th:
            T['ring'] = T['next']
        T0 += 2 * P


def run_chamfer(cov, box, S, P, box_only=False):
    """Shear pre-pass, forward, backward; returns the distances read back
    at every (x, y) in the box as the ownership kernels read them."""
    h, w = cov.shape
    nt = ((w + S - 1) // S + 127) // 128 * 128
    rows = shear_rows(w, h)
    scov = Flat(rows * pitch(w), np.uint8, 0x5A)
    dist = Flat(rows * pitch(w), np.float32, -5)
    shear_cov(cov, scov, box, w, h, S, box_only)
    acc = {}
    kernel(scov, dist, box, w, h, S, P, True, nt, acc)
    kernel(scov, dist, box, w, h, S, P, False, nt, acc)
    out = np.zeros((h, w), np.float32)
    for y in range(box[1], box[3] + 1):
        for x in range(box[0], box[2] + 1):
            out[y, x] = dist.a[shear_at(x, y, S, pitch(w))]
    return out, acc


@pytest.mark.parametrize("seed", range(10))
@pytest.mark.parametrize("S,P", [(4, 8), (8, 4), (16, 2)])
def test_schedule_matches_serial_chamfer(seed, S, P):
    rng = np.random.default_rng(seed)
    h = int(rng.integers(3, 40))
    # ragged widths around multiples of 16 (strips that end past the last column)
    w = int(rng.integers(3, 70)) if seed % 2 else 16 * int(rng.integers(1, 5)) + int(rng.integers(-3, 4))
    cov = (rng.random((h, w)) < rng.uniform(0.3, 0.95)).astype(np.uint8)
    if seed % 3 == 0:
        cov[:, : w // 3] = 0  # a box that does not start at the chunk edge
    if seed % 4 == 1:
        cov[:, -1] = 1  # a box that reaches the last column
    if seed % 4 == 2 and h > 4:
        cov[:2] = 0; cov[-1] = 0  # rows exist above and below the box
    ys, xs = np.nonzero(cov)
    if len(xs) == 0:
        return
    box = (xs.min(), ys.min(), xs.max(), ys.max())
    want = po.coverage_distance(cov)
    got, acc = run_chamfer(cov, box, S, P)
    assert np.array_equal(got, want)
    got, _ = run_chamfer(cov, box, S, P, box_only=True)
    assert np.array_equal(got, want)
    # at every step the strips of one parity (one wavefront) touch one sheared row
    assert all(len(rows) == 1 for rows in acc.values())


def test_sheared_plane_bounds_at_maximum_shear():
    """The last strip of the widest chunk of each strip class stays inside
    shear_rows(w, h) rows, and a strip never crosses its row."""
    for w in (4096, 4095, 4081, 8192, 8177, 16384, 16369):
        for S in (4, 8, 16):
            if S < min_strip(w):
                continue
            h = 5
            xs = ((w - 1) // S) * S
            assert xs + S <= pitch(w)
            assert shear_at(xs + S - 1, h - 1, S, pitch(w)) < shear_rows(w, h) * pitch(w)
