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


def shear_cov(cov, scov, box, w, h, S):
    """chamfer_shear_kernel: the box's rows, columns widened to 16-column
    groups, copied 8 bytes at a time to the sheared plane."""
    x0, y0, x1, y1 = box
    if x1 < 0: return
    cp = pitch(w)
    ga, gb = x0 & ~15, min((x1 + 16) & ~15, cp)
    for y in range(y0, y1 + 1):
        for x in range(ga, gb, 8):
            v = np.array([cov[y, x + k] if x + k < w else 0 for k in range(8)], np.uint8)
            rows = {(xx // S) >> 1 for xx in range(x, x + 8)}
            assert len(rows) == 1  # an 8-byte group never straddles two shear rows
            scov.store(shear_at(x, y, S, cp), v)


def kernel(scov, dist, box, w, h, S, P, FWD, nt, access_rows):
    x0, y0, x1, y1 = box
    if x1 < 0: return
    dp = pitch(w)
    xa = (x0 // S) * S
    nact = (x1 + 1 - xa + S - 1) // S
    rows = y1 - y0 + 1
    kInf = f32(w + h)
    firstv = np.zeros((2, nt), np.float32); lastv = np.zeros((2, nt), np.float32)
    th = []
    for t in range(nt):
        lag = t if FWD else nact - 1 - t
        th.append(dict(t=t, active=t < nact, xs=xa + t * S, xl=xa + min(t, nact - 1) * S, lag=lag,
                       base=lag >> 1, par=lag & 1, lastT=(t == nact - 1),
                       ring=None, prev=np.zeros(S, np.float32)))
    edgePrev = (y0 > 0) if FWD else (y1 < h - 1)

    def load_row(T, r):
        # unconditional, for active and inactive threads alike (strip clamped)
        y = y0 + r if FWD else y1 - r
        if FWD:
            return scov.load(shear_at(T['xl'], y, S, dp), S)
        return dist.load(shear_at(T['xl'], y, S, dp), S)

    def fetch(T, j):
        return [load_row(T, min(max(P * j + k - T['base'], 0), rows - 1)) for k in range(P)]

    def init_of(T, raw, s):
        if T['xs'] + s >= w: return f32(0)
        if not FWD: return raw[s]
        return kInf if raw[s] else f32(0)

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
                hasPrev = (r > 0) or edgePrev
                cur = np.zeros(S, np.float32)
                xs = T['xs']; prev = T['prev']; raw = T['ring'][k]
                if FWD:
                    carryL = snap_last[cb][t - 1] if t > 0 else f32(0)
                    prevL = snap_last[pb][t - 1] if (t > 0 and r > 0) else f32(0)
                    prevR = snap_first[pb][t + 1] if (not T['lastT'] and r > 0) else f32(0)
                    for s in range(S):
                        x = xs + s
                        b = init_of(T, raw, s)
                        if x > 0: b = min(b, f32((carryL if s == 0 else cur[s - 1]) + f32(1)))
                        if hasPrev:
                            b = min(b, f32(prev[s] + f32(1)))
                            if x > 0: b = min(b, f32((prevL if s == 0 else prev[s - 1]) + D))
                            if x < w - 1: b = min(b, f32((prevR if s == S - 1 else prev[s + 1]) + D))
                        cur[s] = b if x < w else f32(0)
                else:
                    carryR = snap_first[cb][t + 1] if not T['lastT'] else f32(0)
                    prevR = snap_first[pb][t + 1] if (not T['lastT'] and r > 0) else f32(0)
                    prevL = snap_last[pb][t - 1] if (t > 0 and r > 0) else f32(0)
                    for s in range(S - 1, -1, -1):
                        x = xs + s
                        b = init_of(T, raw, s)
                        if x < w - 1: b = min(b, f32((carryR if s == S - 1 else cur[s + 1]) + f32(1)))
                        if hasPrev:
                            b = min(b, f32(prev[s] + f32(1)))
                            if x < w - 1: b = min(b, f32((prevR if s == S - 1 else prev[s + 1]) + D))
                            if x > 0: b = min(b, f32((prevL if s == 0 else prev[s - 1]) + D))
                        cur[s] = b if x < w else f32(0)
                assert xs + S <= dp  # the strip stays inside its row
                dist.store(shear_at(xs, y, S, dp), cur)  # the whole strip, as the kernel does
                T['prev'] = cur.copy()
                firstv[cb][t] = cur[0]; lastv[cb][t] = cur[S - 1]
        for T in th:
            T['ring'] = T['next']
        T0 += 2 * P


def run_chamfer(cov, box, S, P):
    """Shear pre-pass, forward, backward; returns the distances read back
    at every (x, y) in the box as the ownership kernels read them."""
    h, w = cov.shape
    nt = ((w + S - 1) // S + 127) // 128 * 128
    rows = shear_rows(w, h)
    scov = Flat(rows * pitch(w), np.uint8, 0x5A)
    dist = Flat(rows * pitch(w), np.float32, -5)
    shear_cov(cov, scov, box, w, h, S)
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
    ys, xs = np.nonzero(cov)
    if len(xs) == 0:
        return
    box = (xs.min(), ys.min(), xs.max(), ys.max())
    got, acc = run_chamfer(cov, box, S, P)
    assert np.array_equal(got, po.coverage_distance(cov))
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
