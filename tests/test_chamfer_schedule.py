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


def kernel(cov, dist, box, w, h, S, P, FWD, nt):
    x0, y0, x1, y1 = box
    if x1 < 0: return
    xa = (x0 // S) * S
    nact = (x1 + 1 - xa + S - 1) // S
    rows = y1 - y0 + 1
    kInf = f32(w + h)
    firstv = np.zeros((2, nt), np.float32); lastv = np.zeros((2, nt), np.float32)
    th = []
    for t in range(nt):
        lag = t if FWD else nact - 1 - t
        th.append(dict(t=t, active=t < nact, xs=xa + t * S, lag=lag, base=lag >> 1, par=lag & 1,
                       lastT=(t == nact - 1), ring=np.zeros((P, S), np.float32), prev=np.zeros(S, np.float32)))
    edgePrev = (y0 > 0) if FWD else (y1 < h - 1)
    def load_row(T, r):
        y = y0 + r if FWD else y1 - r
        out = np.zeros(S, np.float32)
        for s in range(S):
            x = T['xs'] + s
            if FWD: out[s] = kInf if (x < w and cov[y, x]) else f32(0)
            else: out[s] = dist[y, x] if x < w else f32(0)
        return out
    def fetch(T, j):
        # rows of superstep j, clamped into the box (unused slots hold clamped rows)
        return np.stack([load_row(T, min(max(P * j + k - T['base'], 0), rows - 1)) for k in range(P)])
    for T in th:
        if T['active']:
            T['ring'] = fetch(T, 0)
    tmax = 2 * (rows - 1) + nact - 1
    T0 = 0
    while T0 <= tmax:
        for T in th:
            if T['active']:
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
                pb, cb = (r - 1) & 1, r & 1
                hasPrev = (r > 0) or edgePrev
                cur = np.zeros(S, np.float32)
                xs = T['xs']; prev = T['prev']; ring = T['ring']
                if FWD:
                    carryL = snap_last[cb][t - 1] if t > 0 else f32(0)
                    prevL = snap_last[pb][t - 1] if (t > 0 and r > 0) else f32(0)
                    prevR = snap_first[pb][t + 1] if (not T['lastT'] and r > 0) else f32(0)
                    for s in range(S):
                        x = xs + s
                        b = ring[k][s]
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
                        b = ring[k][s]
                        if x < w - 1: b = min(b, f32((carryR if s == S - 1 else cur[s + 1]) + f32(1)))
                        if hasPrev:
                            b = min(b, f32(prev[s] + f32(1)))
                            if x < w - 1: b = min(b, f32((prevR if s == S - 1 else prev[s + 1]) + D))
                            if x > 0: b = min(b, f32((prevL if s == 0 else prev[s - 1]) + D))
                        cur[s] = b if x < w else f32(0)
                for s in range(S):
                    if xs + s < w: dist[y, xs + s] = cur[s]
                T['prev'] = cur.copy()
                firstv[cb][t] = cur[0]; lastv[cb][t] = cur[S - 1]
        for T in th:
            if T['active']:
                T['ring'] = T['next']
        T0 += 2 * P



@pytest.mark.parametrize("seed", range(10))
@pytest.mark.parametrize("S,P", [(4, 8), (8, 4), (16, 4)])
def test_schedule_matches_serial_chamfer(seed, S, P):
    rng = np.random.default_rng(seed)
    h, w = int(rng.integers(3, 40)), int(rng.integers(3, 70))
    cov = (rng.random((h, w)) < rng.uniform(0.3, 0.95)).astype(np.uint8)
    if seed % 3 == 0:
        cov[:, : w // 3] = 0  # a box that does not start at the chunk edge
    ys, xs = np.nonzero(cov)
    if len(xs) == 0:
        return
    box = (xs.min(), ys.min(), xs.max(), ys.max())
    nt = ((w + S - 1) // S + 63) // 64 * 64
    dist = np.full((h, w), -5, np.float32)
    kernel(cov, dist, box, w, h, S, P, True, nt)
    kernel(cov, dist, box, w, h, S, P, False, nt)
    inside = np.zeros((h, w), bool)
    inside[box[1]:box[3] + 1, box[0]:box[2] + 1] = True
    assert np.array_equal(np.where(inside, dist, 0), po.coverage_distance(cov))
