"""The GPU render runs a chunk's seam pairs in rounds: a pair's round is one
past the latest round of an earlier pair it does not commute with, and the
pairs of one round run side by side.  Two pairs commute when they share no
contributor, or when their overlap boxes are disjoint: a pair reads and
rewrites owners only of its own two contributors and only inside its
overlap box.  The rounds therefore give the reference's sequential pair
order's result (SeamBlend.cpp:573-584).  Checked here on the oracle (CPU):
the round-grouped order gives the reference order's owner map on seeded
grid scenes, with fewer rounds than the shared-contributor rule alone,
while an order that breaks the order of pairs that do not commute does
not."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.render_cases import grid_case


def _disjoint(a, b):
    return max(a[0], b[0]) > min(a[2], b[2]) or max(a[1], b[1]) > min(a[3], b[3])


def schedule(pairs, spatial=True):
    """Round of each (i, j, overlap box) pair, as schedule_pairs in
    image360_amd/csrc/render_float.hip (spatial=False: the shared-contributor
    rule alone)."""
    rounds = []
    for k, (i, j, ov) in enumerate(pairs):
        r = 0
        for q in range(k):
            a, b, oq = pairs[q]
            if not ({a, b} & {i, j}):
                continue
            if spatial and _disjoint(ov, oq):
                continue
            r = max(r, rounds[q] + 1)
        rounds.append(r)
    return rounds


def _scene(seed, nx=3, ny=2):
    g, frames = grid_case(seed=seed, nx=nx, ny=ny, fw=120, fh=90)
    c = g.chunks[0]
    warped = []
    for i in c.contributors:
        src = np.clip(frames[i].astype(np.float64) * 65535 + 0.5, 1, 65535).astype(np.uint16)
        hinv = po.invert_homography(g.H[i].astype(np.float32))
        warped.append(po.warp_perspective(src, c.w, c.h, hinv))
    covs = [(x != 0).any(axis=2).astype(np.uint8) for x in warped]
    dists = [po.coverage_distance(cv) for cv in covs]
    boxes = [po.cover_box(cv) for cv in covs]
    pairs = []
    for i in range(len(warped)):
        for j in range(i + 1, len(warped)):
            a, b = boxes[i], boxes[j]
            if a is None or b is None:
                continue
            ov = np.array([max(a[0], b[0]), max(a[1], b[1]), min(a[2], b[2]), min(a[3], b[3])], np.int32)
            if ov[0] <= ov[2] and ov[1] <= ov[3]:
                pairs.append((i, j, ov))
    return warped, dists, boxes, pairs


def _run(warped, dists, boxes, order):
    owner = po.ownership(dists)
    for i, j, ov in order:
        po.refine_seam_dp(warped[i], warped[j], dists[i], dists[j], i, j, boxes[i], boxes[j], ov, owner)
    return owner


@pytest.mark.parametrize("seed,nx,ny", [(7, 3, 2), (8, 3, 2), (9, 3, 2), (7, 4, 1), (8, 4, 1)])
@pytest.mark.parametrize("spatial", [False, True])
def test_round_order_equals_reference_order(seed, nx, ny, spatial):
    warped, dists, boxes, pairs = _scene(seed, nx, ny)
    rounds = schedule(pairs, spatial)
    if spatial or ny > 1:
        assert max(rounds) + 1 < len(pairs)  # some pairs do share a round
    want = _run(warped, dists, boxes, pairs)
    # pairs of one round in reverse order: still the same owner map
    grouped = sorted(range(len(pairs)), key=lambda k: (rounds[k], -k))
    got = _run(warped, dists, boxes, [pairs[k] for k in grouped])
    assert np.array_equal(got, want)


def test_disjoint_overlaps_cut_rounds():
    # a strip of four frames: a frame's pairs with its left and right
    # neighbours overlap on opposite sides of it, so they share a round
    # although they share a contributor
    for seed in (7, 8):
        _, _, _, pairs = _scene(seed, 4, 1)
        assert max(schedule(pairs, True)) < max(schedule(pairs, False))


def test_breaking_shared_order_changes_the_result():
    # the check has teeth: reversing the whole pair order (which reorders
    # pairs sharing a contributor) changes some owner map
    changed = False
    for seed in (7, 8, 9):
        warped, dists, boxes, pairs = _scene(seed)
        want = _run(warped, dists, boxes, pairs)
        changed |= not np.array_equal(_run(warped, dists, boxes, pairs[::-1]), want)
    assert changed


def _line_min(dprev, last_lo, lo, smooth, dt, ranges):
    """Per target lane: minimum over the previous line's candidates of
    dprev + smooth * |t - q|, each range scanned in increasing order with a
    strict <, the ranges' results combined in order with a strict < (one
    range is the serial scan)."""
    out = []
    for lane in range(64):
        t = lo + lane
        best, bp = None, -1
        for p0, p1 in ranges:
            rb, rk = dt(np.finfo(dt).max), -1
            for pk in range(p0, min(p1, len(dprev))):
                cand = dt(dprev[pk] + dt(dt(smooth) * dt(abs(t - (last_lo + pk)))))
                if cand < rb:
                    rb, rk = cand, pk
            if best is None or rb < best:
                best, bp = rb, rk
        out.append((best, bp))
    return out


SPAN, GROUPS = 13, 4  # candidates per lane of the DP kernel, lane groups (rows of a wavefront)


def _line_min_kernel(dprev, last_lo, lo, smooth, dt, combine_le=False):
    """The DP kernel's candidate scan, step for step: the previous line's
    row held in 64 positions with +inf outside its band, GROUPS ranges of
    SPAN candidates each evaluated with no guard, the penalty formed from
    the band shift (lo - last_lo + t - p0) - i, each range reduced as a
    tree (node i absorbs node i + st; the right node's index on a strict <,
    the value by fmin), then the ranges joined in pairs, as the kernel's
    lane swaps do (rows 0-1 and 2-3, then the two halves), the upper range
    winning only on a strict < (combine_le plants a <= there)."""
    row = np.full(64, np.inf, dtype=dt)
    row[: len(dprev)] = dprev

    def join(low, high):
        if high[0] < low[0] or (combine_le and high[0] == low[0]):
            return high
        return low

    out = []
    for t in range(64):
        parts = []
        for g in range(GROUPS):
            p0 = g * SPAN
            d0 = dt(lo - last_lo + t - p0)
            cv = [dt(row[p0 + i] + dt(dt(smooth) * abs(dt(d0 - dt(i))))) for i in range(SPAN)]
            ck = list(range(SPAN))
            st = 1
            while st < SPAN:
                for i in range(0, SPAN - st, 2 * st):
                    ck[i] = ck[i + st] if cv[i + st] < cv[i] else ck[i]
                    cv[i] = np.fmin(cv[i], cv[i + st])
                st *= 2
            parts.append((cv[0], p0 + ck[0]))
        out.append(join(join(parts[0], parts[1]), join(parts[2], parts[3])))
    return out


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("seed", range(6))
def test_split_candidate_scan_keeps_first_minimum(seed, dt):
    """Splitting a line's candidates into 4 ranges of 13 (a band is at most
    49 wide) and joining the partial minima in pairs, lower ranges first,
    selects the same value and the same (first) back-pointer as the
    reference's single serial scan, ties included (integer costs make ties
    frequent)."""
    assert SPAN * GROUPS >= 2 * 24 + 1
    rng = np.random.default_rng(seed)
    for k in range(40):
        last_w = 49 if k < 4 else int(rng.integers(1, 50))
        dprev = rng.integers(0, 6, size=last_w).astype(dt) * dt(2)
        last_lo, lo = int(rng.integers(0, 40)), int(rng.integers(0, 40))
        if k % 3 == 0:
            lo = last_lo  # the common case: the band keeps its offset
        smooth = 2.0 if seed % 2 else 2.0 / 65535.0
        serial = _line_min(dprev, last_lo, lo, smooth, dt, [(0, 64)])
        split = _line_min_kernel(dprev, last_lo, lo, smooth, dt)
        assert serial == split, (k, last_w, last_lo, lo)


def test_kernel_scan_emulation_catches_a_wrong_tie_rule():
    """The emulation is sharp: with a <= planted in the range join, a row
    of tied candidates takes its back-pointer from a later range instead of
    candidate 0."""
    dt = np.float64
    dprev = np.full(40, 4.0, dtype=dt)  # every candidate ties
    serial = _line_min(dprev, 0, 0, 0.0, dt, [(0, 64)])
    assert all(bp == 0 for _, bp in serial)
    assert _line_min_kernel(dprev, 0, 0, 0.0, dt) == serial
    assert _line_min_kernel(dprev, 0, 0, 0.0, dt, combine_le=True) != serial
