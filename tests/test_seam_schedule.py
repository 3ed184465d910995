"""The GPU render runs a chunk's seam pairs in rounds: a pair's round is one
past the latest round of an earlier pair that shares a contributor, and the
pairs of one round run side by side.  That equals the reference's
sequential pair order (SeamBlend.cpp:573-584) because pairs with no
contributor in common commute.  Checked here on the oracle (CPU): the
round-grouped order gives the reference order's owner map on seeded grid
scenes, while an order that breaks the shared-contributor order does not."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.render_cases import grid_case


def schedule(pairs):
    last = {}
    rounds = []
    for i, j in pairs:
        r = max(last.get(i, -1), last.get(j, -1)) + 1
        last[i] = last[j] = r
        rounds.append(r)
    return rounds


def _scene(seed):
    g, frames = grid_case(seed=seed, nx=3, ny=2, fw=120, fh=90)
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


@pytest.mark.parametrize("seed", [7, 8, 9])
def test_round_order_equals_reference_order(seed):
    warped, dists, boxes, pairs = _scene(seed)
    rounds = schedule([(i, j) for i, j, _ in pairs])
    assert max(rounds) + 1 < len(pairs)  # some pairs do share a round
    want = _run(warped, dists, boxes, pairs)
    # pairs of one round in reverse order: still the same owner map
    grouped = sorted(range(len(pairs)), key=lambda k: (rounds[k], -k))
    got = _run(warped, dists, boxes, [pairs[k] for k in grouped])
    assert np.array_equal(got, want)


def test_breaking_shared_order_changes_the_result():
    # the check has teeth: reversing the whole pair order (which reorders
    # pairs sharing a contributor) changes some owner map
    changed = False
    for seed in (7, 8, 9):
        warped, dists, boxes, pairs = _scene(seed)
        want = _run(warped, dists, boxes, pairs)
        changed |= not np.array_equal(_run(warped, dists, boxes, pairs[::-1]), want)
    assert changed
