"""The oracle's erosion brush (no counterpart in the reference: its droplets
change one cell per step, bake/globe/erosion/particle.py:660-667), pinned
by closed forms: which offsets a radius reaches, a lone source's footprint,
exact conservation, shares without a target kept by their source, halo
shares landing on the owner cell, and a vectorised restatement of the rule
on an all-active field."""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import PKG_DIR  # noqa: F401  (import paths)
from oracle import host


def test_footprint_support_per_radius():
    # radius r reaches the offsets with |o| < r: 1 -> the cell, 2 -> 3x3,
    # 3 -> 5x5 (2 sqrt 2 < 3), 4 -> 7x7 without the four corners (3 sqrt 2 > 4)
    assert host.brush_weights(0).shape == (1, 1) and host.brush_weights(1).shape == (1, 1)
    for r, K in ((2, 3), (3, 5), (4, 7)):
        W = host.brush_weights(r)
        assert W.shape == (K, K)
        A = (K - 1) // 2
        for dy in range(-A, A + 1):
            for dx in range(-A, A + 1):
                if (dy, dx) == (0, 0):
                    assert W[A, A] == 0
                    continue
                assert (W[dy + A, dx + A] > 0) == (dx * dx + dy * dy < r * r), (r, dy, dx)
        # the eight symmetries of the square
        for T in (W.T, W[::-1], W[:, ::-1], W[::-1, ::-1]):
            assert np.array_equal(T, W)


def test_radius_two_weights_in_closed_form():
    # 3x3: centre 2, edges 1, corners 2 - sqrt 2; total 14 - 4 sqrt 2
    W = host.brush_weights(2)
    tot = 14.0 - 4.0 * math.sqrt(2.0)
    assert W[0, 1] == W[1, 0] == W[1, 2] == W[2, 1] == round(65536 / tot)
    assert W[0, 0] == round(65536 * (2.0 - math.sqrt(2.0)) / tot)
    # the centre keeps about 2 / total of a source
    assert abs((65536 - W.sum()) - 65536 * 2.0 / tot) <= 8


def test_share_is_sign_symmetric_floor():
    assert host.brush_share(65536, 7) == 7
    assert host.brush_share(-65536, 7) == -7
    assert host.brush_share(65535, 1) == 0 and host.brush_share(-65535, 1) == 0
    assert host.brush_share(3 << 40, 12345) == (3 << 40) * 12345 >> 16


def _window_mask(NE, H, active=None):
    m = np.zeros((1, NE, NE), np.uint8)
    m[0, H:NE - H, H:NE - H] = host.MASK_ACTIVE
    if active is not None:
        m[0] = np.where(active, host.MASK_ACTIVE, 0)
    return m


@pytest.mark.parametrize("r", [2, 3, 4])
def test_lone_source_footprint(r):
    NE, H = 20, 2
    W = host.brush_weights(r)
    A = (W.shape[0] - 1) // 2
    src = np.zeros((1, NE, NE), np.int64)
    k = 1000
    src[0, 9, 10] = k << 16  # a multiple of 2^16: every share is exact
    out = host.brush_spread(src, _window_mask(NE, H), None, r)
    blk = out[0, 9 - A:9 + A + 1, 10 - A:10 + A + 1]
    exp = k * W
    exp[A, A] = (k << 16) - k * int(W.sum())
    assert np.array_equal(blk, exp)
    assert int(out.sum()) == k << 16
    assert np.count_nonzero(out) == np.count_nonzero(exp)
    # a negative source mirrors it
    assert np.array_equal(host.brush_spread(-src, _window_mask(NE, H), None, r), -out)


def test_shares_without_target_stay_with_the_source():
    NE, H, r = 16, 2, 3
    src = np.zeros((1, NE, NE), np.int64)
    src[0, H, H] = 5 << 16  # the window's corner cell: 16 of its 24 targets are outside
    out = host.brush_spread(src, _window_mask(NE, H), None, r)
    W = host.brush_weights(r)
    inside = W[2:, 2:].copy()
    inside[0, 0] = 0
    assert int(out.sum()) == 5 << 16
    assert out[0, H, H] == (5 << 16) - 5 * int(inside.sum())
    assert np.count_nonzero(out[0, :H, :]) == 0 and np.count_nonzero(out[0, :, :H]) == 0


def test_random_sources_conserve_and_match_a_vectorised_spread():
    rng = np.random.default_rng(4)
    NE, H, r = 24, 2, 3
    mask = _window_mask(NE, H)
    src = np.zeros((1, NE, NE), np.int64)
    # sources well inside: every target of every source is active
    inner = rng.integers(-(1 << 40), 1 << 40, size=(NE - 8, NE - 8))
    keep = rng.random(inner.shape) < 0.3
    src[0, 4:NE - 4, 4:NE - 4] = np.where(keep, inner, 0)
    out = host.brush_spread(src, mask, None, r)
    assert int(out.sum()) == int(src.sum())
    # the rule restated on whole arrays: shifted floor shares
    W = host.brush_weights(r)
    A = (W.shape[0] - 1) // 2
    s = src[0].astype(object)
    mag = np.abs(s)
    sgn = np.sign(src[0]).astype(object)
    exp = src[0].astype(object).copy()
    for dy in range(-A, A + 1):
        for dx in range(-A, A + 1):
            if dy == 0 and dx == 0:
                continue
            sh = sgn * ((mag * int(W[dy + A, dx + A])) // 65536)
            exp = exp - sh
            exp = exp + np.roll(np.roll(sh, dy, axis=0), dx, axis=1)
    assert np.array_equal(out[0], np.array(exp.tolist(), dtype=np.int64))


def test_halo_shares_land_on_the_owner():
    from globe_hip.cubesphere import get_grid

    g = get_grid(8, 2, 1000.0)
    NE = g.NE
    mask = np.zeros((6, NE, NE), np.uint8)
    mask[:, 2:-2, 2:-2] = host.MASK_ACTIVE
    own = g.owner
    src = np.zeros((6, NE, NE), np.int64)
    src[0, 2, 5] = 7 << 16  # an edge cell of face 0: its row i = 0, 1 targets are halo cells
    out = host.brush_spread(src, mask, own, 3)
    assert int(out.sum()) == 7 << 16
    # nothing stays in a halo; the halo targets' owners on the neighbouring face received
    inter = np.zeros((6, NE, NE), bool)
    inter[:, 2:-2, 2:-2] = True
    assert np.count_nonzero(out[~inter]) == 0
    W = host.brush_weights(3)
    halo_owner_cells = {int(own[0, i, j]) for i in (0, 1) for j in range(3, 8)}
    got = {int(c) for c in np.flatnonzero(out.reshape(-1)) if c // (NE * NE) != 0}
    assert got and got <= halo_owner_cells
    # each halo target's share reaches its owner (several targets may share one owner)
    expect = {}
    for i in (0, 1):
        for j in range(3, 8):
            o = int(own[0, i, j])
            expect[o] = expect.get(o, 0) + 7 * int(W[i - 2 + 2, j - 5 + 2])
    for c, v in expect.items():
        assert int(out.reshape(-1)[c]) == v


def test_library_weights_are_the_oracles(hiplib):
    from globe_hip.erosion import fast

    for r in range(5):
        assert np.array_equal(fast.brush_weights(r), host.brush_weights(r)), r
