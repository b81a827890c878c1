"""The erosion brush on the GPU (no counterpart in the reference; the
oracle's numpy scatter is the check):

* pg_brush_spread equals the oracle's scatter through the owner table bit
  for bit, on a globe (halo shares folded onto their owners) and on a window
  with holes in its mask, for every radius and LDS tile;
* in the fast mode radius 1 (the identity, through the brush kernels) is
  radius 0 bit for bit; radii 2 and 3 keep the fixed-point mass budget
  exact, give the same bits on every tile, and smooth the terrain's change;
* batched windows with a brush equal one fast-mode pass per window.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from globe_hip import _lib, stubs
from globe_hip.cubesphere import get_grid
from globe_hip.erosion import batch as bt
from globe_hip.erosion import fast, maps

pytestmark = pytest.mark.gpu

FS = _lib.FS


def _sources(shape, mask, rng, frac=0.2, mag=1 << 44):
    src = rng.integers(-mag, mag, size=shape).astype(np.int64)
    src[(rng.random(shape) > frac) | (mask != 1)] = 0
    return src


@pytest.mark.parametrize("radius", [0, 1, 2, 3, 4])
def test_spread_equals_the_oracle_on_a_globe(oracle, cuda, radius):
    g = get_grid(24, 2, 1000.0)
    NE = g.NE
    mask = np.zeros((6, NE, NE), np.uint8)
    mask[:, 2:-2, 2:-2] = 1
    rng = np.random.default_rng(radius)
    src = _sources(mask.shape, mask, rng)
    # sources along every face edge and corner: their footprints reach into the halo
    src[:, 2, 2:-2] = rng.integers(-(1 << 44), 1 << 44, size=(6, NE - 4))
    src[:, 2:-2, NE - 3] = rng.integers(-(1 << 44), 1 << 44, size=(6, NE - 4))
    exp = oracle.brush_spread(src, mask, g.owner, radius)
    d_src = torch.from_numpy(src).to(cuda)
    d_mask = torch.from_numpy(mask).to(cuda)
    d_own = torch.from_numpy(np.ascontiguousarray(g.owner, np.int64)).to(cuda)
    for tile in range(3):
        out = fast.brush_spread(d_src, d_mask, d_own, radius, tile).cpu().numpy()
        assert np.array_equal(out, exp), (radius, tile, int(np.abs(out - exp).max()))
    assert int(exp.sum()) == int(src.sum())


@pytest.mark.parametrize("radius", [2, 3, 4])
def test_spread_equals_the_oracle_on_a_window_with_holes(oracle, cuda, radius):
    NE = 70
    rng = np.random.default_rng(10 + radius)
    mask = np.zeros((1, NE, NE), np.uint8)
    mask[0, 2:-2, 2:-2] = 1
    mask[0, 20:30, 33:41] = 0  # a hole of outside cells
    mask[0, rng.random((NE, NE)) < 0.03] = 0
    src = _sources(mask.shape, mask, rng, frac=0.5)
    exp = oracle.brush_spread(src, mask, None, radius)
    for tile in range(3):
        out = fast.brush_spread(torch.from_numpy(src).to(cuda), torch.from_numpy(mask).to(cuda), None, radius,
                                tile).cpu().numpy()
        assert np.array_equal(out, exp), (radius, tile)
    assert int(exp.sum()) == int(src.sum())
    assert not exp[0][mask[0] != 1].any()


def _state(cuda, N=64, radius=0):
    grid, f = stubs.plain_globe(N, 2)
    p = stubs.workload_params(N, 2, brush_radius=radius)
    st = maps.ErosionState.from_grid(grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"],
                                     p.erosion, device=cuda, mode="fast")
    return p, st


def _budget_closes(s):
    r = s["raw_fx"]
    lhs = (int(r[FS["DEPOSITED_FX"]]) - int(r[FS["ERODED_FX"]]) + int(r[FS["TO_PENDING_FX"]])
           + int(r[FS["LOST_OFFSHORE_FX"]]) + int(r[FS["EXIT_FX"]]) + int(r[FS["VANISHED_FX"]]))
    return lhs == int(r[FS["RELEASED_FX"]])


def _fields(st):
    return [t.cpu().numpy().copy() for t in (st.height, st.sediment, st.discharge, st.momentum, st.pending)]


def test_radius_one_is_radius_zero(cuda):
    outs = []
    for r in (0, 1):
        p, st = _state(cuda, radius=r)
        for it in range(3):
            s = fast.run_iteration(st, p, it)
            assert _budget_closes(s)
        outs.append(_fields(st))
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("radius", [2, 3])
def test_brush_budget_is_exact_and_tile_free(cuda, radius):
    outs = []
    for tile in range(3):
        p, st = _state(cuda, radius=radius)
        st.fast_brush_tile = tile
        for it in range(3):
            s = fast.run_iteration(st, p, it)
            assert _budget_closes(s), (radius, tile, it)
            assert sum(s["deaths"].values()) == s["particles"]
        outs.append(_fields(st))
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            np.testing.assert_array_equal(a, b)
    # the brush changes the result and spreads it: the change reaches more cells, and it is
    # smoother (a smaller mean absolute difference between neighbours, relative to its size)
    p0, s0 = _state(cuda, radius=0)
    h_init = (s0.height + s0.sediment).clone()
    for it in range(3):
        fast.run_iteration(s0, p0, it)
    d0 = (s0.height + s0.sediment - h_init)[s0.interior].cpu().numpy()
    db = (torch.from_numpy(outs[0][0] + outs[0][1]).to(cuda) - h_init)[s0.interior].cpu().numpy()
    assert np.count_nonzero(db) > np.count_nonzero(d0)

    def roughness(d):
        return float(np.abs(np.diff(d, axis=-1)).mean() / max(np.abs(d).mean(), 1e-30))

    assert roughness(db) < roughness(d0)


def test_batched_windows_with_a_brush_equal_per_window(cuda):
    def make():
        wins = stubs.zoom_windows(5, seed=8, side_lo=24, side_hi=80, life_lo=16, life_hi=64, brush_radius=3)
        return bt.WindowBatch([w for w, _ in wins], [p for _, p in wins], device=cuda, chunks=4)

    a, b = make(), make()
    a.brush_tile = 2
    for it in range(3):
        a.step(it)
        b.step_per_window(it)
        torch.cuda.synchronize()
        for n in ("height", "sediment", "discharge", "momentum", "pending"):
            x, y = getattr(a, n), getattr(b, n)
            assert torch.equal(x, y), (it, n, float((x - y).abs().max()))


def test_autotuner_picks_a_brush_tile_and_keeps_results(cuda):
    def make():
        wins = stubs.zoom_windows(4, seed=9, side_lo=32, side_hi=64, life_lo=16, life_hi=48, brush_radius=2)
        return bt.WindowBatch([w for w, _ in wins], [p for _, p in wins], device=cuda, chunks=2)

    a, b = make(), make()
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    tuner = bt.Autotuner(a, candidates=[(cus, 0), (2 * cus, 1)])
    assert tuner.tiles == [0, 1, 2]
    it = 0
    while tuner.best is None:
        tuner.step(a.step, it)
        b.step(it)
        it += 1
    assert it == 5 and tuner.best_tile in (0, 1, 2) and a.brush_tile == tuner.best_tile
    assert all(v > 0 for v in tuner.tile_times.values())
    a.step(it)
    b.step(it)
    torch.cuda.synchronize()
    for n in ("height", "sediment", "discharge", "momentum", "pending"):
        assert torch.equal(getattr(a, n), getattr(b, n)), n
