"""The oracle's routing flood (route.py:30-117), pinned by the reference's
window-lakes test (bake/tests/test_erosion.py:1175-1199), a dug basin's
closed form (filled to its spill plus eps per cell of flood distance),
the filled surface's own invariants (never below the terrain, every reached
cell drains by at least eps, flooding it again changes nothing), and a
brute-force relaxation of the flood's fixed point on small random terrains
(a different algorithm from the heap: a wrong eps, seed or neighbour rule
makes them differ).
"""
from __future__ import annotations

import numpy as np
import pytest

from globe_hip.cubesphere import get_grid
from windows import small_world, window_inputs

OFFS = [(1, 0), (1, 1), (0, 1), (-1, 1), (-1, 0), (-1, -1), (0, -1), (1, -1)]


def _basin(n=40):
    surf = 80.0 - 1.0 * np.repeat(np.arange(n, dtype=np.float64)[:, None], n, 1)
    surf[15:25, 15:25] -= 25.0
    mask = np.ones((n, n), np.uint8)
    mask[0, :] = mask[-1, :] = mask[:, 0] = mask[:, -1] = 0
    return surf[None].copy(), mask[None].copy()


def _identity(shape):
    return np.arange(int(np.prod(shape)), dtype=np.int64).reshape(shape)


def test_dug_basin_fills_to_its_spill(oracle):
    surf, mask = _basin()
    eps = 0.001
    out = oracle.priority_flood_eps(surf, mask, _identity(surf.shape), 1, 38, eps, np.zeros_like(surf))
    inside = np.zeros(surf.shape, bool)
    inside[0, 15:25, 15:25] = True
    # outside the basin every cell already drains down the slope
    assert np.array_equal(out[~inside], surf[~inside])
    # the spill: the rim cell at i = 25 (surface 55) -- the basin fills to it
    # plus eps per cell of flood distance from the spill
    spill = 80.0 - 25.0
    d = out[inside] - spill
    assert d.min() > 0.0 and d.max() <= eps * 11 + 1e-9
    assert np.all(out >= surf)


def _drains(filled, surf, mask, owner, H, N, eps):
    """Every assigned cell (filled above its surface) has an in-neighbour
    whose value + eps it equals: the flood's drainage."""
    F, NE, _ = filled.shape
    fl = filled.reshape(-1)
    for c in np.flatnonzero((filled > surf).reshape(-1)):
        f, rem = divmod(int(c), NE * NE)
        i, j = divmod(rem, NE)
        best = min(fl[owner.reshape(-1)[(f * NE + i + di) * NE + j + dj]] for di, dj in OFFS)
        if not fl[c] == best + eps:
            return False
    return True


def test_flood_of_the_filled_surface_is_itself_and_drains(oracle):
    p = small_world(0)
    kw = window_inputs(48, "tilt", p, seed=3, relief=12.0)
    surf = (kw["height_m"] / kw["height_unit_m"])[None].astype(np.float64)
    surf[0, 20:26, 20:26] -= 4.0
    mask = kw["mask"][None].astype(np.uint8)
    own = _identity(surf.shape)
    eps = 0.002
    out = oracle.priority_flood_eps(surf, mask, own, 4, 48, eps, np.zeros_like(surf))
    assert (out > surf).sum() >= 9
    assert _drains(out, surf, mask, own, 4, 48, eps)
    again = oracle.priority_flood_eps(out, mask, own, 4, 48, eps, np.zeros_like(surf))
    assert np.array_equal(again, out)


def _relax(surf, mask, owner, H, N, eps, base):
    """The flood's fixed point by Bellman-Ford relaxation (pure numpy)."""
    F, NE, _ = surf.shape
    n = surf.size
    idx = np.arange(n)
    f, rem = np.divmod(idx, NE * NE)
    i, j = np.divmod(rem, NE)
    s = surf.reshape(-1)
    m = mask.reshape(-1)
    own = owner.reshape(-1)
    ring = (i == 0) | (j == 0) | (i == NE - 1) | (j == NE - 1)
    visited0 = (m == 0) | ring
    inter = (i >= H) & (i < H + N) & (j >= H) & (j < H + N)
    seed = inter & ~visited0 & (s < 0) & (s <= base.reshape(-1))
    for di, dj in OFFS:
        q = np.where(visited0, 0, idx + di * NE + dj)
        nb = own[q]
        nf, nrem = np.divmod(nb, NE * NE)
        ni, nj = np.divmod(nrem, NE)
        seed |= inter & ~visited0 & ((m[nb] == 0) | (ni < H) | (ni >= H + N) | (nj < H) | (nj >= H + N))
    W = np.where(seed, s, np.inf)
    src = np.flatnonzero(~visited0)
    edges = []
    for di, dj in OFFS:
        nb = own[src + di * NE + dj]
        ok = ~visited0[nb] & ~seed[nb]
        edges.append((src[ok], nb[ok]))
    es = np.concatenate([e[0] for e in edges])
    ed = np.concatenate([e[1] for e in edges])
    for _ in range(n + 1):
        cand = np.full(n, np.inf)
        np.minimum.at(cand, ed, W[es])
        up = cand + eps
        new = np.where(np.isfinite(cand), np.maximum(s, up), np.inf)
        new = np.where(seed, s, np.minimum(W, new))
        if np.array_equal(new, W):
            break
        W = new
    return np.where(np.isfinite(W), W, s).reshape(surf.shape)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_heap_flood_equals_the_relaxed_fixed_point_on_a_window(oracle, seed):
    rng = np.random.default_rng(seed)
    NE, H = 20, 2
    surf = rng.uniform(-1, 5, (1, NE, NE))
    mask = np.ones((1, NE, NE), np.uint8)
    mask[0, :H, :] = mask[0, -H:, :] = mask[0, :, :H] = mask[0, :, -H:] = 0
    mask[0, rng.integers(H, NE - H, 5), rng.integers(H, NE - H, 5)] = 0
    base = np.where(rng.random((1, NE, NE)) < 0.1, -0.5, 0.0)
    own = _identity(surf.shape)
    eps = 0.01
    a = oracle.priority_flood_eps(surf, mask, own, H, NE - 2 * H, eps, base)
    b = _relax(surf, mask, own, H, NE - 2 * H, eps, base)
    assert np.array_equal(a, b)
    assert (a > surf).any()


def test_heap_flood_equals_the_relaxed_fixed_point_on_the_globe(oracle):
    """Cross-face neighbours through the owner table, a sea to seed from."""
    g = get_grid(16, 2, 50.0)
    rng = np.random.default_rng(4)
    p = g.centers
    surf = 3.0 * np.sin(3 * p[..., 0]) * np.cos(2 * p[..., 2]) + rng.uniform(0, 1.5, p.shape[:3])
    mask = np.ones(surf.shape, np.uint8)
    own = np.ascontiguousarray(g.owner, np.int64)
    eps = 0.05
    a = oracle.priority_flood_eps(surf, mask, own, g.H, g.N, eps, np.zeros_like(surf))
    b = _relax(surf, mask, own, g.H, g.N, eps, np.zeros_like(surf))
    assert np.array_equal(a, b)
    inter = (slice(None), slice(2, -2), slice(2, -2))
    assert (a[inter] > surf[inter]).sum() > 20


def test_window_lakes_flag_the_depressions_the_route_fills(oracle):
    """bake/tests/test_erosion.py:1175-1199 on the oracle's step."""
    from windows import window_params

    runs = {}
    for name, on in (("off", False), ("on", True)):
        p = window_params(0, window_lakes=on, flood_every=1 if on else 0)
        kw = window_inputs(32, "tilt", p, seed=5)
        st = oracle.OracleState.window(**kw)
        H = st.H
        c = H + 16
        st.height[0, c - 2:c + 2, c - 2:c + 2] -= 2.0
        oracle.step(st, p, 0)
        runs[name] = st
    off, on_st = runs["off"], runs["on"]
    assert off.lake_flag is None or not off.lake_flag.any()
    flag = on_st.lake_flag
    assert flag is not None and flag[0, c - 1:c + 1, c - 1:c + 1].all()
    assert flag.sum() < 200
    assert np.all(on_st.samp[0][flag[0] > 0, oracle.S_RFLAG] == 2.0)
