"""The oracle's sea mask (maps.py:360-421 refresh_base, on hydro/run.py:64-83
below_sea_components and hydro/lakes.py:19-46 label_components), pinned by
the reference's own closed-basin tests (test_hydro.py:292-309,
test_erosion.py:1109-1136) and by a union-find over the same neighbour
graph -- another algorithm -- on random masks.
"""
from __future__ import annotations

import numpy as np
import pytest

from globe_hip import stubs
from globe_hip.cubesphere import get_grid

OFFS = [(1, 0), (1, 1), (0, 1), (-1, 1), (-1, 0), (-1, -1), (0, -1), (1, -1)]


def _neighbours(g):
    """(M, 8) interior cell ids of every interior cell's D8 neighbours through
    the owner table (hydro/d8.py:71-77)."""
    N, H, NE = g.N, g.H, g.NE
    own = np.asarray(g.owner).reshape(-1)
    f, i, j = np.meshgrid(np.arange(6), np.arange(N), np.arange(N), indexing="ij")
    f, i, j = f.ravel(), i.ravel(), j.ravel()
    out = []
    for di, dj in OFFS:
        o = own[(f * NE + i + H + di) * NE + j + H + dj]
        f2, r = np.divmod(o, NE * NE)
        i2, j2 = np.divmod(r, NE)
        out.append((f2 * N + i2 - H) * N + j2 - H)
    return np.stack(out, 1)


@pytest.mark.parametrize("N,H", [(16, 4), (33, 3), (64, 2)])
def test_neighbour_graph_is_symmetric(N, H):
    """Every neighbour relation through the owner table holds both ways (and
    no cell is its own neighbour): components do not depend on the order a
    search visits them in, which is what lets the device label them by
    union-find."""
    g = get_grid(N, H, 1000.0)
    nb = _neighbours(g)
    M = nb.shape[0]
    c = np.repeat(np.arange(M), 8)
    fwd = set(zip(c.tolist(), nb.ravel().tolist()))
    assert all((b, a) in fwd for a, b in fwd)
    assert not (nb == np.arange(M)[:, None]).any()


def _union_find(mask, nb):
    M = mask.size
    parent = np.arange(M)

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for a in np.flatnonzero(mask):
        for b in nb[a]:
            if mask[b]:
                ra, rb = find(a), find(b)
                if ra != rb:
                    parent[max(ra, rb)] = min(ra, rb)
    return np.array([find(x) if mask[x] else -1 for x in range(M)])


@pytest.mark.parametrize("seed", [0, 1])
def test_labels_are_the_union_find_components(oracle, seed):
    g = get_grid(12, 2, 1000.0)
    rng = np.random.default_rng(seed)
    mask = rng.random(6 * 12 * 12) < 0.45
    labels, n = oracle.label_components(mask, g.owner, g.N, g.H)
    roots = _union_find(mask, _neighbours(g))
    assert np.array_equal(labels < 0, ~mask)
    # the same partition, numbered in order of each component's lowest cell
    lowest = np.full(n, -1)
    for lab in range(n):
        cells = np.flatnonzero(labels == lab)
        lowest[lab] = cells.min()
        assert np.all(roots[cells] == roots[cells[0]])
        assert roots[cells[0]] == cells.min()
    assert np.all(np.diff(lowest) > 0)
    assert n == np.unique(roots[mask]).size and n > 5


def test_a_closed_basin_below_sea_level_is_not_ocean(oracle):
    """test_hydro.py:292-309 on the oracle."""
    g = get_grid(16, 4, 1000.0)
    N = g.N
    surface = np.full((6, N, N), 100.0, np.float32)
    surface[0] = -500.0
    surface[1, 6:10, 6:10] = -300.0

    def ocean(frac):
        below, labels, keep = oracle.below_sea_components(surface, g, frac)
        return (below & keep[np.maximum(labels, 0)]).reshape(6, N, N)

    assert ocean(0.02)[0].all()
    assert not ocean(0.02)[1].any()
    assert ocean(0.99)[0].all()


def _globe(oracle, N=32):
    grid, f = stubs.plain_globe(N, 2)
    p = stubs.workload_params(N, 2)
    return grid, oracle.OracleState.from_grid(grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"],
                                               f["uplift_m"], p.erosion)


def _dry_site(st, size, pad=2):
    land = st.surface()[st.interior] >= 0.0
    w = size + 2 * pad
    for f in range(6):
        for i in range(land.shape[1] - w):
            for j in range(land.shape[2] - w):
                if land[f, i:i + w, j:j + w].all():
                    return f, i + pad, j + pad
    raise AssertionError("no dry site")


def test_refresh_base_separates_a_closed_basin_from_the_sea(oracle):
    """test_erosion.py:1109-1136 on the oracle: a dug basin below sea level
    is given its own floor as a base level; the open ocean keeps 0."""
    g, st = _globe(oracle)
    H = st.H
    face, i0, j0 = _dry_site(st, 5)
    floor = -4.0
    st.height[face, H + i0:H + i0 + 5, H + j0:H + j0 + 5] = floor
    basin = np.zeros(st.height.shape, bool)
    basin[face, H + i0:H + i0 + 5, H + j0:H + j0 + 5] = True
    basin = basin[st.interior].reshape(-1)
    info = oracle.refresh_base(st, 0.02)
    base = st.base[st.interior].reshape(-1)
    now = st.surface()[st.interior].reshape(-1)
    assert info["closed_basins"] >= 1
    assert np.all(base[basin] == floor)
    assert (now[basin] >= base[basin]).all()
    below, labels, keep = oracle.below_sea_components(now.reshape(6, g.N, g.N).astype(np.float32), g, 0.02)
    ocean = below & keep[np.maximum(labels, 0)]
    assert (now[base != 0.0] < 0.0).all()
    assert np.all(base[ocean] == 0.0)
    assert not ocean[basin].any()
    assert info["sea_cells"] == int(ocean.sum())
    assert info["closed_cells"] == int(((now < 0.0) & ~ocean).sum())
    assert info["deepest_basin_floor"] <= floor
    # halo cells mirror their owners exactly
    own = np.asarray(g.owner).reshape(-1)
    assert np.array_equal(st.base.reshape(-1), st.base.reshape(-1)[own])
