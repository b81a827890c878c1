"""The fast mode on the routing surface (particle.py:450-460, 552-560,
626-640, 751-769, 951-955): droplets steer on the epsilon-filled surface
where it stands above the terrain, take the drop on it for the pit test,
cross window lakes without touching the bed, drop their load entering one,
and never erode a lake bed below its water level.

Pinned here by what must hold whatever the float32 rounding: the route
changes the result (the path is live); a flagged lake cell is never eroded
in an iteration; the census's tape replays the traced applying pass bit for
bit; a captured graph, with the flood and the lakes refreshed outside it,
replays the eager steps bit for bit.  Closeness to the oracle's check-mode
semantics on a globe with the flood on: tests/test_gpu_fast.py.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from globe_hip.erosion import fast, maps
from windows import window_inputs, window_params

pytestmark = pytest.mark.gpu


def _window(cuda, lakes=True, flood_every=2, dig=True):
    p = window_params(0, window_lakes=lakes, flood_every=flood_every)
    kw = window_inputs(48, "tilt", p, seed=5, deposit_on_exit=True)
    st = maps.ErosionState.window(**kw, device=cuda, mode="fast")
    if dig:
        H = st.H
        c = H + 24
        st.height[0, c - 3:c + 3, c - 3:c + 3] -= 3.0
    st.fast_chunks = 4
    return p, st


def _state(st):
    return [t.cpu().numpy().copy() for t in (st.height, st.sediment, st.discharge, st.momentum, st.pending)]


def test_route_changes_the_walk(cuda):
    outs = []
    for flood in (0, 2):
        p, st = _window(cuda, lakes=False, flood_every=flood)
        for it in range(3):
            maps.step(st, p, it)
        outs.append(_state(st))
    assert not np.array_equal(outs[0][0], outs[1][0])


def test_lake_beds_are_not_eroded(cuda):
    p, st = _window(cuda)
    for it in range(4):
        maps.refresh_route_and_lakes(st, p)
        flag = st.lake_flag.clone()
        before = (st.height + st.sediment).clone()
        fast.run_iteration(st, p, it)
        st.iteration += 1
        lake = flag > 0
        assert int(lake.sum()) > 4, int(lake.sum())
        after = st.height + st.sediment
        assert bool((after[lake] >= before[lake]).all()), float((before[lake] - after[lake]).max())


@pytest.mark.parametrize("grid_blocks,variant", [(0, 0), (3, 1)])
def test_route_tape_replay_equals_the_traced_applying_pass(cuda, grid_blocks, variant):
    outs = []
    for tape in (False, True):
        p, st = _window(cuda)
        st.fast_tape = tape
        st.fast_grid_blocks = grid_blocks
        st.fast_variant = variant
        for it in range(3):
            maps.step(st, p, it)
        outs.append(_state(st))
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a, b)


def test_route_graph_replays_match_eager(cuda):
    pa, a = _window(cuda)
    pb, b = _window(cuda)
    g = fast.IterationGraph(b, pb)
    for it in range(5):
        maps.step(a, pa, it)
        g.step(it)
        torch.cuda.synchronize()
        for x, y in zip(_state(a), _state(b)):
            np.testing.assert_array_equal(x, y)
        assert torch.equal(a.lake_flag, b.lake_flag) and torch.equal(a.route, b.route)
