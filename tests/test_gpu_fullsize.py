"""The fast mode at the benchmark's full size -- the N_c=1024 earth globe
(6 x 1024^2 cells, 6.3 M droplets per iteration, lifetime 128, 64 uplift
blobs replayed, 16 chunks) -- through the properties that do not depend on
size: the per-iteration mass budget holds exactly in fixed point, no
droplet is left untraced, every droplet dies once, the surface stays finite,
and the result does not depend on the persistent grid, the register budget
or whether the applying pass replays the census's tape or traces again
(bitwise, after two full steps).  The oracle's CPU time at this size is
minutes per iteration; it is compared at N_c=128 elsewhere.
"""
from __future__ import annotations

import hashlib

import pytest
import torch

from globe_hip import _lib, stubs
from globe_hip.erosion import maps

pytestmark = pytest.mark.gpu

N = 1024


@pytest.fixture(scope="module")
def earth():
    grid, f = stubs.earth_globe(N, 2)
    p = stubs.workload_params(N, 2, max_steps=128, uplift_mode="replay", iterations=200)
    return grid, f, p


def _state(earth, cuda, **knobs):
    grid, f, p = earth
    st = maps.ErosionState.from_grid(grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"],
                                     p.erosion, device=cuda, mode="fast")
    st.fast_chunks = 16
    for k, v in knobs.items():
        setattr(st, k, v)
    maps.start_replay(st, p)
    return st


def _digest(st):
    h = hashlib.sha256()
    for t in (st.height, st.sediment, st.discharge, st.momentum, st.pending):
        h.update(t.cpu().numpy().tobytes())
    return h.hexdigest()


def test_full_size_budget_and_deaths(earth, cuda):
    _, _, p = earth
    st = _state(earth, cuda)
    FS = _lib.FS
    for it in range(2):
        print(f"N_c={N} fast step {it}", flush=True)
        s = maps.step(st, p, it)
        r = s["raw_fx"]
        lhs = (int(r[FS["DEPOSITED_FX"]]) - int(r[FS["ERODED_FX"]]) + int(r[FS["TO_PENDING_FX"]])
               + int(r[FS["LOST_OFFSHORE_FX"]]) + int(r[FS["EXIT_FX"]]) + int(r[FS["VANISHED_FX"]]))
        assert lhs == int(r[FS["RELEASED_FX"]]), (it, lhs, int(r[FS["RELEASED_FX"]]))
        assert s["tape_short"] == 0
        assert sum(s["deaths"].values()) == s["particles"] > 6_000_000
        assert 60 < s["moves"] / s["particles"] < 128
    surf = (st.height + st.sediment)[st.interior]
    assert bool(torch.isfinite(surf).all())


def test_full_size_result_is_independent_of_launch_shape(earth, cuda):
    """Two full steps with the autotuner's extremes and with the tape off:
    the same bits (int64 fixed point: no dependence on scheduling)."""
    _, _, p = earth
    digests = []
    for knobs in ({"fast_grid_blocks": 0, "fast_variant": 0},
                  {"fast_grid_blocks": 2560, "fast_variant": 1},
                  {"fast_grid_blocks": 1536, "fast_variant": 2, "fast_tape": False}):
        st = _state(earth, cuda, **knobs)
        for it in range(2):
            maps.step(st, p, it)
        digests.append(_digest(st))
        print(knobs, digests[-1][:16], flush=True)
        del st
        torch.cuda.empty_cache()
    assert len(set(digests)) == 1, digests
