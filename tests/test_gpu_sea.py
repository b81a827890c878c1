"""The sea mask on the GPU (pg_refresh_base: union-find components, exact
areas, basin floors, halo copy) against the oracle's restatement of
maps.py:360-421: base levels bit for bit, the same counts; and ``step``
with the refresh on, in both modes and under graph capture."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from globe_hip import stubs
from globe_hip.cubesphere import get_grid
from globe_hip.erosion import fast, maps

pytestmark = pytest.mark.gpu


def _pair(oracle, cuda, grid, bed_m, p):
    z = np.zeros(bed_m.shape, np.float32)
    args = (grid, bed_m, z, np.ones(bed_m.shape, np.float32), np.ones(bed_m.shape, np.float32),
            np.zeros(bed_m.shape))
    ost = oracle.OracleState.from_grid(*args, p.erosion)
    dst = maps.ErosionState.from_grid(*args, p.erosion, device=cuda)
    return ost, dst


def _cmp(oracle, ost, dst, frac=0.02):
    oi = oracle.refresh_base(ost, frac)
    di = dst.refresh_base(frac)
    b = dst.base.cpu().numpy()
    assert np.array_equal(b, ost.base), int((b != ost.base).sum())
    assert di == oi, (di, oi)
    return oi


def test_closed_basin_globe_matches_oracle(oracle, cuda):
    """test_hydro.py:292-309's globe: face 0 sea, a walled basin on face 1."""
    g = get_grid(16, 4, 1000.0)
    H, N = g.H, g.N
    bed = np.full((6, g.NE, g.NE), 100.0)
    bed[0, H:H + N, H:H + N] = -500.0
    bed[1, H + 6:H + 10, H + 6:H + 10] = -300.0 + np.arange(16).reshape(4, 4)
    p = stubs.workload_params(N, H)
    ost, dst = _pair(oracle, cuda, g, bed, p)
    info = _cmp(oracle, ost, dst)
    assert info["closed_basins"] == 1 and info["closed_cells"] == 16 and info["sea_cells"] == N * N
    assert info["deepest_basin_floor"] == -300.0 / 1000.0
    # the largest body is the sea however small the threshold makes it
    info = _cmp(oracle, ost, dst, 0.99)
    assert info["sea_cells"] == N * N


@pytest.mark.parametrize("N,seed,level,rough", [(64, 0, 0.3, 0.0), (128, 1, 0.45, 0.0), (128, 2, 0.6, 0.0),
                                                 (64, 3, 0.5, 0.4)])
def test_noisy_globes_match_oracle(oracle, cuda, N, seed, level, rough):
    """fBm globes shifted so 30-60 % of the cells are below the waterline
    (one with white noise on top): one ocean and from a few to hundreds of
    closed basins, some of them across face edges and corners."""
    grid, f = stubs.plain_globe(N, 2, seed=seed)
    bed = np.asarray(f["bedrock_m"], np.float64)
    inter = (slice(None), slice(2, 2 + N), slice(2, 2 + N))
    bed = bed + rough * bed[inter].std() * np.random.default_rng(seed).normal(size=bed.shape)
    bed = bed - np.quantile(bed[inter], level)
    p = stubs.workload_params(N, 2)
    ost, dst = _pair(oracle, cuda, grid, bed, p)
    info = _cmp(oracle, ost, dst)
    print(f"N={N} seed={seed}: {info}", flush=True)
    assert info["closed_basins"] >= 3 and info["sea_cells"] > 0


def test_step_refreshes_the_sea_like_the_oracle(oracle, hiplib, cuda):
    """Check mode, sea_mask_every 2 on a globe with a dug basin: the refresh
    runs before the particles of iterations 0 and 2, and with the oracle on
    the device's libm the base level and the state match it bit for bit
    after every step."""
    N = 48
    grid, f = stubs.plain_globe(N, 2)
    bed = np.asarray(f["bedrock_m"], np.float64).copy()
    land = bed[:, 2:2 + N, 2:2 + N] >= 0
    fi, ii, jj = np.argwhere(land)[len(np.argwhere(land)) // 2]
    bed[fi, 2 + ii - 1:2 + ii + 2, 2 + jj - 1:2 + jj + 2] = -50.0
    p = stubs.workload_params(N, 2, sea_mask_every=2)
    ost, dst = _pair(oracle, cuda, grid, bed, p)
    oracle.set_libm(*oracle.device_libm(hiplib))
    try:
        for it in range(3):
            so = oracle.step(ost, p, it)
            sd = maps.step(dst, p, it)
            if it % 2 == 0:
                assert "sea" in so and sd["sea"] == so["sea"]
            assert ("sea" in sd) == (it % 2 == 0)
            assert np.array_equal(dst.base.cpu().numpy(), ost.base), it
            for n in ("height", "sediment", "pending"):
                assert np.array_equal(getattr(dst, n).cpu().numpy(), getattr(ost, n)), (it, n)
    finally:
        oracle.set_libm()


def test_graph_replays_refresh_the_sea_like_eager(cuda):
    """Fast mode, sea_mask_every 2: the refresh runs eagerly before the
    replays it is due on; states bitwise equal to eager steps."""
    N = 48
    grid, f = stubs.plain_globe(N, 2)
    p = stubs.workload_params(N, 2, sea_mask_every=2)
    args = (grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"], p.erosion)
    a = maps.ErosionState.from_grid(*args, device=cuda, mode="fast")
    b = maps.ErosionState.from_grid(*args, device=cuda, mode="fast")
    g = fast.IterationGraph(b, p)
    for it in range(4):
        maps.step(a, p, it)
        g.step(it)
        torch.cuda.synchronize()
        assert a.base_at == b.base_at
        for name in ("height", "sediment", "base", "discharge"):
            assert torch.equal(getattr(a, name), getattr(b, name)), (it, name)
