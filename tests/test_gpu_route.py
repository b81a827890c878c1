"""The routing flood on the GPU against the oracle's heap flood.

Bar: the filled surface bit for bit on every cell the flood assigns (the
device relaxes the flood's fixed point; its values are the flood's own
chains m + eps), the halos refreshed by the linear exchange bit for bit
(both sum the sources in source order); a window's step with the flood and
its lakes bit for bit (no halo exchange there); the reference's window-lakes
test on the device.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from globe_hip import stubs
from globe_hip.erosion import maps
from windows import relief_globe, window_inputs, window_params

pytestmark = pytest.mark.gpu


def _globe_pair(oracle, cuda, N=48, seed=0, mode="check"):
    g, f = relief_globe(N, seed=seed)
    p = stubs.workload_params(N, 2)
    args = (g, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"], p.erosion)
    return oracle.OracleState.from_grid(*args), maps.ErosionState.from_grid(*args, device=cuda, mode=mode), p


def _cmp_route(ost, dst):
    a = dst.route.cpu().numpy()
    b = ost.route
    inter = ost.interior
    assert np.array_equal(a[inter], b[inter])
    assert np.array_equal(a, b)


def test_globe_route_matches_the_heap_flood(oracle, cuda):
    ost, dst, p = _globe_pair(oracle, cuda)
    eps = 0.05 / 50.0
    sweeps = dst.refresh_route(eps)
    ost.refresh_route(eps)
    _cmp_route(ost, dst)
    filled = (ost.route > ost.surface())[ost.interior].sum()
    print(f"globe N=48: {filled} cells filled, {sweeps} sweeps", flush=True)
    assert filled > 50


@pytest.mark.parametrize("seed", [0, 1])
def test_window_route_with_holes_matches(oracle, cuda, seed):
    p = window_params(seed)
    kw = window_inputs(48, "tilt", p, seed=seed, relief=12.0)
    rng = np.random.default_rng(seed)
    kw["height_m"] = kw["height_m"] - 200.0 * (rng.random(kw["height_m"].shape) < 0.05)  # pits
    kw["mask"] = kw["mask"].copy()
    kw["mask"][rng.integers(6, 50, 12), rng.integers(6, 50, 12)] = 0                     # holes
    ost = oracle.OracleState.window(**kw)
    dst = maps.ErosionState.window(**kw, device=cuda)
    ost.refresh_route(0.001)
    dst.refresh_route(0.001)
    assert np.array_equal(dst.route.cpu().numpy(), ost.route)
    assert (ost.route > ost.surface()).sum() > 10


def test_route_at_n256(oracle, cuda):
    grid, f = stubs.earth_globe(256, 2, n_blobs=16)
    p = stubs.workload_params(256, 2)
    args = (grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"], p.erosion)
    ost = oracle.OracleState.from_grid(*args)
    dst = maps.ErosionState.from_grid(*args, device=cuda)
    eps = 0.05 / grid.cell_size_m
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    sweeps = dst.refresh_route(eps)
    ev[1].record()
    torch.cuda.synchronize()
    print(f"N=256 route: {sweeps} sweeps, {ev[0].elapsed_time(ev[1]):.1f} ms", flush=True)
    ost.refresh_route(eps)
    _cmp_route(ost, dst)


def test_window_steps_with_flood_and_lakes_bit_exact(oracle, cuda):
    """Three check-mode window iterations with the flood every 2 and the
    window's lakes: change lists and state bit for bit."""
    p = window_params(0, window_lakes=True, flood_every=2, route_eps=0.05)
    kw = window_inputs(40, "tilt", p, seed=2, relief=10.0, deposit_on_exit=True)
    kw["height_m"] = kw["height_m"].copy()
    kw["height_m"][20:26, 20:26] -= 120.0
    ost = oracle.OracleState.window(**kw)
    dst = maps.ErosionState.window(**kw, device=cuda)
    for it in range(3):
        so = oracle.step(ost, p, it)
        sd = maps.step(dst, p, it)
        assert sd["particles"] == so["particles"] and sd["entries"] == so["entries"], it
        for n in ("height", "sediment", "pending", "discharge", "route"):
            assert np.array_equal(getattr(dst, n).cpu().numpy(), getattr(ost, n)), (it, n)
        assert np.array_equal(dst.lake_flag.cpu().numpy(), ost.lake_flag), it
    assert ost.lake_flag.any()


def test_window_lakes_flag_the_depressions_the_route_fills(cuda):
    """bake/tests/test_erosion.py:1175-1199 on the device's step."""
    from globe_hip.erosion import particle as pk

    runs = {}
    for name, on in (("off", False), ("on", True)):
        p = window_params(0, window_lakes=on, flood_every=1 if on else 0)
        kw = window_inputs(32, "tilt", p, seed=5)
        st = maps.ErosionState.window(**kw, device=cuda)
        H = st.H
        c = H + 16
        st.height[0, c - 2:c + 2, c - 2:c + 2] -= 2.0
        maps.step(st, p, 0)
        runs[name] = st
    off, on_st = runs["off"], runs["on"]
    assert off.lake_flag is None or not off.lake_flag.any().item()
    flag = on_st.lake_flag.cpu().numpy()
    assert flag[0, c - 1:c + 1, c - 1:c + 1].all()
    assert flag.sum() < 200
    samp = on_st.samp.cpu().numpy()
    assert np.all(samp[0][flag[0] > 0, pk.NS - 2] == 2.0)


def test_globe_steps_with_flood_stay_close(oracle, cuda):
    ost, dst, p = _globe_pair(oracle, cuda, N=32, seed=5)
    p.erosion.flood_every = 2
    for it in range(3):
        so = oracle.step(ost, p, it)
        sd = maps.step(dst, p, it)
        assert sd["particles"] == so["particles"]
    h_o = ost.height + ost.sediment
    h_d = (dst.height + dst.sediment).cpu().numpy()
    inter = ost.interior
    assert np.abs(h_d[inter] - h_o[inter]).max() <= 1e-6 * np.abs(h_o[inter]).max()


def test_differentiable_mode_refuses_the_route(cuda):
    """The fast mode steers on the route (tests/test_gpu_fast_route.py); the
    differentiable mode, which replays frozen walks on the terrain, does not
    take one."""
    from globe_hip.erosion import adjoint

    p = window_params(0, flood_every=1)
    kw = window_inputs(32, "tilt", p, seed=1)
    st = maps.ErosionState.window(**kw, device=cuda, mode="fast")
    maps.step(st, p, 0)
    assert st.route is not None
    p0 = window_params(0)
    with pytest.raises(NotImplementedError):
        adjoint.GradRun(st, p0, 2)
