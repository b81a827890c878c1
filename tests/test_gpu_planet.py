"""The planet-wide workload at N_c = 8192 on one GPU: 6 x 8192^2 = 402.7 M
cells, fBm bedrock of 10 octaves, latitudinal rain, one droplet per cell
(402.7 M droplets), lifetime 256, the fast mode -- inputs built on the
device, one erosion iteration.  The fixed-point mass budget closes exactly,
every droplet dies exactly once, and the peak of device memory stays inside
the part's 288 GB."""
from __future__ import annotations

import time

import numpy as np
import pytest
import torch

from globe_hip import _lib, stubs
from globe_hip.erosion import fast, maps

pytestmark = pytest.mark.gpu

FS = _lib.FS
HBM_BYTES = 288e9


def test_planet_n8192_iteration(cuda):
    N, H = 8192, 2
    torch.cuda.empty_cache()
    torch.cuda.reset_peak_memory_stats()
    t0 = time.time()
    grid, f, metric, metric_inv = stubs.planet_globe(N, H, octaves=10, device=cuda)
    p = stubs.workload_params(N, H, max_steps=256)
    st = maps.ErosionState.from_grid(grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"],
                                     p.erosion, device=cuda, mode="fast", metric=metric, metric_inv=metric_inv)
    del f, metric, metric_inv
    torch.cuda.synchronize()
    land = float((st.height[st.interior] >= 0).double().mean())
    print(f"planet N={N}: inputs on the device in {time.time() - t0:.1f}s, land {100 * land:.1f} %", flush=True)
    assert abs(land - 0.30) < 1e-3
    t1 = time.time()
    s = fast.run_iteration(st, p, 0)
    torch.cuda.synchronize()
    dt = time.time() - t1
    peak = torch.cuda.max_memory_allocated()
    print(f"planet N={N}: one iteration {dt:.1f}s, {s['particles']} droplets, {s['moves']} moves "
          f"({s['moves'] / dt:.3g} moves/s), tape {'yes' if getattr(st._fast, 'tape', None) is not None else 'no (traced twice)'}; "
          f"peak device memory {peak / 1e9:.1f} GB of {HBM_BYTES / 1e9:.0f}", flush=True)
    r = s["raw_fx"]
    lhs = (int(r[FS["DEPOSITED_FX"]]) - int(r[FS["ERODED_FX"]]) + int(r[FS["TO_PENDING_FX"]])
           + int(r[FS["LOST_OFFSHORE_FX"]]) + int(r[FS["EXIT_FX"]]) + int(r[FS["VANISHED_FX"]]))
    assert lhs == int(r[FS["RELEASED_FX"]])
    assert s["particles"] == 6 * N * N
    assert sum(s["deaths"].values()) == s["particles"]
    assert s["tape_short"] == 0
    assert s["moves"] > 50 * s["particles"]
    assert peak < HBM_BYTES
    h = st.height[st.interior]
    assert bool(torch.isfinite(h).all())
    assert np.isfinite(float(st.sediment.sum()))


def test_planet_n8192_adjoint_in_pieces(cuda):
    """The differentiable mode at N_c = 8192: one iteration, float32, the
    hypsometry loss, the tape recorded and reversed in pieces sized to the
    memory left (the whole tape would take ~5.8 TB).  The erodibility
    gradient is finite and reaches the land; dL/d(erodibility scalar)
    agrees with central differences of the frozen-trajectory map (relative
    step 1e-3) within 1 %; device memory stays inside the part."""
    from globe_hip.erosion import adjoint as adj

    N, H = 8192, 2
    torch.cuda.empty_cache()
    torch.cuda.reset_peak_memory_stats()
    grid, f, metric, metric_inv = stubs.planet_globe(N, H, octaves=10, device=cuda)
    p = stubs.workload_params(N, H, max_steps=256)
    st = maps.ErosionState.from_grid(grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"],
                                     p.erosion, device=cuda, mode="fast", metric=metric, metric_inv=metric_inv)
    del f, metric, metric_inv
    run = adj.GradRun(st, p, iterations=1, real="f32", checkpoint_every=1, tape_batch="auto")
    erod = torch.ones(tuple(st.height.shape), dtype=torch.float32, device=cuda)
    theta = adj.theta_nominal(run)
    t0 = time.time()
    out = run.forward(erod, theta, record=True)
    s = (out["height"] + out["sediment"]).detach().requires_grad_(True)
    del out
    torch.cuda.synchronize()
    t_fwd = time.time() - t0
    land = (s[st.interior].reshape(-1) * st.height_unit_m > 0).detach()
    loss = adj.hypsometry_loss(s, run, land=land)
    loss.backward()
    g_s = s.grad
    del s
    torch.cuda.empty_cache()
    t1 = time.time()
    g_erod, g_theta, lam = run.backward(g_s, erod, theta)
    torch.cuda.synchronize()
    t_bwd = time.time() - t1
    n_pieces = -(-run.P // (run.tape_batch or run.P))
    peak = torch.cuda.max_memory_allocated()
    print(f"planet adjoint N={N}: forward {t_fwd:.1f}s, backward {t_bwd:.1f}s in {n_pieces} tape pieces of "
          f"{run.tape_batch} droplets ({run.tape_bytes / 1e9:.1f} GB); peak device memory {peak / 1e9:.1f} GB",
          flush=True)
    assert n_pieces > 1
    assert bool(torch.isfinite(g_erod).all()) and bool(torch.isfinite(g_theta).all())
    inter = torch.zeros(st.height.shape, dtype=torch.bool, device=cuda)
    inter[st.interior] = True
    assert int(((g_erod != 0) & inter).sum()) > 0.1 * int(land.sum())
    del lam, g_s, g_erod
    run._tapes = {}
    run._rev_work = None
    torch.cuda.empty_cache()
    h = 1e-3 * float(theta[0])
    vals = []
    for sgn in (1, -1):
        th = theta.clone()
        th[0] += sgn * h
        o = run.replay(erod, th)
        vals.append(float(adj.hypsometry_loss(o["height"] + o["sediment"], run, land=land)))
        del o
    fd = (vals[0] - vals[1]) / (2 * h)
    print(f"planet adjoint N={N}: dL/d erodibility adjoint {float(g_theta[0]):.6g}, central difference {fd:.6g}",
          flush=True)
    assert abs(float(g_theta[0]) - fd) <= 1e-2 * abs(fd)
    assert peak < torch.cuda.get_device_properties(cuda).total_memory
