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
