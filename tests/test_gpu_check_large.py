"""Check mode at scale: one iteration of the N_c=512 plain globe (1,572,864
droplets, 768 chunks of 2048, ~1e8 change-list entries) against the oracle
with the correctly rounded tan / atan the device uses.  Bar: every chunk's
per-droplet entry counts and an order-sensitive checksum of each droplet's
cell sequence and volumes identical, the stats identical, and the state
after the iteration bit for bit.  (~2 minutes: the reference's apply is a
serial walk, run on one lane.)"""
from __future__ import annotations

import numpy as np
import pytest

from globe_hip import stubs
from globe_hip.erosion import maps

pytestmark = pytest.mark.gpu


def _digest(store, to_np):
    def diag(state, cl_cell, cl_vol, cl_count, cap, sp_death):
        n = int(cl_count.shape[0])
        cnt = to_np(cl_count).astype(np.int64).copy()
        cells = to_np(cl_cell)[: n * cap].reshape(n, cap).astype(np.int64)
        vols = to_np(cl_vol)[: n * cap].reshape(n, cap).view(np.int32).astype(np.int64)
        live = np.arange(cap)[None, :] < cnt[:, None]
        w = (np.arange(cap, dtype=np.int64) * 2654435761 + 97) % 1000003
        h = ((cells * w[None, :] + vols * (w[None, :] ^ 12345)) * live).sum(axis=1)
        store.append((cnt, h, to_np(sp_death).copy()))
    return diag


def test_globe_512_first_iteration_bit_exact(oracle, hiplib, cuda):
    N = 512
    grid, f = stubs.plain_globe(N, 2)
    p = stubs.workload_params(N, 2)
    args = (grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"], p.erosion)
    ost = oracle.OracleState.from_grid(*args)
    dst = maps.ErosionState.from_grid(*args, device=cuda)
    o, d = [], []
    oracle.set_libm(*oracle.device_libm(hiplib))
    try:
        print(f"globe N={N}: oracle iteration (correctly rounded tan / atan)", flush=True)
        so = oracle.run_iteration(ost, p, 0, diag=_digest(o, lambda x: x))
    finally:
        oracle.set_libm()
    print(f"globe N={N}: check-mode iteration on the device", flush=True)
    sd = maps.run_iteration(dst, p, 0, diag=_digest(d, lambda t: t.cpu().numpy()))
    assert len(o) == len(d)
    bad = sum(int((a[0] != b[0]).sum() + (a[1] != b[1]).sum() + (a[2] != b[2]).sum()) for a, b in zip(o, d))
    print(f"{len(o)} chunks, {so['particles']} droplets, {so['entries']} entries; droplets differing: {bad}", flush=True)
    assert bad == 0
    for k in ("particles", "entries", "deaths", "clamped"):
        assert sd[k] == so[k], k
    for n in ("height", "sediment", "pending", "discharge", "momentum"):
        assert np.array_equal(getattr(dst, n).cpu().numpy(), getattr(ost, n)), n
