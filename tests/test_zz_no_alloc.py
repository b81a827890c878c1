"""The allocation-free parallel apply (pg_apply_changes_ordered_noalloc).

Called through the C ABI on two chunks in a row -- the second smaller -- its
results must equal pg_apply_changes_ordered's (the same apply with a
caller-allocated workspace) to the bit: acc, stockpiles, tracks, packed
record, the chunk's 4-tuple and the status words.  Sizes beyond its scratch
are refused on the host with an error.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from globe_hip import _lib, stubs
from globe_hip.erosion import maps
from globe_hip.erosion import particle as pk

pytestmark = pytest.mark.gpu

MAX_SWEEPS = 256


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _chunk_lists(st, p, c0, m):
    """Trace droplets [c0, c0 + m) of iteration 0's spawn on st's record."""
    ep = p.erosion
    dev = st.device
    rng = p.rng("erosion", 0)
    n_cells = int(np.count_nonzero(st._mask_host[st.interior] == 1))
    P = int(round(ep.particles_per_cell * n_cells))
    sp_face, sp_x, sp_y, volume0 = maps.spawn_particles(st, P, rng)
    st.pack()
    chunk = maps.chunk_size(ep, P)
    max_steps = maps.max_steps_of(ep, st.N)
    cap = pk.change_list_cap(max_steps, False)
    disc_sat, mom_scale = maps.discharge_scales(ep, st, volume0, P)
    prm = maps.trace_params(ep, st, volume0, P, chunk, disc_sat, mom_scale, max_steps, 0.0)
    cl = dict(cell=torch.empty(m * cap, dtype=torch.int32, device=dev),
              delta=torch.empty(m * cap, dtype=torch.float64, device=dev),
              vol=torch.empty(m * cap, dtype=torch.float32, device=dev),
              mom=torch.empty((m * cap, 2), dtype=torch.float32, device=dev),
              count=torch.zeros(m, dtype=torch.int64, device=dev))
    lo, hi = torch.zeros_like(cl["count"]), torch.zeros_like(cl["count"])
    death = torch.zeros(m, dtype=torch.int8, device=dev)
    z = torch.zeros(m, dtype=torch.float64, device=dev)
    pk.trace_particles(sp_face[c0:c0 + m], sp_x[c0:c0 + m], sp_y[c0:c0 + m], z, z, volume0, st.samp, False, st.mask,
                       st.N, st.H, True, prm, None, False, cl["cell"], cl["delta"], cl["vol"], cl["mom"], cl["count"],
                       lo, hi, death)
    return cl, cap


def _apply(L, st, cl, cap, m, ep, work, noalloc):
    """One apply on copies of st's mutable arrays; returns them with out4 and status."""
    unit = st.height_unit_m
    acc, pend, samp = st.acc.clone(), st.pending.clone(), st.samp.clone()
    disch, mom = st.disch_track.clone(), st.mom_track.clone()
    out4 = torch.zeros(4, dtype=torch.float64, device=st.device)
    status = torch.zeros(4, dtype=torch.int32, device=st.device)
    common = [_p(cl["cell"]), _p(cl["delta"]), _p(cl["vol"]), _p(cl["mom"]), _p(cl["count"]), m, cap, _p(st.height),
              _p(st.sediment), _p(acc), _p(pend), _p(samp), _p(disch), _p(mom), _p(st.mask), st.height.numel(),
              float(ep.iter_erode) / unit, float(ep.iter_deposit) / unit, float(ep.fan_slope), 0,
              float(ep.dep_floor_m) / unit, float(ep.offshore_writeoff), None, 0, _p(out4), _p(status)]
    if noalloc:
        rc = L.pg_apply_changes_ordered_noalloc(*common, MAX_SWEEPS, _stream())
    else:
        rc = L.pg_apply_changes_ordered(*common, MAX_SWEEPS, 40, _p(work), work.numel(), _stream())
    _lib.check(rc, "noalloc" if noalloc else "ordered")
    torch.cuda.synchronize()
    return dict(acc=acc, pending=pend, samp=samp, disch=disch, mom=mom, out4=out4, status=status)


def test_noalloc_apply_equals_the_workspace_apply(cuda):
    L = _lib.lib()
    assert L.pg_ordered_scratch_bytes() >= L.pg_ordered_workspace_bytes(2048, 161, MAX_SWEEPS)
    N = 128
    grid, f = stubs.earth_globe(N, 2, n_blobs=16)
    p = stubs.workload_params(N, 2, max_steps=128)
    st = maps.ErosionState.from_grid(grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"],
                                     p.erosion, device=cuda)
    for c0, m in ((0, 2048), (30000, 700)):
        print(f"chunk of {m} droplets from {c0}", flush=True)
        cl, cap = _chunk_lists(st, p, c0, m)
        work = torch.empty(L.pg_ordered_workspace_bytes(m, cap, MAX_SWEEPS), dtype=torch.uint8, device=cuda)
        ref = _apply(L, st, cl, cap, m, p.erosion, work, noalloc=False)
        got = _apply(L, st, cl, cap, m, p.erosion, None, noalloc=True)
        assert int(ref["status"][0]) == 0 and int(ref["status"][1]) > 1
        for k in ref:
            assert torch.equal(ref[k].view(torch.uint8), got[k].view(torch.uint8)), k
        print(f"  identical; {int(got['status'][1])} sweeps, clamped {float(got['out4'][0]):.0f}", flush=True)


def test_noalloc_refuses_sizes_beyond_its_scratch(cuda):
    L = _lib.lib()
    one = torch.zeros(1, dtype=torch.float64, device=cuda)
    st4 = torch.zeros(4, dtype=torch.int32, device=cuda)
    rc = L.pg_apply_changes_ordered_noalloc(*([None] * 5), 1 << 20, 161, *([None] * 8), 1 << 20, 0.0, 0.0, 0.0, 0, 0.0,
                                            0.0, None, 0, _p(one), _p(st4), MAX_SWEEPS, _stream())
    assert rc != 0
    assert b"maximum" in L.pg_last_error()
