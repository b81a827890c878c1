"""The check mode's parallel apply (pg_apply_changes_ordered) on the GPU.

It must give the serial apply's results (apply_changes, bake/globe/erosion/
particle.py:821-1042) to the bit: the same acc, stockpiles, tracks, packed
record and the chunk's 4-tuple.  Checked two ways on seeded inputs:

* against the one-lane serial kernel on the device, over full steps
  (windows of each synthetic kind, plain and earth globes): every state array
  and every statistic identical;
* against the oracle (the C restatement of the serial apply) on the earth
  globe's first iteration at N_c=256 and at the benchmark's N_c=1024, with the oracle on the correctly rounded
  tan / atan the device uses: change lists, statistics and state identical.

* on hand-built change lists whose outcome was worked out by hand from the
  reference's rules (tests/apply_cases.py): both device applies give exactly
  those numbers.

Error path: standalone, a sweep limit too small to converge is reported as
an error and nothing is applied (the kernels still exit normally); in the
chunk loop such a chunk, and any chunk with lake parks, goes to the serial
apply on the device, with the same results.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import apply_cases
from globe_hip import stubs
from globe_hip.erosion import maps
from globe_hip.erosion import particle as pk
from windows import window_inputs, window_params

pytestmark = pytest.mark.gpu

STATE = ("height", "sediment", "pending", "discharge", "momentum", "samp", "acc")
STATS = ("particles", "entries", "deaths", "clamped", "to_pending", "deficit_out", "lost_offshore", "pending_total")


def _same_state(a, b):
    for n in STATE:
        x, y = getattr(a, n), getattr(b, n)
        # bitwise, -0.0 and NaN included
        assert torch.equal(x.contiguous().view(torch.uint8), y.contiguous().view(torch.uint8)), n


def _same_stats(sa, sb):
    for k in STATS:
        if k in sa:
            assert sa[k] == sb[k], (k, sa[k], sb[k])


@pytest.mark.parametrize("mode,N", [("tilt", 48), ("closed", 40), ("shelf", 40)])
def test_windows_ordered_equals_serial(cuda, mode, N):
    p = window_params(0)
    kw = window_inputs(N, mode, p, seed=1, deposit_on_exit=True)
    ser = maps.ErosionState.window(**kw, device=cuda)
    ord_ = maps.ErosionState.window(**kw, device=cuda)
    ser.check_apply = "serial"
    for it in range(3):
        print(f"window {mode} N={N} iteration {it}", flush=True)
        sa = maps.run_iteration(ser, p, it)
        sb = maps.run_iteration(ord_, p, it)
        maps.apply_uplift(ser)
        maps.apply_uplift(ord_)
        _same_stats(sa, sb)
        _same_state(ser, ord_)
        print(f"  sweeps max {sb['sweeps_max']} total {sb['sweeps_total']}", flush=True)


@pytest.mark.parametrize("N,kind,iters", [(64, "earth", 3), (128, "plain", 3), (256, "earth", 1)])
def test_globe_steps_ordered_equals_serial(cuda, N, kind, iters):
    grid, f = stubs.earth_globe(N, 2, n_blobs=16) if kind == "earth" else stubs.plain_globe(N, 2)
    p = stubs.workload_params(N, 2, max_steps=128 if kind == "earth" else 64)
    args = (grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"], p.erosion)
    ser = maps.ErosionState.from_grid(*args, device=cuda)
    ord_ = maps.ErosionState.from_grid(*args, device=cuda)
    ser.check_apply = "serial"
    for it in range(iters):
        print(f"globe {kind} N={N} step {it}", flush=True)
        sa = maps.step(ser, p, it)
        sb = maps.step(ord_, p, it)
        _same_stats(sa, sb)
        _same_state(ser, ord_)
        print(f"  {sb['particles']} droplets, {sb['entries']} entries, clamped {sb['clamped']}; "
              f"sweeps max {sb['sweeps_max']} total {sb['sweeps_total']}", flush=True)


@pytest.mark.parametrize("N", [256, 1024])
def test_earth_first_iteration_matches_oracle(oracle, hiplib, cuda, N):
    """The earth workload's first iteration (N_c=1024: the benchmark's own
    size and settings, 6.3 M droplets in 3072 chunks, ~5e8 entries)."""
    grid, f = stubs.earth_globe(N, 2)
    p = stubs.workload_params(N, 2, max_steps=128)
    args = (grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"], p.erosion)
    ost = oracle.OracleState.from_grid(*args)
    dst = maps.ErosionState.from_grid(*args, device=cuda)
    oracle.set_libm(*oracle.device_libm(hiplib))
    try:
        print(f"earth N={N}: oracle iteration", flush=True)
        so = oracle.run_iteration(ost, p, 0)
    finally:
        oracle.set_libm()
    print(f"earth N={N}: device iteration (parallel apply)", flush=True)
    sd = maps.run_iteration(dst, p, 0)
    print(f"  {sd['particles']} droplets, {sd['entries']} entries, clamped {sd['clamped']}, "
          f"sweeps max {sd['sweeps_max']}", flush=True)
    for k in ("particles", "entries", "deaths", "clamped", "to_pending", "deficit_out", "lost_offshore"):
        assert sd[k] == so[k], (k, sd[k], so[k])
    for n in ("height", "sediment", "pending", "discharge", "momentum"):
        assert np.array_equal(getattr(dst, n).cpu().numpy(), getattr(ost, n)), n
    assert np.array_equal(dst.samp.cpu().numpy(), ost.samp)


def _earth(cuda, N=64, **erosion):
    grid, f = stubs.earth_globe(N, 2, n_blobs=16)
    p = stubs.workload_params(N, 2, max_steps=128, **erosion)
    args = (grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"], p.erosion)
    return p, (lambda: maps.ErosionState.from_grid(*args, device=cuda))


def test_sweep_limit_alone_is_an_error_and_applies_nothing(cuda):
    """The standalone parallel apply with a sweep limit too small: status bit
    2, and no state array changes (the kernels still exit normally)."""
    p, mk = _earth(cuda)
    st = mk()
    from test_zz_no_alloc import _chunk_lists

    cl, cap = _chunk_lists(st, p, 0, 2048)
    ep = p.erosion
    out4 = torch.zeros(4, dtype=torch.float64, device=cuda)
    status = torch.zeros(4, dtype=torch.int32, device=cuda)
    work = torch.empty(pk.ordered_workspace_bytes(2048, cap, 1), dtype=torch.uint8, device=cuda)
    before = [t.clone() for t in (st.acc, st.pending, st.samp, st.disch_track, st.mom_track)]
    unit = st.height_unit_m
    pk.apply_changes_ordered(cl["cell"], cl["delta"], cl["vol"], cl["mom"], cl["count"], cap, st.height, st.sediment,
                             st.acc, st.pending, st.samp, st.disch_track, st.mom_track, st.mask,
                             maps.cell_units(ep, "iter_erode", unit), maps.slope_limited_deposit(ep, unit),
                             float(ep.fan_slope), False, maps.cell_units(ep, "dep_floor_m", unit),
                             float(ep.offshore_writeoff), None, False, out4, status, work, 1, 1)
    torch.cuda.synchronize()
    assert int(status[0]) & 2
    assert "did not converge" in pk.ordered_status_error(status.cpu().numpy())
    for a, b in zip(before, (st.acc, st.pending, st.samp, st.disch_track, st.mom_track)):
        assert torch.equal(a, b)


def test_sweep_limit_in_the_chunk_loop_falls_back_to_the_serial_apply(cuda):
    """In the chunk loop a chunk that does not converge within its sweeps is
    applied by the serial apply: with a limit of one sweep every chunk but
    the trivial ones goes that way, and the iteration is the unlimited one
    bit for bit."""
    p, mk = _earth(cuda)
    a, b = mk(), mk()
    a.check_max_sweeps = 1
    for it in range(2):
        sa = maps.step(a, p, it)
        sb = maps.step(b, p, it)
        print(f"step {it}: serial chunks {sa['serial_chunks']} / {sb['serial_chunks']}", flush=True)
        assert sa["serial_chunks"] > 0 and sb["serial_chunks"] == 0
        _same_stats(sa, sb)
        _same_state(a, b)


def test_lake_parks_fall_back_to_the_serial_apply(cuda):
    """Lake parks (volume -3 entries: a load a lake keeps, particle.py:681-700,
    912-933) are outside the parallel apply; the chunk loop hands every chunk
    holding one to the serial apply on the device.  Lakes laid on the
    epsilon-filled route of an earth globe with the lake balance's parking on:
    the one-call loop and the serial apply agree bit for bit, lakes included."""
    p, mk = _earth(cuda, lake_fill=True, lake_balance=True)
    a, b = mk(), mk()
    b.check_apply = "serial"
    for st in (a, b):
        st.refresh_route(0.05 / st.height_unit_m)
        lake = ((st.route - (st.height + st.sediment)) > 1e-4) & (st.mask == pk.MASK_ACTIVE)
        st.lake_flag = torch.zeros(st.height.shape, dtype=torch.uint8, device=cuda)
        st.lake_flag[lake] = 2
        st.lake_id = torch.full(st.height.shape, -1, dtype=torch.int32, device=cuda)
        st.lake_id[lake] = 0
        st.lake_room = torch.full((1,), 1e3, dtype=torch.float64, device=cuda)
    assert int(a.lake_flag.sum()) > 0
    for it in range(2):
        sa = maps.run_iteration(a, p, it)
        sb = maps.run_iteration(b, p, it)
        print(f"iteration {it}: serial chunks {sa['serial_chunks']}, lake load {float(a.lake_load.sum()):.4g}", flush=True)
        _same_stats(sa, sb)
        _same_state(a, b)
        assert torch.equal(a.lake_load, b.lake_load) and torch.equal(a.lake_room, b.lake_room)
    assert sa["serial_chunks"] > 0 and float(a.lake_load.sum()) > 0.0


@pytest.mark.parametrize("apply", ["serial", "ordered"])
@pytest.mark.parametrize("case", apply_cases.CASES, ids=lambda c: c.name)
def test_hand_built_lists_on_device(cuda, case, apply):
    print(f"hand-built list {case.name}, {apply} apply", flush=True)
    (cl_cell, cl_delta, cl_vol, cl_mom, cl_count), f, cap = apply_cases.inputs(case)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    t = {k: d(v) for k, v in f.items()}
    lists = [d(cl_cell), d(cl_delta), d(cl_vol), d(cl_mom), d(cl_count)]
    prm = case.params()
    out4 = torch.zeros(4, dtype=torch.float64, device=cuda)
    common = (*lists, cap, t["height"], t["sediment"], t["acc"], t["pending"], t["samp"], t["disch_track"],
              t["mom_track"], t["mask"], prm["iter_erode"], prm["iter_deposit"], prm["fan_slope"], prm["use_route"],
              prm["dep_floor"], prm["offshore_writeoff"], t["erode_cap"], prm["use_erode_cap"])
    if apply == "serial":
        pk.apply_changes(*common, None, None, None, out4)
    else:
        status = torch.zeros(4, dtype=torch.int32, device=cuda)
        work = torch.empty(pk.ordered_workspace_bytes(len(case.droplets), cap, 64), dtype=torch.uint8, device=cuda)
        pk.apply_changes_ordered(*common, out4, status, work, 64)
        assert status[0].item() == 0, status
    torch.cuda.synchronize()
    apply_cases.check(case, {k: v.cpu().numpy() for k, v in t.items()}, out4.cpu().numpy())
