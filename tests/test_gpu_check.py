"""Check mode on the GPU against the oracle, same seeded inputs.

The HIP trace (float64 droplets) and serial apply must reproduce the
oracle's change lists and the state after an iteration.  Bar: droplet cell
sequences, entry counts, death causes and volumes bit for bit; float64
deltas and state bit for bit except where a droplet crossed a cube face and
glibc's tan/atan (the oracle's, like the reference's) misround in the last
bit where the device's correctly rounded ones do not -- those differ by a
few ulp, so the float comparison is bit-exact for >= 99.9 % of values and
within 1e-9 relative for all.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from globe_hip import stubs
from globe_hip.erosion import maps
from windows import window_inputs, window_params

pytestmark = pytest.mark.gpu


def _capture(store, to_np):
    def diag(state, cl_cell, cl_vol, cl_count, cap, sp_death):
        n = int(cl_count.shape[0])
        cnt = to_np(cl_count).copy()
        cells = to_np(cl_cell).copy()
        vols = to_np(cl_vol).copy()
        store.append((cnt, cells, vols, to_np(sp_death).copy(), cap, n))
    return diag


def _cmp_chunks(dev, ora):
    """Number of droplets whose change list (cells, volumes, count, death)
    differs between the device and the oracle."""
    assert len(dev) == len(ora)
    bad = 0
    for (c1, e1, v1, d1, cap, n), (c2, e2, v2, d2, cap2, n2) in zip(dev, ora):
        assert cap == cap2 and n == n2
        for p in range(n):
            if c1[p] != c2[p] or d1[p] != d2[p]:
                bad += 1
                continue
            sl = slice(p * cap, p * cap + int(c1[p]))
            if not (np.array_equal(e1[sl], e2[sl]) and np.array_equal(v1[sl], v2[sl])):
                bad += 1
    return bad


def _cmp_float(name, a, b, exact_frac=0.999, rtol=1e-9):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    same = np.mean(a == b)
    scale = np.maximum(np.abs(b), 1e-300)
    bad = np.abs(a - b) > rtol * np.maximum(scale, np.abs(b).max() * 1e-6)
    assert not bad.any(), (name, int(bad.sum()), float(np.abs(a - b).max()))
    assert same >= exact_frac, (name, same)
    return same


def _run_pair(oracle, ostate, dstate, params, key, allow=0):
    o_chunks, d_chunks = [], []
    so = oracle.run_iteration(ostate, params, key, diag=_capture(o_chunks, lambda x: x))
    sd = maps.run_iteration(dstate, params, key, diag=_capture(d_chunks, lambda t: t.cpu().numpy()))
    bad = _cmp_chunks(d_chunks, o_chunks)
    print(f"droplets whose change list differs: {bad} of {so['particles']}", flush=True)
    assert bad <= allow, bad
    if bad:
        return so, sd
    for k in ("particles", "entries", "deaths", "clamped"):
        assert sd[k] == so[k], (k, sd[k], so[k])
    for k in ("to_pending", "deficit_out", "lost_offshore", "volume"):
        assert sd[k] == pytest.approx(so[k], rel=1e-9, abs=1e-15), k
    return so, sd


def _cmp_state(ostate, dstate, exact_frac=0.999, rtol=1e-9):
    fr = {}
    for n in ("height", "sediment", "pending", "discharge", "momentum"):
        fr[n] = _cmp_float(n, getattr(dstate, n).cpu().numpy(), getattr(ostate, n), exact_frac, rtol)
    return fr


@pytest.mark.parametrize("mode,N", [("tilt", 48), ("closed", 40), ("shelf", 40)])
def test_window_iterations_match_oracle(oracle, cuda, mode, N):
    p = window_params(0)
    kw = window_inputs(N, mode, p, seed=1, deposit_on_exit=True)
    ost = oracle.OracleState.window(**kw)
    dst = maps.ErosionState.window(**kw, device=cuda)
    for it in range(3):
        print(f"window {mode} N={N} iteration {it}", flush=True)
        _run_pair(oracle, ost, dst, p, it)
        # the window pass: particles, then uplift as given (no datum, no halos)
        oracle.apply_uplift(ost)
        maps.apply_uplift(dst)
        _cmp_state(ost, dst)
    if mode == "shelf":
        assert ost.pending.sum() > 0  # the full shelf parks what it cannot place


@pytest.mark.parametrize("N,libm", [(64, "glibc"), (128, "glibc"), (128, "correctly-rounded")])
def test_globe_first_iteration_matches_oracle(oracle, hiplib, cuda, N, libm):
    """With the oracle on glibc's tan/atan (the reference's), face-crossing
    droplets may differ (glibc misrounds the last bit of ~0.1-0.25 % of
    tan/atan values): at most 0.1 % of droplets, and >= 99 % of cells stay
    bit-identical.  With the oracle switched to the
    correctly rounded tan/atan the device uses, everything is bit-identical."""
    grid, f = stubs.plain_globe(N, 2)
    p = stubs.workload_params(N, 2)
    ost = oracle.OracleState.from_grid(grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"], p.erosion)
    dst = maps.ErosionState.from_grid(grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"], p.erosion, device=cuda)
    print(f"globe N={N} iteration 0, oracle libm {libm}", flush=True)
    cr = libm == "correctly-rounded"
    if cr:
        oracle.set_libm(*oracle.device_libm(hiplib))
    # glibc: measured 52 of 98,304 droplets at N=128 (0.05 %); bound 0.1 %
    allow = 0 if cr else -(-int(round(p.erosion.particles_per_cell * 6 * N * N)) // 1000)
    try:
        so, sd = _run_pair(oracle, ost, dst, p, 0, allow=allow)
    finally:
        oracle.set_libm()
    if cr:
        fr = _cmp_state(ost, dst, exact_frac=1.0, rtol=0.0)
    else:
        # a droplet whose face crossing rounds differently walks another path:
        # the cells it touches differ, everything else stays bit-identical
        fr = {}
        for n in ("height", "sediment", "pending", "discharge", "momentum"):
            a = getattr(dst, n).cpu().numpy().ravel()
            b = np.asarray(getattr(ost, n)).ravel()
            fr[n] = float(np.mean(a == b))
            assert fr[n] >= 0.99, (n, fr[n])
    print("bit-identical fractions", fr, flush=True)


@pytest.mark.parametrize("N,iters", [(64, 3), (256, 8)])
def test_globe_steps_match_oracle(oracle, hiplib, cuda, N, iters):
    """Whole steps (particles, uplift, datum hold, halo exchange) on the
    earth globe, with the oracle on the device's correctly rounded tan /
    atan and the halo sums in source order: the state and every step's
    statistics bit for bit, iteration after iteration."""
    grid, f = stubs.earth_globe(N, 2, n_blobs=16)
    p = stubs.workload_params(N, 2, max_steps=128, uplift_mode="replay", iterations=iters)
    args = (grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"], p.erosion)
    ost = oracle.OracleState.from_grid(*args)
    dst = maps.ErosionState.from_grid(*args, device=cuda)
    oracle.set_libm(*oracle.device_libm(hiplib))
    try:
        io, idv = oracle.start_replay(ost, p), maps.start_replay(dst, p)
        assert float(idv["datum_shift"]) == float(io["datum_shift"])
        for it in range(iters):
            print(f"globe N={N} step {it}", flush=True)
            so = oracle.step(ost, p, it)
            sd = maps.step(dst, p, it)
            for k in ("particles", "entries", "deaths", "clamped", "to_pending", "deficit_out", "lost_offshore"):
                assert sd[k] == so[k], (it, k, sd[k], so[k])
            assert float(sd["datum_shift"]) == float(so["datum_shift"]), it
            for n in ("height", "sediment", "pending", "discharge", "momentum"):
                assert np.array_equal(getattr(dst, n).cpu().numpy(), getattr(ost, n)), (it, n)
    finally:
        oracle.set_libm()


@pytest.mark.parametrize("inputs,rate", [("plain", 50.0), ("tilt", 5.0)])
def test_window_lateral_cuts_match_oracle(oracle, cuda, inputs, rate):
    """Lateral erosion (particle.py:699-740; volume -2 entries applied under
    the erosion caps, particle.py:934-1018): change lists and state bit for
    bit over two iterations."""
    from test_oracle_lateral import _plain_window_inputs

    p = window_params(0, lateral_rate=rate, max_steps=60)
    kw = _plain_window_inputs(p) if inputs == "plain" else window_inputs(48, "tilt", p, seed=4, deposit_on_exit=True)
    ost = oracle.OracleState.window(**kw)
    dst = maps.ErosionState.window(**kw, device=cuda)
    cuts = []
    for it in range(2):
        print(f"lateral {inputs} rate {rate} iteration {it}", flush=True)
        d_chunks = []
        oracle.run_iteration(ost, p, it, diag=_capture(cuts, lambda x: x))
        maps.run_iteration(dst, p, it, diag=_capture(d_chunks, lambda t: t.cpu().numpy()))
        assert _cmp_chunks(d_chunks, cuts[-len(d_chunks):]) == 0
        _cmp_state(ost, dst, exact_frac=1.0, rtol=0.0)
    n_cut = sum(int((v[q * cap: q * cap + int(c[q])] == -2.0).sum())
                for c, _, v, _, cap, n in cuts for q in range(n))
    assert n_cut > 100, n_cut


def test_droplets_per_wave_keep_results(cuda):
    """The trace's droplets per wave change only the launch shape: three
    steps at 64, 8 and 1 droplets per wave, and under the autotuner, give the
    same bits."""
    N = 64
    grid, f = stubs.earth_globe(N, 2, n_blobs=16)
    p = stubs.workload_params(N, 2, max_steps=128, uplift_mode="replay", iterations=3)
    args = (grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"], p.erosion)
    outs = []
    for lanes in (64, 8, 1, None):
        st = maps.ErosionState.from_grid(*args, device=cuda)
        tuner = None
        if lanes is None:
            tuner = maps.CheckAutotuner(st, candidates=(64, 8, 1))
        else:
            st.check_lanes = lanes
        for it in range(3):
            if tuner is not None:
                tuner.step(lambda k: maps.step(st, p, k), it)
            else:
                maps.step(st, p, it)
        if tuner is not None:
            assert tuner.best in (64, 8, 1) and st.check_lanes == tuner.best and len(tuner.times) == 3
        outs.append([t.cpu().numpy() for t in (st.height, st.sediment, st.discharge, st.momentum, st.pending)])
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert np.array_equal(a, b)
