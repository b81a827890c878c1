"""The oracle's particle pass, pinned by the reference's own invariants
(bake/tests/test_erosion.py) and by hand-checked single-droplet traces.

These pin the C restatement of trace_particles / apply_changes and the numpy
restatement of run_iteration / step without numba (not installable here):
a dropped term, a wrong sign or a transposed index in the restatement breaks
mass conservation, the offshore split, the change-list layout or one of the
closed-form traces below.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from windows import window_inputs, window_params


def _window(oracle, N, mode, seed=0, deposit_on_exit=True, **erosion):
    p = window_params(0, **erosion)
    return p, oracle.OracleState.window(**window_inputs(N, mode, p, seed=seed, deposit_on_exit=deposit_on_exit))


def _wstep(oracle, st, p, it):
    """The window iteration: particles, then uplift as given (maps.py:1577-1587)."""
    s = oracle.run_iteration(st, p, it)
    oracle.apply_uplift(st)
    return s


def test_conservation_closed_window(oracle):
    """test_erosion.py:191-206: height + sediment + pending is conserved."""
    p, st = _window(oracle, 48, "closed")
    assert np.all(st.surface()[st.interior] > 0)
    m0 = st.total_mass()
    clamped = 0
    for it in range(20):
        s = _wstep(oracle, st, p, it)
        assert s["particles"] > 0
        clamped += s["clamped"]
    m1 = st.total_mass()
    assert abs(m1 - m0) <= 1e-6 * abs(m0), (m0, m1)
    assert clamped > 0
    assert st.sediment[st.interior].max() > 0
    assert np.all(st.sediment >= 0) and np.all(st.pending >= 0)


@pytest.mark.parametrize("writeoff", [0.25, 1.0])
def test_offshore_loss_is_accounted(oracle, writeoff):
    """test_erosion.py:209-248: what leaves is exactly lost_offshore, and the
    sea stockpile is (1 - w) / w of it."""
    p, st = _window(oracle, 48, "tilt", offshore_writeoff=writeoff)
    w = writeoff
    m0 = st.total_mass()
    lost = parked = 0.0
    for it in range(25):
        s = _wstep(oracle, st, p, it)
        lost += s["lost_offshore"]
        assert s["deficit_out"] <= 1e-12 * abs(m0)
        sea_pending = float(st.pending[st.surface() < 0].sum())
        parked += sea_pending
        assert abs(sea_pending - (1.0 - w) / w * s["lost_offshore"]) <= 1e-9 * (sea_pending + s["lost_offshore"] + 1e-300)
    upl = float(st.uplift[st.interior][st.mask[st.interior] == 1].sum()) * 25
    m1 = st.total_mass()
    assert abs((m1 + lost) - (m0 + upl)) <= 1e-6 * abs(m0)
    assert s["deaths"]["ocean"] > 0
    if w == 1.0:
        assert lost > 0 and parked == 0.0
    else:
        assert parked > 0


def test_seafloor_stockpile_is_bounded(oracle):
    """test_erosion.py:251-289: a full shelf's stockpile decays geometrically.

    The kept share never exceeds (1 - w) / w of the write-off (exact, up to
    rounding); it falls short by the dust the kernel writes off whole (kept
    shares under 1e-4 cells).  The reference's scenario runs with thermal
    erosion and a routing surface, which this restatement leaves out; without
    them more stockpiles decay to dust, so the dust slack is 10 % here, not
    the reference's 1 %."""
    p, st = _window(oracle, 48, "shelf")
    w = p.erosion.offshore_writeoff
    sea0 = st.surface() < 0
    m0 = st.total_mass()
    n = 40
    lost_k = np.zeros(n)
    pend_k = np.zeros(n)
    for it in range(n):
        s = _wstep(oracle, st, p, it)
        assert s["lost_offshore"] > 0.0, it
        lost_k[it] = s["lost_offshore"]
        pend_k[it] = float(st.pending[st.surface() < 0].sum())
        gap = (1.0 - w) / w * lost_k[it] - pend_k[it]
        assert -1e-9 * (pend_k[it] + lost_k[it]) <= gap <= 0.1 * (1.0 - w) / w * lost_k[it] + 1e-9, (it, gap)
    assert np.all(st.surface()[sea0] < 0.0)
    upl = float(st.uplift[st.interior][st.mask[st.interior] == 1].sum()) * n
    assert abs((st.total_mass() + lost_k.sum()) - (m0 + upl)) <= 1e-6 * abs(m0)
    fresh = lost_k / w - np.concatenate([[0.0], pend_k[:-1]])
    assert fresh.max() > 0.0
    assert pend_k.max() <= (1.0 - w) / w * fresh.max() * (1.0 + 1e-9)
    assert pend_k[-1] < 0.25 * np.sum((1.0 - w) * fresh)


def test_change_list_layout_via_diag(oracle):
    """test_erosion.py:292-328: steps first, then negative final deposits;
    the hook changes nothing."""
    p, st = _window(oracle, 48, "tilt")
    _, ref = _window(oracle, 48, "tilt")
    seen = []

    def diag(state, cl_cell, cl_vol, cl_count, cap, sp_death):
        assert cl_count.shape == sp_death.shape
        for q in range(cl_count.shape[0]):
            row = cl_vol[q * cap: q * cap + cl_count[q]]
            if row.size == 0:
                continue
            neg = np.flatnonzero(row < 0.0)
            if neg.size:
                assert np.all(row[neg[0]:] < 0.0)
            assert np.all(row[:neg[0] if neg.size else row.size] >= 0.0)
        seen.append(int(cl_count.sum()))

    for it in range(3):
        a = oracle.run_iteration(st, p, it, diag=diag)
        b = oracle.run_iteration(ref, p, it)
        assert a["entries"] == b["entries"] and a["deaths"] == b["deaths"]
    assert seen and any(n > 0 for n in seen)
    assert np.array_equal(st.height, ref.height) and np.array_equal(st.sediment, ref.sediment)


def test_exit_without_deposit_loses_mass_only(oracle):
    """test_erosion.py:331-339."""
    p, st = _window(oracle, 48, "closed", deposit_on_exit=False)
    m0 = st.total_mass()
    for it in range(10):
        _wstep(oracle, st, p, it)
    assert st.total_mass() < m0


def test_run_iteration_stats(oracle):
    """test_erosion.py:939-949."""
    p, st = _window(oracle, 32, "closed")
    s = oracle.run_iteration(st, p, (0, 1))
    assert s["particles"] == round(p.erosion.particles_per_cell * 32 * 32)
    assert sum(s["deaths"].values()) == s["particles"]
    assert s["deaths"]["ocean"] == 0
    assert st.disch_track.max() == 0.0
    assert st.discharge.max() > 0.0
    assert np.isfinite(st.momentum).all()


def test_determinism_and_thread_independence(oracle):
    """test_erosion.py:443-453 and test_refine.py:165-184: identical runs,
    independent of the number of trace threads."""
    outs = []
    for nt in (1, 3, 0):
        p, st = _window(oracle, 40, "tilt")
        for it in range(3):
            oracle.run_iteration(st, p, it, nthreads=nt)
        outs.append((st.height.copy(), st.sediment.copy(), st.pending.copy()))
    for a in outs[1:]:
        for x, y in zip(outs[0], a):
            assert np.array_equal(x, y)


def test_change_list_cap_formula(oracle):
    """test_erosion.py:1743-1754."""
    for max_steps in (40, 500):
        for lateral in (False, True):
            cap = oracle.change_list_cap(max_steps, lateral)
            assert cap >= (2 * max_steps if lateral else max_steps) + 2 * oracle.SPREAD + 1
            assert cap == (2 * max_steps if lateral else max_steps) + 2 * 8 + 1 + 8 + 8


# ---------------------------------------------------------------------------
# hand-checked single droplets on a flat-metric window
# ---------------------------------------------------------------------------
def _one_droplet(oracle, h, x, y, max_steps=40, **over):
    """Trace one droplet of volume 1 at (x, y) over the surface h (NE x NE,
    cell units, H = 2, mask 1 inside) and return (cells, deltas, vols, death)."""
    H = 2
    NE = h.shape[0]
    N = NE - 2 * H
    p = window_params(0)
    ep = p.erosion
    mask = np.zeros((1, NE, NE), np.uint8)
    mask[0, H:H + N, H:H + N] = 1
    st = oracle.OracleState.window(h, np.zeros_like(h), np.zeros_like(h), np.zeros((NE, NE, 2)),
                                   np.zeros((NE, NE), np.float32), np.ones((NE, NE), np.float32),
                                   np.ones((NE, NE), np.float32), np.zeros_like(h), mask[0],
                                   np.stack([np.ones((NE, NE)), np.zeros((NE, NE)), np.ones((NE, NE))], -1),
                                   np.stack([np.ones((NE, NE)), np.zeros((NE, NE)), np.ones((NE, NE))], -1), H, 1.0)
    st.pack()
    prm = oracle.TraceParams(dt=1.2, density=1.0, friction=0.25, deposition_rate=0.1, evap_rate=0.001, k_mom=0.0,
                             mom_scale=1.0, k_disc=0.0, disc_saturation=32.0, disc_exponent=0.5, slope_gain=2.0,
                             slope_sat=0.0, erodibility=0.2, cover_depth=0.1, min_volume=5e-5, max_steps=max_steps,
                             max_erode=1.0, conc_scale=0.0, pit_steps=3, deposit_on_exit=0, ocean_rate=0.3,
                             ocean_steps=64, fan_room=1.0, dep_floor=0.02, lake_trap=0.9, lake_park=0.0, lateral_rate=0.0)
    for k, v in over.items():
        setattr(prm, k, v)
    cap = oracle.change_list_cap(max_steps, False)
    cl_cell = np.zeros(cap, np.int32)
    cl_delta = np.zeros(cap)
    cl_vol = np.zeros(cap, np.float32)
    cl_mom = np.zeros((cap, 2), np.float32)
    cnt, lo, hi = np.zeros(1, np.int64), np.zeros(1, np.int64), np.zeros(1, np.int64)
    death = np.zeros(1, np.int8)
    arr = lambda v, dt: np.array([v], dt)
    sp = [arr(0, np.int64), arr(x, np.float64), arr(y, np.float64), arr(0.0, np.float64), arr(0.0, np.float64)]
    P = oracle._p
    oracle.lib().or_trace_particles(1, *[P(a) for a in sp], 1.0, P(st.samp), 0, P(st.mask), N, H, 0, ctypes.byref(prm),
                                    P(np.zeros((1, 1, 1), np.float32)), 0, P(cl_cell), P(cl_delta), P(cl_vol), P(cl_mom),
                                    P(cnt), P(lo), P(hi), P(death), 1)
    n = int(cnt[0])
    return cl_cell[:n], cl_delta[:n], cl_vol[:n], int(death[0]), NE


def test_flat_plane_stops_at_once(oracle):
    """On a flat plane with no inertia, momentum or slope the droplet has no
    direction: it stops where it is born (DEATH_STOP), writing nothing."""
    h = np.full((20, 20), 5.0)
    cells, deltas, vols, death, _ = _one_droplet(oracle, h, 7.3, 8.6)
    assert death == 5 and cells.size == 0


def test_tilted_plane_steps_one_cell_downhill(oracle):
    """A plane falling along +j: each step is exactly one cell along +j (the
    metric is flat), the droplet erodes on the way and leaves the window."""
    NE, H = 24, 2
    j = np.arange(NE)
    h = np.broadcast_to(50.0 - 0.5 * j[None, :], (NE, NE)).copy()
    cells, deltas, vols, death, ne = _one_droplet(oracle, h, 9.5, 2.5)
    ei, ej = (cells % (ne * ne)) // ne, cells % ne
    assert death == 2  # exit
    assert np.all(ei == 9 + H)
    assert np.array_equal(ej, np.arange(2, 2 + ej.size) + H)
    assert np.all(deltas < 0.0)  # erosion only (the load never exceeds capacity on a steady slope)
    assert np.all(np.diff(vols) < 0)  # evaporation


def test_pit_kills_after_pit_steps_uphill(oracle):
    """A conical pit: a droplet rolling in oscillates across the bottom and
    dies of DEATH_PIT once it has climbed pit_steps + 1 steps in a row --
    or comes to rest (DEATH_STOP) -- but never leaves."""
    NE, H = 24, 2
    i, j = np.meshgrid(np.arange(NE), np.arange(NE), indexing="ij")
    h = 0.3 * np.hypot(i - 11.5, j - 11.5)
    cells, deltas, vols, death, _ = _one_droplet(oracle, h, 3.2, 4.1, pit_steps=3, friction=0.0)
    assert death in (3, 5)
    assert cells.size > 0 and vols[-1] == -1.0  # the final deposit lands inside
