"""The oracle's amounts, pinned by values fixed independently of it.

* apply_changes (particle.py:821-1042): hand-built change lists whose
  outcome was worked out by hand from the rules the reference states
  (tests/apply_cases.py), bitwise;
* trace_particles (particle.py:275-818): single droplets on planes where
  the per-step amounts have a closed form -- the load's geometric approach
  to its fixed point on a steady slope, the erosion cap, the flat deposit
  capped at dep_floor, the final deposit equal to everything eroded;
* uplift, replay and the datum hold (maps.py:1088-1213, 1413-1508): the
  properties of the reference's own tests (test_erosion.py:545-642,
  806-933, 1201-1223) restated on the oracle: mean-free, capped, a replay
  that ends at bedrock plus one constant, a hold that leaves exactly the
  requested number of land cells and carries the base level with it.

A plausible slip in the restatement (a dropped evaporation term, the
hardness or cover weight inverted, the ceiling taken from the wrong path
cell, a sign in the fan slope, the mean taken over halos) moves one of
these numbers.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import apply_cases
from windows import window_params


# ---------------------------------------------------------------------------
# apply_changes on hand-built lists
# ---------------------------------------------------------------------------
def _oracle_apply(oracle, case):
    (cl_cell, cl_delta, cl_vol, cl_mom, cl_count), f, cap = apply_cases.inputs(case)
    prm = case.params()
    out = np.zeros(4)
    lake_load = np.zeros_like(f["height"])
    lake_id = np.full(f["height"].shape, -1, np.int32)
    lake_room = np.zeros(1)
    P = oracle._p
    oracle.lib().or_apply_changes(P(cl_cell), P(cl_delta), P(cl_vol), P(cl_mom), P(cl_count), cl_count.size, cap,
                                  P(f["height"]), P(f["sediment"]), P(f["acc"]), P(f["pending"]), P(f["samp"]),
                                  P(f["disch_track"]), P(f["mom_track"]), P(f["mask"]), prm["iter_erode"],
                                  prm["iter_deposit"], prm["fan_slope"], prm["use_route"], prm["dep_floor"],
                                  prm["offshore_writeoff"], P(f["erode_cap"]), prm["use_erode_cap"], P(lake_load),
                                  P(lake_id), P(lake_room), P(out))
    return f, out


@pytest.mark.parametrize("case", apply_cases.CASES, ids=lambda c: c.name)
def test_apply_hand_built_lists(oracle, case):
    f, out = _oracle_apply(oracle, case)
    apply_cases.check(case, f, out)


def test_apply_cases_conserve_mass(oracle):
    """Independently of the worked numbers: what the lists request, less
    what was lost, written off, left with a droplet that never stopped, or
    cancelled, equals what acc and pending hold -- for the cases without a
    dropped surplus (seafloor_fan leaves with its 0.125)."""
    for case in apply_cases.CASES:
        f, out = _oracle_apply(oracle, case)
        requested = sum(d for drop in case.droplets for (c, d, v, m) in drop if c not in case.inactive)
        placed = float(f["acc"].sum() + f["pending"].sum())
        left = 0.125 if case.name == "seafloor_fan" else 0.0
        assert placed == requested + out[2] - out[3] - left, (case.name, placed, requested, out)


# ---------------------------------------------------------------------------
# single droplets with closed-form amounts
# ---------------------------------------------------------------------------
TRACE = dict(dt=1.2, density=1.0, friction=0.25, deposition_rate=0.1, evap_rate=0.001, k_mom=0.0, mom_scale=1.0,
             k_disc=0.0, disc_saturation=32.0, disc_exponent=0.5, slope_gain=2.0, slope_sat=0.0, erodibility=0.2,
             cover_depth=0.1, min_volume=5e-5, max_steps=40, max_erode=1.0, conc_scale=0.0, pit_steps=3,
             deposit_on_exit=0, ocean_rate=0.3, ocean_steps=64, fan_room=1.0, dep_floor=0.02, lake_trap=0.9,
             lake_park=0.0, lateral_rate=0.0)


def _trace(oracle, h, x, y, hardness=0.0, sed_thick=0.0, sed0=0.0, dir0=0.0, **over):
    """One droplet of volume 1 from (x, y) over h (NE x NE cell units,
    H = 2, flat metric, base 0, uniform hardness and sediment cover).
    Returns the change list (cells, deltas, vols) and the death cause."""
    H, NE = 2, h.shape[0]
    N = NE - 2 * H
    mask = np.zeros((NE, NE), np.uint8)
    mask[H:H + N, H:H + N] = 1
    one, zero = np.ones((NE, NE)), np.zeros((NE, NE))
    st = oracle.OracleState.window(h, np.full((NE, NE), sed_thick), zero, np.zeros((NE, NE, 2)),
                                   np.full((NE, NE), hardness, np.float32), one.astype(np.float32),
                                   one.astype(np.float32), zero, mask, np.stack([one, zero, one], -1),
                                   np.stack([one, zero, one], -1), H, 1.0)
    st.pack()
    prm = oracle.TraceParams(**{**TRACE, **over})
    cap = oracle.change_list_cap(prm.max_steps, False)
    cl_cell, cl_delta = np.zeros(cap, np.int32), np.zeros(cap)
    cl_vol, cl_mom = np.zeros(cap, np.float32), np.zeros((cap, 2), np.float32)
    cnt, lo, hi, death = np.zeros(1, np.int64), np.zeros(1, np.int64), np.zeros(1, np.int64), np.zeros(1, np.int8)
    a = lambda v, dt: np.array([v], dt)
    sp = [a(0, np.int64), a(x, np.float64), a(y, np.float64), a(sed0, np.float64), a(dir0, np.float64)]
    P = oracle._p
    oracle.lib().or_trace_particles(1, *[P(v) for v in sp], 1.0, P(st.samp), 0, P(st.mask), N, H, 0, ctypes.byref(prm),
                                    P(np.zeros((1, 1, 1), np.float32)), 0, P(cl_cell), P(cl_delta), P(cl_vol),
                                    P(cl_mom), P(cnt), P(lo), P(hi), P(death), 1)
    n = int(cnt[0])
    return cl_cell[:n], cl_delta[:n], cl_vol[:n], int(death[0])


def _plane(NE, slope, flat_from=None):
    """h = 50 - slope * j along +j (flat from column flat_from on)."""
    j = np.arange(NE, dtype=np.float64)
    if flat_from is not None:
        j = np.minimum(j, flat_from)
    return np.broadcast_to(50.0 - slope * j[None, :], (NE, NE)).copy()


@pytest.mark.parametrize("hardness,cover,k_e", [(0.0, 0.0, 0.1), (0.5, 0.0, 0.05),
                                                (0.5, 0.0625, 0.1 * (0.625 + 0.375 * 0.5)), (0.5, 0.25, 0.1)])
def test_steady_slope_load_approaches_its_fixed_point(oracle, hardness, cover, k_e):
    """On a plane of slope 0.5 each step moves one cell downhill with
    dh = 0.5, so c_eq = erodibility * dh = 0.1 at every step; with the
    volume decaying by ev = 1 - dt * evap_rate the concentration obeys
    sed_{k+1} = ((1 - k_e) sed_k + k_e c_eq) / ev, i.e.

        sed_k = s* (1 - r^k),  s* = k_e c_eq / (ev - 1 + k_e),  r = (1 - k_e) / ev,

    and the entry of step k erodes k_e (c_eq - sed_k) ev^k.  The erosion
    rate k_e is deposition_rate weighted by the cover w = min(thickness /
    cover_depth, 1): k_e = rate (w + (1 - w)(1 - hardness))."""
    NE = 24
    cells, deltas, vols, death = _trace(oracle, _plane(NE, 0.5), 9.5, 2.5, hardness=hardness, sed_thick=cover)
    assert death == 2 and cells.size == NE - 2 * 2 - 3   # one entry per cell from j = 2 to N - 2, then the exit
    ev = 1.0 - 1.2 * 0.001
    c_eq = 0.2 * 0.5
    s_star = k_e * c_eq / (ev - 1.0 + k_e)
    k = np.arange(cells.size)
    sed = s_star * (1.0 - ((1.0 - k_e) / ev) ** k)
    np.testing.assert_allclose(deltas, -k_e * (c_eq - sed) * ev ** k, rtol=1e-12, atol=0)
    np.testing.assert_allclose(vols, (ev ** k).astype(np.float32), rtol=2e-7, atol=0)


def test_erosion_cap_binds_every_step(oracle):
    """With max_erode below k_e c_eq the cap binds at every step and each
    entry erodes exactly max_erode: the load after k steps is k max_erode
    and the uncapped request k_e (c_eq ev^k - k max_erode) stays above it."""
    cells, deltas, vols, death = _trace(oracle, _plane(24, 0.5), 9.5, 2.5, max_erode=0.001)
    assert death == 2 and cells.size == 17
    np.testing.assert_allclose(deltas, -0.001, rtol=1e-13, atol=0)


def test_final_deposit_is_everything_eroded(oracle):
    """deposit_on_exit: a droplet that leaves lays its load sed * vol (which
    evaporation leaves unchanged) on the newest ring cell -- the sum of what
    it eroded -- and zero-load entries on the older ones (SPREAD = 8)."""
    cells, deltas, vols, death = _trace(oracle, _plane(24, 0.5), 9.5, 2.5, deposit_on_exit=1)
    steps = vols >= 0.0
    fin = np.flatnonzero(vols == -1.0)
    assert death == 2 and fin.size == 8 and np.all(fin == np.arange(steps.sum(), steps.sum() + 8))
    # the ring's newest cell is the one the droplet left from (pushed, then
    # the exit check fires before that step's entry), then the step cells
    assert cells[fin[0]] == cells[steps][-1] + 1
    np.testing.assert_array_equal(cells[fin[1:]], cells[steps][::-1][:7])
    np.testing.assert_allclose(deltas[fin[0]], -deltas[steps].sum(), rtol=1e-12)
    assert np.all(deltas[fin[1:]] == 0.0)


def test_flat_run_out_deposits_dep_floor_per_step(oracle):
    """A slope that turns flat: on the flat c_eq = 0, so the request is
    deposition_rate * load; it is capped by dcap = max(h_prev - h_here,
    dep_floor): 0.5 (the last slope drop) on the first flat cell, then
    dep_floor = 0.02 on every later cell until half the load is below it."""
    NE, J0 = 32, 10
    cells, deltas, vols, death = _trace(oracle, _plane(NE, 0.5, flat_from=J0), 9.5, 2.5, deposition_rate=0.5,
                                        erodibility=0.8)
    ej = cells % NE
    flat = np.flatnonzero(ej >= J0)
    slope = np.flatnonzero(ej < J0)
    assert death == 2 and flat.size == NE - 2 - 1 - J0 and slope.size == J0 - 4
    assert np.all(deltas[slope] < 0) and np.all(deltas[flat] > 0)
    load = -deltas[slope].sum()
    np.testing.assert_allclose(deltas[flat[0]], 0.5 * load, rtol=1e-12)   # uncapped: 0.5 * load < 0.5
    load -= deltas[flat[0]]
    n_capped = 0
    for i in flat[1:]:
        if 0.5 * load >= 0.02:
            assert abs(deltas[i] - 0.02) <= 1e-15, (i, deltas[i])
            n_capped += 1
        else:   # the running load is a difference of sums: a few ulps of cancellation
            np.testing.assert_allclose(deltas[i], 0.5 * load, rtol=1e-10)
        load -= deltas[i]
    assert n_capped >= 4


@pytest.mark.parametrize("floor,slope,sed0,rooms", [
    (-10.0, 0.5, 5.0, (1.0, 1.0)),      # drops of 0.5 < fan_room: every step may take fan_room
    (-10.0, 2.0, 20.0, (1.0, 2.0)),     # first step fan_room (no previous cell), then the 2.0 drop
    (-0.5, 0.0, 5.0, (0.48, 0.48)),     # shallow flat shelf: the waterline b - dep_floor - h binds
])
def test_seafloor_walk_fills_to_its_room(oracle, floor, slope, sed0, rooms):
    """A droplet born under water carrying a load (a re-injected stockpile)
    walks the seafloor in deposit-only mode: each step lays ocean_rate * load,
    at most room = min(max(h_prev - h_here, fan_room), b - dep_floor - h_here)
    (fan_room on the first step).  Evaporation leaves the load sed * vol
    unchanged, so the deposits obey d_k = min(0.3 u_k, room_k), u_{k+1} =
    u_k - d_k from u_0 = sed0; every entry is a volume-0 seafloor step."""
    NE = 24
    j = np.arange(NE, dtype=np.float64)
    h = np.broadcast_to(floor - slope * j[None, :], (NE, NE)).copy()
    cells, deltas, vols, death = _trace(oracle, h, 9.5, 2.5, sed0=sed0, dir0=np.pi / 2)
    assert death == 2 and cells.size == NE - 2 * 2 - 3
    assert np.all(vols == 0.0) and np.all(cells % NE == np.arange(cells.size) + 4)
    u = sed0
    for k, d in enumerate(deltas):
        room = rooms[0] if k == 0 else rooms[1]
        want = min(0.3 * u, room)
        assert abs(d - want) <= 1e-13 * max(abs(want), 1.0), (k, d, want)
        u -= d
    assert deltas[0] == pytest.approx(rooms[0], rel=1e-15)   # the cap binds on the first step of each case


# ---------------------------------------------------------------------------
# uplift, replay, datum
# ---------------------------------------------------------------------------
def _globe(oracle, N=32, scale=1.0, core=None, **erosion):
    from globe_hip import stubs
    grid, f = stubs.earth_globe(N, 2, n_blobs=8)
    u = f["uplift_m"] * scale
    if core is not None:
        u[(2,) + core] = 5.0
    p = stubs.workload_params(N, 2, **erosion)
    st = oracle.OracleState.from_grid(grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], u, p.erosion)
    st.exchange_halos()
    return p, st, f


def test_uplift_is_mean_free_on_the_sphere(oracle):
    """test_erosion.py:545-570: the global uplift adds no net mass; only the
    datum moves; a window keeps its uplift as given."""
    p, st, _ = _globe(oracle, scale=300.0)
    inter = st.interior
    assert st.uplift[inter].mean() > 0.0
    m0, h0 = st.total_mass(), st.height.copy()
    oracle.apply_uplift(st)
    assert abs(st.total_mass() - m0) <= 1e-9 * max(abs(m0), 1.0)
    d, u = (st.height - h0)[inter], st.uplift[inter]
    np.testing.assert_allclose(d - d.mean(), u - u.mean(), atol=1e-15, rtol=0)
    assert d.max() > 0.0 > d.min()
    # a window: no datum, the uplift is mass
    wp = window_params(0)
    from windows import window_inputs
    w = oracle.OracleState.window(**window_inputs(32, "closed", wp, seed=0))
    w.uplift[...] = 0.01
    m0 = w.total_mass()
    oracle.apply_uplift(w)
    up = float(w.uplift[w.interior][w.mask[w.interior] == 1].sum())
    assert abs((w.total_mass() - m0) - up) <= 1e-9 * max(abs(m0), 1.0)


def test_uplift_cap_bounds_the_rise_and_stays_mean_free(oracle):
    """test_erosion.py:573-642: under the cap a cell gets exactly its own
    uplift less the datum, the core is pinned at the cap less the datum,
    the change is mass-free, and cap None is the uncapped rule bit for bit."""
    p, st, _ = _globe(oracle, scale=300.0, core=(slice(10, 20), slice(12, 22)), uplift_max_m=2.0)
    cap = oracle.uplift_cap(st, p.erosion)
    assert cap == pytest.approx(2.0 / st.height_unit_m)
    inter = st.interior
    u = st.uplift[inter]
    assert (u > cap).sum() >= 100 and (u < cap).sum() > 0.9 * u.size
    a = st.height.copy()
    m0 = st.total_mass()
    oracle.apply_uplift(st, cap)
    d = (st.height - a)[inter]
    mean = float(np.minimum(u, cap).mean())
    assert abs(st.total_mass() - m0) <= 1e-9 * max(abs(m0), 1.0)
    assert d.max() <= cap - mean + 1e-15
    np.testing.assert_allclose(d[u < cap], u[u < cap] - mean, atol=1e-15, rtol=0)
    np.testing.assert_allclose(d[u > cap], cap - mean, atol=1e-15, rtol=0)
    _, s1, _ = _globe(oracle, scale=300.0, core=(slice(10, 20), slice(12, 22)))
    _, s2, _ = _globe(oracle, scale=300.0, core=(slice(10, 20), slice(12, 22)))
    oracle.apply_uplift(s1)
    oracle.apply_uplift(s2, None)
    assert s1.height.tobytes() == s2.height.tobytes()
    p0, _, _ = _globe(oracle, uplift_max_m=0.0)
    assert oracle.uplift_cap(s1, p0.erosion) is None


@pytest.mark.parametrize("cap_m", [0.0, 2.0])
def test_replay_without_surface_processes_ends_at_bedrock(oracle, cap_m):
    """test_erosion.py:806-850: start_replay lowers the crust by the total
    the run will apply; replaying uplift and the datum hold for that many
    iterations (the particle and wasting passes off) ends at bedrock plus
    one global constant, capped or not."""
    n = 12
    p, st, f = _globe(oracle, scale=300.0, core=(slice(10, 20), slice(12, 22)), uplift_mode="replay",
                      iterations=n, uplift_max_m=cap_m)
    inter = st.interior
    bed = np.asarray(f["bedrock_m"], np.float64)[inter] / st.height_unit_m
    cap = oracle.uplift_cap(st, p.erosion)
    assert (cap is None) == (cap_m == 0.0)
    info = oracle.start_replay(st, p)
    d0 = st.height[inter] - bed
    assert info is not None and info["replay_max"] > 0.0 and np.ptp(d0) > 1e-6   # a real replay
    land = np.mean(st.surface()[inter] >= 0)
    assert abs(land - p.world.land_fraction) < 2.0 / d0.size
    for _ in range(n):
        oracle.apply_uplift(st, cap)
        oracle.hold_datum(st, oracle.held_land_fraction(st, p))
        st.exchange_halos()
    d = st.height[inter] - bed
    assert np.ptp(d) < 1e-9, np.ptp(d)
    # stack mode has no replay
    ps, ss, _ = _globe(oracle, uplift_mode="stack")
    assert oracle.start_replay(ss, ps) is None and oracle.uplift_replay(ss, ps.erosion) is None


def test_replay_is_a_no_op_on_a_window(oracle):
    """test_erosion.py:921-933."""
    from windows import window_inputs
    wp = window_params(0)
    wp.erosion.uplift_mode = "replay"
    w = oracle.OracleState.window(**window_inputs(32, "closed", wp, seed=0))
    h0 = w.height.copy()
    assert oracle.uplift_replay(w, wp.erosion) is None
    assert oracle.start_replay(w, wp) is None
    assert w.height.tobytes() == h0.tobytes()


@pytest.mark.parametrize("lf", [0.0, 0.3, 0.71, 1.0])
def test_hold_datum_leaves_the_requested_land_count(oracle, lf):
    """maps.py:1413-1451: after the hold exactly round(lf * M) interior
    cells are at or above sea level (the surface is continuous noise: no
    ties), and everything registered to the terrain moves rigidly with it:
    the route by q, a closed basin's (negative) base by q, the open sea's
    base (0) not at all (test_erosion.py:1201-1223)."""
    p, st, _ = _globe(oracle)
    inter = st.interior
    M = st.height[inter].size
    st.route = st.surface() + 0.25
    st.base = np.zeros_like(st.height)
    st.base[inter][:, :4, :4] = -3.0
    r0, b0, s0 = st.route.copy(), st.base.copy(), st.surface().copy()
    q = oracle.hold_datum(st, lf)
    n_land = int(np.sum(st.surface()[inter] >= 0.0))
    assert n_land == int(round(lf * M)), (n_land, lf * M)
    np.testing.assert_array_equal(st.surface(), s0 - q)
    np.testing.assert_array_equal(st.route, r0 - q)
    neg = b0 < 0.0
    np.testing.assert_array_equal(st.base[neg], b0[neg] - q)
    assert np.all(st.base[~neg] == 0.0)
    if 0.0 < lf < 1.0:
        # the new sea level sits on a surface value: one cell is exactly at 0
        assert np.sum(st.surface()[inter] == 0.0) >= 1
