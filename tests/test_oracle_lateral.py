"""The oracle's lateral erosion (particle.py:699-740, ``erosion.lateral_rate``):
a droplet that erodes while turning also cuts the bank one cell across the
flow on the outside of the bend, an entry of volume -2.

Pinned by the reference's own test (bake/tests/test_erosion.py:1757-1791:
no -2 entries at rate 0, a real share of them at rate 50, and the ground
moves), and by what the formula fixes and the change lists expose: a lateral
entry follows the eroding step's own entry, its amount is
rate x |sin(turn)| x that step's erosion unless a cap cut it (never more),
with the turn read off the two path entries' momenta (direction x volume);
and the bank cell lies within a cell of the current cell plus the outward
normal (sgn(turn) * (s_b, -s_a)) -- a flipped side, a dropped factor or a
wrong rate fail one of these.  Plus the closed window's mass conservation
with the cuts on.
"""
from __future__ import annotations

import numpy as np
import pytest
from scipy import ndimage

from windows import window_inputs, window_params


def _plain_window_inputs(p, n=64, seed=3):
    """The reference test's ``_plain_window`` and ``_run_window`` inputs (a
    0.02 m-per-cell incline with smoothed noise, 1 m of loose sediment on
    hard rock, a one-cell outside ring, 30-m height unit) as
    ``ErosionState.window`` arguments (flat metric, halo 1)."""
    surf = 40.0 - 0.02 * np.repeat(np.arange(n, dtype=np.float64)[:, None], n, 1)
    surf += 0.12 * ndimage.gaussian_filter(np.random.default_rng(seed).normal(size=(n, n)), 4.0)
    mask = np.ones((n, n), np.uint8)
    mask[0, :] = mask[-1, :] = mask[:, 0] = mask[:, -1] = 0
    metric = np.zeros((n, n, 3), np.float32)
    metric[..., 0] = 1.0
    metric[..., 2] = 1.0
    z = np.zeros((n, n))
    return dict(height_m=surf, sediment_m=np.full((n, n), 1.0), discharge=z, momentum=np.zeros((n, n, 2)),
                hardness=np.ones((n, n), np.float32), precip=np.ones((n, n), np.float32),
                evap=np.ones((n, n), np.float32), uplift_m=z, mask=mask, metric=metric, metric_inv=metric, H=1,
                height_unit_m=30.0)


def _lists(oracle, rate, its=1, ppc=0.3):
    p = window_params(0, lateral_rate=rate, max_steps=60)
    st = oracle.OracleState.window(**_plain_window_inputs(p))
    got = []

    def hook(cl_cell, cl_delta, cl_vol, cl_mom, cl_count, cap):
        for q in range(cl_count.size):
            sl = slice(q * cap, q * cap + int(cl_count[q]))
            got.append((cl_cell[sl].copy(), cl_delta[sl].copy(), cl_vol[sl].copy(), cl_mom[sl].copy()))

    for it in range(its):
        oracle.run_iteration(st, p, it, particles_per_cell=ppc, rng_stage="refine", diag_lists=hook)
    return got, st


def test_lateral_cuts_appear_with_the_rate_and_move_ground(oracle):
    """bake/tests/test_erosion.py:1757-1791 on the oracle."""
    l0, s0 = _lists(oracle, 0.0)
    l1, s1 = _lists(oracle, 50.0)
    n0 = sum(v.size for _, _, v, _ in l0)
    lat0 = sum(int((v == -2.0).sum()) for _, _, v, _ in l0)
    n1 = sum(v.size for _, _, v, _ in l1)
    lat1 = sum(int((v == -2.0).sum()) for _, _, v, _ in l1)
    assert lat0 == 0 and n0 > 100
    assert lat1 > 0.01 * n1, (lat1, n1)
    assert not np.array_equal(s0.surface(), s1.surface())


def test_lateral_amount_and_side_follow_the_bend(oracle):
    rate = 0.5
    lists, st = _lists(oracle, rate)
    NE = st.NE
    n_lat = n_match = n_side = 0
    for cells, delta, vol, mom in lists:
        for k in np.flatnonzero(vol == -2.0):
            n_lat += 1
            j = k - 1
            # the eroding step's own entry comes right before its cut
            assert vol[j] > 0.0 and delta[j] < 0.0, (vol[j], delta[j])
            prev = [i for i in range(j) if vol[i] > 0.0]
            if not prev:
                continue
            i = prev[-1]
            d0 = mom[i].astype(np.float64) / float(vol[i])
            d1 = mom[j].astype(np.float64) / float(vol[j])
            cross = d0[0] * d1[1] - d0[1] * d1[0]
            turn = abs(cross)
            ratio = delta[k] / delta[j]
            # the turn read back from float32 momenta and volumes is good to
            # ~4e-7 absolute; a cap only ever cuts the amount down
            tol = rate * (1e-6 + 1e-4 * turn)
            assert ratio <= rate * turn + tol, (ratio, rate * turn)
            if abs(ratio - rate * turn) > tol:
                continue
            n_match += 1
            if turn < 1e-5:
                continue
            n_side += 1
            sgn = 1.0 if cross > 0 else -1.0
            ci, cj = divmod(int(cells[j]) % (NE * NE), NE)
            bi, bj = divmod(int(cells[k]) % (NE * NE), NE)
            # bank = floor(x + sgn s_b), floor(y - sgn s_a) with the cell = floor(x), floor(y)
            assert abs((bi - ci) - sgn * d1[1]) <= 1.0 + 1e-5 and abs((bj - cj) + sgn * d1[0]) <= 1.0 + 1e-5, \
                ((bi - ci, bj - cj), sgn, d1)
    assert n_lat > 50, n_lat
    assert n_match >= 0.9 * n_lat, (n_match, n_lat)
    assert n_side >= 0.5 * n_match, (n_side, n_match)


def test_closed_window_conserves_mass_with_lateral_cuts(oracle):
    """test_erosion.py:191-206's invariant with the cuts on: each cut is an
    erosion that goes into the droplet's load."""
    p = window_params(0, lateral_rate=20.0)
    st = oracle.OracleState.window(**window_inputs(48, "closed", p, seed=2, deposit_on_exit=True))
    m0 = st.total_mass()
    cuts = []

    def hook(cl_cell, cl_delta, cl_vol, cl_mom, cl_count, cap):
        for q in range(cl_count.size):
            cuts.append(int((cl_vol[q * cap: q * cap + int(cl_count[q])] == -2.0).sum()))

    for it in range(6):
        oracle.run_iteration(st, p, it, diag_lists=hook)
    m1 = st.total_mass()
    assert sum(cuts) > 100, sum(cuts)
    assert abs(m1 - m0) <= 1e-6 * abs(m0), (m0, m1)
