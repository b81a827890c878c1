"""The oracle's side passes -- mass wasting (thermal + creep), the slope
limits, the Laplace-Beltrami smoother and isostasy -- pinned by the
reference's own tests of them (bake/tests/test_erosion.py:645-720, 1599-1619;
bake/tests/test_cubesphere.py:314-336) and by closed forms: a spike on a flat
plane (the D8 table, the diagonal distance, the 2 x max cap, sediment before
bedrock), exact conservation between active cells (the opposite-direction
index), the submerged receiver's room, mirror symmetry, the Laplacian of
quadratics under a constant metric (the metric components' roles), and the
flexural response's width (the step count and the kappa / cell^2 scaling).
"""
from __future__ import annotations

import numpy as np
import pytest

from globe_hip.config import ErosionParams
from globe_hip.cubesphere import get_grid


def _flat_state(oracle, NE, H=2, height=None, sediment=None, metric=(1.0, 0.0, 1.0)):
    """One-face state on a plane with a constant metric (cell units)."""
    F = 1
    h = np.zeros((F, NE, NE)) if height is None else np.asarray(height, np.float64).reshape(F, NE, NE).copy()
    s = np.zeros((F, NE, NE)) if sediment is None else np.asarray(sediment, np.float64).reshape(F, NE, NE).copy()
    g = np.empty((F, NE, NE, 3), np.float32)
    g[...] = metric
    det = float(metric[0]) * metric[2] - float(metric[1]) ** 2
    gi = np.empty_like(g)
    gi[...] = (metric[2] / det, -metric[1] / det, metric[0] / det)
    z = np.zeros((F, NE, NE))
    return oracle.OracleState(height=h, sediment=s, discharge=z.copy(), momentum=np.zeros((F, NE, NE, 2)),
                              hardness=np.zeros((F, NE, NE), np.float32), precip=np.ones((F, NE, NE), np.float32),
                              evap=np.ones((F, NE, NE), np.float32), uplift=z.copy(), mask=np.ones((F, NE, NE), np.uint8),
                              metric=g, metric_inv=gi, N=NE - 2 * H, H=H, spherical=False, height_unit_m=1.0)


# ---------------------------------------------------------------------------
# thermal erosion / creep
# ---------------------------------------------------------------------------
def test_spike_sheds_a_quarter_capped_share_to_each_neighbour(oracle):
    """A spike of 10 on a flat plane, talus 0, rate 0.5: each D8 outflow is
    0.5 x 0.5 x 10 = 2.5, their sum 20 exceeds 2 x the largest (5), so all
    scale by 5 / 20: the spike loses 5 (bedrock: it has no sediment) and
    each neighbour gains 0.625 of sediment."""
    NE = 9
    h = np.zeros((NE, NE))
    h[4, 4] = 10.0
    st = _flat_state(oracle, NE, height=h)
    oracle.mass_wasting_pass(st, np.zeros((1, NE, NE)), 0.5, 0.0)
    assert st.height[0, 4, 4] == 5.0 and st.sediment[0, 4, 4] == 0.0
    ring = np.ones((3, 3), bool)
    ring[1, 1] = False
    assert np.all(st.sediment[0, 3:6, 3:6][ring] == 0.625)
    st.sediment[0, 3:6, 3:6] = 0.0
    assert not st.sediment.any() and np.count_nonzero(st.height) == 1


def test_talus_and_diagonal_distance(oracle):
    """Talus 1, spike 10: axis neighbours get 0.25 (10 - 1), diagonals
    0.25 (10 - sqrt 2), all scaled by 2 x the larger / their sum."""
    NE = 9
    h = np.zeros((NE, NE))
    h[4, 4] = 10.0
    st = _flat_state(oracle, NE, height=h)
    oracle.mass_wasting_pass(st, np.ones((1, NE, NE)), 0.5, 0.0)
    ax, dg = 0.25 * 9.0, 0.25 * (10.0 - np.sqrt(2.0))
    sc = 2.0 * ax / (4 * ax + 4 * dg)
    assert st.sediment[0, 5, 4] == pytest.approx(sc * ax, rel=1e-15)
    assert st.sediment[0, 4, 3] == pytest.approx(sc * ax, rel=1e-15)
    assert st.sediment[0, 5, 5] == pytest.approx(sc * dg, rel=1e-15)
    assert st.sediment[0, 3, 5] == pytest.approx(sc * dg, rel=1e-15)
    assert st.height[0, 4, 4] == pytest.approx(10.0 - sc * (4 * ax + 4 * dg), rel=1e-15)


def test_sediment_goes_first_and_thermal_max_caps(oracle):
    NE = 9
    h = np.zeros((NE, NE))
    s = np.zeros((NE, NE))
    h[4, 4], s[4, 4] = 8.0, 2.0
    st = _flat_state(oracle, NE, height=h, sediment=s)
    oracle.mass_wasting_pass(st, np.zeros((1, NE, NE)), 0.5, 1.5)  # cap 1.5 < the uncapped 5
    assert st.sediment[0, 4, 4] == pytest.approx(0.5, rel=1e-15)   # 2 - 1.5, bedrock untouched
    assert st.height[0, 4, 4] == 8.0
    assert st.sediment[0, 4, 5] == pytest.approx(1.5 / 8, rel=1e-15)


def test_below_talus_nothing_moves(oracle):
    rng = np.random.default_rng(1)
    NE = 20
    h = rng.uniform(0, 0.5, (NE, NE))
    st = _flat_state(oracle, NE, height=h)
    h0 = st.height.copy()
    oracle.mass_wasting_pass(st, np.ones((1, NE, NE)), 0.5, 0.0)
    assert np.array_equal(st.height, h0) and not st.sediment.any()


def test_conserved_between_active_cells_and_mirror_symmetric(oracle):
    """Every outflow is some active neighbour's inflow (land receivers take
    all), and the pass commutes with a mirror of the plane."""
    rng = np.random.default_rng(2)
    NE = 24
    h = 5.0 + rng.uniform(0, 6, (NE, NE))
    s = rng.uniform(0, 0.5, (NE, NE))
    talus = np.full((1, NE, NE), 0.6) + rng.uniform(0, 0.3, (1, NE, NE))
    st = _flat_state(oracle, NE, height=h, sediment=s, metric=(1.2, 0.15, 0.9))
    st.mask[0, 0, :] = st.mask[0, -1, :] = st.mask[0, :, 0] = st.mask[0, :, -1] = 0
    m0 = (st.height + st.sediment)[0, 1:-1, 1:-1].sum()
    oracle.mass_wasting_pass(st, talus, 0.5, 0.0)
    m1 = (st.height + st.sediment)[0, 1:-1, 1:-1].sum()
    assert m1 == pytest.approx(m0, rel=1e-13)
    moved = np.abs(st.height + st.sediment - h[None] - s[None]).max()
    assert moved > 0.1
    # mirror i -> NE-1-i: the metric's cross term changes sign
    st2 = _flat_state(oracle, NE, height=h[::-1], sediment=s[::-1], metric=(1.2, -0.15, 0.9))
    st2.mask[...] = st.mask[:, ::-1]
    oracle.mass_wasting_pass(st2, talus[:, ::-1].copy(), 0.5, 0.0)
    np.testing.assert_allclose(st2.height[:, ::-1], st.height, rtol=0, atol=1e-13)
    np.testing.assert_allclose(st2.sediment[:, ::-1], st.sediment, rtol=0, atol=1e-13)


def test_submerged_receiver_fills_to_dep_floor_only(oracle):
    """A cliff of 10 above a sea cell at -1: the sea cell takes at most
    -dep_floor - (-1) and the cliff loses exactly what was received."""
    NE = 7
    h = np.full((NE, NE), 10.0)
    h[3, 3] = -1.0
    st = _flat_state(oracle, NE, height=h)
    st.mask[0, 0, :] = st.mask[0, -1, :] = st.mask[0, :, 0] = st.mask[0, :, -1] = 0
    m0 = (st.height + st.sediment)[0, 1:-1, 1:-1].sum()
    oracle.mass_wasting_pass(st, np.zeros((1, NE, NE)), 0.5, 0.0, dep_floor=0.02)
    sea = st.height[0, 3, 3] + st.sediment[0, 3, 3]
    assert sea == pytest.approx(-0.02, abs=1e-12)
    assert (st.height + st.sediment)[0, 1:-1, 1:-1].sum() == pytest.approx(m0, rel=1e-13)


def test_creep_leaves_the_sea_alone(oracle):
    """Creep (talus 0) freezes submerged cells: a coast next to the sea
    creeps along the land only."""
    NE = 10
    h = np.full((NE, NE), 3.0)
    h[:, :4] = -5.0
    h[5, 6] = 6.0
    st = _flat_state(oracle, NE, height=h)
    p = ErosionParams(thermal_rate=0.0, creep_rate=0.3, thermal_max=1e9)
    oracle.thermal_erosion(st, p)
    assert np.array_equal(st.height[0, :, :4], h[:, :4]) and not st.sediment[0, :, :4].any()
    assert st.sediment[0, 5, 5] > 0.0


# ---------------------------------------------------------------------------
# slope limits (bake/tests/test_erosion.py:1599-1619)
# ---------------------------------------------------------------------------
def test_slope_erode_cap_is_soillibs_downhill_limit(oracle):
    h = np.full((7, 7), 5.0)
    h[2, 2] = 1.0                    # a one-cell pit
    h[4, 5] = 3.0                    # (4, 4) drains to it, 2 below
    h[4, 3] = 9.0                    # uphill of (4, 4): ignored
    st = _flat_state(oracle, 7, height=h)
    out = oracle.slope_erode_cap(st, 0.25, 0.02)
    k = 0.25 * np.sqrt(2.0)
    assert out[0, 2, 2] == 0.0
    assert out[0, 4, 4] == pytest.approx(k * 2.0)
    assert out[0, 0, 3] == pytest.approx(k * 0.02)
    st.mask[0, 3, 3] = 0
    out = oracle.slope_erode_cap(st, 0.25, 0.02)
    assert out[0, 3, 3] == 0.0


def test_tilted_plane_cap_is_the_gradient_norm(oracle):
    i, j = np.meshgrid(np.arange(9.0), np.arange(9.0), indexing="ij")
    st = _flat_state(oracle, 9, height=0.3 * i - 0.4 * j)
    out = oracle.slope_erode_cap(st, 0.1, 0.0)
    assert out[0, 1:-1, 1:-1] == pytest.approx(np.full((7, 7), 0.1 * np.sqrt(2.0) * 0.5), rel=1e-14)


# ---------------------------------------------------------------------------
# the Laplace-Beltrami smoother
# ---------------------------------------------------------------------------
def _raw_laplacian(oracle, d, ginv, sq):
    import ctypes

    F, NE = d.shape[0], d.shape[1]
    out = np.empty_like(d)
    oracle.lib().or_laplacian(F, NE, d.ctypes.data_as(ctypes.c_void_p), ginv.ctypes.data_as(ctypes.c_void_p),
                              sq.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p))
    return out


@pytest.mark.parametrize("fn,expect", [
    (lambda i, j: i * i, lambda a, b, c: 2 * a),
    (lambda i, j: j * j, lambda a, b, c: 2 * c),
    (lambda i, j: i * j, lambda a, b, c: 2 * b),
    (lambda i, j: 3 * i - 2 * j + 1, lambda a, b, c: 0.0),
])
def test_laplacian_of_quadratics_under_a_constant_metric(oracle, fn, expect):
    """L f = g^ab d_a d_b f for a constant metric: the i-, j- and cross
    components of the inverse metric each have their own closed form."""
    NE = 10
    a, b, c = 1.3, 0.4, 0.7
    i, j = np.meshgrid(np.arange(NE, dtype=np.float64), np.arange(NE, dtype=np.float64), indexing="ij")
    d = np.ascontiguousarray(fn(i, j)[None], np.float64)
    ginv = np.empty((1, NE, NE, 3))
    ginv[...] = (a, b, c)
    sq = np.full((1, NE, NE), 1.7)
    out = _raw_laplacian(oracle, d, ginv, sq)
    assert out[0, 1:-1, 1:-1] == pytest.approx(np.full((NE - 2, NE - 2), expect(a, b, c)), abs=1e-12)
    assert not out[0, 0].any() and not out[0, :, -1].any()


def test_laplacian_eigenfunction(oracle):
    """bake/tests/test_cubesphere.py:314-336: spherical harmonics are
    eigenfunctions of the sphere's Laplacian, -l(l+1)/R^2, to 2 %, with no
    seam artefacts at the face edges."""
    g = get_grid(64, 4, 50.0)
    R = g.R_planet
    for name, fn, l in (
        ("Y20", lambda p: 3 * p[..., 2] ** 2 - 1, 2),
        ("Y11", lambda p: p[..., 0], 1),
        ("Y33", lambda p: p[..., 0] ** 3 - 3 * p[..., 0] * p[..., 1] ** 2, 3),
    ):
        d = np.ascontiguousarray(fn(g.centers), np.float64)
        oracle.exchange(g, d)
        L = oracle.laplacian(g, d)[:, 4:-4, 4:-4]
        f = d[:, 4:-4, 4:-4]
        mask = np.abs(f) > 0.3 * np.abs(f).max()
        ratio = L[mask] / f[mask]
        expect = -l * (l + 1) / R ** 2
        assert np.abs(ratio / expect - 1).max() < 0.02, (name, ratio.min(), ratio.max(), expect)
        res = np.abs(L - expect * f)
        inner = res[:, 8:-8, 8:-8].max()
        edge = max(res[:, [0, -1], :].max(), res[:, :, [0, -1]].max())
        assert edge < 5 * inner + 1e-12, (name, edge / inner)


# ---------------------------------------------------------------------------
# isostasy (bake/tests/test_erosion.py:645-720)
# ---------------------------------------------------------------------------
def _iso_state(oracle, N=32, H=4):
    """A globe of N=32 cells of 50 m with flat land, as the reference's tiny
    world (flexure 0.2 km = 4 cells)."""
    g = get_grid(N, H, 50.0)
    F, NE = 6, g.NE
    ones = np.ones((F, NE, NE), np.float32)
    st = oracle.OracleState.from_grid(g, np.full((F, NE, NE), 100.0), 0.0 * ones, ones, ones, np.zeros((F, NE, NE)))
    p = ErosionParams(isostasy=0.8, flexure_km=0.2)
    return st, p


def _point_load(st, face, i, j, size=4):
    st.iso_acc = np.zeros_like(st.height)
    H = st.H
    st.iso_acc[face, H + i:H + i + size, H + j:H + j + size] = -1.0
    return size * size


def test_isostasy_is_regional_and_mean_free(oracle):
    st, p = _iso_state(oracle)
    n = _point_load(st, 0, 14, 14)
    h0, s0 = st.height.copy(), st.sediment.copy()
    info = oracle.apply_isostasy(st, p)
    dh = (st.height - h0)[st.interior]
    assert abs(dh.mean()) < 1e-9 * dh.max()
    peak = np.unravel_index(np.argmax(dh), dh.shape)
    assert peak[0] == 0 and 13 <= peak[1] <= 18 and 13 <= peak[2] <= 18
    assert int((dh > 0.01 * dh.max()).sum()) > 4 * n
    assert abs((dh - dh.min()).sum() - 0.8 * n) < 0.15 * 0.8 * n
    assert np.array_equal(st.sediment, s0)
    assert not st.iso_acc.any()
    assert info["diffusion_steps"] > 0 and info["sigma_cells"] == 4.0


def test_isostasy_width_is_the_flexural_sigma(oracle):
    """n diffusion steps of kappa add 2 kappa n ~ sigma^2 of variance per
    axis (cell units; the metric is ~identity at a face centre)."""
    st, p = _iso_state(oracle)
    _point_load(st, 0, 15, 15, size=2)
    h0 = st.height.copy()
    info = oracle.apply_isostasy(st, p)
    dh = (st.height - h0)[0, st.H:-st.H, st.H:-st.H]
    w = dh - dh.min()
    i = np.arange(dh.shape[0]) - 15.5
    var_i = float((w.sum(1) * i ** 2).sum() / w.sum())
    expect = 2.0 * info["kappa"] * info["diffusion_steps"]
    assert expect == pytest.approx(info["sigma_cells"] ** 2, rel=0.1)
    assert var_i == pytest.approx(expect, rel=0.2)


def test_isostasy_has_no_seam(oracle):
    st, p = _iso_state(oracle)
    _point_load(st, 0, 0, 12)
    h0 = st.height.copy()
    oracle.apply_isostasy(st, p)
    d = st.height - h0
    oracle.exchange(st.grid, d, linear=False)
    H = st.H
    f = d[0]
    jump_edge = np.abs(f[H, H:-H] - f[H - 1, H:-H]).max()
    jump_in = np.abs(np.diff(f[H:-H, H:-H], axis=0)).max()
    assert jump_edge < 4.0 * jump_in


def test_isostasy_smoother_is_stable_on_white_noise(oracle):
    st, p = _iso_state(oracle)
    rng = np.random.default_rng(3)
    st.iso_acc = np.zeros_like(st.height)
    st.iso_acc[st.interior] = rng.standard_normal(st.iso_acc[st.interior].shape)
    peak = float(np.abs(st.iso_acc[st.interior]).max())
    h0 = st.height.copy()
    info = oracle.apply_isostasy(st, p)
    dh = (st.height - h0)[st.interior]
    assert np.isfinite(dh).all()
    assert info["diffusion_steps"] >= 20
    assert float(np.abs(dh).max()) <= 0.8 * peak * 1.05


def test_kappa_is_below_the_monotone_bound(oracle):
    """kappa x lam_max <= 1 (monotone explicit step): lam_max is at least
    the Rayleigh quotient of any field, e.g. the face checkerboard."""
    g = get_grid(32, 4, 50.0)
    kappa = oracle.smoothing_kappa(g)
    H, N = g.H, g.N
    d = np.zeros((6, g.NE, g.NE))
    ii, jj = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    d[:, H:H + N, H:H + N] = ((ii + jj) % 2 * 2 - 1)[None]
    oracle.exchange(g, d, linear=True)
    Lv = oracle.laplacian(g, d).astype(np.float64) * g.cell_size_m ** 2
    sl = (slice(None), slice(H + 2, H + N - 2), slice(H + 2, H + N - 2))
    rq = -float((Lv[sl] * d[sl]).sum() / (d[sl] ** 2).sum())
    assert 0.9 / kappa >= 0.99 * rq
    assert kappa * rq <= 0.9 * 1.01
    assert 0.9 / kappa < 1.5 * rq
