"""The side passes on the GPU against the oracle, same inputs.

Bars: the mass-wasting passes (thermal + creep), the slope limit and the
smoother's Laplacian / step are elementwise float64 chains in the
reference's order, so bit for bit.  Quantities behind a reduction or the
halo exchange (kappa's norms, the rebound's mean, the einsum halo weights)
agree to rounding: kappa to 1e-12 relative, the isostatic rebound to 1e-12
of its largest value.  Whole steps with every implemented side pass stay
within 1e-6 of the oracle's relief (the halo exchange is not bit-pinned), and
the fast mode's captured graph replays its eager iterations bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from globe_hip import stubs
from globe_hip.cubesphere import get_grid
from globe_hip.erosion import fast, maps
from globe_hip.erosion import particle as pk
from windows import window_inputs, window_params

pytestmark = pytest.mark.gpu


def _relief_globe(N=48, H=2, seed=0):
    """A globe of 50 m cells with ~60 cells of relief and a sea: slopes well
    above the talus, so every branch of the thermal passes is taken."""
    g = get_grid(N, H, 50.0)
    rng = np.random.default_rng(seed)
    p = g.centers
    bed = 1500.0 * (np.sin(4 * p[..., 0] + 1.0) * np.cos(3 * p[..., 1]) + 0.6 * np.sin(7 * p[..., 2])) + 0.0
    bed = bed + 250.0 * rng.standard_normal(bed.shape)
    F, NE = 6, g.NE
    hard = rng.uniform(0, 1, (F, NE, NE)).astype(np.float32)
    ones = np.ones((F, NE, NE), np.float32)
    return g, dict(bedrock_m=bed, hardness=hard, precip=ones, evap=ones, uplift_m=np.zeros((F, NE, NE)))


def _pair(oracle, cuda, g, f, ep, mode="check", sed_seed=1):
    args = (g, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"], ep)
    ost = oracle.OracleState.from_grid(*args)
    dst = maps.ErosionState.from_grid(*args, device=cuda, mode=mode)
    sed = np.random.default_rng(sed_seed).uniform(0, 2.0, ost.height.shape)
    ost.sediment[...] = sed
    dst.sediment.copy_(torch.from_numpy(sed))
    return ost, dst


def _params(**erosion):
    p = stubs.workload_params(48, 2)
    for k, v in erosion.items():
        setattr(p.erosion, k, v)
    return p


def test_thermal_and_creep_bit_exact(oracle, cuda):
    g, f = _relief_globe()
    p = _params(thermal_rate=0.5, creep_rate=0.3, thermal_max=60.0, dep_floor_m=1.0)
    ost, dst = _pair(oracle, cuda, g, f, p.erosion)
    assert (ost.height < 0).mean() > 0.1  # a sea to fill and to freeze for creep
    s_before = ost.surface().copy()
    oracle.thermal_erosion(ost, p)
    maps.thermal_erosion(dst, p)
    assert np.abs(ost.surface() - s_before).max() > 1.0
    assert np.array_equal(dst.height.cpu().numpy(), ost.height)
    assert np.array_equal(dst.sediment.cpu().numpy(), ost.sediment)


def test_thermal_passes_bit_exact(oracle, cuda):
    """Each pass on its own, with a mask that has outside and frozen cells."""
    g, f = _relief_globe(N=32, seed=4)
    p = _params(thermal_rate=0.5)
    ost, dst = _pair(oracle, cuda, g, f, p.erosion)
    rng = np.random.default_rng(9)
    m = rng.choice(np.array([0, 1, 1, 1, 1, 2], np.uint8), ost.mask.shape)
    ost.mask[...] = m
    dst.mask.copy_(torch.from_numpy(m))
    talus = oracle.talus_of(p.erosion, ost.hardness)
    F, NE = ost.F, ost.NE
    import ctypes

    L, P = oracle.lib(), lambda a: a.ctypes.data_as(ctypes.c_void_p)
    ot, osc = np.zeros_like(ost.height), np.zeros_like(ost.height)
    L.or_thermal_pass_a(F, NE, P(ost.height), P(ost.sediment), P(ost.metric), P(talus), P(m), 0.5, 1.2, P(ot), P(osc))
    ors = np.ones_like(ost.height)
    L.or_thermal_pass_r(F, NE, P(ost.height), P(ost.sediment), P(ost.metric), P(talus), P(m), 0.5, P(osc), P(ors), 0.02)
    onh, ons = ost.height.copy(), ost.sediment.copy()
    L.or_thermal_pass_b(F, NE, P(ost.height), P(ost.sediment), P(ost.metric), P(talus), P(m), 0.5, P(osc), P(ors),
                        P(onh), P(ons))
    dt = torch.empty_like(dst.height)
    pk.talus(dst.hardness, p.erosion.talus_slope_soft, p.erosion.talus_slope_hard, dt)
    assert np.array_equal(dt.cpu().numpy(), talus)
    tot, sc, rs = torch.empty_like(dst.height), torch.empty_like(dst.height), torch.empty_like(dst.height)
    pk.thermal_pass_a(dst.height, dst.sediment, dst.metric, dt, dst.mask, 0.5, 1.2, tot, sc)
    pk.thermal_pass_r(dst.height, dst.sediment, dst.metric, dt, dst.mask, 0.5, sc, rs, 0.02)
    nh, ns = torch.empty_like(dst.height), torch.empty_like(dst.height)
    pk.thermal_pass_b(dst.height, dst.sediment, dst.metric, dt, dst.mask, 0.5, tot, sc, rs, nh, ns)
    for a, b, n in ((tot, ot, "out_total"), (sc, osc, "scale"), (rs, ors, "rscale"), (nh, onh, "new_h"), (ns, ons, "new_s")):
        assert np.array_equal(a.cpu().numpy(), b), n
    assert (osc > 0).mean() > 0.2 and (ors < 1).any() and (osc == 1.0).any() and ((osc > 0) & (osc < 1)).any()


def test_slope_erode_cap_bit_exact(oracle, cuda):
    g, f = _relief_globe(N=32, seed=2)
    p = _params()
    ost, dst = _pair(oracle, cuda, g, f, p.erosion)
    ost.mask[0, 5, 5] = 0
    dst.mask[0, 5, 5] = 0
    out = torch.empty_like(dst.height)
    pk.slope_erode_cap(dst.height, dst.sediment, dst.mask, 0.25, 0.02, out)
    ref = oracle.slope_erode_cap(ost, 0.25, 0.02)
    assert np.array_equal(out.cpu().numpy(), ref)
    assert (ref == 0).sum() > 0 and ref.max() > 0


def test_laplacian_and_smoothing_step_bit_exact(oracle, cuda):
    """Same exchanged input on both sides: the raw operator, the float32
    Laplacian x c2 and one smoothing step are the reference's chains."""
    g = get_grid(48, 2, 50.0)
    rng = np.random.default_rng(5)
    d = np.zeros((6, g.NE, g.NE))
    d[:, 2:-2, 2:-2] = rng.standard_normal((6, 48, 48))
    oracle.exchange(g, d, linear=True)
    import ctypes

    raw = np.empty_like(d)
    ginv = np.ascontiguousarray(g.metric_inv, np.float64)
    sq = np.ascontiguousarray(g.sqrt_det_metric, np.float64)
    oracle.lib().or_laplacian(6, g.NE, *[a.ctypes.data_as(ctypes.c_void_p) for a in (d, ginv, sq, raw)])
    dd = torch.from_numpy(d).to(cuda)
    gi = torch.from_numpy(np.ascontiguousarray(g.metric_inv, np.float32)).to(cuda)
    sqd = torch.from_numpy(np.ascontiguousarray(g.sqrt_det_metric, np.float32)).to(cuda)
    out = torch.empty_like(dd)
    cs2 = float(g.cell_size_m) ** 2
    pk.laplacian(dd, gi, sqd, cs2, 0, 0.0, 0.0, out)
    assert np.array_equal(out.cpu().numpy(), raw)
    c2 = 50.0 * 50.0
    pk.laplacian(dd, gi, sqd, cs2, 1, 0.0, c2, out)
    assert np.array_equal(out.cpu().numpy(), oracle.laplacian(g, d).astype(np.float64) * c2)
    kappa = 0.0917
    pk.laplacian(dd, gi, sqd, cs2, 2, kappa, c2, out)
    assert np.array_equal(out.cpu().numpy(), d + kappa * oracle.laplacian(g, d).astype(np.float64) * c2)


def test_kappa_and_step_count_match(oracle, cuda):
    g, f = _relief_globe(N=32)
    p = _params(isostasy=0.8, flexure_km=0.2)
    ost, dst = _pair(oracle, cuda, g, f, p.erosion)
    kd = maps.smoothing_kappa(dst)
    ko = oracle.smoothing_kappa(g)
    assert kd == pytest.approx(ko, rel=1e-12)
    assert maps.isostasy_steps(dst, p.erosion)[2] == oracle.isostasy_steps(ost, p.erosion)[2] > 0


def _iso_pair(oracle, cuda, N=32, H=4):
    g = get_grid(N, H, 50.0)
    F, NE = 6, g.NE
    ones = np.ones((F, NE, NE), np.float32)
    args = (g, np.full((F, NE, NE), 100.0), 0.0 * ones, ones, ones, np.zeros((F, NE, NE)))
    p = _params(isostasy=0.8, flexure_km=0.2)
    return oracle.OracleState.from_grid(*args), maps.ErosionState.from_grid(*args, device=cuda), p


def _load(st, acc):
    if isinstance(st.height, torch.Tensor):
        st.iso_acc = torch.from_numpy(acc.copy()).to(st.height.device)
    else:
        st.iso_acc = acc.copy()


def test_isostasy_matches_oracle_and_the_reference_properties(oracle, cuda):
    ost, dst, p = _iso_pair(oracle, cuda)
    H = ost.H
    acc = np.zeros_like(ost.height)
    acc[0, H + 14:H + 18, H + 14:H + 18] = -1.0
    acc[3, H:H + 4, H + 10:H + 14] = 0.5          # a load at a face edge
    _load(ost, acc)
    _load(dst, acc)
    h0 = ost.height.copy()
    io = oracle.apply_isostasy(ost, p)
    idv = maps.apply_isostasy(dst, p)
    assert idv["diffusion_steps"] == io["diffusion_steps"] > 0
    do = (ost.height - h0)
    dd = dst.height.cpu().numpy() - h0
    scale = np.abs(do).max()
    assert np.abs(dd - do).max() <= 1e-12 * scale, np.abs(dd - do).max() / scale
    for k in ("rebound_max", "rebound_min", "load_abs_cells"):
        assert idv[k] == pytest.approx(io[k], rel=1e-12, abs=1e-15), k
    # bake/tests/test_erosion.py:652-670 on the device's result
    dh = dd[dst.interior]
    assert abs(dh.mean()) < 1e-9 * dh.max()
    peak = np.unravel_index(np.argmax(dh), dh.shape)
    assert peak[0] == 0 and 13 <= peak[1] <= 18 and 13 <= peak[2] <= 18
    assert np.array_equal(dst.sediment.cpu().numpy(), ost.sediment)
    assert not dst.iso_acc.any().item()


def test_isostasy_stable_on_white_noise(oracle, cuda):
    ost, dst, p = _iso_pair(oracle, cuda)
    rng = np.random.default_rng(3)
    acc = np.zeros_like(ost.height)
    acc[ost.interior] = rng.standard_normal(acc[ost.interior].shape)
    _load(dst, acc)
    h0 = dst.height.clone()
    info = maps.apply_isostasy(dst, p)
    dh = (dst.height - h0).cpu().numpy()[dst.interior]
    assert np.isfinite(dh).all() and info["diffusion_steps"] >= 20
    assert float(np.abs(dh).max()) <= 0.8 * float(np.abs(acc[ost.interior]).max()) * 1.05


def test_steps_with_side_passes_stay_close(oracle, hiplib, cuda):
    """Three check-mode steps with thermal, creep, slope limits and isostasy
    every 2 iterations against the oracle's steps (the oracle on the device's
    correctly rounded tan / atan / sincos).  Before the first rebound the
    steps are bit for bit; the rebound's norms and mean are device
    reductions in another order than numpy's, so after it the surfaces
    differ by rounding, and a droplet at a bifurcation of its path may then
    walk another way: the third step is bounded by the share of cells that
    moved beyond rounding."""
    g, f = _relief_globe(N=32, seed=7)
    p = _params(thermal_rate=0.5, creep_rate=0.05, thermal_max=60.0, isostasy=0.8, flexure_km=0.2, isostasy_every=2,
                slope_limit_erode=0.05, slope_limit_deposit=0.5)
    ost, dst = _pair(oracle, cuda, g, f, p.erosion)
    inter = ost.interior
    gaps = []
    oracle.set_libm(*oracle.device_libm(hiplib))
    try:
        for it in range(3):
            so = oracle.step(ost, p, it)
            sd = maps.step(dst, p, it)
            assert sd["particles"] == so["particles"]
            assert ("isostasy" in sd) == ("isostasy" in so) == (it == 1)
            h_o = (ost.height + ost.sediment)[inter]
            h_d = (dst.height + dst.sediment).cpu().numpy()[inter]
            scale = np.abs(h_o).max()
            d = np.abs(h_d - h_o)
            gaps.append((float(d.max() / scale), float((d > 1e-6 * scale).mean()),
                         abs(float(h_d.mean() - h_o.mean())) / scale))
            print(f"side-pass globe step {it}: max gap {gaps[-1][0]:.3g} x peak, cells beyond 1e-6 "
                  f"{100 * gaps[-1][1]:.3f} %, mean gap {gaps[-1][2]:.3g}", flush=True)
            if it == 0:
                assert np.array_equal(h_d, h_o)
    finally:
        oracle.set_libm()
    assert gaps[1][0] <= 1e-9
    assert gaps[2][1] <= 0.01 and gaps[2][2] <= 1e-6


def test_slope_limit_check_mode_bit_exact(oracle, cuda):
    """bake/tests/test_erosion.py:1622-1642's window with the limits on:
    the device's change lists and state equal the oracle's, and no cell
    loses more than its cap."""
    p = window_params(0, erodibility=1.0, k_disc=10.0, disc_exponent=0.0, max_erode=1e9, iter_erode=1e9,
                      slope_limit_erode=0.02, slope_limit_deposit=0.2)
    kw = window_inputs(48, "tilt", p, seed=11, relief=12.0)
    ost = oracle.OracleState.window(**kw)
    dst = maps.ErosionState.window(**kw, device=cuda)
    s0 = ost.surface().copy()
    cap = oracle.slope_erode_cap(ost, 0.02, float(p.erosion.slope_limit_exit))
    so = oracle.run_iteration(ost, p, 0)
    sd = maps.run_iteration(dst, p, 0)
    assert sd["clamped"] == so["clamped"] and sd["entries"] == so["entries"]
    for n in ("height", "sediment", "pending", "discharge"):
        assert np.array_equal(getattr(dst, n).cpu().numpy(), getattr(ost, n)), n
    loss = s0 - ost.surface()
    act = ost.mask == 1
    assert np.all(loss[act] <= cap[act] + 1e-9)


def test_slope_limit_fast_mode_bounds_erosion(cuda):
    """The same bound in the fast mode, whose per-cell erosion scale is a
    float32 factor applied to float32 requests: exact to their rounding
    (1e-6 relative), where the reference's float64 apply is exact to 1e-9."""
    p = window_params(0, erodibility=1.0, k_disc=10.0, disc_exponent=0.0, max_erode=1e9, iter_erode=1e9,
                      slope_limit_erode=0.02)
    kw = window_inputs(48, "tilt", p, seed=11, relief=12.0)
    dst = maps.ErosionState.window(**kw, device=cuda, mode="fast")
    s0 = (dst.height + dst.sediment).cpu().numpy()
    cap = torch.empty_like(dst.height)
    pk.slope_erode_cap(dst.height, dst.sediment, dst.mask, 0.02, float(p.erosion.slope_limit_exit), cap)
    maps.run_iteration(dst, p, 0)
    loss = s0 - (dst.height + dst.sediment).cpu().numpy()
    act = dst.mask.cpu().numpy() == 1
    c = cap.cpu().numpy()
    assert np.all(loss[act] <= c[act] * (1.0 + 1e-6) + 1e-9), float((loss[act] - c[act]).max())
    p2 = window_params(0, erodibility=1.0, k_disc=10.0, disc_exponent=0.0, max_erode=1e9, iter_erode=1e9)
    d2 = maps.ErosionState.window(**kw, device=cuda, mode="fast")
    maps.run_iteration(d2, p2, 0)
    loss2 = s0 - (d2.height + d2.sediment).cpu().numpy()
    assert (loss2[act] > c[act] + 1e-6).sum() > 10  # without it the same droplets take more


def test_fast_graph_with_side_passes_replays_eager(cuda):
    """The captured iteration (two graphs: with and without the isostatic
    rebound) reproduces eager fast-mode steps bit for bit."""
    g, f = _relief_globe(N=48, seed=3)
    p = _params(thermal_rate=0.5, creep_rate=0.05, thermal_max=60.0, isostasy=0.8, flexure_km=0.2, isostasy_every=2,
                slope_limit_erode=0.05)
    args = (g, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"], p.erosion)
    a = maps.ErosionState.from_grid(*args, device=cuda, mode="fast")
    b = maps.ErosionState.from_grid(*args, device=cuda, mode="fast")
    for it in range(4):
        maps.step(a, p, it)
    gr = fast.IterationGraph(b, p)
    for it in range(4):
        gr.step(it)
    torch.cuda.synchronize()
    assert b.iteration == a.iteration == 4
    for n in ("height", "sediment", "discharge", "pending", "iso_acc"):
        assert torch.equal(getattr(a, n), getattr(b, n)), n
