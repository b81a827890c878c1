"""The step's side functions on the device against the oracle, at the
N_c=1024 earth globe (the benchmark's grid):

* halo exchange (field.py:104-124; maps.py:329-340): the device gathers
  (scalar cubic weights, the momentum rotated into the destination face, the
  route's linear weights) on random fields, bit for bit against the oracle
  (both sum the sources in source order), and against the reference's own
  formula, numpy's einsum, whose summation order numpy's SIMD kernels
  choose: they agree to the forward error bound of any order of a 16-term
  (vector: 16 x 2-term) float64 dot product, |device - einsum| <= 2 gamma_n
  sum_s |w_s x_s|, gamma_n = n u / (1 - n u), u = 2^-53, n = 16 (scalar);
  for the vector the device's 4 roundings per term plus 16 terms and the
  einsum's 2 per term plus 32 terms give gamma_20 + gamma_33;
* hold_datum (maps.py:1446-1508): the device radix select returns exactly
  numpy's np.partition order statistic, for any land fraction, with ties,
  and shifts height, route and the negative base by it bitwise;
* start_replay and the replayed uplift (maps.py:1088-1213): heights bitwise
  equal to the oracle's (the lowered crust, its halo exchange, the datum
  hold, one capped mean-free uplift).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from globe_hip import stubs
from globe_hip.erosion import maps

pytestmark = pytest.mark.gpu

U = 2.0 ** -53


def _gamma(n):
    return n * U / (1.0 - n * U)


@pytest.fixture(scope="module")
def earth1024():
    return stubs.earth_globe(1024, 2)


def test_halo_exchange_matches_the_einsum(oracle, cuda, earth1024):
    grid, _ = earth1024
    hm = grid.halo
    tables = maps.HaloTables(grid, cuda)
    rng = np.random.default_rng(7)
    shape = (6, grid.NE, grid.NE)
    print(f"halo tables: {hm.src.shape[0]} halo cells x {hm.src.shape[1]} sources", flush=True)
    for linear in (False, True):
        x = rng.standard_normal(shape) * 10.0 ** rng.uniform(-3, 3, shape)
        ref = x.copy()
        oracle.exchange(grid, ref, linear=linear)
        d = torch.from_numpy(x).to(cuda)
        tables.exchange(d, linear=linear)
        dev = d.cpu().numpy()
        assert np.array_equal(dev, ref), linear      # halos and the untouched interior
        w = hm.w_lin if linear else hm.w
        flat = x.reshape(-1)
        ein = np.einsum("ks,ks->k", w, flat[hm.src])  # field.py:104-124 as the reference writes it
        bound = 2.0 * _gamma(16) * np.abs(w * flat[hm.src]).sum(1)
        err = np.abs(dev.reshape(-1)[hm.dst] - ein)
        assert np.all(err <= bound), (linear, float((err / np.maximum(bound, 1e-300)).max()))
        print(f"  {'linear' if linear else 'cubic'}: max |dev - einsum| / bound "
              f"{float((err / np.maximum(bound, 1e-300)).max()):.3g}", flush=True)
    v = rng.standard_normal(shape + (2,))
    ref = v.copy()
    oracle.exchange_vector(grid, ref)
    d = torch.from_numpy(v).to(cuda)
    tables.exchange(d, vector=True)
    dev = d.cpu().numpy()
    assert np.array_equal(dev, ref)
    fv = v.reshape(-1, 2)[hm.src]                                  # (K, S, 2)
    ein = np.einsum("ks,ksab,ksb->ka", hm.w, hm.rot, fv)
    mag = np.abs(hm.w)[..., None] * (np.abs(hm.rot) * np.abs(fv)[:, :, None, :]).sum(-1)
    bound = (_gamma(20) + _gamma(33)) * mag.sum(1)   # device: 4 roundings + 16 terms; einsum: 2 + 32 terms
    err = np.abs(dev.reshape(-1, 2)[hm.dst] - ein)
    assert np.all(err <= bound), float((err / np.maximum(bound, 1e-300)).max())
    print(f"  vector: max |dev - einsum| / bound {float((err / np.maximum(bound, 1e-300)).max()):.3g}", flush=True)


@pytest.mark.parametrize("lf", [0.0, 0.3, 0.5, 0.999, 1.0])
@pytest.mark.parametrize("ties", [False, True])
def test_hold_datum_is_the_exact_order_statistic(cuda, earth1024, lf, ties):
    grid, f = earth1024
    p = stubs.workload_params(1024, 2)
    st = maps.ErosionState.from_grid(grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"],
                                     p.erosion, device=cuda)
    if ties:   # a quantized surface: many equal values around every quantile
        st.height.copy_(torch.round(st.height * 64.0) / 64.0)
    st.route = st.height + 0.5
    st.base.zero_()
    st.base[0, 10:20, 10:20] = -2.0
    h0, r0, b0 = st.height.cpu().numpy(), st.route.cpu().numpy(), st.base.cpu().numpy()
    surf = (h0 + st.sediment.cpu().numpy())[st.interior].ravel()
    M = surf.size
    n_land = min(max(int(round(lf * M)), 0), M)
    if n_land == 0:
        want = float(surf.max()) + 1.0
    elif n_land == M:
        want = float(surf.min())
    else:
        want = float(np.partition(surf, M - n_land)[M - n_land])
    q = maps.hold_datum(st, lf)
    print(f"lf {lf} ties {ties}: q {q!r} (np.partition {want!r})", flush=True)
    assert q == want
    assert np.array_equal(st.height.cpu().numpy(), h0 - q)
    assert np.array_equal(st.route.cpu().numpy(), r0 - q)
    b1 = st.base.cpu().numpy()
    assert np.array_equal(b1[b0 < 0], b0[b0 < 0] - q) and np.all(b1[b0 >= 0] == 0.0)


@pytest.mark.parametrize("cap_m", [0.0, 2.0])
def test_start_replay_and_uplift_match_the_oracle(oracle, cuda, earth1024, cap_m):
    grid, f = earth1024
    u = f["uplift_m"] * 3000.0                      # ~3 m/it blobs: the cap binds on their cores
    p = stubs.workload_params(1024, 2, uplift_mode="replay", iterations=200, uplift_max_m=cap_m)
    args = (grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], u, p.erosion)
    ost = oracle.OracleState.from_grid(*args)
    dst = maps.ErosionState.from_grid(*args, device=cuda)
    cap_o, cap_d = oracle.uplift_cap(ost, p.erosion), maps.uplift_cap(dst, p.erosion)
    assert cap_o == cap_d and (cap_o is None) == (cap_m == 0.0)
    if cap_o is not None:
        assert int((ost.uplift[ost.interior] > cap_o).sum()) > 1000
    io = oracle.start_replay(ost, p)
    idv = maps.start_replay(dst, p)
    print(f"cap {cap_m}: replay max {io['replay_max']!r} min {io['replay_min']!r} shift {io['datum_shift']!r}",
          flush=True)
    for k in ("replay_max", "replay_min", "land_fraction", "datum_shift"):
        assert float(idv[k]) == float(io[k]), (k, idv[k], io[k])
    assert np.array_equal(dst.height.cpu().numpy(), ost.height)   # interior and halos
    # one replayed uplift
    oracle.apply_uplift(ost, cap_o)
    maps.apply_uplift(dst, cap_d)
    assert np.array_equal(dst.height.cpu().numpy(), ost.height)
