"""The host-side geometry tables the kernels consume (metric, halo gather
tables, owner), pinned by the closed forms of bake/tests/test_cubesphere.py:
halo accuracy on analytic fields, monotone bilinear weights, exact vector
rotation, owner = nearest interior cell, metric = J^T J of the mapping."""
from __future__ import annotations

import numpy as np
import pytest

from globe_hip.cubesphere import Grid, jacobian_v, to_sphere_v


def _exchange(g, data, linear=False, vector=False):
    """The reference's einsum gather (bake/globe/field.py:104-124) on the
    product's tables (a CPU check of the tables, not a product path)."""
    hm = g.halo
    flat = data.reshape(6 * g.NE * g.NE, -1) if data.ndim == 4 else data.reshape(-1)
    w = hm.w_lin if linear else hm.w
    if vector:
        flat[hm.dst] = np.einsum("ks,ksab,ksb->ka", w, hm.rot, flat[hm.src])
    else:
        flat[hm.dst] = np.einsum("ks,ks->k", w, flat[hm.src])
    return data


def _analytic_error(N, fn, linear=False):
    g = Grid(N, 4, 50.0)
    ref = np.asarray(fn(g.centers), np.float64)
    d = ref.copy()
    d[:, g.halo_mask] = np.nan
    _exchange(g, d, linear)
    assert not np.isnan(d).any()
    err = np.abs(d - ref)
    return float(err[:, g.edge_halo_mask].max()), float(err[:, g.corner_mask].max())


def test_halo_exchange_analytic():
    fn = lambda p: p[..., 0] * p[..., 1] + p[..., 2] ** 2
    e64, c64 = _analytic_error(64, fn)
    e128, c128 = _analytic_error(128, fn)
    assert e64 < 3e-4 and c64 < 1e-3
    assert e128 < e64 / 3 and c128 < c64 / 3


def test_linear_weights_are_convex():
    for N in (32, 64):
        hm = Grid(N, 4, 50.0).halo
        assert (hm.w_lin >= 0).all() and np.abs(hm.w_lin.sum(1) - 1).max() < 1e-5
        assert np.abs(hm.w.sum(1) - 1).max() < 1e-9


def test_halo_exchange_linear_field():
    e, c = _analytic_error(32, lambda p: 0.3 * p[..., 0] - 0.7 * p[..., 1] + 0.2 * p[..., 2])
    assert max(e, c) < 2e-3


def _rotation_field(g, omega):
    w3 = np.cross(np.asarray(omega, float), g.centers)
    U, V = g._uv_grid
    out = np.empty((6, g.NE, g.NE, 2))
    for f in range(6):
        ju, jv = jacobian_v(np.full(U.shape, f), U, V)
        guu, guv, gvv = (ju * ju).sum(-1), (ju * jv).sum(-1), (jv * jv).sum(-1)
        ru, rv = (ju * w3[f]).sum(-1), (jv * w3[f]).sum(-1)
        det = guu * gvv - guv * guv
        out[f, ..., 0] = (gvv * ru - guv * rv) / det * g.N
        out[f, ..., 1] = (guu * rv - guv * ru) / det * g.N
    return out


@pytest.mark.parametrize("omega", [(0, 0, 1), (1, 0, 0), (0.3, -0.5, 0.8)])
def test_vector_halo_rotation(omega):
    g = Grid(32, 4, 50.0)
    ref = _rotation_field(g, omega)
    d = ref.copy()
    d[:, g.halo_mask] = np.nan
    _exchange(g, d, vector=True)
    err = np.abs(d - ref).max(-1) / np.abs(ref).max()
    assert err[:, g.edge_halo_mask].max() < 1e-5 and err[:, g.corner_mask].max() < 1e-4


def test_owner_is_the_nearest_interior_cell():
    g = Grid(32, 4, 50.0)
    H, N, NE = g.H, g.N, g.NE
    own = g.owner
    ident = np.arange(6 * NE * NE).reshape(6, NE, NE)
    assert np.array_equal(own[g.interior], ident[g.interior])
    of, oi, oj = g.unflat_index(own)
    assert (oi[:, g.halo_mask] >= H).all() and (oi[:, g.halo_mask] < H + N).all()
    assert (of[:, g.halo_mask] != np.arange(6)[:, None]).all()
    p = g.centers[:, g.halo_mask]
    q = g.centers.reshape(-1, 3)[own[:, g.halo_mask]]
    d = np.arccos(np.clip(np.sum(p * q, -1), -1, 1)) / (np.pi / 2 / N)
    assert d.max() < 0.75


def test_metric_is_the_mapping_jacobian():
    g = Grid(16, 2, 1000.0)
    U, V = g._uv_grid
    h = 1e-7
    f = 3
    pu = (to_sphere_v(np.full(U.shape, f), U + h, V) - to_sphere_v(np.full(U.shape, f), U - h, V)) / (2 * h)
    pv = (to_sphere_v(np.full(U.shape, f), U, V + h) - to_sphere_v(np.full(U.shape, f), U, V - h)) / (2 * h)
    k = (g.R_planet / (g.N * g.cell_size_m)) ** 2
    gm = np.stack([(pu * pu).sum(-1), (pu * pv).sum(-1), (pv * pv).sum(-1)], -1) * k
    assert np.abs(gm - g.metric[f]).max() < 1e-5
    # identity at the face centre, and metric_inv its inverse
    c = g.N // 2 + g.H
    assert np.abs(g.metric[0, c, c] - [1.0, 0.0, 1.0]).max() < 0.05
    m, mi = g.metric.astype(np.float64), g.metric_inv.astype(np.float64)
    prod00 = m[..., 0] * mi[..., 0] + m[..., 1] * mi[..., 1]
    assert np.abs(prod00 - 1).max() < 1e-5
