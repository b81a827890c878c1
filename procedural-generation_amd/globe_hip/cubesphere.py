"""Equi-angular cube-sphere geometry tables (host side, built once per grid).

Conventions are the reference's (bake/globe/cubesphere.py:13-23, 37-48):
faces 0..5 = +X, -X, +Y, -Y, +Z, -Z with bases (right, up, normal);
``s = tan((u - 0.5) pi/2)``, ``p = normalize(normal + s right + t up)``;
the extended array of a face is ``NE = N + 2H`` cells with the interior at
``[H, H + N)``.  What the device kernels consume from here:

* ``metric`` / ``metric_inv`` (6, NE, NE, 3) float32 -- the dimensionless
  metric of cell displacements (bake/globe/cubesphere.py:488-502);
* ``halo`` -- gather tables (16 interior sources per halo cell, cubic
  Lagrange on edge blocks, local quadratic least squares on corner blocks,
  plus per-source 2x2 rotations for vector fields;
  bake/globe/cubesphere.py:540-662);
* ``owner`` -- the interior cell owning each extended cell
  (bake/globe/cubesphere.py:664-674).

Everything is numpy: these tables are inputs of the particle pass, computed
once per resolution, never per iteration.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from functools import cached_property

import numpy as np

#: (right, up, normal) per face, bake/globe/cubesphere.py:37-48
BASES = np.array(
    [
        [[0.0, 0.0, -1.0], [0.0, -1.0, 0.0], [1.0, 0.0, 0.0]],
        [[0.0, 0.0, 1.0], [0.0, -1.0, 0.0], [-1.0, 0.0, 0.0]],
        [[1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, 1.0, 0.0]],
        [[1.0, 0.0, 0.0], [0.0, 0.0, -1.0], [0.0, -1.0, 0.0]],
        [[1.0, 0.0, 0.0], [0.0, -1.0, 0.0], [0.0, 0.0, 1.0]],
        [[-1.0, 0.0, 0.0], [0.0, -1.0, 0.0], [0.0, 0.0, -1.0]],
    ],
    dtype=np.float64,
)
HALF_PI = math.pi / 2.0
TWO_OVER_PI = 2.0 / math.pi


def planet_radius(N: int, cell_size_m: float) -> float:
    """Radius at which a face edge (a quarter great circle) is N cells."""
    return N * cell_size_m * 4.0 / (2.0 * math.pi)


def to_sphere_v(face, u, v) -> np.ndarray:
    """(face, u, v) arrays -> (..., 3) unit vectors."""
    face = np.asarray(face, dtype=np.int64)
    s = np.tan((np.asarray(u, dtype=np.float64) - 0.5) * HALF_PI)
    t = np.tan((np.asarray(v, dtype=np.float64) - 0.5) * HALF_PI)
    b = BASES[face]
    q = b[..., 2, :] + s[..., None] * b[..., 0, :] + t[..., None] * b[..., 1, :]
    return q / np.linalg.norm(q, axis=-1, keepdims=True)


def face_of_v(p) -> np.ndarray:
    """Face of each unit vector: largest |component| with its sign, ties
    resolved X > Y > Z."""
    p = np.asarray(p, dtype=np.float64)
    a = np.abs(p)
    xf = (a[..., 0] >= a[..., 1]) & (a[..., 0] >= a[..., 2])
    yf = ~xf & (a[..., 1] >= a[..., 2])
    axis = np.where(xf, 0, np.where(yf, 1, 2))
    comp = np.take_along_axis(p, axis[..., None], axis=-1)[..., 0]
    return (2 * axis + (comp < 0.0)).astype(np.int64)


def project_to_face_v(face, p):
    """(u, v) of unit vectors on a given face's parametrisation."""
    p = np.asarray(p, dtype=np.float64)
    b = BASES[np.asarray(face, dtype=np.int64)]
    d = np.sum(p * b[..., 2, :], axis=-1)
    s = np.sum(p * b[..., 0, :], axis=-1) / d
    t = np.sum(p * b[..., 1, :], axis=-1) / d
    return np.arctan(s) * TWO_OVER_PI + 0.5, np.arctan(t) * TWO_OVER_PI + 0.5


def from_sphere_v(p):
    """Unit vectors -> (face, u, v), u and v in the closed interval [0, 1]."""
    face = face_of_v(p)
    u, v = project_to_face_v(face, p)
    return face, u, v


def jacobian_v(face, u, v):
    """dp/du and dp/dv (each (..., 3)) on the unit sphere."""
    face = np.asarray(face, dtype=np.int64)
    s = np.tan((np.asarray(u, dtype=np.float64) - 0.5) * HALF_PI)
    t = np.tan((np.asarray(v, dtype=np.float64) - 0.5) * HALF_PI)
    b = BASES[face]
    right, up, normal = b[..., 0, :], b[..., 1, :], b[..., 2, :]
    q = normal + s[..., None] * right + t[..., None] * up
    ql = np.linalg.norm(q, axis=-1, keepdims=True)
    p = q / ql
    dq_du = right * (HALF_PI * (1.0 + s * s))[..., None]
    dq_dv = up * (HALF_PI * (1.0 + t * t))[..., None]
    ju = (dq_du - p * np.sum(p * dq_du, axis=-1, keepdims=True)) / ql
    jv = (dq_dv - p * np.sum(p * dq_dv, axis=-1, keepdims=True)) / ql
    return ju, jv


def transfer_vector_v(face_src, u_src, v_src, ab, face_dst, u_dst, v_dst) -> np.ndarray:
    """Tangent vector components (du, dv) on the source face -> components on
    the destination face at the same point (2x2 normal equations)."""
    ju, jv = jacobian_v(face_src, u_src, v_src)
    ab = np.asarray(ab, dtype=np.float64)
    w = ab[..., :1] * ju + ab[..., 1:] * jv
    ku, kv = jacobian_v(face_dst, u_dst, v_dst)
    guu, guv, gvv = np.sum(ku * ku, -1), np.sum(ku * kv, -1), np.sum(kv * kv, -1)
    ru, rv = np.sum(ku * w, -1), np.sum(kv * w, -1)
    det = guu * gvv - guv * guv
    return np.stack([(gvv * ru - guv * rv) / det, (guu * rv - guv * ru) / det], axis=-1)


def lagrange4_weights(t) -> np.ndarray:
    """Cubic Lagrange weights of nodes -1, 0, 1, 2 at offset ``t`` from node 0."""
    t = np.asarray(t, dtype=np.float64)
    tm1, tm2, tp1 = t - 1.0, t - 2.0, t + 1.0
    return np.stack([-t * tm1 * tm2 / 6.0, tp1 * tm1 * tm2 / 2.0, -tp1 * t * tm2 / 2.0, tp1 * t * tm1 / 6.0], axis=-1)


def linear_weights_on_4(t) -> np.ndarray:
    """Two-node (linear) weights placed on the 4-node stencil -1..2."""
    t = np.asarray(t, dtype=np.float64)
    lo = np.clip(np.floor(t).astype(np.int64), -1, 1)
    frac = np.clip(t - lo, 0.0, 1.0)
    w = np.zeros(t.shape + (4,), dtype=np.float64)
    np.put_along_axis(w, (lo + 1)[..., None], (1.0 - frac)[..., None], axis=-1)
    np.put_along_axis(w, (lo + 2)[..., None], frac[..., None], axis=-1)
    return w


@dataclass(frozen=True)
class HaloMap:
    """Gather tables for halo cells (K halo cells, S = 16 sources each):
    ``dst`` (K,), ``src`` (K, S) flat indices into (6, NE, NE); ``w`` / ``w_lin``
    (K, S) weights (rows sum to 1); ``rot`` (K, S, 2, 2) source-face ->
    destination-face component matrices; ``nearest`` (K,)."""

    dst: np.ndarray
    src: np.ndarray
    w: np.ndarray
    w_lin: np.ndarray
    rot: np.ndarray
    nearest: np.ndarray


class Grid:
    """Geometry tables of one cube-sphere resolution (N cells per face edge,
    halo H, nominal cell size in metres)."""

    HALO_SOURCES = 16

    def __init__(self, N: int, H: int = 4, cell_size_m: float = 50.0, R_planet: float | None = None):
        if N < 2 * H:
            raise ValueError("N must be at least 2H")
        self.N, self.H = int(N), int(H)
        self.NE = self.N + 2 * self.H
        self.cell_size_m = float(cell_size_m)
        self.R_planet = float(R_planet) if R_planet is not None else planet_radius(N, cell_size_m)

    # -- indexing ----------------------------------------------------------
    def cell_uv(self, ei, ej):
        return (np.asarray(ei) - self.H + 0.5) / self.N, (np.asarray(ej) - self.H + 0.5) / self.N

    def uv_cell(self, u, v):
        return np.asarray(u) * self.N - 0.5 + self.H, np.asarray(v) * self.N - 0.5 + self.H

    def flat_index(self, face, ei, ej):
        return (np.asarray(face) * self.NE + np.asarray(ei)) * self.NE + np.asarray(ej)

    def unflat_index(self, flat):
        flat = np.asarray(flat)
        f, rem = np.divmod(flat, self.NE * self.NE)
        i, j = np.divmod(rem, self.NE)
        return f, i, j

    @property
    def interior(self):
        return (slice(None), slice(self.H, self.H + self.N), slice(self.H, self.H + self.N))

    @cached_property
    def _uv_grid(self):
        e = (np.arange(self.NE) - self.H + 0.5) / self.N
        return np.meshgrid(e, e, indexing="ij")

    # -- per-cell geometry -------------------------------------------------
    @cached_property
    def centers(self) -> np.ndarray:
        U, V = self._uv_grid
        return np.stack([to_sphere_v(np.full(U.shape, f), U, V) for f in range(6)])

    @cached_property
    def interior_centers(self) -> np.ndarray:
        return self.centers[self.interior]

    @cached_property
    def metric_uv(self) -> np.ndarray:
        """(g_uu, g_uv, g_vv) of the (u, v) parametrisation, float64."""
        U, V = self._uv_grid
        g = np.empty((6, self.NE, self.NE, 3), dtype=np.float64)
        for f in range(6):
            ju, jv = jacobian_v(np.full(U.shape, f), U, V)
            g[f, ..., 0] = np.sum(ju * ju, -1)
            g[f, ..., 1] = np.sum(ju * jv, -1)
            g[f, ..., 2] = np.sum(jv * jv, -1)
        return g

    @cached_property
    def metric(self) -> np.ndarray:
        """Dimensionless metric of cell displacements, float32 (physical
        length of (di, dj) is cell_size_m * sqrt(g_ii di^2 + 2 g_ij di dj + g_jj dj^2))."""
        scale = (self.R_planet / (self.N * self.cell_size_m)) ** 2
        return (self.metric_uv * scale).astype(np.float32)

    @cached_property
    def metric_inv(self) -> np.ndarray:
        g = self.metric.astype(np.float64)
        det = g[..., 0] * g[..., 2] - g[..., 1] ** 2
        return np.stack([g[..., 2] / det, -g[..., 1] / det, g[..., 0] / det], axis=-1).astype(np.float32)

    @cached_property
    def sqrt_det_metric(self) -> np.ndarray:
        """sqrt(det g) of the cell metric, float32 (cubesphere.py:504-507):
        the area of a cell in units of cell_size_m^2."""
        g = self.metric.astype(np.float64)
        return np.sqrt(g[..., 0] * g[..., 2] - g[..., 1] ** 2).astype(np.float32)

    def latitude(self) -> np.ndarray:
        """Latitude (radians) of every extended cell centre, +Z the pole."""
        return np.arcsin(np.clip(self.centers[..., 2], -1.0, 1.0))

    # -- halo tables -------------------------------------------------------
    @cached_property
    def halo_mask(self) -> np.ndarray:
        m = np.ones((self.NE, self.NE), dtype=bool)
        m[self.H:self.H + self.N, self.H:self.H + self.N] = False
        return m

    @cached_property
    def corner_mask(self) -> np.ndarray:
        H, NE = self.H, self.NE
        m = np.zeros((NE, NE), dtype=bool)
        for a in (slice(0, H), slice(NE - H, NE)):
            for b in (slice(0, H), slice(NE - H, NE)):
                m[a, b] = True
        return m

    @cached_property
    def edge_halo_mask(self) -> np.ndarray:
        return self.halo_mask & ~self.corner_mask

    def _edge_tables(self):
        N, H, S = self.N, self.H, self.HALO_SOURCES
        ei, ej = np.nonzero(self.edge_halo_mask)
        u, v = self.cell_uv(ei, ej)
        nodes = np.arange(-1, 3)
        parts = []
        for f in range(6):
            faces = np.full(ei.size, f)
            sf, su, sv = from_sphere_v(to_sphere_v(faces, u, v))
            fi, fj = self.uv_cell(su, sv)
            # the 4x4 stencil stays inside the source face's interior
            i0 = np.clip(np.floor(fi).astype(np.int64), H + 1, H + N - 3)
            j0 = np.clip(np.floor(fj).astype(np.int64), H + 1, H + N - 3)
            ti, tj = fi - i0, fj - j0
            ii = i0[:, None] + nodes
            jj = j0[:, None] + nodes
            src = self.flat_index(sf[:, None, None], ii[:, :, None], jj[:, None, :]).reshape(-1, S)
            w = (lagrange4_weights(ti)[:, :, None] * lagrange4_weights(tj)[:, None, :]).reshape(-1, S)
            wl = (linear_weights_on_4(ti)[:, :, None] * linear_weights_on_4(tj)[:, None, :]).reshape(-1, S)
            near = src[np.arange(src.shape[0]), np.argmax(wl, axis=1)]
            e1 = transfer_vector_v(sf, su, sv, np.tile([1.0, 0.0], (ei.size, 1)), faces, u, v)
            e2 = transfer_vector_v(sf, su, sv, np.tile([0.0, 1.0], (ei.size, 1)), faces, u, v)
            R = np.stack([e1, e2], axis=-1)  # columns: images of the source basis vectors
            rot = np.broadcast_to(R[:, None], (ei.size, S, 2, 2)).copy()
            parts.append((self.flat_index(f, ei, ej), src, w, wl, rot, near))
        return parts

    def _corner_tables(self):
        from scipy.spatial import cKDTree

        N, H, S = self.N, self.H, self.HALO_SOURCES
        ci, cj = np.nonzero(self.corner_mask)
        cu, cv = self.cell_uv(ci, cj)
        # candidate sources: interior cells within H + 3 cells of a face corner
        M = H + 3
        blocks = []
        for f in range(6):
            for a0 in (H, H + N - M):
                for b0 in (H, H + N - M):
                    ii, jj = np.meshgrid(np.arange(a0, a0 + M), np.arange(b0, b0 + M), indexing="ij")
                    blocks.append(np.stack([np.full(ii.size, f), ii.ravel(), jj.ravel()]))
        cf, cii, cjj = np.concatenate(blocks, axis=1)
        cflat = self.flat_index(cf, cii, cjj)
        cpos = self.centers[cf, cii, cjj]
        cuu, cvv = self.cell_uv(cii, cjj)
        tree = cKDTree(cpos)
        parts = []
        for f in range(6):
            faces = np.full(ci.size, f)
            p = to_sphere_v(faces, cu, cv)
            _, idx = tree.query(p, k=S)
            q = cpos[idx]
            # inverse-distance-squared over the 4 nearest (monotone variant)
            d2 = np.sum((q[:, :4] - p[:, None, :]) ** 2, axis=2)
            wl = np.zeros((ci.size, S))
            wl[:, :4] = (1.0 / d2) / np.sum(1.0 / d2, axis=1, keepdims=True)
            # quadratic least squares in gnomonic tangent-plane coordinates at p
            ref = np.where(np.abs(p[:, :1]) < 0.9, [[1.0, 0.0, 0.0]], [[0.0, 1.0, 0.0]])
            t1 = ref - p * np.sum(ref * p, axis=1, keepdims=True)
            t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
            t2 = np.cross(p, t1)
            g = q / np.sum(q * p[:, None, :], axis=2, keepdims=True) - p[:, None, :]
            x = np.sum(g * t1[:, None, :], axis=2)
            y = np.sum(g * t2[:, None, :], axis=2)
            A = np.stack([np.ones_like(x), x, y, x * x, x * y, y * y], axis=2)
            w = np.linalg.pinv(A)[:, 0, :]
            ff = np.repeat(faces[:, None], S, axis=1)
            uu = np.repeat(cu[:, None], S, axis=1)
            vv = np.repeat(cv[:, None], S, axis=1)
            sfc, suc, svc = cf[idx], cuu[idx], cvv[idx]
            r1 = transfer_vector_v(sfc, suc, svc, np.broadcast_to([1.0, 0.0], (ci.size, S, 2)), ff, uu, vv)
            r2 = transfer_vector_v(sfc, suc, svc, np.broadcast_to([0.0, 1.0], (ci.size, S, 2)), ff, uu, vv)
            rot = np.stack([r1, r2], axis=-1)
            parts.append((self.flat_index(f, ci, cj), cflat[idx], w, wl, rot, cflat[idx[:, 0]]))
        return parts

    @cached_property
    def halo(self) -> HaloMap:
        parts = self._edge_tables() + self._corner_tables()
        cat = [np.concatenate([p[k] for p in parts]) for k in range(6)]
        return HaloMap(dst=cat[0].astype(np.int64), src=cat[1].astype(np.int64), w=cat[2], w_lin=cat[3],
                       rot=cat[4], nearest=cat[5].astype(np.int64))

    @cached_property
    def owner(self) -> np.ndarray:
        """(6, NE, NE) int64 flat index of the owning interior cell."""
        own = np.arange(6 * self.NE * self.NE, dtype=np.int64)
        hm = self.halo
        own[hm.dst] = hm.nearest
        return own.reshape(6, self.NE, self.NE)

    def describe(self) -> str:
        return f"Grid(N={self.N}, H={self.H}, cell={self.cell_size_m} m, R={self.R_planet / 1000:.2f} km)"


_CACHE: dict = {}


def get_grid(N: int, H: int = 4, cell_size_m: float = 50.0, R_planet: float | None = None) -> Grid:
    key = (int(N), int(H), float(cell_size_m), None if R_planet is None else float(R_planet))
    if key not in _CACHE:
        _CACHE[key] = Grid(N, H, cell_size_m, R_planet)
    return _CACHE[key]
