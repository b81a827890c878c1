"""The device geometry (double-double tan/atan, to_sphere, from_sphere,
transfer_velocity) on the GPU: against the reference's golden vectors,
against the same code compiled for the host, and against mpmath.

glibc's tan/atan -- which wrote the golden file -- are not correctly rounded
on every input; the device's are.  So the golden records match bit for bit
except a handful (measured 8 of 5000 for to_sphere, 4 of 5000 for
from_sphere), each off by at most one ulp per coordinate, and at every one
of them the device value is the correctly rounded one.
"""
from __future__ import annotations

import ctypes
import math

import mpmath as mp
import numpy as np
import pytest
import torch

from globe_hip import _lib

pytestmark = pytest.mark.gpu


def _eval(op, inp: np.ndarray, nout: int, with_int: bool, dev):
    x = torch.from_numpy(np.ascontiguousarray(inp, np.float64)).to(dev)
    n = inp.shape[0]
    out = torch.empty((n, nout), dtype=torch.float64, device=dev)
    iout = torch.empty(n, dtype=torch.int32, device=dev) if with_int else None
    _lib.check(_lib.lib().pg_geom_eval(op, n, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                       ctypes.c_void_p(iout.data_ptr() if with_int else 0),
                                       ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "pg_geom_eval")
    torch.cuda.synchronize()
    return out.cpu().numpy(), (iout.cpu().numpy() if with_int else None)


def _ulps(a, b):
    return np.abs(a - b) / np.spacing(np.abs(b))


def test_device_geometry_vs_golden_vectors(vectors, cuda):
    rec = vectors
    inp = np.stack([rec["face"].astype(np.float64), rec["u"], rec["v"]], 1)
    xyz, _ = _eval(2, inp, 3, False, cuda)
    ref = np.stack([rec["x"], rec["y"], rec["z"]], 1)
    rows = np.any(xyz != ref, axis=1)
    assert rows.sum() <= 10, int(rows.sum())
    assert _ulps(xyz, ref).max() <= 2
    uv, face = _eval(3, ref, 2, True, cuda)
    np.testing.assert_array_equal(face, rec["face2"])
    refuv = np.stack([rec["u2"], rec["v2"]], 1)
    rows2 = np.any(uv != refuv, axis=1)
    assert rows2.sum() <= 10, int(rows2.sum())
    # u = atan(s) 2/pi + 0.5: one ulp of atan is an absolute 2^-53 or so
    assert np.abs(uv - refuv).max() <= 2.0 ** -52
    # the same code compiled for the host gives the same bits
    L = _lib.lib()
    o3 = (ctypes.c_double * 3)()
    for k in np.flatnonzero(rows)[:10].tolist() + list(range(0, 5000, 97)):
        L.pg_geom_to_sphere(int(rec["face"][k]), float(rec["u"][k]), float(rec["v"][k]), o3)
        assert tuple(o3) == tuple(xyz[k])


def test_device_tan_atan_correctly_rounded(cuda):
    mp.mp.prec = 200
    rng = np.random.default_rng(7)
    xs = (rng.random(4000) - 0.5) * 1.7
    t, _ = _eval(0, xs, 1, False, cuda)
    ys = np.concatenate([(rng.random(2000) - 0.5) * 4, rng.standard_cauchy(2000) * 30])
    a, _ = _eval(1, ys, 1, False, cuda)
    sc, _ = _eval(5, rng.random(2000) * 2 * math.pi, 2, False, cuda)
    for x, v in zip(xs, t[:, 0]):
        assert v == float(mp.tan(mp.mpf(float(x)))), x
    for y, v in zip(ys, a[:, 0]):
        assert v == float(mp.atan(mp.mpf(float(y)))), y
    # and glibc agrees with them nearly everywhere
    diff = np.mean([math.tan(float(x)) != v for x, v in zip(xs, t[:, 0])])
    assert diff < 0.01


def test_device_transfer_preserves_velocity(cuda):
    """cubesphere.py:204-220 / test_cubesphere.py:130-159: the 3-D velocity
    of a droplet crossing a face edge is preserved."""
    rng = np.random.default_rng(3)
    n = 2000
    face = rng.integers(0, 6, n).astype(np.float64)
    side = rng.integers(0, 4, n)
    u = rng.random(n)
    v = rng.random(n)
    eps = rng.random(n) * 0.004 + 1e-6
    u = np.where(side == 0, -eps, np.where(side == 1, 1 + eps, u))
    v = np.where(side == 2, -eps, np.where(side == 3, 1 + eps, v))
    a = (rng.random(n) - 0.5) * 1e-3
    b = (rng.random(n) - 0.5) * 1e-3
    out, f2 = _eval(4, np.stack([face, u, v, a, b], 1), 4, True, cuda)
    assert np.all(f2 != face.astype(np.int32))
    from globe_hip.cubesphere import jacobian_v, to_sphere_v

    ju, jv = jacobian_v(face.astype(np.int64), u, v)
    w1 = a[:, None] * ju + b[:, None] * jv
    ku, kv = jacobian_v(f2.astype(np.int64), out[:, 0], out[:, 1])
    w2 = out[:, 2:3] * ku + out[:, 3:4] * kv
    assert np.abs(w1 - w2).max() <= 1e-9 * np.abs(w1).max()
    p1 = to_sphere_v(face.astype(np.int64), u, v)
    p2 = to_sphere_v(f2.astype(np.int64), out[:, 0], out[:, 1])
    assert np.abs(p1 - p2).max() < 1e-12
