"""The fast mode's single-precision face transfer (``transfer_velocity_f``,
the float32 restatement of cubesphere.py:204-220 that the fast trace runs
when a droplet crosses a face edge) against the double-precision,
correctly rounded ``transfer_velocity`` (itself pinned bit for bit by the
reference's golden vectors): same destination face, position within float
precision, direction within float precision of the transferred vector.

The host build runs without a GPU; the ``gpu`` test evaluates the device
build of both over the same inputs (``pg_geom_eval`` ops 4 and 6).
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest


def _crossings(n=4000, seed=0):
    """Positions just across a face edge or corner (as a droplet lands after
    a one-cell step on a 64..8192-cell face) with random directions."""
    rng = np.random.default_rng(seed)
    face = rng.integers(0, 6, n)
    u = rng.uniform(0.0, 1.0, n)
    v = rng.uniform(0.0, 1.0, n)
    over = rng.uniform(1e-6, 0.02, n)
    side = rng.integers(0, 4, n)
    u = np.where(side == 0, -over, np.where(side == 1, 1.0 + over, u))
    v = np.where(side == 2, -over, np.where(side == 3, 1.0 + over, v))
    corner = rng.random(n) < 0.1
    over2 = rng.uniform(1e-6, 0.02, n)
    v = np.where(corner & (side < 2), np.where(rng.random(n) < 0.5, -over2, 1.0 + over2), v)
    a = rng.normal(size=n)
    b = rng.normal(size=n)
    # float32 inputs, so both versions see the same point
    return face, u.astype(np.float32), v.astype(np.float32), a.astype(np.float32), b.astype(np.float32)


def _check(fd, out_d, ff, out_f, a, b):
    same = fd == ff
    # a point within float rounding of a cube edge may land on either face
    assert same.mean() >= 0.999, same.mean()
    du = np.abs(out_f[same, :2] - out_d[same, :2])
    assert du.max() <= 2e-6, du.max()
    scale = np.hypot(a, b)[same]
    dab = np.abs(out_f[same, 2:] - out_d[same, 2:]).max(axis=1) / scale
    assert dab.max() <= 2e-5, dab.max()


def test_float_transfer_matches_double_on_the_host(hiplib):
    face, u, v, a, b = _crossings()
    n = face.size
    fd = np.empty(n, np.int32)
    ff = np.empty(n, np.int32)
    out_d = np.empty((n, 4))
    out_f = np.empty((n, 4))
    o4 = (ctypes.c_double * 4)()
    f4 = (ctypes.c_float * 4)()
    fo = ctypes.c_int()
    for k in range(n):
        assert hiplib.pg_geom_transfer_velocity(int(face[k]), float(u[k]), float(v[k]), float(a[k]), float(b[k]),
                                                ctypes.byref(fo), o4) == 0
        fd[k] = fo.value
        out_d[k] = o4[:]
        assert hiplib.pg_geom_transfer_velocity_f32(int(face[k]), float(u[k]), float(v[k]), float(a[k]), float(b[k]),
                                                    ctypes.byref(fo), f4) == 0
        ff[k] = fo.value
        out_f[k] = f4[:]
    assert (fd != face).mean() > 0.5  # most crossings change face
    _check(fd, out_d, ff, out_f, a, b)
    # a position on the face is returned as it is
    assert hiplib.pg_geom_transfer_velocity_f32(3, 0.25, 0.75, 0.5, -0.5, ctypes.byref(fo), f4) == 0
    assert fo.value == 3 and list(f4) == [0.25, 0.75, 0.5, -0.5]


@pytest.mark.gpu
def test_float_transfer_matches_double_on_the_device(cuda):
    from test_gpu_geometry import _eval

    face, u, v, a, b = _crossings(20000, seed=1)
    inp = np.stack([face.astype(np.float64), u, v, a, b], 1)
    out_d, fd = _eval(4, inp, 4, True, cuda)
    out_f, ff = _eval(6, inp, 4, True, cuda)
    _check(fd, out_d, ff, out_f, a, b)
