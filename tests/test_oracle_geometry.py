"""The oracle's cube-sphere geometry and the library's host-side geometry,
pinned without a GPU.

* oracle to_sphere / from_sphere: bit for bit the reference's golden vector
  file (bake/tests/test_cubesphere.py:379-402);
* Jacobian against central finite differences (test_cubesphere.py:115-127);
* transfer_velocity preserves the 3-D velocity and the point
  (test_cubesphere.py:130-159);
* the library's double-double tan / atan / sin / cos are correctly rounded
  (checked against mpmath), and differ from glibc only where glibc is not.
"""
from __future__ import annotations

import ctypes
import math

import mpmath as mp
import numpy as np
import pytest


def test_oracle_reproduces_golden_vectors(oracle, vectors):
    for r in vectors:
        assert oracle.to_sphere(int(r["face"]), float(r["u"]), float(r["v"])) == (r["x"], r["y"], r["z"])
        assert oracle.from_sphere(float(r["x"]), float(r["y"]), float(r["z"])) == (r["face2"], r["u2"], r["v2"])


def test_oracle_jacobian_matches_finite_differences(oracle):
    rng = np.random.default_rng(4)
    for _ in range(200):
        f = int(rng.integers(0, 6))
        u, v = rng.random(2) * 1.2 - 0.1
        h = 1e-6
        J = np.array(oracle.jacobian(f, u, v))
        du = (np.array(oracle.to_sphere(f, u + h, v)) - np.array(oracle.to_sphere(f, u - h, v))) / (2 * h)
        dv = (np.array(oracle.to_sphere(f, u, v + h)) - np.array(oracle.to_sphere(f, u, v - h))) / (2 * h)
        assert np.abs(J[:3] - du).max() < 1e-6 and np.abs(J[3:] - dv).max() < 1e-6


def test_oracle_transfer_preserves_velocity(oracle):
    rng = np.random.default_rng(5)
    for _ in range(300):
        f = int(rng.integers(0, 6))
        side = int(rng.integers(0, 4))
        u, v = rng.random(2)
        e = float(rng.random() * 0.003 + 1e-7)
        u = -e if side == 0 else (1 + e if side == 1 else u)
        v = -e if side == 2 else (1 + e if side == 3 else v)
        a, b = (rng.random(2) - 0.5) * 1e-3
        f2, u2, v2, a2, b2 = oracle.transfer_velocity(f, u, v, a, b)
        assert f2 != f and 0.0 <= u2 <= 1.0 and 0.0 <= v2 <= 1.0
        J, K = np.array(oracle.jacobian(f, u, v)), np.array(oracle.jacobian(f2, u2, v2))
        w1 = a * J[:3] + b * J[3:]
        w2 = a2 * K[:3] + b2 * K[3:]
        assert np.abs(w1 - w2).max() <= 1e-9 * np.abs(w1).max()
        assert np.abs(np.array(oracle.to_sphere(f, u, v)) - np.array(oracle.to_sphere(f2, u2, v2))).max() < 1e-12


def test_oracle_inside_the_face_is_unchanged(oracle):
    assert oracle.transfer_velocity(2, 0.3, 0.7, 0.01, -0.02) == (2, 0.3, 0.7, 0.01, -0.02)


def test_oracle_exact_edges_and_corners(oracle):
    """Points on cube edges / corners map to the higher-priority face with u
    or v exactly 1.0 (cubesphere.py:96-112, test_cubesphere.py:65-102)."""
    n_one = 0
    for x in (-1, 1):
        for y in (-1, 1):
            for z in (-1, 1):
                p = np.array([x, y, z]) / np.sqrt(3)
                f, u, v = oracle.from_sphere(*p)
                assert 0.0 <= u <= 1.0 and 0.0 <= v <= 1.0
                assert np.abs(np.array(oracle.to_sphere(f, u, v)) - p).max() < 1e-12
                n_one += u == 1.0 or v == 1.0
    assert n_one > 0


# -- the library's own host-side geometry (the code the kernels run) -------
def test_library_geometry_vs_golden_vectors(hiplib, vectors):
    """Bit for bit except where glibc (which wrote the file) misrounds
    tan/atan; there the library's value is the correctly rounded one."""
    mp.mp.prec = 200
    o3 = (ctypes.c_double * 3)()
    o2 = (ctypes.c_double * 2)()
    fo = ctypes.c_int()
    bad_to = bad_from = 0
    for r in vectors:
        hiplib.pg_geom_to_sphere(int(r["face"]), float(r["u"]), float(r["v"]), o3)
        if tuple(o3) != (r["x"], r["y"], r["z"]):
            bad_to += 1
            assert max(abs(a - b) / math.ulp(b) for a, b in zip(o3, (r["x"], r["y"], r["z"]))) <= 2
            # the tan arguments: the library's tan is correctly rounded there
            for w in (r["u"], r["v"]):
                t = (float(w) - 0.5) * (math.pi / 2.0)
                assert hiplib.pg_geom_tan(t) == float(mp.tan(mp.mpf(t)))
        hiplib.pg_geom_from_sphere(float(r["x"]), float(r["y"]), float(r["z"]), ctypes.byref(fo), o2)
        assert fo.value == r["face2"]
        if (o2[0], o2[1]) != (r["u2"], r["v2"]):
            bad_from += 1
            # u = atan(s) 2/pi + 0.5: one ulp of atan is an absolute 2^-53 or so
            assert max(abs(a - b) for a, b in zip(o2, (r["u2"], r["v2"]))) <= 2.0 ** -52
    assert bad_to <= 10 and bad_from <= 10, (bad_to, bad_from)


@pytest.mark.parametrize("name", ["tan", "atan", "sincos"])
def test_library_math_is_correctly_rounded(hiplib, name):
    mp.mp.prec = 200
    rng = np.random.default_rng(11)
    if name == "tan":
        xs = (rng.random(3000) - 0.5) * 1.7
        for x in xs:
            assert hiplib.pg_geom_tan(float(x)) == float(mp.tan(mp.mpf(float(x))))
    elif name == "atan":
        xs = np.concatenate([(rng.random(1500) - 0.5) * 4, rng.standard_cauchy(1500) * 50])
        for x in xs:
            assert hiplib.pg_geom_atan(float(x)) == float(mp.atan(mp.mpf(float(x))))
    else:
        s, c = ctypes.c_double(), ctypes.c_double()
        for x in rng.random(2000) * 2 * math.pi:
            hiplib.pg_geom_sincos(float(x), ctypes.byref(s), ctypes.byref(c))
            assert s.value == float(mp.sin(mp.mpf(float(x)))) and c.value == float(mp.cos(mp.mpf(float(x))))


def test_glibc_misrounds_where_library_differs(hiplib):
    """Where the library and glibc disagree, glibc is the one off by an ulp
    (this is why face crossings are bit-exact only up to libm rounding)."""
    mp.mp.prec = 200
    rng = np.random.default_rng(12)
    n_diff = 0
    for x in (rng.random(20000) - 0.5) * 1.6:
        a, g = hiplib.pg_geom_tan(float(x)), math.tan(float(x))
        if a != g:
            n_diff += 1
            assert a == float(mp.tan(mp.mpf(float(x)))) and abs(a - g) <= math.ulp(a)
    assert n_diff < 200
