"""Test setup: import paths, the ``gpu`` marker, shared fixtures.

``-m "not gpu"`` tests run anywhere (the oracle, host logic, the library's
exports and host-side geometry); ``-m gpu`` tests need an MI355X and call the
HIP library through its C ABI.
"""
from __future__ import annotations

import os
import struct
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "procedural-generation_amd")
for p in (ROOT, PKG_DIR):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")

REC = np.dtype([("face", "<i4"), ("u", "<f8"), ("v", "<f8"), ("x", "<f8"), ("y", "<f8"), ("z", "<f8"),
                ("face2", "<i4"), ("u2", "<f8"), ("v2", "<f8")])


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running")


def load_vectors():
    """The reference's golden cube-sphere vectors (tests/golden/make_fixtures.py)."""
    with open(os.path.join(GOLDEN, "cubesphere_vectors.bin"), "rb") as fh:
        assert fh.read(4) == b"CSV2"
        (K,) = struct.unpack("<i", fh.read(4))
        rec = np.frombuffer(fh.read(), dtype=REC)
    assert rec.shape[0] == K
    return rec


@pytest.fixture(scope="session")
def vectors():
    return load_vectors()


@pytest.fixture(scope="session")
def oracle():
    """The compiled oracle (built on demand with make)."""
    lib = os.path.join(ROOT, "oracle", "liboracle.so")
    src = os.path.join(ROOT, "oracle", "oracle.c")
    if not os.path.exists(lib) or os.path.getmtime(lib) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    from oracle import host

    return host


@pytest.fixture(scope="session")
def hiplib():
    from globe_hip import _lib

    return _lib.lib()


@pytest.fixture(scope="session")
def cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("a gpu test ran without a visible GPU")
    return torch.device("cuda:0")
