"""Checkpoint and FaceField files in the reference's layout, on the CPU
(bake/globe/erosion/run.py:99-225, bake/globe/field.py:312-342)."""
from __future__ import annotations

import json
import types

import numpy as np
import torch

from globe_hip import cubesphere
from globe_hip.erosion import checkpoint as ck
from globe_hip.field import FaceField


def _state(seed, F=6, NE=12, unit=50.0, it=0):
    g = torch.Generator().manual_seed(seed)
    r = lambda *s: torch.rand(*s, generator=g, dtype=torch.float64)
    return types.SimpleNamespace(height=r(F, NE, NE), sediment=r(F, NE, NE), discharge=r(F, NE, NE),
                                 momentum=r(F, NE, NE, 2), pending=r(F, NE, NE), lake_load=r(F, NE, NE),
                                 route=None, iteration=it, height_unit_m=unit)


def test_checkpoint_round_trip_and_layout(tmp_path):
    a = _state(1, it=7)
    p = ck.save_checkpoint(tmp_path, a, "h1")
    assert p.name == "erosion_iter0007.npz"
    meta = json.loads(p.with_suffix(".json").read_text())
    assert set(meta) == {"iteration", "params_hash", "height_unit_m", "time"} and meta["iteration"] == 7
    with np.load(p) as z:
        assert sorted(z.files) == sorted(ck.ARRAYS)
        assert np.array_equal(z["momentum"], a.momentum.numpy())
    b = _state(2)
    found = ck.find_checkpoint(tmp_path, "h1")
    assert found is not None and found[0] == p
    ck.load_checkpoint(b, *found)
    for n in ck.ARRAYS:
        assert torch.equal(getattr(a, n), getattr(b, n)), n
    assert b.iteration == 7


def test_checkpoint_family_pruning_and_limits(tmp_path):
    for it in (2, 4, 6, 8):
        ck.save_checkpoint(tmp_path, _state(it, it=it), "fam")
    ck.save_checkpoint(tmp_path, _state(9, it=5), "other")
    names = sorted(q.name for q in tmp_path.glob("*.npz"))
    assert names == ["erosion_iter0005.npz", "erosion_iter0006.npz", "erosion_iter0008.npz"]
    assert ck.find_checkpoint(tmp_path, "fam")[1]["iteration"] == 8
    assert ck.find_checkpoint(tmp_path, "fam", max_iteration=7)[1]["iteration"] == 6
    assert ck.find_checkpoint(tmp_path, "fam", max_iteration=3) is None
    assert ck.find_checkpoint(tmp_path, "nope") is None


def test_facefield_files_round_trip(tmp_path):
    grid = cubesphere.Grid(8, 2, 1000.0)
    rng = np.random.default_rng(0)
    interior = rng.random((6, 8, 8)).astype(np.float32)
    f = FaceField.from_interior(grid, interior, name="height", exchange=False)
    paths = f.save(tmp_path)
    assert [q.name for q in paths] == [f"height.f{k}.npy" for k in range(6)]
    assert FaceField.exists(tmp_path, "height")
    g = FaceField.load(tmp_path, "height", grid, exchange=False)
    assert np.array_equal(g.interior, interior) and g.data.dtype == np.float32
    mom = FaceField.from_interior(grid, rng.random((6, 8, 8, 2)), name="momentum", exchange=False)
    mom.save(tmp_path)
    assert FaceField.load(tmp_path, "momentum", grid, exchange=False).is_vector
