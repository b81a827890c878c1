"""The exact mode over the benchmark workload's first iterations, against the
oracle's own run of them (tests/golden/earth1024_oracle_digests.json, written
by tests/golden/make_oracle_digests.py with the oracle on the correctly
rounded tan / atan the device uses): after start_replay and after every
step -- particles over 3072+ chunks, uplift, datum hold, halo exchange --
the SHA-256 of height, sediment, pending, discharge and momentum and the
step's statistics are identical.  The per-cell tolerance at the N_c=1024
earth globe is therefore zero, iteration after iteration.
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

from globe_hip import stubs
from globe_hip.erosion import maps

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "earth1024_oracle_digests.json")


def _digest(t) -> str:
    return hashlib.sha256(np.ascontiguousarray(t.cpu().numpy(), np.float64).tobytes()).hexdigest()


@pytest.mark.timeout(900)
def test_exact_mode_reproduces_the_oracle_run(cuda):
    with open(FIXTURE) as fh:
        rec = json.load(fh)
    grid, f = stubs.earth_globe(1024, 2)
    p = stubs.workload_params(1024, 2, max_steps=128, uplift_mode="replay", iterations=200)
    st = maps.ErosionState.from_grid(grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"],
                                     p.erosion, device=cuda)
    info = maps.start_replay(st, p)
    assert float(info["datum_shift"]) == rec["start_replay"]["datum_shift"]
    for n in rec["arrays"]:
        assert _digest(getattr(st, n)) == rec["start_replay"][n], ("start_replay", n)
    for row in rec["iterations"]:
        it = row["iteration"]
        s = maps.step(st, p, it)
        print(f"iteration {it}: {s['entries']} entries, sweeps max {s['sweeps_max']}", flush=True)
        for k in ("particles", "entries", "clamped"):
            assert int(s[k]) == row[k], (it, k, s[k], row[k])
        for k in ("to_pending", "deficit_out", "lost_offshore", "datum_shift"):
            assert float(s[k]) == row[k], (it, k, float(s[k]), row[k])
        for n in rec["arrays"]:
            assert _digest(getattr(st, n)) == row[n], (it, n)
