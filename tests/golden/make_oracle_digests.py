"""Write the oracle's run of the benchmark workload as per-iteration digests:
tests/golden/earth1024_oracle_digests.json.

The workload is the earth globe at N_c=1024 (stubs.earth_globe, lifetime
128, the uplift replayed over 200 iterations); the oracle runs its
start_replay and then `step` (particles over 2048-droplet chunks with the
serial apply, uplift, datum hold, halo exchange) for ITERATIONS iterations,
with its tan / atan / sincos switched to the correctly rounded double-double
ones of the library under test (pg_geom_tan / pg_geom_atan / pg_geom_sincos,
host code).  After each
iteration the SHA-256 of the state arrays' bytes and the step's statistics
are recorded; the device test replays the same iterations in exact mode and
compares.  The state itself (0.5 GB) is not kept.

Usage: python tests/golden/make_oracle_digests.py [ITERATIONS]   (about 30 s
per iteration on 8 cores)
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "procedural-generation_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

N = 1024
ARRAYS = ("height", "sediment", "pending", "discharge", "momentum")
STATS = ("particles", "entries", "clamped", "to_pending", "deficit_out", "lost_offshore", "datum_shift")
OUT = os.path.join(HERE, "earth1024_oracle_digests.json")


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, np.float64).tobytes()).hexdigest()


def workload():
    from globe_hip import stubs

    grid, f = stubs.earth_globe(N, 2)
    p = stubs.workload_params(N, 2, max_steps=128, uplift_mode="replay", iterations=200)
    args = (grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"], p.erosion)
    return args, p


def main(iterations: int) -> None:
    from globe_hip import _lib
    from oracle import host

    args, p = workload()
    st = host.OracleState.from_grid(*args)
    L = _lib.lib()
    host.set_libm(*host.device_libm(L))
    info = host.start_replay(st, p)
    rec = {"workload": "earth globe N_c=1024 H=2, lifetime 128, uplift replayed over 200 iterations",
           "libm": "correctly rounded (pg_geom_tan / pg_geom_atan / pg_geom_sincos)", "arrays": list(ARRAYS),
           "start_replay": {"datum_shift": float(info["datum_shift"]), **{n: digest(getattr(st, n)) for n in ARRAYS}},
           "iterations": []}
    t0 = time.time()
    for it in range(iterations):
        s = host.step(st, p, it)
        row = {"iteration": it, **{k: (float(s[k]) if k in ("to_pending", "deficit_out", "lost_offshore", "datum_shift")
                                       else int(s[k])) for k in STATS},
               **{n: digest(getattr(st, n)) for n in ARRAYS}}
        rec["iterations"].append(row)
        print(f"iteration {it}: {row['entries']} entries, {time.time() - t0:.0f} s", flush=True)
        with open(OUT, "w") as fh:
            json.dump(rec, fh, indent=1)
    host.set_libm()


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 20)
