"""Build liberode_hip.so for gfx950 with hipcc (in-tree, next to the package).

Usage: python build.py [--jobs N]

Flags:
* ``--offload-arch=gfx950``: MI355X only;
* ``-ffp-contract=off``: no implicit fused multiply-adds anywhere -- the
  check mode's float64 arithmetic must round where the reference's does, and
  the fast mode writes its FMAs explicitly;
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 ALU
  instructions (their forms have given wrong results on this hardware pool);
* ``-fPIC -shared``: a C-ABI shared library loaded with ctypes.
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
INCLUDE = os.path.join(ROOT, "include")
OUT_DIR = os.path.join(os.path.dirname(HERE), "globe_hip")
LIB = os.path.join(OUT_DIR, "liberode_hip.so")
BUILD_DIR = os.path.join(HERE, "_build")
SOURCES = ["pg_api.hip", "pg_check.hip", "pg_ordered.hip", "pg_fields.hip", "pg_fast.hip", "pg_grad.hip", "pg_wasting.hip", "pg_route.hip", "pg_sea.hip", "pg_brush.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-I", INCLUDE, "-I", HERE]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(HERE, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return obj


def build(jobs: int = 4, verbose: bool = False) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(HERE, s))]
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args()
    try:
        build(a.jobs, verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
