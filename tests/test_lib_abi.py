"""The C-ABI library loads and exports every entry point include/*.h
declares; the ctypes binding covers them all.  No compute call: this runs
without a GPU."""
from __future__ import annotations

import glob
import os
import re

from conftest import ROOT


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(pg_[a-z0-9_]+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported(hiplib):
    names = _declared()
    assert len(names) >= 25
    missing = [n for n in sorted(names) if not hasattr(hiplib, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    from globe_hip import _lib

    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert _declared() <= bound, sorted(_declared() - bound)


def test_host_entry_points_and_errors(hiplib):
    from globe_hip import _lib

    assert hiplib.pg_abi_version() == 9
    assert hiplib.pg_change_list_cap(64, 0) == 64 + 16 + 1 + 8 + 8
    assert hiplib.pg_change_list_cap(64, 1) == 128 + 16 + 1 + 8 + 8
    assert hiplib.pg_select_work_bytes() > 0
    # argument errors come back as a status and a message, not a crash
    assert hiplib.pg_pack_samples(None, 0, 0, None, None, None, 0, None, None, None, None, None, None, None, None, None) != 0
    assert b"bad sizes" in hiplib.pg_last_error()
    assert hiplib.pg_geom_to_sphere(7, 0.5, 0.5, None) != 0
    assert b"bad arguments" in hiplib.pg_last_error()
