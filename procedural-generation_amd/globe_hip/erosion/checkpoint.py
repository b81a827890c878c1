"""Erosion checkpoints in the reference's file layout
(bake/globe/erosion/run.py:99-225): ``erosion_iterNNNN.npz`` with the state
arrays (height, sediment, discharge, momentum, pending, lake_load; cell
units, float64; plus route, base, iso_acc and the lakes' lake_flag, lake_id
and lake_room when the state has them) beside ``erosion_iterNNNN.json`` with ``iteration``,
``params_hash``, ``height_unit_m`` and ``time``; the two newest checkpoints of
a parameter family are kept.  A device state is written from and read back
into its tensors, so a run of this path and a run of the reference resume
from each other's files.

The parameter hash is the caller's (the reference computes it from its
WorldParams, stage hashes and kernel version, erosion/run.py:68-97).
"""
from __future__ import annotations

import json
import re
import time
from pathlib import Path

import numpy as np
import torch

ARRAYS = ("height", "sediment", "discharge", "momentum", "pending", "lake_load")
_CKPT_RE = re.compile(r"erosion_iter(\d+)\.npz$")


def _host(t):
    return t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)


def save_checkpoint(directory, state, params_hash: str, keep: int = 2) -> Path:
    """Write state's arrays at state.iteration (atomically: a .tmp file
    renamed into place, then the .json), then prune this family to `keep`."""
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    it = int(state.iteration)
    p = d / f"erosion_iter{it:04d}.npz"
    tmp = p.with_suffix(".npz.tmp")
    arrays = {}
    for n in ARRAYS:
        t = getattr(state, n, None)
        if t is None or (isinstance(t, torch.Tensor) and t.numel() == 0):
            continue
        arrays[n] = _host(t)
    if getattr(state, "route", None) is not None:
        arrays["route"] = _host(state.route)
    if getattr(state, "base_at", None) is not None:  # run.py:110-111: the sea mask's base level is state
        arrays["base"] = _host(state.base)
    if getattr(state, "iso_acc", None) is not None:  # run.py:112-113: a rebound not yet applied is state
        arrays["iso_acc"] = _host(state.iso_acc)
    if getattr(state, "lake_flag", None) is not None:  # run.py:125-128: the lakes and the room each has left
        arrays["lake_flag"] = _host(state.lake_flag)
        arrays["lake_id"] = _host(state.lake_id)
        arrays["lake_room"] = _host(state.lake_room)
    with open(tmp, "wb") as fh:
        np.savez(fh, **arrays)
    tmp.replace(p)
    meta = {"iteration": it, "params_hash": params_hash, "height_unit_m": float(state.height_unit_m),
            "time": time.strftime("%Y-%m-%dT%H:%M:%S")}
    with open(p.with_suffix(".json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    _prune(d, params_hash, keep)
    return p


def _prune(d: Path, params_hash: str, keep: int) -> None:
    mine = []
    for q in d.glob("erosion_iter*.npz"):
        m = _CKPT_RE.search(q.name)
        jq = q.with_suffix(".json")
        if not m or not jq.exists():
            continue
        try:
            if json.loads(jq.read_text()).get("params_hash") != params_hash:
                continue
        except json.JSONDecodeError:
            continue
        mine.append((int(m.group(1)), q))
    for _, q in sorted(mine)[: max(0, len(mine) - keep)]:
        q.unlink(missing_ok=True)
        q.with_suffix(".json").unlink(missing_ok=True)


def find_checkpoint(directory, params_hash: str, max_iteration: int | None = None):
    """(path, meta) of the newest checkpoint of this parameter family with
    iteration <= max_iteration, or None."""
    d = Path(directory)
    if not d.exists():
        return None
    best = None
    for p in d.glob("erosion_iter*.npz"):
        if not _CKPT_RE.search(p.name):
            continue
        jp = p.with_suffix(".json")
        if not jp.exists():
            continue
        try:
            meta = json.loads(jp.read_text())
        except json.JSONDecodeError:
            continue
        if meta.get("params_hash") != params_hash:
            continue
        it = int(meta.get("iteration", -1))
        if max_iteration is not None and it > max_iteration:
            continue
        if best is None or it > best[1]["iteration"]:
            best = (p, meta)
    return best


def _load_lakes(state, z) -> None:
    """run.py:221-224: lake_flag, lake_id, lake_room as written, or no lakes.
    The flag is copied into the state's own buffer when it has one (a
    captured iteration graph keeps reading that pointer)."""
    if "lake_flag" not in z.files:
        state.lake_flag = None
        return
    flag = torch.from_numpy(np.ascontiguousarray(z["lake_flag"], np.uint8))
    if tuple(flag.shape) != tuple(state.height.shape):
        raise ValueError(f"lake_flag: checkpoint shape {tuple(flag.shape)}, state {tuple(state.height.shape)}")
    if getattr(state, "lake_flag", None) is None:
        state.lake_flag = torch.zeros(tuple(state.height.shape), dtype=torch.uint8, device=state.height.device)
    state.lake_flag.copy_(flag)
    if "lake_id" in z.files:
        lid = torch.from_numpy(np.ascontiguousarray(z["lake_id"], np.int32))
        if getattr(state, "lake_id", None) is None or tuple(state.lake_id.shape) != tuple(lid.shape):
            state.lake_id = torch.empty(tuple(lid.shape), dtype=torch.int32, device=state.height.device)
        state.lake_id.copy_(lid)
        room = torch.from_numpy(np.ascontiguousarray(z["lake_room"], np.float64))
        if getattr(state, "lake_room", None) is None or tuple(state.lake_room.shape) != tuple(room.shape):
            state.lake_room = torch.empty(tuple(room.shape), dtype=torch.float64, device=state.height.device)
        state.lake_room.copy_(room)


def load_checkpoint(state, path, meta: dict) -> None:
    """Copy a checkpoint's arrays into state's tensors (absent pending /
    lake_load read as zero) and set state.iteration."""
    if abs(float(meta.get("height_unit_m", state.height_unit_m)) - float(state.height_unit_m)) > 1e-9 * float(state.height_unit_m):
        raise ValueError("checkpoint written with another cell size")
    with np.load(path) as z:
        for n in ARRAYS:
            t = getattr(state, n, None)
            if t is None or (isinstance(t, torch.Tensor) and t.numel() == 0):
                continue
            src = z[n] if n in z.files else np.zeros(tuple(t.shape))
            if tuple(src.shape) != tuple(t.shape):
                raise ValueError(f"{n}: checkpoint shape {src.shape}, state {tuple(t.shape)}")
            t.copy_(torch.from_numpy(np.ascontiguousarray(src)).to(t.dtype))
        # run.py:203: the routing surface is the file's, or none (a fresh
        # state has none: the next step floods again, as the continuous run's
        # schedule says)
        if "route" in z.files:
            r = torch.from_numpy(np.ascontiguousarray(z["route"], np.float64))
            if tuple(r.shape) != tuple(state.height.shape):
                raise ValueError(f"route: checkpoint shape {tuple(r.shape)}, state {tuple(state.height.shape)}")
            if getattr(state, "route", None) is None:
                state.route = torch.empty_like(state.height)
            state.route.copy_(r)
        else:
            state.route = None
        _load_lakes(state, z)
        if "iso_acc" in z.files:  # run.py:209
            acc = np.ascontiguousarray(z["iso_acc"], np.float64)
            if tuple(acc.shape) != tuple(state.height.shape):
                raise ValueError(f"iso_acc: checkpoint shape {acc.shape}, state {tuple(state.height.shape)}")
            if getattr(state, "iso_acc", None) is None:
                state.iso_acc = torch.zeros_like(state.height)
            state.iso_acc.copy_(torch.from_numpy(acc))
        elif getattr(state, "iso_acc", None) is not None:
            state.iso_acc.zero_()
        if "base" in z.files:  # run.py:204-208: a resume sees the mask the continuous run had
            b = np.ascontiguousarray(z["base"], np.float64)
            if tuple(b.shape) != tuple(state.base.shape):
                raise ValueError(f"base: checkpoint shape {b.shape}, state {tuple(state.base.shape)}")
            state.base.copy_(torch.from_numpy(b))
            state.base_at = int(meta["iteration"])
        else:
            state.base_at = None
    state.iteration = int(meta["iteration"])
