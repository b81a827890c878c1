"""Device operators of the droplet loop, named and shaped like the reference's
numba kernels (bake/globe/erosion/particle.py) but taking PyTorch-ROCm
tensors and running the HIP kernels of liberode_hip.so on the current stream.

Every operator validates dtypes, shapes, devices and contiguity on the host
before launching and raises ValueError on a mismatch: the kernels index with
the shapes they are given and have no bounds checks of their own.
"""
from __future__ import annotations

import ctypes

import torch

from .._lib import TraceParams, check, lib

MASK_OUTSIDE, MASK_ACTIVE, MASK_FROZEN = 0, 1, 2
DEATH_NAMES = ("age", "ocean", "exit", "pit", "evap", "stop")
NS = 16
SPREAD = 8
LAKE_PARKS = 8
#: bumped whenever a kernel change alters results (the reference's is 13,
#: bake/globe/erosion/particle.py:60; this is the HIP build's own counter)
KERNEL_VERSION = 1


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: torch.Tensor | None):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _need(t: torch.Tensor, dtype, shape=None, name="array"):
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name}: expected a torch tensor, got {type(t).__name__}")
    if t.device.type != "cuda":
        raise ValueError(f"{name}: must live on the GPU (got {t.device})")
    if t.dtype != dtype:
        raise ValueError(f"{name}: dtype {t.dtype}, expected {dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: must be contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    return t


def change_list_cap(max_steps: int, lateral: bool) -> int:
    """particle.py:258-272."""
    return int(lib().pg_change_list_cap(int(max_steps), int(bool(lateral))))


def pack_samples(samp, height, sediment, route, use_route, discharge, momentum, evap, hardness, metric, metric_inv, base, lake_flag):
    """particle.py:164-191 on the device."""
    F, NE, _ = height.shape
    e = (F, NE, NE)
    _need(samp, torch.float32, e + (NS,), "samp")
    for t, n in ((height, "height"), (sediment, "sediment"), (discharge, "discharge"), (base, "base")):
        _need(t, torch.float64, e, n)
    _need(momentum, torch.float64, e + (2,), "momentum")
    _need(evap, torch.float32, e, "evap")
    _need(hardness, torch.float32, e, "hardness")
    _need(metric, torch.float32, e + (3,), "metric")
    _need(metric_inv, torch.float32, e + (3,), "metric_inv")
    if use_route:
        _need(route, torch.float64, e, "route")
    if lake_flag is not None:
        _need(lake_flag, torch.uint8, e, "lake_flag")
    check(lib().pg_pack_samples(_ptr(samp), F, NE, _ptr(height), _ptr(sediment), _ptr(route if use_route else None),
                                int(bool(use_route)), _ptr(discharge), _ptr(momentum), _ptr(evap), _ptr(hardness),
                                _ptr(metric), _ptr(metric_inv), _ptr(base), _ptr(lake_flag), _stream()), "pack_samples")


def trace_particles(sp_face, sp_x, sp_y, sp_sed, sp_dir, volume0, samp, use_route, mask, N, H, spherical,
                    params: TraceParams, roots, use_roots, cl_cell, cl_delta, cl_vol, cl_mom, cl_count, cl_lo, cl_hi,
                    sp_death):
    """particle.py:275-818 (check mode: float64 droplets, change list out)."""
    P = sp_face.shape[0]
    F, NE = samp.shape[0], samp.shape[1]
    if NE != N + 2 * H:
        raise ValueError(f"samp width {NE} != N + 2H = {N + 2 * H}")
    _need(sp_face, torch.int64, (P,), "sp_face")
    for t, n in ((sp_x, "sp_x"), (sp_y, "sp_y"), (sp_sed, "sp_sed"), (sp_dir, "sp_dir")):
        _need(t, torch.float64, (P,), n)
    _need(samp, torch.float32, (F, NE, NE, NS), "samp")
    _need(mask, torch.uint8, (F, NE, NE), "mask")
    cap = change_list_cap(params.max_steps, params.lateral_rate > 0.0)
    for t, dt, n in ((cl_cell, torch.int32, "cl_cell"), (cl_delta, torch.float64, "cl_delta"), (cl_vol, torch.float32, "cl_vol")):
        _need(t, dt, None, n)
        if t.numel() < P * cap:
            raise ValueError(f"{n}: {t.numel()} entries, need P * cap = {P * cap}")
    _need(cl_mom, torch.float32, None, "cl_mom")
    if cl_mom.numel() < 2 * P * cap:
        raise ValueError("cl_mom too small")
    for t, n in ((cl_count, "cl_count"), (cl_lo, "cl_lo"), (cl_hi, "cl_hi")):
        _need(t, torch.int64, None, n)
        if t.numel() < P:
            raise ValueError(f"{n} too small")
    _need(sp_death, torch.int8, None, "sp_death")
    if use_roots:
        _need(roots, torch.float32, (F, NE, NE), "roots")
    if P == 0:
        return
    check(lib().pg_trace_particles(P, _ptr(sp_face), _ptr(sp_x), _ptr(sp_y), _ptr(sp_sed), _ptr(sp_dir), float(volume0),
                                   _ptr(samp), int(bool(use_route)), _ptr(mask), int(N), int(H), int(bool(spherical)),
                                   ctypes.byref(params), _ptr(roots if use_roots else None), int(bool(use_roots)),
                                   _ptr(cl_cell), _ptr(cl_delta), _ptr(cl_vol), _ptr(cl_mom), _ptr(cl_count), _ptr(cl_lo),
                                   _ptr(cl_hi), _ptr(sp_death), _stream()), "trace_particles")


def apply_changes(cl_cell, cl_delta, cl_vol, cl_mom, cl_count, cap, height, sediment, acc, pending, samp, disch_track,
                  mom_track, mask, iter_erode, iter_deposit, fan_slope, use_route, dep_floor, offshore_writeoff,
                  erode_cap, use_erode_cap, lake_load, lake_id, lake_room, out4):
    """particle.py:821-1042: serial apply on the device; the four returned
    quantities land in the device tensor ``out4`` (float64[4])."""
    P = cl_count.shape[0]
    e = tuple(height.shape)
    for t, n in ((height, "height"), (sediment, "sediment"), (acc, "acc"), (pending, "pending"), (disch_track, "disch_track")):
        _need(t, torch.float64, e, n)
    _need(mom_track, torch.float64, e + (2,), "mom_track")
    _need(samp, torch.float32, e + (NS,), "samp")
    _need(mask, torch.uint8, e, "mask")
    _need(out4, torch.float64, (4,), "out4")
    _need(cl_count, torch.int64, (P,), "cl_count")
    if cl_cell.numel() < P * cap:
        raise ValueError("change list smaller than P * cap")
    if use_erode_cap:
        _need(erode_cap, torch.float64, e, "erode_cap")
    if lake_id is not None:
        _need(lake_id, torch.int32, e, "lake_id")
        _need(lake_load, torch.float64, e, "lake_load")
        _need(lake_room, torch.float64, None, "lake_room")
    check(lib().pg_apply_changes(_ptr(cl_cell), _ptr(cl_delta), _ptr(cl_vol), _ptr(cl_mom), _ptr(cl_count), P, int(cap),
                                 _ptr(height), _ptr(sediment), _ptr(acc), _ptr(pending), _ptr(samp), _ptr(disch_track),
                                 _ptr(mom_track), _ptr(mask), float(iter_erode), float(iter_deposit), float(fan_slope),
                                 int(bool(use_route)), float(dep_floor), float(offshore_writeoff),
                                 _ptr(erode_cap if use_erode_cap else None), int(bool(use_erode_cap)), _ptr(lake_load),
                                 _ptr(lake_id), _ptr(lake_room), _ptr(out4), _stream()), "apply_changes")


def chunk_stats(cl_count, sp_death, P, counts7):
    """Entries and death causes of a chunk, added to ``counts7`` (int64[7])."""
    _need(counts7, torch.int64, (7,), "counts7")
    check(lib().pg_chunk_stats(_ptr(cl_count), _ptr(sp_death), int(P), _ptr(counts7), _stream()), "chunk_stats")


def fold_changes(height, sediment, acc, mask):
    """particle.py:1079-1101."""
    e = tuple(height.shape)
    for t, n in ((height, "height"), (sediment, "sediment"), (acc, "acc")):
        _need(t, torch.float64, e, n)
    _need(mask, torch.uint8, e, "mask")
    check(lib().pg_fold_changes(height.numel(), _ptr(height), _ptr(sediment), _ptr(acc), _ptr(mask), _stream()), "fold_changes")


def ema_update(discharge, momentum, disch_track, mom_track, mask, ema):
    """particle.py:1230-1242."""
    e = tuple(discharge.shape)
    _need(discharge, torch.float64, e, "discharge")
    _need(disch_track, torch.float64, e, "disch_track")
    _need(momentum, torch.float64, e + (2,), "momentum")
    _need(mom_track, torch.float64, e + (2,), "mom_track")
    _need(mask, torch.uint8, e, "mask")
    check(lib().pg_ema_update(discharge.numel(), _ptr(discharge), _ptr(momentum), _ptr(disch_track), _ptr(mom_track),
                              _ptr(mask), float(ema), _stream()), "ema_update")


def spawn_cells(cdf, u, out=None):
    """particle.py:1245-1260: index of the first cdf entry > u * cdf[-1]."""
    _need(cdf, torch.float64, None, "cdf")
    _need(u, torch.float64, None, "u")
    if cdf.numel() == 0:
        raise ValueError("spawn_cells: empty cdf")
    if out is None:
        out = torch.empty(u.numel(), dtype=torch.int64, device=u.device)
    _need(out, torch.int64, (u.numel(),), "out")
    check(lib().pg_spawn_cells(_ptr(cdf), cdf.numel(), _ptr(u), u.numel(), _ptr(out), _stream()), "spawn_cells")
    return out


def apply_uplift(height, uplift, mask, cap, mean):
    """maps.py:1132-1134: height += min(uplift, cap) - mean on active cells."""
    e = tuple(height.shape)
    _need(height, torch.float64, e, "height")
    _need(uplift, torch.float64, e, "uplift")
    _need(mask, torch.uint8, e, "mask")
    check(lib().pg_apply_uplift(height.numel(), _ptr(height), _ptr(uplift), _ptr(mask), 0.0 if cap is None else float(cap),
                                int(cap is not None), float(mean), _stream()), "apply_uplift")
