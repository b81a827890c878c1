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


def ordered_workspace_bytes(P: int, cap: int, max_sweeps: int) -> int:
    """Device scratch bytes of ``apply_changes_ordered`` for chunks of up to P droplets."""
    n = int(lib().pg_ordered_workspace_bytes(int(P), int(cap), int(max_sweeps)))
    if n < 0:
        raise RuntimeError(lib().pg_last_error().decode(errors="replace"))
    return n


def apply_changes_ordered(cl_cell, cl_delta, cl_vol, cl_mom, cl_count, cap, height, sediment, acc, pending, samp,
                          disch_track, mom_track, mask, iter_erode, iter_deposit, fan_slope, use_route, dep_floor,
                          offshore_writeoff, erode_cap, use_erode_cap, out4, status, work, max_sweeps,
                          launched_sweeps=40):
    """particle.py:821-1042 computed in parallel with the serial apply's
    results to the bit (fixed-point sweeps over the serial order).  ``status``
    (int32[4], zeroed per iteration by the caller) collects the error bits
    (1: lake parks, unsupported; 2: no convergence within ``max_sweeps``)
    and the sweep counts; ``work`` is a uint8 scratch tensor of
    ``ordered_workspace_bytes`` bytes.  The first ``launched_sweeps`` sweeps
    are plain launches, the rest one cooperative launch."""
    P = cl_count.shape[0]
    e = tuple(height.shape)
    for t, n in ((height, "height"), (sediment, "sediment"), (acc, "acc"), (pending, "pending"), (disch_track, "disch_track")):
        _need(t, torch.float64, e, n)
    _need(mom_track, torch.float64, e + (2,), "mom_track")
    _need(samp, torch.float32, e + (NS,), "samp")
    _need(mask, torch.uint8, e, "mask")
    _need(out4, torch.float64, (4,), "out4")
    _need(status, torch.int32, (4,), "status")
    _need(cl_count, torch.int64, (P,), "cl_count")
    _need(work, torch.uint8, None, "work")
    for t, dt, n in ((cl_cell, torch.int32, "cl_cell"), (cl_delta, torch.float64, "cl_delta"), (cl_vol, torch.float32, "cl_vol")):
        _need(t, dt, None, n)
        if t.numel() < P * cap:
            raise ValueError(f"{n}: {t.numel()} entries, need P * cap = {P * cap}")
    _need(cl_mom, torch.float32, None, "cl_mom")
    if cl_mom.numel() < 2 * P * cap:
        raise ValueError("cl_mom too small")
    if use_erode_cap:
        _need(erode_cap, torch.float64, e, "erode_cap")
    n_cells = int(height.numel())
    check(lib().pg_apply_changes_ordered(_ptr(cl_cell), _ptr(cl_delta), _ptr(cl_vol), _ptr(cl_mom), _ptr(cl_count), P,
                                         int(cap), _ptr(height), _ptr(sediment), _ptr(acc), _ptr(pending), _ptr(samp),
                                         _ptr(disch_track), _ptr(mom_track), _ptr(mask), n_cells, float(iter_erode),
                                         float(iter_deposit), float(fan_slope), int(bool(use_route)), float(dep_floor),
                                         float(offshore_writeoff), _ptr(erode_cap if use_erode_cap else None),
                                         int(bool(use_erode_cap)), _ptr(out4), _ptr(status), int(max_sweeps),
                                         int(launched_sweeps), _ptr(work),
                                         int(work.numel()), _stream()), "apply_changes_ordered")


def check_chunks(sp_face, sp_x, sp_y, sp_sed, sp_dir, volume0, chunk, samp, use_route, mask, N, H, spherical,
                 params: TraceParams, roots, use_roots, cl_cell, cl_delta, cl_vol, cl_mom, cl_count, cl_lo, cl_hi,
                 sp_death, height, sediment, acc, pending, disch_track, mom_track, iter_erode, iter_deposit, fan_slope,
                 dep_floor, offshore_writeoff, erode_cap, use_erode_cap, outs, counts7, status, work, drop_out,
                 max_sweeps, launched_sweeps=40, moves=None, timing=False, lake_load=None, lake_id=None,
                 lake_room=None):
    """The check mode's chunk loop (maps.py:953-1029) in one call: per chunk
    of ``chunk`` droplets trace_particles, the parallel apply and the chunk
    statistics, enqueued without a host round trip.  ``outs`` (float64
    [n_chunks, 4]) receives each chunk's 4-tuple, summed from ``drop_out``
    (float64 [P, 4] scratch).  ``moves`` (int64 [1], optional) accumulates the
    steps traced; ``timing`` records HIP events around each trace launch
    (read them with ``check_trace_time``).  A chunk with lake parks, or one
    that does not converge within ``max_sweeps``, is applied by the serial
    apply (same results); ``status[3]`` counts them.  ``lake_load`` /
    ``lake_id`` / ``lake_room`` as in ``apply_changes``."""
    P = sp_face.shape[0]
    F, NE = samp.shape[0], samp.shape[1]
    e = (F, NE, NE)
    if NE != N + 2 * H:
        raise ValueError(f"samp width {NE} != N + 2H = {N + 2 * H}")
    _need(sp_face, torch.int64, (P,), "sp_face")
    for t, n in ((sp_x, "sp_x"), (sp_y, "sp_y"), (sp_sed, "sp_sed"), (sp_dir, "sp_dir")):
        _need(t, torch.float64, (P,), n)
    _need(samp, torch.float32, e + (NS,), "samp")
    _need(mask, torch.uint8, e, "mask")
    for t, n in ((height, "height"), (sediment, "sediment"), (acc, "acc"), (pending, "pending"), (disch_track, "disch_track")):
        _need(t, torch.float64, e, n)
    _need(mom_track, torch.float64, e + (2,), "mom_track")
    cap = change_list_cap(params.max_steps, params.lateral_rate > 0.0)
    for t, dt, n in ((cl_cell, torch.int32, "cl_cell"), (cl_delta, torch.float64, "cl_delta"), (cl_vol, torch.float32, "cl_vol")):
        _need(t, dt, None, n)
        if t.numel() < chunk * cap:
            raise ValueError(f"{n}: {t.numel()} entries, need chunk * cap = {chunk * cap}")
    _need(cl_mom, torch.float32, None, "cl_mom")
    if cl_mom.numel() < 2 * chunk * cap:
        raise ValueError("cl_mom too small")
    for t, n in ((cl_count, "cl_count"), (cl_lo, "cl_lo"), (cl_hi, "cl_hi")):
        _need(t, torch.int64, None, n)
        if t.numel() < chunk:
            raise ValueError(f"{n} too small")
    _need(sp_death, torch.int8, None, "sp_death")
    if sp_death.numel() < chunk:
        raise ValueError("sp_death too small")
    n_chunks = -(-P // chunk) if P else 0
    _need(outs, torch.float64, (n_chunks, 4), "outs")
    _need(counts7, torch.int64, (7,), "counts7")
    _need(status, torch.int32, (4,), "status")
    _need(work, torch.uint8, None, "work")
    _need(drop_out, torch.float64, None, "drop_out")
    if moves is not None:
        _need(moves, torch.int64, (1,), "moves")
    if drop_out.numel() < 4 * P:
        raise ValueError("drop_out too small")
    if use_roots:
        _need(roots, torch.float32, e, "roots")
    if use_erode_cap:
        _need(erode_cap, torch.float64, e, "erode_cap")
    if lake_id is not None:
        _need(lake_id, torch.int32, e, "lake_id")
        _need(lake_load, torch.float64, e, "lake_load")
        _need(lake_room, torch.float64, None, "lake_room")
    check(lib().pg_check_chunks(P, int(chunk), _ptr(sp_face), _ptr(sp_x), _ptr(sp_y), _ptr(sp_sed), _ptr(sp_dir),
                                float(volume0), _ptr(samp), int(bool(use_route)), _ptr(mask), int(N), int(H),
                                int(bool(spherical)), ctypes.byref(params), _ptr(roots if use_roots else None),
                                int(bool(use_roots)), _ptr(cl_cell), _ptr(cl_delta), _ptr(cl_vol), _ptr(cl_mom),
                                _ptr(cl_count), _ptr(cl_lo), _ptr(cl_hi), _ptr(sp_death), _ptr(height), _ptr(sediment),
                                _ptr(acc), _ptr(pending), _ptr(disch_track), _ptr(mom_track), float(iter_erode),
                                float(iter_deposit), float(fan_slope), float(dep_floor), float(offshore_writeoff),
                                _ptr(erode_cap if use_erode_cap else None), int(bool(use_erode_cap)),
                                _ptr(lake_load if lake_id is not None else None), _ptr(lake_id),
                                _ptr(lake_room if lake_id is not None else None), _ptr(outs),
                                _ptr(counts7), _ptr(status), int(max_sweeps), int(launched_sweeps), _ptr(work),
                                int(work.numel()), _ptr(drop_out), _ptr(moves), int(bool(timing)), _stream()),
          "check_chunks")


def set_trace_lanes(lanes: int) -> None:
    """Droplets per wave of the check-mode trace (8, 16, 32, 64; the same
    results at every setting)."""
    check(lib().pg_check_set_lanes(int(lanes)), "pg_check_set_lanes")


def check_trace_time():
    """(ms, launches) of the trace launches timed by ``check_chunks`` since the last call."""
    ms = ctypes.c_double(0.0)
    n = ctypes.c_int64(0)
    check(lib().pg_check_trace_time(ctypes.byref(ms), ctypes.byref(n)), "check_trace_time")
    return ms.value, n.value


def ordered_status_error(status) -> str | None:
    """The error an ordered apply's status words report, or None."""
    bits = int(status[0])
    if bits & 1:
        return "the change list holds lake parks (volume -3), which the parallel apply does not handle"
    if bits & 2:
        return f"the parallel apply did not converge within its sweep limit (status {list(map(int, status))})"
    return None


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


# ---------------------------------------------------------------------------
# side passes: mass wasting, slope limit, the isostasy smoother
# ---------------------------------------------------------------------------
#: particle.py:99
DEP_FLOOR = 0.02


def _stencil_args(height, sediment, metric, talus, mask):
    F, NE, _ = height.shape
    e = (F, NE, NE)
    _need(height, torch.float64, e, "height")
    _need(sediment, torch.float64, e, "sediment")
    _need(metric, torch.float32, e + (3,), "metric")
    _need(talus, torch.float64, e, "talus")
    _need(mask, torch.uint8, e, "mask")
    if NE < 3:
        raise ValueError("thermal passes need NE >= 3")
    return F, NE, e


def talus(hardness, soft, hard, out):
    """maps.py:1062: soft + (hard - soft) * hardness, in float32, widened."""
    _need(hardness, torch.float32, None, "hardness")
    _need(out, torch.float64, tuple(hardness.shape), "talus")
    check(lib().pg_talus(hardness.numel(), _ptr(hardness), float(soft), float(hard), _ptr(out), _stream()), "talus")


def thermal_pass_a(height, sediment, metric, talus, mask, thermal_rate, thermal_max, out_total, scale):
    """particle.py:1128-1160 (out_total may be None)."""
    F, NE, e = _stencil_args(height, sediment, metric, talus, mask)
    _need(scale, torch.float64, e, "scale")
    if out_total is not None:
        _need(out_total, torch.float64, e, "out_total")
    check(lib().pg_thermal_pass_a(F, NE, _ptr(height), _ptr(sediment), _ptr(metric), _ptr(talus), _ptr(mask),
                                  float(thermal_rate), float(thermal_max), _ptr(out_total), _ptr(scale), _stream()),
          "thermal_pass_a")


def thermal_pass_r(height, sediment, metric, talus, mask, thermal_rate, scale, rscale, dep_floor):
    """particle.py:1163-1187."""
    F, NE, e = _stencil_args(height, sediment, metric, talus, mask)
    _need(scale, torch.float64, e, "scale")
    _need(rscale, torch.float64, e, "rscale")
    check(lib().pg_thermal_pass_r(F, NE, _ptr(height), _ptr(sediment), _ptr(metric), _ptr(talus), _ptr(mask),
                                  float(thermal_rate), _ptr(scale), _ptr(rscale), float(dep_floor), _stream()),
          "thermal_pass_r")


def thermal_pass_b(height, sediment, metric, talus, mask, thermal_rate, out_total, scale, rscale, new_height, new_sediment):
    """particle.py:1190-1227 (out_total is not read, as in the reference)."""
    F, NE, e = _stencil_args(height, sediment, metric, talus, mask)
    for t, n in ((scale, "scale"), (rscale, "rscale"), (new_height, "new_height"), (new_sediment, "new_sediment")):
        _need(t, torch.float64, e, n)
    check(lib().pg_thermal_pass_b(F, NE, _ptr(height), _ptr(sediment), _ptr(metric), _ptr(talus), _ptr(mask),
                                  float(thermal_rate), _ptr(scale), _ptr(rscale), _ptr(new_height), _ptr(new_sediment),
                                  _stream()), "thermal_pass_b")


def mass_wasting_pass(height, sediment, metric, talus, mask, rate, cap, dep_floor, work):
    """maps.py:1073-1085: the three passes, written back into height /
    sediment; work: float64 scratch of 4 x height.numel()."""
    F, NE, e = _stencil_args(height, sediment, metric, talus, mask)
    _need(work, torch.float64, None, "work")
    if work.numel() < 4 * height.numel():
        raise ValueError("mass_wasting_pass: work needs 4 x the cells")
    check(lib().pg_mass_wasting_pass(F, NE, _ptr(height), _ptr(sediment), _ptr(metric), _ptr(talus), _ptr(mask),
                                     float(rate), float(cap), float(dep_floor), _ptr(work), _stream()),
          "mass_wasting_pass")


def creep_mask(height, sediment, base, mask, out):
    """maps.py:1067: FROZEN where the surface is below base, else mask."""
    e = tuple(height.shape)
    for t, n in ((height, "height"), (sediment, "sediment"), (base, "base")):
        _need(t, torch.float64, e, n)
    _need(mask, torch.uint8, e, "mask")
    _need(out, torch.uint8, e, "out")
    check(lib().pg_creep_mask(height.numel(), _ptr(height), _ptr(sediment), _ptr(base), _ptr(mask), _ptr(out), _stream()),
          "creep_mask")


def slope_erode_cap(height, sediment, mask, coef, exit_slope, out):
    """particle.py:1045-1076."""
    F, NE, _ = height.shape
    e = (F, NE, NE)
    _need(height, torch.float64, e, "height")
    _need(sediment, torch.float64, e, "sediment")
    _need(mask, torch.uint8, e, "mask")
    _need(out, torch.float64, e, "out")
    check(lib().pg_slope_erode_cap(F, NE, _ptr(height), _ptr(sediment), _ptr(mask), float(coef), float(exit_slope),
                                   _ptr(out), _stream()), "slope_erode_cap")


def laplacian(data, metric_inv, sqrt_det, cell_size_sq, mode, kappa, c2, out):
    """field.py:179-188 / 372-404 (mode 0: the raw operator; 1: float32
    Laplacian x c2; 2: data + kappa x Laplacian x c2)."""
    F, NE, _ = data.shape
    e = (F, NE, NE)
    _need(data, torch.float64, e, "data")
    _need(metric_inv, torch.float32, e + (3,), "metric_inv")
    _need(sqrt_det, torch.float32, e, "sqrt_det")
    _need(out, torch.float64, e, "out")
    check(lib().pg_laplacian(F, NE, _ptr(data), _ptr(metric_inv), _ptr(sqrt_det), float(cell_size_sq), int(mode),
                             float(kappa), float(c2), _ptr(out), _stream()), "laplacian")


def interior_sums(data, mask, N, H, what, work, out4):
    """{sum of v / v^2 / |v|, count, max, min} over interior cells (mask
    None: all; else mask == ACTIVE), into the device double[4] out4."""
    F = data.shape[0]
    _need(data, torch.float64, None, "data")
    if mask is not None:
        _need(mask, torch.uint8, tuple(data.shape), "mask")
    _need(out4, torch.float64, (4,), "out4")
    check(lib().pg_interior_sums(_ptr(data), _ptr(mask), F, int(N), int(H), int(what), _ptr(work), _ptr(out4), _stream()),
          "interior_sums")


def rebound(height, data, mask, sums):
    e = tuple(height.shape)
    _need(height, torch.float64, e, "height")
    _need(data, torch.float64, e, "data")
    _need(mask, torch.uint8, e, "mask")
    check(lib().pg_rebound(height.numel(), _ptr(height), _ptr(data), _ptr(mask), _ptr(sums), _stream()), "rebound")


def surface(height, sediment, out):
    e = tuple(height.shape)
    for t, n in ((height, "height"), (sediment, "sediment"), (out, "out")):
        _need(t, torch.float64, e, n)
    check(lib().pg_surface(height.numel(), _ptr(height), _ptr(sediment), _ptr(out), _stream()), "surface")


def iso_accumulate(iso_acc, height, sediment, s0):
    e = tuple(height.shape)
    for t, n in ((iso_acc, "iso_acc"), (height, "height"), (sediment, "sediment"), (s0, "s0")):
        _need(t, torch.float64, e, n)
    check(lib().pg_iso_accumulate(height.numel(), _ptr(iso_acc), _ptr(height), _ptr(sediment), _ptr(s0), _stream()),
          "iso_accumulate")


def scale(x, a, out):
    """out = a * x (a == 0: zeros, x may be None)."""
    _need(out, torch.float64, None, "out")
    if x is not None:
        _need(x, torch.float64, tuple(out.shape), "x")
    check(lib().pg_scale(out.numel(), _ptr(x), float(a), _ptr(out), _stream()), "scale")


def divide(x, v, out):
    _need(x, torch.float64, None, "x")
    _need(out, torch.float64, tuple(x.shape), "out")
    check(lib().pg_divide(x.numel(), _ptr(x), float(v), _ptr(out), _stream()), "divide")


# ---------------------------------------------------------------------------
# the routing surface (route.py:30-117)
# ---------------------------------------------------------------------------
def refresh_base(height, sediment, owner, area_fx, threshold_fx, N, H, base, work) -> dict:
    """The sea mask's base levels (pg_refresh_base; host-synchronous)."""
    info = (ctypes.c_double * 5)()
    check(lib().pg_refresh_base(_ptr(height), _ptr(sediment), _ptr(owner), _ptr(area_fx), float(threshold_fx), int(N),
                                int(H), _ptr(base), _ptr(work), int(work.numel() * work.element_size()), info,
                                _stream()), "pg_refresh_base")
    return {"closed_cells": int(info[0]), "closed_basins": int(info[1]), "sea_cells": int(info[2]),
            "deepest_basin_floor": float(info[3]), "rounds": int(info[4])}


def flood_init(height, sediment, base, cls, H, N, s_out, W, flag):
    F, NE, _ = height.shape
    e = (F, NE, NE)
    for t, n in ((height, "height"), (sediment, "sediment"), (base, "base"), (s_out, "s"), (W, "W")):
        _need(t, torch.float64, e, n)
    _need(cls, torch.uint8, e, "cls")
    _need(flag, torch.int32, (1,), "flag")
    check(lib().pg_flood_init(F, NE, int(H), int(N), _ptr(height), _ptr(sediment), _ptr(base), _ptr(cls), _ptr(s_out),
                              _ptr(W), _ptr(flag), _stream()), "flood_init")


def flood_sweeps(s, cls, W, eps, spec_cell, spec_ptr, spec_src, sweeps, flag):
    F, NE, _ = s.shape
    n_spec = 0 if spec_cell is None else int(spec_cell.numel())
    if n_spec:
        _need(spec_cell, torch.int64, (n_spec,), "spec_cell")
        _need(spec_ptr, torch.int32, (n_spec + 1,), "spec_ptr")
        _need(spec_src, torch.int64, None, "spec_src")
    check(lib().pg_flood_sweeps(F, NE, _ptr(s), _ptr(cls), _ptr(W), float(eps), _ptr(spec_cell if n_spec else None),
                                _ptr(spec_ptr if n_spec else None), _ptr(spec_src if n_spec else None), n_spec,
                                int(sweeps), _ptr(flag), _stream()), "flood_sweeps")


def flood_finish(s, W, route):
    e = tuple(s.shape)
    for t, n in ((s, "s"), (W, "W"), (route, "route")):
        _need(t, torch.float64, e, n)
    check(lib().pg_flood_finish(s.numel(), _ptr(s), _ptr(W), _ptr(route), _stream()), "flood_finish")
