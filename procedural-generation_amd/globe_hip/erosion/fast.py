"""Fast mode of the particle pass: one C-ABI call per iteration, everything
on the device (see ``pg_fast_iterate`` in include/pg_erode.h).

The spawn stream is still ``params.rng(stage, *iteration_key)``: the host
seeds numpy's PCG64 exactly as the reference does and hands its 128-bit
state to the device, which draws the same doubles (jump-ahead per thread),
sorts them, and spawns through an exact fixed-point prefix sum of the rain
weights -- so the droplets' cells and offsets are the reference's.  What
differs from the check mode is the trace (float32, every droplet of a chunk
against the same frozen terrain) and the apply (int64 fixed-point atomics
with order-independent caps).
"""
from __future__ import annotations

import ctypes
import time

import numpy as np
import torch

from .._lib import FAST_NSTATS, FS, FastFields, FastParams, check, lib
from ..config import cell_units
from . import particle as pk

#: default frozen-terrain chunks per iteration (the reference splits an
#: iteration into at least 16, bake/globe/erosion/maps.py:781-787)
DEFAULT_CHUNKS = 16
HEIGHT_BITS = 48
TRACK_BITS = 32
#: largest move tape allocated by default: 64 GiB, under a quarter of the part's 288 GB of HBM, and never
#: more than half of the device's free memory (a planet at N_c=8192 would need more: it traces twice instead)
TAPE_MAX_BYTES = 64 << 30


class FastContext:
    """Per-state device workspace of the fast mode."""

    def __init__(self, state, n_particles: int):
        self.n_particles = int(n_particles)
        dev = state.device
        nbytes = int(lib().pg_fast_workspace_bytes(state.F, state.NE, self.n_particles))
        if nbytes < 0:
            check(-1, "pg_fast_workspace_bytes")
        self.work = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        self.stats = torch.zeros(FAST_NSTATS, dtype=torch.int64, device=dev)
        self.fields = FastFields(
            F=state.F, N=state.N, H=state.H, spherical=int(bool(state.spherical)),
            deposit_on_exit=int(bool(state.deposit_on_exit)), height=state.height.data_ptr(),
            sediment=state.sediment.data_ptr(), discharge=state.discharge.data_ptr(), momentum=state.momentum.data_ptr(),
            pending=state.pending.data_ptr(), base=state.base.data_ptr(), precip=state.precip.data_ptr(),
            evap=state.evap.data_ptr(), hardness=state.hardness.data_ptr(), metric=state.metric.data_ptr(),
            metric_inv=state.metric_inv.data_ptr(), mask=state.mask.data_ptr())
        self._ptrs = self._pointers(state)
        fp = FastParams(n_particles=self.n_particles, chunks=1, height_bits=HEIGHT_BITS, track_bits=TRACK_BITS)
        check(lib().pg_fast_setup(ctypes.byref(self.fields), ctypes.byref(fp), ctypes.c_void_p(self.work.data_ptr()),
                                  nbytes, pk._stream()), "pg_fast_setup")

    @staticmethod
    def _pointers(state):
        return tuple(t.data_ptr() for t in (state.height, state.sediment, state.discharge, state.momentum, state.pending,
                                            state.base, state.precip, state.evap, state.hardness, state.metric,
                                            state.metric_inv, state.mask))

    def valid_for(self, state, n_particles: int) -> bool:
        return n_particles == self.n_particles and self._pointers(state) == self._ptrs

    def tape_for(self, state, chunks: int, max_steps: int, grid_blocks: int, variant: int):
        """The move tape for this chunking and trace grid (allocated on first
        use, grown when a larger one is needed), or None when
        ``state.fast_tape`` is False."""
        if not bool(getattr(state, "fast_tape", True)):
            return None
        need = int(lib().pg_fast_tape_bytes(self.n_particles, state.F, state.NE, int(chunks), int(max_steps),
                                            int(grid_blocks), int(variant)))
        if need < 0:
            check(-1, "pg_fast_tape_bytes")
        if need > int(getattr(state, "fast_tape_max_bytes", TAPE_MAX_BYTES)):
            return None  # too large for this chunking: the applying pass traces again (same results)
        if getattr(self, "tape", None) is None or self.tape.numel() < need:
            self.tape = None
            # at most half of what the device has free; without a tape the applying pass traces again
            if need > torch.cuda.mem_get_info(state.device)[0] // 2:
                return None
            try:
                self.tape = torch.empty(need, dtype=torch.uint8, device=state.device)
            except torch.cuda.OutOfMemoryError:
                self.tape = None
                return None
        return self.tape


def _fast_context(state, n_particles: int) -> FastContext:
    ctx = state._fast
    if ctx is None or not ctx.valid_for(state, n_particles):
        ctx = FastContext(state, n_particles)
        state._fast = ctx
    return ctx


def pcg64_state(rng: np.random.Generator):
    st = rng.bit_generator.state
    if st.get("bit_generator") != "PCG64":
        raise ValueError(f"the fast mode draws numpy's PCG64 stream; params.rng gave {st.get('bit_generator')}")
    s, inc = int(st["state"]["state"]), int(st["state"]["inc"])
    m = (1 << 64) - 1
    return s & m, s >> 64, inc & m, inc >> 64


def decode_stats(raw: np.ndarray, hscale: float) -> dict:
    """Device statistics -> the reference's stats keys (cell units)."""
    raw = np.asarray(raw, np.int64)
    P = int(raw[FS["PARTICLES"]])
    deaths = {n: int(raw[FS["DEATHS"] + i]) for i, n in enumerate(pk.DEATH_NAMES)}
    fx = lambda k: float(raw[FS[k]]) / hscale
    vol0 = float(np.array([raw[FS["VOLUME0_BITS"]]], np.int64).view(np.float64)[0])
    return {"particles": P, "entries": int(raw[FS["ENTRIES"]]), "steps_mean": int(raw[FS["ENTRIES"]]) / max(P, 1),
            "moves": int(raw[FS["MOVES"]]), "volume": vol0, "deaths": deaths, "clamped": int(raw[FS["CLAMPED"]]),
            "to_pending": fx("TO_PENDING_FX"), "released": fx("RELEASED_FX"), "deficit_out": fx("EXIT_FX"),
            "lost_offshore": fx("LOST_OFFSHORE_FX"), "vanished": fx("VANISHED_FX"), "eroded": fx("ERODED_FX"),
            "deposited": fx("DEPOSITED_FX"), "deficit_cancelled": fx("DEFICIT_FX"), "lateral": fx("LATERAL_FX"),
            "eligible": int(raw[FS["ELIGIBLE"]]),
            "cdf_serial": bool(raw[FS["CDF_SERIAL"]]), "tape_short": int(raw[FS["TAPE_SHORT"]]), "raw_fx": raw.copy()}


def run_iteration(state, params, iteration_key, n_particles=None, particles_per_cell=None, rng_stage="erosion",
                  log=None, chunks=None, sync_stats: bool | None = None) -> dict:
    """Fast-mode particle pass of one iteration; same arguments and stats keys
    as the reference's run_iteration (bake/globe/erosion/maps.py:839-1047),
    plus ``moves`` (droplet-steps executed).  With ``sync_stats=False`` the
    statistics stay on the device (``state._fast.stats``) and nothing waits
    (default: ``state.fast_sync_stats``, True when unset)."""
    from .maps import (_check_scope, _eparams, chunk_size, discharge_scales, lake_park_of, max_steps_of,
                       slope_limit_cap, slope_limited_deposit, trace_params)

    ep = _eparams(params)
    _check_scope(ep, state)
    if state.roots is not None or lake_park_of(state, ep) > 0.0:
        raise NotImplementedError("roots and lake parks are outside the fast mode: use mode='check'")
    key = iteration_key if isinstance(iteration_key, (tuple, list)) else (int(iteration_key),)
    rng = params.rng(rng_stage, *key)
    t0 = time.time()
    n_cells = int(np.count_nonzero(state._mask_host[state.interior] == pk.MASK_ACTIVE))
    if n_particles is None:
        ppc = ep.particles_per_cell if particles_per_cell is None else particles_per_cell
        n_particles = max(1, int(round(ppc * n_cells)))
    ctx = _fast_context(state, n_particles)
    unit = state.height_unit_m
    max_steps = max_steps_of(ep, state.N)
    # in cells of upstream rain (disc_saturation_cells / momentum_saturation_cells) the scales are computed on
    # the device from the iteration's spawn; otherwise they are the parameters'
    cells = float(ep.disc_saturation_cells) > 0.0 or float(ep.momentum_saturation_cells) > 0.0
    disc_sat, mom_scale = (float(ep.disc_saturation), 1.0) if cells else discharge_scales(ep, state, 1.0, n_particles,
                                                                                          n_rain=1)
    prm = trace_params(ep, state, 1.0, 1, 1, disc_sat, mom_scale, max_steps, 0.0)
    ecap = slope_limit_cap(state, ep)
    n_chunks = int(chunks if chunks is not None else getattr(state, "fast_chunks", DEFAULT_CHUNKS))
    tape = ctx.tape_for(state, n_chunks, max_steps, int(getattr(state, "fast_grid_blocks", 0)),
                        int(getattr(state, "fast_variant", 0)))
    brush = int(getattr(ep, "brush_radius", 0))
    owner = state.owner_device() if brush >= 1 else None
    slo, shi, ilo, ihi = pcg64_state(rng)
    fp = FastParams(rng_state_lo=slo, rng_state_hi=shi, rng_inc_lo=ilo, rng_inc_hi=ihi, n_particles=int(n_particles),
                    chunks=n_chunks,
                    iter_deposit=slope_limited_deposit(ep, unit), iter_erode=cell_units(ep, "iter_erode", unit),
                    conc_chunk=float(getattr(state, "fast_conc_chunk", -1)) if getattr(state, "fast_conc_chunk", -1) >= 0
                    else float(chunk_size(ep, n_particles)), fan_slope=float(ep.fan_slope),
                    offshore_writeoff=float(ep.offshore_writeoff),
                    ema=float(ep.ema), min_volume_frac=float(ep.min_volume_frac), height_bits=HEIGHT_BITS,
                    track_bits=TRACK_BITS, grid_blocks=int(getattr(state, "fast_grid_blocks", 0)),
                    time_trace=int(bool(getattr(state, "fast_time_trace", False))),
                    variant=int(getattr(state, "fast_variant", 0)),
                    rng_dev=(state._fast_rng_dev.data_ptr() if getattr(state, "_fast_rng_dev", None) is not None else None),
                    erode_cap=None if ecap is None else ecap.data_ptr(),
                    tape=None if tape is None else tape.data_ptr(), tape_bytes=0 if tape is None else tape.numel(),
                    route=None if state.route is None else state.route.data_ptr(),
                    lake_flag=None if state.route is None or state.lake_flag is None else state.lake_flag.data_ptr(),
                    disc_saturation_cells=float(ep.disc_saturation_cells),
                    momentum_saturation_cells=float(ep.momentum_saturation_cells),
                    inflow_volume=float(state.inflow_volume), brush_radius=brush,
                    brush_owner=(None if owner is None else owner.data_ptr()),
                    brush_tile=int(getattr(state, "fast_brush_tile", 0)))
    check(lib().pg_fast_iterate(ctypes.byref(ctx.fields), ctypes.byref(prm), ctypes.byref(fp),
                                ctypes.c_void_p(ctx.work.data_ptr()), ctx.work.numel(),
                                ctypes.c_void_p(ctx.stats.data_ptr()), pk._stream()), "pg_fast_iterate")
    if sync_stats is None:
        sync_stats = bool(getattr(state, "fast_sync_stats", True))
    if not sync_stats:
        return {"device_stats": ctx.stats, "seconds": time.time() - t0}
    stats = decode_stats(ctx.stats.cpu().numpy(), float(2 ** HEIGHT_BITS))
    stats["pending_total"] = float(state.pending[state.interior].sum().item())
    stats["seconds"] = time.time() - t0
    stats["chunk"] = -(-stats["particles"] // fp.chunks) if stats["particles"] else 0
    if cells and stats["eligible"] > 0:
        # the device's scales, restated from the spawn it reports (maps.py:911-921)
        vol = stats["volume"] * float(n_particles)
        rpc = max(vol - float(state.inflow_volume), 0.0) / stats["eligible"]
        rpc = rpc if rpc > 0.0 else vol / stats["eligible"]
        if float(ep.disc_saturation_cells) > 0.0:
            disc_sat = float(ep.disc_saturation_cells) * rpc
        if float(ep.momentum_saturation_cells) > 0.0:
            mom_scale = stats["volume"] / (float(ep.momentum_saturation_cells) * rpc)
    stats["disc_saturation"] = disc_sat
    stats["mom_scale"] = mom_scale
    if log is not None:
        log(f"  particles {stats['particles']} volume {stats['volume']:.3f} mean steps {stats['steps_mean']:.1f} "
            f"deaths {stats['deaths']} clamped {stats['clamped']} pending {stats['pending_total']:.1f} ({stats['seconds']:.2f}s)")
    return stats


def trace_time():
    """(ms, launches) of the trace launches recorded since the last call
    (states with ``fast_time_trace = True``); waits for them."""
    ms = ctypes.c_double()
    n = ctypes.c_int64()
    check(lib().pg_fast_trace_time(ctypes.byref(ms), ctypes.byref(n)), "pg_fast_trace_time")
    return ms.value, n.value


def trace_time_passes():
    """((census ms, launches), (applying-pass ms, launches)) of the trace
    launches recorded since the last call, split by pass; waits for them."""
    ms = (ctypes.c_double * 2)()
    n = (ctypes.c_int64 * 2)()
    check(lib().pg_fast_trace_time_passes(ms, n), "pg_fast_trace_time_passes")
    return (ms[0], n[0]), (ms[1], n[1])


def brush_time():
    """(ms, launches) of the brush launches recorded since the last call
    (states with ``fast_time_trace = True``); waits for them."""
    ms = ctypes.c_double()
    n = ctypes.c_int64()
    check(lib().pg_fast_brush_time(ctypes.byref(ms), ctypes.byref(n)), "pg_fast_brush_time")
    return ms.value, n.value


def brush_weights(radius: int) -> np.ndarray:
    """The library's integer brush weights of a radius, (2A+1, 2A+1) in
    units of 2^-16 (the centre 0: it keeps the remainder)."""
    w = (ctypes.c_int32 * 81)()
    A = lib().pg_brush_weights(int(radius), w, 81)
    if A < 0:
        check(-1, "pg_brush_weights")
    K = 2 * A + 1
    return np.array(w[:K * K], np.int64).reshape(K, K)


def brush_spread(src: torch.Tensor, mask: torch.Tensor, owner: torch.Tensor | None, radius: int, tile: int = 0):
    """pg_brush_spread on device tensors: src int64 (F, NE, NE) fixed-point
    amounts, mask uint8, owner int64 or None; returns the spread amounts."""
    F, NE = int(src.shape[0]), int(src.shape[1])
    dst = torch.empty_like(src)
    valid = torch.empty(src.shape, dtype=torch.uint8, device=src.device)
    check(lib().pg_brush_spread(src.data_ptr(), dst.data_ptr(), mask.data_ptr(),
                                None if owner is None else owner.data_ptr(), F, NE, int(radius), int(tile),
                                valid.data_ptr(), pk._stream()), "pg_brush_spread")
    return dst


def _signed(w: int) -> int:
    return w - (1 << 64) if w >= (1 << 63) else w


class IterationGraph:
    """A whole fast-mode erosion iteration -- particle pass, mass wasting,
    uplift, isostasy, datum hold, halo exchange (bake/globe/erosion/maps.py:
    1516-1617) -- captured once as a HIP graph and replayed per iteration
    (two graphs when isostasy is on: with and without its application).  The
    sea mask, the routing flood and a window's lakes, when due, run eagerly
    before the replay, in place.

    The only per-iteration input from the host is the iteration's PCG64
    state (``params.rng(stage, *key)``), copied into a device buffer before
    each replay; the kernels read it from there.  Everything else (droplet
    counts, the stockpiles to re-inject, the datum shift) stays on the
    device, so a replay needs no host synchronisation.  Statistics of the
    last replay: ``self.stats`` (device int64, ``decode_stats``)."""

    def __init__(self, state, params, rng_stage: str = "erosion", n_particles=None):
        from . import maps

        if state.mode != "fast":
            raise ValueError("graph capture runs the fast mode")
        self.state, self.params, self.rng_stage = state, params, rng_stage
        ep = maps._eparams(params)
        dev = state.device
        state.fast_sync_stats = False
        state.fast_time_trace = False
        self.rng_dev = torch.zeros(4, dtype=torch.int64, device=dev)
        state._fast_rng_dev = self.rng_dev
        # everything the iteration allocates lazily, before capture
        n_cells = int(np.count_nonzero(state._mask_host[state.interior] == pk.MASK_ACTIVE))
        if n_particles is None:
            n_particles = max(1, int(round(ep.particles_per_cell * n_cells)))
        self.n_particles = n_particles
        ctx = _fast_context(state, n_particles)
        state.uplift_mean(maps.uplift_cap(state, ep))
        state.select_work()
        if state.spherical:
            state.halo_tables()
        maps.prepare_side_passes(state, params)
        # the routing surface and lake flags exist (their buffers) before the capture
        if int(ep.flood_every) > 0:
            maps.refresh_route_and_lakes(state, params, due=True)
        torch.cuda.synchronize()
        self._load_rng((0,))
        # with isostasy on, a second graph holds the iterations that apply it
        self.iso_every = max(int(getattr(ep, "isostasy_every", 10)), 1) \
            if state.spherical and float(getattr(ep, "isostasy", 0.0)) > 0.0 else 0
        it0 = state.iteration
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.result = maps.step(state, params, (0,), n_particles=n_particles, isostasy_due=False, sea_due=False,
                                    route_due=False)
        self.graph_iso = None
        if self.iso_every:
            state.iteration = it0
            self.graph_iso = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph_iso):
                self.result_iso = maps.step(state, params, (0,), n_particles=n_particles, isostasy_due=True,
                                            sea_due=False, route_due=False)
        state.iteration = it0
        self.stats = ctx.stats

    def _load_rng(self, key):
        rng = self.params.rng(self.rng_stage, *key)
        words = [_signed(w) for w in pcg64_state(rng)]
        self.rng_dev.copy_(torch.tensor(words, dtype=torch.int64), non_blocking=False)

    def step(self, iteration_key) -> None:
        """Run one iteration (asynchronous)."""
        from . import maps

        key = iteration_key if isinstance(iteration_key, (tuple, list)) else (int(iteration_key),)
        ep = maps._eparams(self.params)
        if maps.sea_refresh_due(self.state, ep):
            # host-synchronous, so outside the graph: the replay reads the new base
            self.state.refresh_base(maps.ocean_min_fraction(self.params))
        if maps.route_refresh_due(self.state, ep) or maps.lakes_refresh_due(self.state, ep):
            # the flood and the lakes likewise, in place into the buffers the graph reads
            maps.refresh_route_and_lakes(self.state, self.params)
        self._load_rng(key)
        if self.graph_iso is not None and (self.state.iteration + 1) % self.iso_every == 0:
            self.graph_iso.replay()
        else:
            self.graph.replay()
        self.state.iteration += 1
