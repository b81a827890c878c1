"""Batched windows: many single-face windows eroded together on one GPU.

The reference erodes zoom tiles one process per tile (each an
``ErosionState.window`` driven by ``maps.run_iteration``,
bake/globe/zoom/bake.py:777-923 and zoom/planet.py:324-437).  Here the
windows -- each with its own size, seed, droplet count and lifetime, all
sharing the other erosion parameters and the cell size -- live in one device
arena (every field's windows back to back), and one iteration is

  - per window: pack and spawn from the window's own ``params.rng`` stream
    into the window's slice of the arena (``pg_batch_iterate``);
  - per chunk: the census trace as ONE persistent launch over every
    window's share of the chunk (a device work queue across windows; lanes
    pick up the geometry and droplet scalars of the window they draw from),
    recording each move on the batch's move tape; the per-cell erosion scale;
    the applying pass replaying the tape (or, without a tape, tracing again
    in one launch); the resolve over the whole arena;
  - fold / EMA over the arena, then the windows' uplift.

Windows never interact and every accumulation is int64 fixed point, so the
result is bit-identical to running each window alone in fast mode
(``step_per_window``, the one-launch-per-window form it is measured against).
``BatchGraph`` captures the batched iteration in a HIP graph;
``Autotuner`` picks the persistent grid from timed real iterations (results
do not depend on it).
"""
from __future__ import annotations

import math

import ctypes

import numpy as np
import torch

from .._lib import FAST_NSTATS, FS, FastFields, FastParams, check, lib
from ..config import cell_units
from . import fast
from . import maps
from . import particle as pk


def _signed(w: int) -> int:
    return w - (1 << 64) if w >= (1 << 63) else w


class WindowBatch:
    """windows: ``ErosionState.window`` keyword arguments (metres, one dict
    per window); params: one WorldParams per window (seeds and max_steps may
    differ, everything else must agree)."""

    def __init__(self, windows: list[dict], params: list, device="cuda", chunks: int = fast.DEFAULT_CHUNKS,
                 rng_stage: str = "erosion"):
        if len(windows) != len(params) or not windows:
            raise ValueError("one WorldParams per window")
        self.W = len(windows)
        self.params = list(params)
        self.rng_stage = rng_stage
        self.chunks = int(chunks)
        dev = torch.device(device)
        eps = [maps._eparams(p) for p in params]
        ref = {k: v for k, v in vars(eps[0]).items() if k != "max_steps"}
        for e in eps[1:]:
            if {k: v for k, v in vars(e).items() if k != "max_steps"} != ref:
                raise ValueError("batched windows share their erosion parameters (max_steps aside)")
        hu = {float(w["height_unit_m"]) for w in windows}
        H = {int(w["H"]) for w in windows}
        doe = {bool(w.get("deposit_on_exit", False)) for w in windows}
        if len(hu) != 1 or len(H) != 1 or len(doe) != 1:
            raise ValueError("batched windows share the cell size, halo width and deposit_on_exit")
        self.unit, self.H = hu.pop(), H.pop()
        NE = [int(np.asarray(w["height_m"]).shape[0]) for w in windows]
        self.NE = NE
        off = np.concatenate([[0], np.cumsum([n * n for n in NE])]).astype(np.int64)
        self.off = off
        n = int(off[-1])

        def arena(key, dtype, comp=(), scale=None):
            a = np.empty((n,) + comp, dtype)
            for w, win in enumerate(windows):
                x = np.asarray(win[key], np.float64 if scale else dtype)
                a[off[w]:off[w + 1]] = (x / scale if scale else x).reshape((-1,) + comp)
            return torch.from_numpy(a).to(dev)

        u = self.unit
        self.height = arena("height_m", np.float64, scale=u)
        self.sediment = arena("sediment_m", np.float64, scale=u)
        self.discharge = arena("discharge", np.float64)
        self.momentum = arena("momentum", np.float64, (2,))
        self.hardness = arena("hardness", np.float32)
        self.precip = arena("precip", np.float32)
        self.evap = arena("evap", np.float32)
        self.uplift = arena("uplift_m", np.float64, scale=u)
        self.mask = arena("mask", np.uint8)
        self.metric = arena("metric", np.float32, (3,))
        self.metric_inv = arena("metric_inv", np.float32, (3,))
        self.pending = torch.zeros(n, dtype=torch.float64, device=dev)
        self.base = torch.zeros(n, dtype=torch.float64, device=dev)
        empty = torch.empty(0, device=dev)
        self.states = []
        for w in range(self.W):
            s = slice(int(off[w]), int(off[w + 1]))
            v = lambda t, c=(): t[s].view((1, NE[w], NE[w]) + c)
            st = maps.ErosionState(height=v(self.height), sediment=v(self.sediment), discharge=v(self.discharge),
                                   momentum=v(self.momentum, (2,)), hardness=v(self.hardness), precip=v(self.precip),
                                   evap=v(self.evap), uplift=v(self.uplift), mask=v(self.mask),
                                   metric=v(self.metric, (3,)), metric_inv=v(self.metric_inv, (3,)),
                                   N=NE[w] - 2 * self.H, H=self.H, spherical=False, height_unit_m=u,
                                   deposit_on_exit=doe_val(windows[w]), mode="fast", pending=v(self.pending),
                                   base=v(self.base), samp=empty, acc=empty, disch_track=empty, mom_track=empty,
                                   lake_load=empty)
            st.fast_chunks = self.chunks
            st.fast_sync_stats = False
            self.states.append(st)
        self.P = [max(1, int(round(eps[w].particles_per_cell *
                                   np.count_nonzero(st._mask_host[st.interior] == pk.MASK_ACTIVE))))
                  for w, st in enumerate(self.states)]
        self.max_steps = [maps.max_steps_of(eps[w], st.N) for w, st in enumerate(self.states)]
        st0 = self.states[0]
        # discharge scales in cells of upstream rain are computed per window on the device
        cells = float(eps[0].disc_saturation_cells) > 0.0 or float(eps[0].momentum_saturation_cells) > 0.0
        disc_sat, mom_scale = (float(eps[0].disc_saturation), 1.0) if cells else \
            maps.discharge_scales(eps[0], st0, 1.0, self.P[0], n_rain=1)
        # a window's step (maps.py:1516-1617 with spherical False): the routing flood and the window's lakes,
        # the particles, mass wasting, uplift -- isostasy, the sea mask, lake balance and glacial carving are
        # globe passes a window never runs
        # prm.max_steps: the largest lifetime (each window's own comes from pg_batch_setup), the tape's bound
        self.prm = maps.trace_params(eps[0], st0, 1.0, 1, 1, disc_sat, mom_scale, max(self.max_steps), 0.0)
        self.ep = eps[0]
        self.fields = (FastFields * self.W)(*[self._fields(st) for st in self.states])
        self.NE_arr = (ctypes.c_int64 * self.W)(*NE)
        self.P_arr = (ctypes.c_int64 * self.W)(*self.P)
        self.steps_arr = (ctypes.c_int64 * self.W)(*self.max_steps)
        nbytes = int(lib().pg_batch_workspace_bytes(self.W, self.NE_arr, self.P_arr))
        if nbytes < 0:
            check(-1, "pg_batch_workspace_bytes")
        self.work = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        self.stats = torch.zeros(FAST_NSTATS, dtype=torch.int64, device=dev)
        self.win_stats = torch.zeros((self.W, FAST_NSTATS), dtype=torch.int64, device=dev)
        self.grid_blocks = 0
        self.variant = 0
        self.time_trace = False
        self.rng_dev = None
        self.use_tape = True  # the census records every move and the applying pass replays it (same results)
        self.brush_tile = 0  # the brush stencil's LDS tile (same results)
        self.tape = None
        self.iteration = 0
        # the slope limit's per-cell cap over the arena: every window's side buffer is its view
        self.ecap = None
        if float(getattr(self.ep, "slope_limit_erode", 0.0)) > 0.0:
            self.ecap = torch.empty_like(self.height)
            for w, st in enumerate(self.states):
                st._side["ecap"] = self.ecap[int(off[w]):int(off[w + 1])].view(1, NE[w], NE[w])
        check(lib().pg_batch_setup(self.fields, self.W, self._fparams((0,)), self.steps_arr,
                                   ctypes.c_void_p(self.work.data_ptr()), nbytes, pk._stream()), "pg_batch_setup")

    @staticmethod
    def _fields(st) -> FastFields:
        return FastFields(F=1, N=st.N, H=st.H, spherical=0, deposit_on_exit=int(bool(st.deposit_on_exit)),
                          height=st.height.data_ptr(), sediment=st.sediment.data_ptr(),
                          discharge=st.discharge.data_ptr(), momentum=st.momentum.data_ptr(),
                          pending=st.pending.data_ptr(), base=st.base.data_ptr(), precip=st.precip.data_ptr(),
                          evap=st.evap.data_ptr(), hardness=st.hardness.data_ptr(), metric=st.metric.data_ptr(),
                          metric_inv=st.metric_inv.data_ptr(), mask=st.mask.data_ptr())

    def _fparams(self, key):
        ep = self.ep
        out = (FastParams * self.W)()
        for w in range(self.W):
            slo, shi, ilo, ihi = fast.pcg64_state(self.params[w].rng(self.rng_stage, *key))
            st = self.states[w]
            out[w] = FastParams(rng_state_lo=slo, rng_state_hi=shi, rng_inc_lo=ilo, rng_inc_hi=ihi,
                                n_particles=self.P[w], chunks=self.chunks,
                                iter_deposit=maps.slope_limited_deposit(ep, self.unit),
                                iter_erode=cell_units(ep, "iter_erode", self.unit),
                                conc_chunk=float(maps.chunk_size(ep, self.P[w])), fan_slope=float(ep.fan_slope),
                                offshore_writeoff=float(ep.offshore_writeoff), ema=float(ep.ema),
                                min_volume_frac=float(ep.min_volume_frac), height_bits=fast.HEIGHT_BITS,
                                track_bits=fast.TRACK_BITS, grid_blocks=int(self.grid_blocks),
                                time_trace=int(self.time_trace), variant=int(self.variant),
                                rng_dev=None if self.rng_dev is None else self.rng_dev.data_ptr() + 32 * w,
                                route=None if st.route is None else st.route.data_ptr(),
                                lake_flag=None if st.route is None or st.lake_flag is None else st.lake_flag.data_ptr(),
                                disc_saturation_cells=float(ep.disc_saturation_cells),
                                momentum_saturation_cells=float(ep.momentum_saturation_cells),
                                inflow_volume=float(st.inflow_volume),
                                brush_radius=int(getattr(ep, "brush_radius", 0)), brush_tile=int(self.brush_tile))
        return out

    def flood_due(self) -> bool:
        """Whether this iteration refreshes the windows' routing surfaces
        (every ``flood_every``; maps.py:1548-1549)."""
        flood = int(self.ep.flood_every)
        return flood > 0 and (self.iteration % flood == 0 or any(st.route is None for st in self.states))

    def refresh_routes(self) -> None:
        """Every window's routing flood and, with ``window_lakes``, its lakes
        (host-synchronous: outside a captured graph)."""
        for w, st in enumerate(self.states):
            maps.refresh_route_and_lakes(st, self.params[w], due=True)

    def tape_for(self, fp):
        """The batch's move tape for the current grid and variant (allocated
        on first use, grown when a larger one is needed; allocate it before a
        graph capture), or None when ``use_tape`` is off or it would exceed
        fast.TAPE_MAX_BYTES."""
        if not self.use_tape:
            return None
        need = int(lib().pg_batch_tape_bytes(self.fields, fp, self.W, int(self.prm.max_steps), int(self.grid_blocks),
                                             int(self.variant)))
        if need < 0:
            check(-1, "pg_batch_tape_bytes")
        if need > fast.TAPE_MAX_BYTES:
            return None
        if self.tape is None or self.tape.numel() < need:
            self.tape = None
            self.tape = torch.empty(need, dtype=torch.uint8, device=self.height.device)
        return self.tape

    def step(self, iteration_key, refresh: bool = True) -> None:
        """One batched iteration of every window (asynchronous but for the
        routing flood when it is due; ``refresh=False`` leaves the flood to
        the caller, as a captured graph does)."""
        key = iteration_key if isinstance(iteration_key, (tuple, list)) else (int(iteration_key),)
        if refresh and self.flood_due():
            self.refresh_routes()
        fp = self._fparams(key)
        if self.ecap is not None:
            for st in self.states:
                maps.slope_limit_cap(st, self.ep)
            fp[0].erode_cap = self.ecap.data_ptr()
        tape = self.tape_for(fp)
        if tape is not None:
            fp[0].tape = tape.data_ptr()
            fp[0].tape_bytes = tape.numel()
        check(lib().pg_batch_iterate(self.fields, self.W, ctypes.byref(self.prm), fp,
                                     ctypes.c_void_p(self.work.data_ptr()), self.work.numel(),
                                     ctypes.c_void_p(self.stats.data_ptr()), ctypes.c_void_p(self.win_stats.data_ptr()),
                                     pk._stream()), "pg_batch_iterate")
        for w, st in enumerate(self.states):
            maps.thermal_erosion(st, self.params[w])  # nothing at rate 0
        pk.apply_uplift(self.height, self.uplift, self.mask, None, 0.0)
        self.iteration += 1

    def step_per_window(self, iteration_key) -> None:
        """The same iteration as one fast-mode pass per window, one after
        another (the reference's one-tile-at-a-time form, on the GPU).  Each
        window's persistent grid is sized to its chunk (1.5 lanes per droplet
        of a chunk, at most the census's natural occupancy): a small window
        does not launch a full-chip grid that would idle."""
        key = iteration_key if isinstance(iteration_key, (tuple, list)) else (int(iteration_key),)
        cus = torch.cuda.get_device_properties(0).multi_processor_count
        if self.flood_due():
            self.refresh_routes()
        for w, st in enumerate(self.states):
            per_chunk = self.P[w] / max(self.chunks, 1)
            st.fast_grid_blocks = int(min(4 * cus, max(cus // 4, math.ceil(1.5 * per_chunk / 256))))
            st.fast_variant = 0
            st.fast_time_trace = self.time_trace
            fast.run_iteration(st, self.params[w], key, chunks=self.chunks, sync_stats=False)
            maps.thermal_erosion(st, self.params[w])
            maps.apply_uplift(st)
        self.iteration += 1

    def moves(self, per_window: bool = False) -> int:
        """Droplet moves of the last iteration (synchronises)."""
        if per_window:
            return int(sum(int(st._fast.stats[FS["MOVES"]]) for st in self.states))
        return int(self.stats[FS["MOVES"]])


def doe_val(win: dict) -> bool:
    return bool(win.get("deposit_on_exit", False))


class BatchGraph:
    """The batched iteration captured once as a HIP graph; each replay reads
    its windows' PCG64 states from a device buffer filled from the host."""

    def __init__(self, batch: WindowBatch):
        self.batch = batch
        dev = batch.height.device
        batch.rng_dev = torch.zeros((batch.W, 4), dtype=torch.int64, device=dev)
        batch.time_trace = False
        if int(batch.ep.flood_every) > 0 and batch.flood_due():
            batch.refresh_routes()  # the routes and lake flags exist before the capture
        batch.tape_for(batch._fparams((0,)))  # allocated before the capture
        for w, st in enumerate(batch.states):  # side buffers too
            maps.prepare_side_passes(st, batch.params[w])
        torch.cuda.synchronize()
        self._load((0,))
        it0 = batch.iteration
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            batch.step((0,), refresh=False)
        batch.iteration = it0

    def _load(self, key):
        b = self.batch
        words = [[_signed(x) for x in fast.pcg64_state(b.params[w].rng(b.rng_stage, *key))] for w in range(b.W)]
        b.rng_dev.copy_(torch.tensor(words, dtype=torch.int64), non_blocking=False)

    def step(self, iteration_key) -> None:
        key = iteration_key if isinstance(iteration_key, (tuple, list)) else (int(iteration_key),)
        b = self.batch
        if b.flood_due():
            b.refresh_routes()  # in place into the buffers the graph reads
        self._load(key)
        self.graph.replay()
        b.iteration += 1


class Autotuner:
    """Picks the persistent trace grid and the trace kernel's register-budget
    variant (0: the kernels' natural budget, 1 / 2: one / two more waves per
    SIMD) of a batch or of a fast-mode state, and with a brush its stencil's
    LDS tile (64 x 4, 32 x 8, 16 x 16), from real iterations: the first
    len(candidates) iterations each run one (grid, variant) candidate with
    the trace launches timed (HIP events on the launch stream) and
    normalised by the iteration's droplet moves (the iterations differ);
    with a brush the next len(tiles) iterations each run one tile at the
    fastest (grid, variant), timed by the brush launches.  Results are
    independent of all three (fixed-point accumulation), so no iteration is
    wasted.  ``best`` is set once every candidate has run."""

    def __init__(self, target, candidates=None, tiles=None):
        self.target = target
        if candidates is None:
            cus = torch.cuda.get_device_properties(0).multi_processor_count
            candidates = [(g * cus, v) for v in (0, 1, 2) for g in (4 + v, 2 * (4 + v))]
        self.candidates = [(int(c[0]), int(c[1])) if isinstance(c, (tuple, list)) else (int(c), 0)
                           for c in candidates]
        self.candidates = [c for c in self.candidates if c[0] > 0 and c[1] in (0, 1, 2)]
        if tiles is None:
            brushed = isinstance(target, WindowBatch) and int(getattr(target.ep, "brush_radius", 0)) > 1
            tiles = (0, 1, 2) if brushed else ()
        self.tiles = [int(t) for t in tiles if int(t) in (0, 1, 2)]
        self.times: dict = {}
        self.tile_times: dict = {}
        self.best = None
        self.best_tile = None
        self._gv = None

    def _set(self, cand, timed: bool):
        grid, variant = cand
        t = self.target
        if isinstance(t, WindowBatch):
            t.grid_blocks, t.variant, t.time_trace = grid, variant, timed
        else:
            t.fast_grid_blocks, t.fast_variant, t.fast_time_trace = grid, variant, timed

    def _set_tile(self, tile):
        t = self.target
        if isinstance(t, WindowBatch):
            t.brush_tile = tile
        else:
            t.fast_brush_tile = tile

    def step(self, run, iteration_key) -> None:
        """run(iteration_key) performs one iteration of the target."""
        if self.best is not None:
            run(iteration_key)
            return
        if self._gv is None:
            cand = self.candidates[len(self.times)]
            fast.trace_time()
            self._set(cand, True)
            run(iteration_key)
            ms, _ = fast.trace_time()
            fast.brush_time()
            t = self.target
            moves = t.moves() if isinstance(t, WindowBatch) else int(t._fast.stats[FS["MOVES"]])
            self.times[cand] = ms / max(moves, 1) * 1e6  # ns per move
            if len(self.times) < len(self.candidates):
                self._set((0, 0), False)
                return
            self._gv = min(self.times, key=self.times.get)
            if not self.tiles:
                self.finish()
            else:
                self._set(self._gv, False)
            return
        tile = self.tiles[len(self.tile_times)]
        fast.brush_time()
        self._set(self._gv, True)
        self._set_tile(tile)
        run(iteration_key)
        ms, n = fast.brush_time()
        fast.trace_time()
        self.tile_times[tile] = ms / max(n, 1)  # ms per brush launch
        if len(self.tile_times) == len(self.tiles):
            self.finish()
        else:
            self._set(self._gv, False)

    def finish(self):
        """Keep the fastest candidates measured so far."""
        if self.times:
            self.best = self._gv if self._gv is not None else min(self.times, key=self.times.get)
            self._set(self.best, False)
        if self.tile_times:
            self.best_tile = min(self.tile_times, key=self.tile_times.get)
            self._set_tile(self.best_tile)
        return self.best
