"""Differentiable erosion: gradients of a terrain misfit with respect to a
per-cell erodibility field and six scalar erosion parameters.

The reference has no adjoint; this module is the differentiable counterpart
of its erosion loop (bake/globe/erosion/maps.py:1516-1617, plain profile).
An iteration in the differentiable mode (``pg_grad_iterate``) is the fast
mode with one frozen-terrain chunk and only the trace-time caps: droplets
steer on the packed record of the state they start from, exchange mass with
float32 or float64 arithmetic, and everything they take or lay down is
summed in int64 fixed point; then uplift, the datum hold and the halo
exchange follow as in ``maps.step``.

Trajectories are frozen: the cells a droplet visits, where it stops and
why, the discharge it reads and the datum's order statistic are held at
the nominal run's; the surface, the erodibility factor and the six scalars
act through the exchange.  ``GradRun.replay`` evaluates exactly that
frozen map at other parameters (finite differences use it), and
``GradRun.backward`` is its adjoint: per iteration, in reverse, the halo
exchange's transpose, the datum's, and the hand-written reverse kernel
(``pg_grad_reverse``) over the iteration's tape.  Iterations are
checkpointed every ``checkpoint_every``; a segment's tapes are recomputed
from its checkpoint when the backward pass reaches it (the last segment's
are kept from the forward pass).

The six scalars, in kernel units (``THETA_NAMES``): erodibility,
deposition_rate, evap_rate, k_disc, ocean_rate, dep_floor (cell units).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .._lib import FAST_NSTATS, FastFields, FastParams, GradArgs, check, lib
from . import particle as pk

THETA_NAMES = ("erodibility", "deposition_rate", "evap_rate", "k_disc", "ocean_rate", "dep_floor")
#: fixed point of the surface change (units of 2^-bits cells): float64 runs
#: resolve 16x finer than float32 ones (finite differences of the frozen map
#: see the quantisation as noise); range +-2^(63-bits) cells per iteration
HEIGHT_BITS = {"f32": 48, "f64": 52}
TRACK_BITS = 32

_SNAP = ("height", "sediment", "discharge", "momentum", "pending", "base")


def _snapshot(src) -> dict:
    if isinstance(src, dict):
        return {n: src[n].clone() for n in _SNAP}
    return {n: getattr(src, n).clone() for n in _SNAP}


class GradRun:
    """One differentiable erosion run of ``iterations`` iterations from the
    current contents of ``state`` (a fast-mode ErosionState, which it uses
    for the static fields and as scratch).

    real: "f32" or "f64" exchange arithmetic.  replay_of: another GradRun
    run with ``record=True`` -- this run then follows its trajectories
    (steering records, move counts, datum cells) instead of its own.

    feedback: True differentiates the exchange's dependence on the surface
    the droplets read (h0, h1, their cell's and the previous cell's
    height) -- the exact derivative of the frozen-trajectory map, whose
    sensitivity grows with the horizon (every droplet of the frozen chunk
    answers the same surface change).  False holds that surface at the
    nominal run's (a replay then reads the nominal surface too), so the
    surface adjoint passes each particle pass unchanged: the derivative of
    the frozen-terrain map, stable at any horizon.

    tape_batch: record each iteration's tape in pieces of this many droplets
    (None: one tape of the whole spawn; "auto": pieces sized at the backward
    pass to the device memory then free).  The droplets of an iteration do
    not interact, so the backward pass then re-traces the iteration one
    piece at a time from the iteration's starting state and reverses each
    piece into the same fixed-point accumulators: the gradients are the
    whole tape's bit for bit, at one piece's memory (what the N_c=8192
    planet needs).  The forward pass records no tapes; the backward pass
    keeps the starting states of one checkpoint segment.

    store: where the stored states live -- the starting state, the
    checkpoints, a segment's starting states at the backward pass and a
    recorded run's nominal states: "device", or "host" (pinned host memory,
    copied to the device when read; the gradients are the same bits).  At
    N_c=8192 a state is 22.6 GB: "host" lets a run keep checkpoints every
    few iterations beyond what device memory holds."""

    def __init__(self, state, params, iterations: int, real: str = "f32", checkpoint_every: int = 10,
                 first_key: int = 0, rng_stage: str = "erosion", n_particles=None, replay_of: "GradRun | None" = None,
                 feedback: bool = True, tape_batch: "int | None" = None, store: str = "device"):
        from .maps import (_check_scope, _eparams, discharge_scales, max_steps_of, require_plain_side_passes,
                           trace_params, uplift_cap)

        if state.mode != "fast":
            raise ValueError("the differentiable mode runs on a fast-mode state")
        if real not in ("f32", "f64"):
            raise ValueError("real must be 'f32' or 'f64'")
        ep = _eparams(params)
        _check_scope(ep, state)
        require_plain_side_passes(ep, "the differentiable mode")
        if float(getattr(ep, "lateral_rate", 0.0)) > 0.0 or state.roots is not None:
            raise NotImplementedError("lateral cuts and roots are outside the differentiable mode")
        if state.route is not None:
            raise NotImplementedError("the routing surface is outside the differentiable mode (it steers on the terrain)")
        if float(ep.disc_saturation_cells) > 0.0 or float(ep.momentum_saturation_cells) > 0.0:
            raise NotImplementedError("cell-count discharge scales are outside the differentiable mode")
        if float(state.pending.abs().max().item()) != 0.0:
            raise ValueError("the differentiable mode starts without stockpiles (state.pending must be 0)")
        self.state, self.params, self.ep = state, params, ep
        self.T = int(iterations)
        self.ce = max(1, int(checkpoint_every))
        self.real = real
        self.dtype = torch.float32 if real == "f32" else torch.float64
        self.keys = [int(first_key) + t for t in range(self.T)]
        self.rng_stage = rng_stage
        self.replay_of = replay_of
        self.feedback = bool(feedback)
        if store not in ("device", "host"):
            raise ValueError("store must be 'device' or 'host'")
        self.store = store
        if int(getattr(ep, "brush_radius", 0)) > 1:
            raise NotImplementedError("the differentiable mode has no brush")
        dev = state.device
        n_cells = int(np.count_nonzero(state._mask_host[state.interior] == pk.MASK_ACTIVE))
        if n_particles is None:
            n_particles = max(1, int(round(ep.particles_per_cell * n_cells)))
        self.P = int(n_particles)
        self.max_steps = max_steps_of(ep, state.N)
        disc_sat, mom_scale = discharge_scales(ep, state, 1.0, self.P, n_rain=1)
        self.prm = trace_params(ep, state, 1.0, 1, 1, disc_sat, mom_scale, self.max_steps, 0.0)
        self.cap = uplift_cap(state, ep)
        self.uplift_mean = state.uplift_mean(self.cap)
        self.datum = bool(state.spherical and hasattr(params, "world"))
        if self.datum:
            from .maps import held_land_fraction

            M = state.F * state.N * state.N
            n_land = min(max(int(round(float(held_land_fraction(state, params)) * M)), 0), M)
            if n_land in (0, M):
                raise ValueError("the differentiable datum hold needs 0 < land fraction < 1")
            self.kth = M - n_land
            inter = torch.zeros(state.height.shape, dtype=torch.bool, device=dev)
            inter[state.interior] = True
            self.interior_flat = inter.reshape(-1)
            state.select_work()
        if state.spherical:
            self.halo = state.halo_tables()
        nbytes = int(lib().pg_fast_workspace_bytes(state.F, state.NE, self.P))
        if nbytes < 0:
            check(-1, "pg_fast_workspace_bytes")
        self.work = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        self.stats = torch.zeros(FAST_NSTATS, dtype=torch.int64, device=dev)
        self.height_bits = HEIGHT_BITS[real]
        fp = FastParams(n_particles=self.P, chunks=1, height_bits=self.height_bits, track_bits=TRACK_BITS)
        check(lib().pg_fast_setup(ctypes.byref(self._fields(state)), ctypes.byref(fp), ctypes.c_void_p(self.work.data_ptr()),
                                  nbytes, pk._stream()), "pg_fast_setup")
        self._tapes: dict[int, torch.Tensor] = {}
        self._auto_batch = tape_batch == "auto"
        self.set_tape_batch(None if self._auto_batch else tape_batch)
        self._taped: dict[int, int] = {}  # slot -> iteration whose tape it holds
        self.s0 = self._keep(state)
        self.nominal = None
        self.moves = None
        self.idx: list = [None] * self.T
        self.checkpoints: dict[int, dict] = {}

    # -- stored states ---------------------------------------------------------
    def _keep(self, src) -> dict:
        """A stored copy of a state (device clone, or pinned host copy)."""
        get = (lambda n: src[n]) if isinstance(src, dict) else (lambda n: getattr(src, n))
        if self.store == "device":
            return {n: get(n).clone() for n in _SNAP}
        out = {}
        for n in _SNAP:
            v = get(n)
            out[n] = torch.empty(v.shape, dtype=v.dtype, pin_memory=True)
            out[n].copy_(v)
        return out

    def _fetch(self, kept: dict) -> dict:
        """A stored state for reading on the device (the stored tensors
        themselves when they are on the device)."""
        if self.store == "device":
            return kept
        dev = self.state.device
        return {n: v.to(dev, non_blocking=True) for n, v in kept.items()}

    def _restore(self, kept: dict) -> dict:
        """A new device copy of a stored state, to run on."""
        return _snapshot(kept) if self.store == "device" else self._fetch(kept)

    def _tape_bytes_for(self, droplets: int) -> int:
        tb = ctypes.c_int64()
        check(lib().pg_grad_tape_bytes(int(droplets), self.max_steps, 1 if self.real == "f64" else 0, ctypes.byref(tb)),
              "pg_grad_tape_bytes")
        return int(tb.value)

    def set_tape_batch(self, tape_batch) -> None:
        """Droplets per tape piece (None: the whole spawn); drops the tapes held."""
        if tape_batch is not None and int(tape_batch) < 1:
            raise ValueError("tape_batch must be >= 1 droplet")
        self.tape_batch = None if tape_batch is None or int(tape_batch) >= self.P else int(tape_batch)
        self.tape_bytes = self._tape_bytes_for(self.tape_batch or self.P)
        self._tapes = {}
        self._taped = {}

    def _size_auto_batch(self) -> None:
        """tape_batch="auto": one piece as large as the device's free memory
        allows once the backward pass's other buffers are set aside."""
        st = self.state
        # free on the device plus what torch's cache holds unused
        free = int(torch.cuda.mem_get_info(st.device)[0]) + int(torch.cuda.memory_reserved(st.device)
                                                                 - torch.cuda.memory_allocated(st.device))
        n = st.height.numel()
        snap = sum(t.numel() * t.element_size() for t in self.s0.values())
        held = (self.ce - 1) * snap if self.store == "device" else snap  # the segment's states on the device
        reserve = int(lib().pg_grad_reverse_work_bytes(st.F, st.N, st.H)) + 6 * 8 * n + held + (2 << 30)
        per = self._tape_bytes_for(2) - self._tape_bytes_for(1)
        fit = (free - reserve) // max(per, 1)
        if fit < 1:
            raise MemoryError(f"no room for a tape piece: {free / 1e9:.1f} GB free, {reserve / 1e9:.1f} GB set aside")
        self.set_tape_batch(min(fit, self.P))

    # -- plumbing ------------------------------------------------------------
    def _fields(self, s) -> FastFields:
        st = self.state
        get = (lambda n: s[n]) if isinstance(s, dict) else (lambda n: getattr(s, n))
        return FastFields(F=st.F, N=st.N, H=st.H, spherical=int(bool(st.spherical)),
                          deposit_on_exit=int(bool(st.deposit_on_exit)), height=get("height").data_ptr(),
                          sediment=get("sediment").data_ptr(), discharge=get("discharge").data_ptr(),
                          momentum=get("momentum").data_ptr(), pending=get("pending").data_ptr(),
                          base=get("base").data_ptr(), precip=st.precip.data_ptr(), evap=st.evap.data_ptr(),
                          hardness=st.hardness.data_ptr(), metric=st.metric.data_ptr(),
                          metric_inv=st.metric_inv.data_ptr(), mask=st.mask.data_ptr())

    def _fparams(self, t: int) -> FastParams:
        from .fast import pcg64_state

        ep = self.ep
        slo, shi, ilo, ihi = pcg64_state(self.params.rng(self.rng_stage, self.keys[t]))
        return FastParams(rng_state_lo=slo, rng_state_hi=shi, rng_inc_lo=ilo, rng_inc_hi=ihi, n_particles=self.P,
                          chunks=1, conc_chunk=0.0, fan_slope=float(ep.fan_slope),
                          offshore_writeoff=float(ep.offshore_writeoff), ema=float(ep.ema),
                          min_volume_frac=float(ep.min_volume_frac), height_bits=self.height_bits,
                          track_bits=TRACK_BITS,
                          grid_blocks=int(getattr(self.state, "fast_grid_blocks", 0)))

    def _tape(self, slot: int) -> torch.Tensor:
        if slot not in self._tapes:
            self._tapes[slot] = torch.empty(self.tape_bytes, dtype=torch.uint8, device=self.state.device)
        return self._tapes[slot]

    def _args(self, surf, erod, theta, tape=None, fixed=None, moves_out=None) -> GradArgs:
        th = [float(v) for v in (theta.tolist() if torch.is_tensor(theta) else theta)]
        if len(th) != 6:
            raise ValueError("theta holds 6 values: " + ", ".join(THETA_NAMES))
        if erod is not None:
            pk._need(erod, torch.float32, tuple(self.state.height.shape), "erodibility field")
        return GradArgs(real=1 if self.real == "f64" else 0, feedback=1 if self.feedback else 0, surf=surf.data_ptr(),
                        erod=None if erod is None else erod.data_ptr(), theta=(ctypes.c_double * 6)(*th),
                        tape=None if tape is None else tape.data_ptr(), tape_bytes=0 if tape is None else tape.numel(),
                        fixed_moves=None if fixed is None else fixed.data_ptr(),
                        moves_out=None if moves_out is None else moves_out.data_ptr())

    def _steer(self, t: int, target: dict):
        """(steering state, fixed moves, datum cell) of iteration t."""
        if self.replay_of is None:
            return target, None, self.idx[t]
        r = self.replay_of
        return r._fetch(r.nominal[t]), r.moves[t], r.idx[t]

    def _iterate(self, target: dict, t: int, erod, theta, tape=None, moves_out=None, steer=None, fixed=None,
                 idx=None):
        st = self.state
        steer = target if steer is None else steer
        src = target if self.feedback else steer
        surf = (src["height"] + src["sediment"]).to(self.dtype)
        ga = self._args(surf, erod, theta, tape, fixed, moves_out)
        check(lib().pg_grad_iterate(ctypes.byref(self._fields(steer)), ctypes.byref(self._fields(target)),
                                    ctypes.byref(self.prm), ctypes.byref(self._fparams(t)), ctypes.byref(ga),
                                    ctypes.c_void_p(self.work.data_ptr()), self.work.numel(),
                                    ctypes.c_void_p(self.stats.data_ptr()), pk._stream()), "pg_grad_iterate")
        pk.apply_uplift(target["height"], st.uplift, st.mask, self.cap, self.uplift_mean)
        if self.datum:
            x = (target["height"] + target["sediment"]).reshape(-1)
            if idx is None:
                q = torch.zeros(1, dtype=torch.float64, device=st.device)
                check(lib().pg_select_kth_surface(pk._ptr(target["height"]), pk._ptr(target["sediment"]), st.F, st.N,
                                                  st.H, self.kth, pk._ptr(q), pk._ptr(st.select_work()), pk._stream()),
                      "hold_datum")
                idx = torch.argmax(((x == q) & self.interior_flat).to(torch.uint8)).reshape(1)
            q = x.index_select(0, idx).contiguous()
            check(lib().pg_datum_shift(target["height"].numel(), pk._ptr(target["height"]), None,
                                       pk._ptr(target["base"]), pk._ptr(q), pk._stream()), "datum_shift")
        if st.spherical:
            for n in ("height", "sediment", "discharge"):
                self.halo.exchange(target[n])
            self.halo.exchange(target["momentum"], vector=True)
        return idx

    # -- forward / replay / backward -----------------------------------------
    def forward(self, erod, theta, record: bool = False, tapes: str = "last") -> dict:
        """Run all iterations; returns the final state's tensors.  Keeps
        checkpoints and the datum cells; with record=True also every
        iteration's starting state and move counts (what a replay follows);
        tapes: "last" (the last segment's, for the backward pass), "all" or
        "none"."""
        T, ce = self.T, self.ce
        target = self._restore(self.s0)
        self.checkpoints = {}
        self._taped = {}
        if record:
            self.nominal, self.moves = [], []
        last = ((T - 1) // ce) * ce
        if self.tape_batch is not None or self._auto_batch:
            tapes = "none"  # the backward pass re-traces piece by piece
        for t in range(T):
            if t % ce == 0:
                # the first checkpoint is the starting state itself (nothing writes either)
                self.checkpoints[t] = self.s0 if t == 0 else self._keep(target)
            mv = None
            if record:
                # a nominal state is read only: the first is the starting state itself
                self.nominal.append(self.s0 if t == 0 else self._keep(target))
                mv = torch.zeros(self.P, dtype=torch.int32, device=self.state.device)
                self.moves.append(mv)
            tape = None
            if tapes == "all" or (tapes == "last" and t >= last):
                slot = t if tapes == "all" else t - last
                tape = self._tape(slot)
                self._taped[slot] = t
            steer, fixed, idx = self._steer(t, target)
            if self.replay_of is None:
                idx = None  # a nominal run takes its own order statistic
            self.idx[t] = self._iterate(target, t, erod, theta, tape=tape, moves_out=mv, steer=steer, fixed=fixed,
                                        idx=idx)
        self._tapes_all = tapes == "all"
        return target

    def replay(self, erod, theta) -> dict:
        """The frozen-trajectory map at (erod, theta): this run's recorded
        trajectories (or replay_of's) with the exchange re-evaluated."""
        src = self.replay_of if self.replay_of is not None else self
        if src.nominal is None:
            raise ValueError("replay needs a forward pass with record=True")
        target = self._restore(self.s0)
        for t in range(self.T):
            self._iterate(target, t, erod, theta, steer=src._fetch(src.nominal[t]), fixed=src.moves[t], idx=src.idx[t])
        return target

    def backward(self, g_surface: torch.Tensor, erod, theta):
        """Adjoint of forward(): g_surface = dL/d(final surface) (F, NE, NE)
        float64.  Returns (dL/d erod (float64, extended cells), dL/d theta
        (float64[6]), dL/d initial surface)."""
        st = self.state
        T, ce = self.T, self.ce
        lam = g_surface.to(torch.float64).contiguous().clone()
        g_erod = torch.zeros(st.height.shape, dtype=torch.float64, device=st.device)
        g_theta = torch.zeros(6, dtype=torch.float64, device=st.device)
        segs = [(a, min(a + ce, T)) for a in range(0, T, ce)]
        if self._auto_batch:
            self._tapes = {}
            self._size_auto_batch()
        if self.tape_batch is not None or self._auto_batch:
            return self._backward_pieces(lam, segs, erod, theta, g_erod, g_theta)
        for a, b in reversed(segs):
            slot_of = (lambda t: t) if self._tapes_all else (lambda t: t - a)
            if not all(self._taped.get(slot_of(t)) == t for t in range(a, b)):
                target = self._restore(self.checkpoints[a])
                for t in range(a, b):
                    steer, fixed, _ = self._steer(t, target)
                    self._iterate(target, t, erod, theta, tape=self._tape(slot_of(t)), steer=steer, fixed=fixed,
                                  idx=self.idx[t])
                    self._taped[slot_of(t)] = t
            for t in range(b - 1, a - 1, -1):
                lam = self._reverse(lam, self._tape(slot_of(t)), self.idx[t], erod, theta, g_erod, g_theta)
        return g_erod, g_theta, lam

    def _backward_pieces(self, lam, segs, erod, theta, g_erod, g_theta):
        """backward() with the tape in pieces of tape_batch droplets."""
        step = self.tape_batch or self.P
        pieces = [(lo, min(lo + step, self.P)) for lo in range(0, self.P, step)]
        tape = self._tape(0)
        for a, b in reversed(segs):
            # the segment's starting states: the checkpoint itself (read only), then copies
            starts = [self.checkpoints[a]]
            if b - a > 1:
                target = self._restore(self.checkpoints[a])
                for t in range(a, b - 1):
                    steer, fixed, _ = self._steer(t, target)
                    self._iterate(target, t, erod, theta, steer=steer, fixed=fixed, idx=self.idx[t])
                    starts.append(self._keep(target))
                del target
            for t in range(b - 1, a - 1, -1):
                start = self._fetch(starts[t - a])
                steer, fixed, _ = self._steer(t, start)
                src = start if self.feedback else steer
                surf = (src["height"] + src["sediment"]).to(self.dtype)

                def record(k, t=t, steer=steer, fixed=fixed, surf=surf):
                    # piece k of iteration t's tape: a range records the tape only and writes no target
                    ga = self._args(surf, erod, theta, tape, fixed)
                    ga.d_lo, ga.d_hi = pieces[k]
                    ga.reuse_spawn = 1 if k else 0
                    check(lib().pg_grad_iterate(ctypes.byref(self._fields(steer)), ctypes.byref(self._fields(steer)),
                                                ctypes.byref(self.prm), ctypes.byref(self._fparams(t)),
                                                ctypes.byref(ga), ctypes.c_void_p(self.work.data_ptr()),
                                                self.work.numel(), ctypes.c_void_p(self.stats.data_ptr()),
                                                pk._stream()), "pg_grad_iterate")

                lam = self._reverse(lam, tape, self.idx[t], erod, theta, g_erod, g_theta, record=record,
                                    n_pieces=len(pieces))
                del starts[t - a:]
        return g_erod, g_theta, lam

    def _reverse(self, lam, tape, idx, erod, theta, g_erod, g_theta, record=None, n_pieces: int = 1):
        """One iteration's reverse sweep.  With record: the tape in n_pieces
        pieces, record(k) writing piece k; the pieces accumulate in the
        fixed-point sums and are flushed once, after the last."""
        st = self.state
        if st.spherical:
            ht = self.halo
            ht.adjoint(lam)
        if self.datum:
            total = lam.sum().reshape(1)
            lam.view(-1).index_add_(0, idx, -total)
        out = lam.clone()
        surf = st.height  # unused by the reverse sweep; any device pointer of the right size
        ga = self._args(surf, erod, theta, tape=tape)
        if getattr(self, "_rev_work", None) is None:
            nbytes = int(lib().pg_grad_reverse_work_bytes(st.F, st.N, st.H))
            if nbytes < 0:
                check(-1, "pg_grad_reverse_work_bytes")
            self._rev_work = torch.empty(nbytes, dtype=torch.uint8, device=st.device)
        ga.rev_work = self._rev_work.data_ptr()
        ga.rev_work_bytes = self._rev_work.numel()
        for k in range(n_pieces):
            if record is not None:
                record(k)
            # bit 0: keep the earlier pieces' sums; bit 1: more pieces follow, flush after the last
            ga.rev_flags = (1 if k else 0) | (2 if k < n_pieces - 1 else 0)
            check(lib().pg_grad_reverse(st.F, st.N, st.H, ctypes.byref(self.prm), ctypes.byref(ga), pk._ptr(lam),
                                        pk._ptr(out), pk._ptr(g_erod), pk._ptr(g_theta), pk._stream()),
                  "pg_grad_reverse")
        return out


def theta_nominal(run: GradRun) -> torch.Tensor:
    """The six scalars of the run's parameters, kernel units (CPU float64)."""
    p = run.prm
    return torch.tensor([p.erodibility, p.deposition_rate, p.evap_rate, p.k_disc, p.ocean_rate, p.dep_floor],
                        dtype=torch.float64)


class ErosionFunction(torch.autograd.Function):
    """surface_T = erosion(erod, theta) with the hand-written adjoint.

    erod: (F, NE, NE) float32 on the run's device; theta: float64[6] on the
    CPU.  Returns the final surface (height + sediment, cell units, float64,
    extended cells)."""

    @staticmethod
    def forward(ctx, erod, theta, run: GradRun, tapes: str = "last"):
        out = run.forward(erod.detach().contiguous(), theta.detach(), tapes=tapes)
        ctx.run = run
        ctx.save_for_backward(erod, theta)
        return out["height"] + out["sediment"]

    @staticmethod
    def backward(ctx, g):
        erod, theta = ctx.saved_tensors
        g_erod, g_theta, _ = ctx.run.backward(g, erod.detach().contiguous(), theta.detach())
        return g_erod.to(erod.dtype), g_theta.to(theta.device), None, None


def erosion(erod, theta, run: GradRun, tapes: str = "last"):
    return ErosionFunction.apply(erod, theta, run, tapes)


def beta25_cdf(x: torch.Tensor) -> torch.Tensor:
    """CDF of Beta(2, 5) on [0, 1]: 1 - (1-x)^6 - 6 x (1-x)^5."""
    x = x.clamp(0.0, 1.0)
    return 1.0 - (1.0 - x) ** 6 - 6.0 * x * (1.0 - x) ** 5


class _SmoothCDF(torch.autograd.Function):
    """cdf_k = mean_i sigmoid((z_k - h_i) / tau), evaluated over the cells in
    slices (a levels x cells matrix of the planet's 121 M land cells would
    take 62 GB); the backward pass recomputes the slices."""

    SLICE = 1 << 22

    @staticmethod
    def forward(ctx, h, z, tau):
        cdf = torch.zeros_like(z)
        for a in range(0, h.numel(), _SmoothCDF.SLICE):
            cdf += torch.sigmoid((z[:, None] - h[None, a:a + _SmoothCDF.SLICE]) / tau).sum(dim=1)
        ctx.save_for_backward(h, z)
        ctx.tau = tau
        return cdf / max(h.numel(), 1)

    @staticmethod
    def backward(ctx, g):
        h, z = ctx.saved_tensors
        tau, n = ctx.tau, max(h.numel(), 1)
        gh = torch.empty_like(h)
        for a in range(0, h.numel(), _SmoothCDF.SLICE):
            sg = torch.sigmoid((z[:, None] - h[None, a:a + _SmoothCDF.SLICE]) / tau)
            gh[a:a + _SmoothCDF.SLICE] = -(g[:, None] * sg * (1.0 - sg)).sum(dim=0) / (tau * n)
        return gh, None, None


def land_hypsometry_loss(h_m: torch.Tensor, land: torch.Tensor, bins: int = 64, zmax_m: float = 4000.0,
                         mean_m: float = 800.0, weight: float = 0.1, tau_m: float | None = None):
    """L2 misfit of the land hypsometric CDF (``bins`` levels z_k = k zmax /
    bins, k = 1..bins) to Beta(2, 5) on [0, zmax] + weight * (mean land
    elevation - mean_m)^2; h_m in metres, land a boolean mask.  The CDF is
    the sigmoid-smoothed fraction of land cells at or below each level
    (width tau_m, default half a bin)."""
    hl = h_m[land]
    z = torch.arange(1, bins + 1, dtype=h_m.dtype, device=h_m.device) * (zmax_m / bins)
    tau = tau_m if tau_m is not None else 0.5 * zmax_m / bins
    cdf = _SmoothCDF.apply(hl, z, float(tau))
    target = beta25_cdf(z / zmax_m)
    return ((cdf - target) ** 2).sum() + weight * (hl.mean() - mean_m) ** 2


def hypsometry_loss(surface: torch.Tensor, run: GradRun, land=None, **kw):
    """land_hypsometry_loss of a surface (cell units, extended cells) over
    the interior; the land set (surface > 0) is taken from the surface given
    unless passed, and is not differentiated."""
    st = run.state
    h = surface[st.interior].reshape(-1) * st.height_unit_m
    if land is None:
        land = (h > 0).detach()
    return land_hypsometry_loss(h, land, **kw)
