"""ORACLE -- test infrastructure only.

A numpy restatement of the reference's erosion iteration
(bake/globe/erosion/maps.py) around the C restatement of its kernels
(oracle.c, loaded from liboracle.so).  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may import this module, and only as the
checker or as the timed CPU baseline; the product never does.

What is restated (each function cites the reference lines it follows):
``OracleState`` (maps.py:56-191 from_grid / window, 309-340 pack + halos),
``spawn_weights`` / ``spawn_particles`` (maps.py:793-833), ``run_iteration``
(maps.py:839-1047: pending re-injection, chunk loop of trace + serial apply,
fold, EMA, stats), ``apply_uplift`` / ``uplift_replay`` / ``start_replay``
(maps.py:1088-1213), ``hold_datum`` (maps.py:1446-1508), the slope limits
(maps.py:923-931 with particle.py:1045-1076), ``thermal_erosion``
(maps.py:1050-1085 with particle.py:1104-1227), ``apply_isostasy`` and its
smoothing step (maps.py:1325-1410 with field.py:179-188, 372-404) and
``step`` (maps.py:1516-1617: particles -> mass wasting -> uplift ->
isostasy -> datum -> halos), the routing flood (route.py:30-117) and a
window's lakes without the evaporation balance (maps.py:629-672).  The
remaining side passes (sea mask, the planet's lake balance, glacial, ice,
strata) raise NotImplementedError here.

One function has no counterpart in the reference: ``brush_spread``, the
fast mode's erosion brush (the reference changes one cell per step,
particle.py:660-667), restated here as a plain scatter through the owner
table and pinned by closed forms in tests/test_oracle_brush.py (a lone
source's footprint is the weight table, Lague's max(0, r - |o|) normalised;
the sum over the grid is preserved; shares that have no target stay with
their source; halo shares land on the owner).

Parity pins: the reference's own invariants (mass conservation on a closed
window, the offshore split, the change-list layout, the stockpile bound,
uplift mean-free, datum held) are ported in tests/test_oracle_*.py.  The
numpy reductions (pairwise mean in apply_uplift, einsum in the halo exchange)
are the same numpy calls the reference makes, so they agree by construction
but are not pinned against the reference's numpy version.
"""
from __future__ import annotations

import ctypes
import math
import os
import time
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

MASK_OUTSIDE, MASK_ACTIVE, MASK_FROZEN = 0, 1, 2
DEATH_NAMES = ("age", "ocean", "exit", "pit", "evap", "stop")
SPREAD, LAKE_PARKS, NS = 8, 8, 16
S_SURF, S_ROUTE, S_DISCH, S_MOMA, S_MOMB, S_EVAP, S_HARD, S_SED, S_G, S_GINV, S_RFLAG, S_BASE = 0, 1, 2, 3, 4, 5, 6, 7, 8, 11, 14, 15
MIN_CHUNKS = 16  # maps.py:783
STAGE_SEED_OFFSETS = {"erosion": 3000, "refine": 6000, "stub": 9000}


class TraceParams(ctypes.Structure):
    """Mirror of ``or_trace_params`` in oracle.c (particle.py:292-318 order)."""

    _fields_ = [(n, ctypes.c_double) for n in (
        "dt", "density", "friction", "deposition_rate", "evap_rate", "k_mom", "mom_scale", "k_disc", "disc_saturation",
        "disc_exponent", "slope_gain", "slope_sat", "erodibility", "cover_depth", "min_volume")] + [
        ("max_steps", ctypes.c_int64), ("max_erode", ctypes.c_double), ("conc_scale", ctypes.c_double),
        ("pit_steps", ctypes.c_int64), ("deposit_on_exit", ctypes.c_int64), ("ocean_rate", ctypes.c_double),
        ("ocean_steps", ctypes.c_int64)] + [(n, ctypes.c_double) for n in ("fan_room", "dep_floor", "lake_trap", "lake_park", "lateral_rate")]


_lib = None


def lib():
    """The compiled oracle (built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: run `make -C {HERE}`")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, dbl, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int
        L.or_trace_particles.argtypes = [i64, vp, vp, vp, vp, vp, dbl, vp, i32, vp, i64, i64, i32, ctypes.POINTER(TraceParams),
                                         vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, i32]
        L.or_apply_changes.argtypes = [vp, vp, vp, vp, vp, i64, i64, vp, vp, vp, vp, vp, vp, vp, vp, dbl, dbl, dbl, i32, dbl,
                                       dbl, vp, i32, vp, vp, vp, vp]
        L.or_pack_samples.argtypes = [vp, i64, i64, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp]
        L.or_fold_changes.argtypes = [i64, vp, vp, vp, vp]
        L.or_ema_update.argtypes = [i64, vp, vp, vp, vp, vp, dbl]
        L.or_spawn_cells.argtypes = [vp, i64, vp, i64, vp]
        L.or_set_libm.argtypes = [vp, vp]
        L.or_set_sincos.argtypes = [vp]
        L.or_change_list_cap.argtypes = [i64, i32]
        L.or_change_list_cap.restype = i64
        d = ctypes.POINTER(ctypes.c_double)
        L.or_to_sphere.argtypes = [i32, dbl, dbl, d, d, d]
        L.or_from_sphere.argtypes = [dbl, dbl, dbl, ctypes.POINTER(ctypes.c_int), d, d]
        L.or_jacobian.argtypes = [i32, dbl, dbl, d]
        L.or_transfer_velocity.argtypes = [i32, dbl, dbl, dbl, dbl, ctypes.POINTER(ctypes.c_int), d, d, d, d]
        L.or_thermal_pass_a.argtypes = [i64, i64, vp, vp, vp, vp, vp, dbl, dbl, vp, vp]
        L.or_thermal_pass_r.argtypes = [i64, i64, vp, vp, vp, vp, vp, dbl, vp, vp, dbl]
        L.or_thermal_pass_b.argtypes = [i64, i64, vp, vp, vp, vp, vp, dbl, vp, vp, vp, vp]
        L.or_slope_erode_cap.argtypes = [i64, i64, vp, vp, vp, dbl, dbl, vp]
        L.or_laplacian.argtypes = [i64, i64, vp, vp, vp, vp]
        L.or_priority_flood_eps.argtypes = [i64, i64, i64, i64, vp, vp, vp, dbl, vp, vp, vp, vp]
        L.or_label_components.argtypes = [vp, vp, i64, i64, vp, vp]
        L.or_label_components.restype = i64
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------------------
# scalar geometry (cubesphere.py:70-220), for tests
# ---------------------------------------------------------------------------
def to_sphere(face, u, v):
    x, y, z = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().or_to_sphere(int(face), float(u), float(v), ctypes.byref(x), ctypes.byref(y), ctypes.byref(z))
    return x.value, y.value, z.value


def from_sphere(x, y, z):
    f, u, v = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    lib().or_from_sphere(float(x), float(y), float(z), ctypes.byref(f), ctypes.byref(u), ctypes.byref(v))
    return f.value, u.value, v.value


def jacobian(face, u, v):
    j = (ctypes.c_double * 6)()
    lib().or_jacobian(int(face), float(u), float(v), j)
    return tuple(j)


def transfer_velocity(face, u, v, a, b):
    f = ctypes.c_int()
    o = [ctypes.c_double() for _ in range(4)]
    lib().or_transfer_velocity(int(face), float(u), float(v), float(a), float(b), ctypes.byref(f), *[ctypes.byref(x) for x in o])
    return (f.value,) + tuple(x.value for x in o)


def set_libm(tan_fn=None, atan_fn=None, sincos_fn=None):
    """Use other tan / atan (the geometry) and sincos(x, &sin, &cos) (a
    loaded droplet's initial direction) -- C function pointers, e.g. ctypes
    function objects; None restores glibc's."""
    as_ptr = lambda fn: ctypes.cast(fn, ctypes.c_void_p) if fn is not None else None
    lib().or_set_libm(as_ptr(tan_fn), as_ptr(atan_fn))
    lib().or_set_sincos(as_ptr(sincos_fn))


def device_libm(hiplib):
    """The correctly rounded tan / atan / sincos of the library under test
    (its host-side pg_geom_* functions) as set_libm's arguments."""
    return hiplib.pg_geom_tan, hiplib.pg_geom_atan, hiplib.pg_geom_sincos


def change_list_cap(max_steps: int, lateral: bool) -> int:
    return int(lib().or_change_list_cap(int(max_steps), int(bool(lateral))))


# ---------------------------------------------------------------------------
# parameters
# ---------------------------------------------------------------------------
def _eparams(params):
    return params.erosion if hasattr(params, "erosion") else params


def cell_units(ep, name, unit):
    return float(getattr(ep, name)) / float(unit)


def rng_of(params, stage, *key):
    """WorldParams.rng restated (config.py:830-842)."""
    if hasattr(params, "rng"):
        return params.rng(stage, *key)
    raise TypeError("params needs an rng(stage, *key) method")


# ---------------------------------------------------------------------------
# state (maps.py:56-191)
# ---------------------------------------------------------------------------
@dataclass
class OracleState:
    height: np.ndarray
    sediment: np.ndarray
    discharge: np.ndarray
    momentum: np.ndarray
    hardness: np.ndarray
    precip: np.ndarray
    evap: np.ndarray
    uplift: np.ndarray
    mask: np.ndarray
    metric: np.ndarray
    metric_inv: np.ndarray
    N: int
    H: int
    spherical: bool
    height_unit_m: float
    grid: object = None
    base_at: object = None
    route: np.ndarray | None = None
    lake_flag: np.ndarray | None = None
    disch_track: np.ndarray = None
    mom_track: np.ndarray = None
    samp: np.ndarray = None
    acc: np.ndarray = None
    pending: np.ndarray = None
    lake_load: np.ndarray = None
    lake_id: np.ndarray = None
    lake_room: np.ndarray = None
    base: np.ndarray = None
    iteration: int = 0
    land_target: float | None = None
    deposit_on_exit: bool = False
    roots: np.ndarray | None = None
    inflow_volume: float = 0.0
    iso_acc: np.ndarray | None = None

    def __post_init__(self):
        F, NE = self.height.shape[0], self.height.shape[1]
        z = lambda: np.zeros((F, NE, NE), np.float64)
        if self.disch_track is None:
            self.disch_track = z()
        if self.mom_track is None:
            self.mom_track = np.zeros((F, NE, NE, 2), np.float64)
        if self.samp is None:
            self.samp = np.zeros((F, NE, NE, NS), np.float32)
        if self.acc is None:
            self.acc = z()
        if self.pending is None:
            self.pending = z()
        if self.lake_load is None:
            self.lake_load = z()
        if self.lake_id is None:
            self.lake_id = np.full((F, NE, NE), -1, np.int32)
        if self.lake_room is None:
            self.lake_room = np.zeros(1, np.float64)
        if self.base is None:
            self.base = z()
        assert NE == self.N + 2 * self.H

    @classmethod
    def from_grid(cls, grid, bedrock_m, hardness, precip, evap, uplift_m, eparams=None, mask=None):
        """maps.py:152-191 (heights to cell units; unit = grid.cell_size_m)."""
        unit = float(grid.cell_size_m)
        arr = lambda x, dt: np.ascontiguousarray(np.asarray(getattr(x, "data", x)), dtype=dt)
        F, NE = 6, grid.NE
        m = np.ones((F, NE, NE), np.uint8) if mask is None else np.ascontiguousarray(mask, np.uint8)
        return cls(height=np.ascontiguousarray(arr(bedrock_m, np.float64) / unit), sediment=np.zeros((F, NE, NE)),
                   discharge=np.zeros((F, NE, NE)), momentum=np.zeros((F, NE, NE, 2)),
                   hardness=arr(hardness, np.float32), precip=arr(precip, np.float32), evap=arr(evap, np.float32),
                   uplift=np.ascontiguousarray(arr(uplift_m, np.float64) / unit), mask=m,
                   metric=np.ascontiguousarray(grid.metric, np.float32), metric_inv=np.ascontiguousarray(grid.metric_inv, np.float32),
                   N=grid.N, H=grid.H, spherical=True, height_unit_m=unit, grid=grid)

    @classmethod
    def window(cls, height_m, sediment_m, discharge, momentum, hardness, precip, evap, uplift_m, mask, metric, metric_inv,
               H, height_unit_m, deposit_on_exit=False):
        """maps.py:193-240 (F = 1, spherical False)."""
        NE = height_m.shape[0]
        a = lambda x, dt, extra=(): np.ascontiguousarray(np.asarray(x), dtype=dt).reshape((1, NE, NE) + tuple(extra))
        return cls(height=a(height_m, np.float64) / height_unit_m, sediment=a(sediment_m, np.float64) / height_unit_m,
                   discharge=a(discharge, np.float64), momentum=a(momentum, np.float64, (2,)), hardness=a(hardness, np.float32),
                   precip=a(precip, np.float32), evap=a(evap, np.float32), uplift=a(uplift_m, np.float64) / height_unit_m,
                   mask=a(mask, np.uint8), metric=a(metric, np.float32, (3,)), metric_inv=a(metric_inv, np.float32, (3,)),
                   N=NE - 2 * H, H=H, spherical=False, height_unit_m=float(height_unit_m), deposit_on_exit=deposit_on_exit)

    @property
    def F(self):
        return self.height.shape[0]

    @property
    def NE(self):
        return self.height.shape[1]

    @property
    def interior(self):
        return (slice(None), slice(self.H, self.H + self.N), slice(self.H, self.H + self.N))

    def surface(self):
        return self.height + self.sediment

    def total_mass(self) -> float:
        """maps.py:322-326."""
        act = self.mask[self.interior] == MASK_ACTIVE
        return float(np.sum((self.height + self.sediment + self.pending + self.lake_load)[self.interior][act]))

    def pack(self):
        """maps.py:309-314 -> particle.py:164-191."""
        use_route = self.route is not None
        if self.lake_flag is None:
            self.lake_flag = np.zeros(self.height.shape, np.uint8)
        route = self.route if use_route else self.height
        lib().or_pack_samples(_p(self.samp), self.F, self.NE, _p(self.height), _p(self.sediment), _p(route), int(use_route),
                              _p(self.discharge), _p(self.momentum), _p(self.evap), _p(self.hardness), _p(self.metric),
                              _p(self.metric_inv), _p(self.base), _p(self.lake_flag))

    def owner_table(self):
        """maps.py:342-349."""
        n = self.F * self.NE * self.NE
        own = self.grid.owner if self.spherical else np.arange(n, dtype=np.int64).reshape(self.height.shape)
        return np.ascontiguousarray(own, dtype=np.int64)

    def refresh_route(self, eps):
        """maps.py:351-358."""
        self.route = priority_flood_eps(self.surface(), self.mask, self.owner_table(), self.H, self.N, float(eps),
                                        self.base)
        if self.spherical:
            exchange(self.grid, self.route, linear=True)

    def refresh_lakes_window(self, min_depth, lake_evap=0.0):
        """maps.py:629-672 without the evaporation balance."""
        if float(lake_evap) > 0.0:
            raise NotImplementedError("the window lakes' evaporation balance is outside the oracle's scope")
        flag = np.zeros(self.height.shape, dtype=np.uint8)
        if self.route is None:
            self.lake_flag = flag
            return {"lake_cells": 0}
        lake = ((self.route - self.surface()) > float(min_depth)) & (self.mask == MASK_ACTIVE)
        inter = self.interior
        flag[inter] = lake[inter]
        self.lake_flag = flag
        return {"lake_cells": int(lake[inter].sum())}

    def exchange_halos(self):
        """maps.py:329-340 with field.py:104-124 (numpy einsum, cubic weights;
        the vector momentum rotated into the destination face first)."""
        if not self.spherical:
            return
        for arr in (self.height, self.sediment, self.discharge):
            exchange(self.grid, arr)
        exchange_vector(self.grid, self.momentum)
        if self.route is not None:
            exchange(self.grid, self.route, linear=True)


# ---------------------------------------------------------------------------
# spawning (maps.py:793-833)
# ---------------------------------------------------------------------------
def spawn_weights(state):
    inter = np.zeros(state.height.shape, bool)
    inter[state.interior] = True
    ok = inter & (state.mask == MASK_ACTIVE) & (state.height + state.sediment >= state.base) & (state.precip > 0)
    idx = np.flatnonzero(ok)
    return idx, state.precip.reshape(-1)[idx].astype(np.float64)


def spawn_cells(cdf, u):
    out = np.empty(u.shape[0], np.int64)
    cdf = np.ascontiguousarray(cdf, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    lib().or_spawn_cells(_p(cdf), cdf.shape[0], _p(u), u.shape[0], _p(out))
    return out


def spawn_particles(state, n_particles, rng):
    idx, w = spawn_weights(state)
    if idx.size == 0 or n_particles <= 0:
        return None
    total = float(w.sum())
    if total <= 0.0:
        return None
    cdf = np.cumsum(w)
    u = np.sort(rng.random(n_particles))
    off = rng.random((n_particles, 2))
    flat = idx[spawn_cells(cdf, u)]
    NE = state.NE
    f = flat // (NE * NE)
    rem = flat - f * NE * NE
    ei = rem // NE
    ej = rem - ei * NE
    x = (ei - state.H).astype(np.float64) + off[:, 0]
    y = (ej - state.H).astype(np.float64) + off[:, 1]
    return f.astype(np.int64), x, y, total / n_particles


def max_steps_of(ep, N):
    return int(ep.max_steps) if ep.max_steps > 0 else 2 * int(N)


def chunk_size(ep, n_particles):
    return max(1, min(int(ep.chunk), -(-int(n_particles) // MIN_CHUNKS)))


def _check_plain(ep, state):
    pass


# ---------------------------------------------------------------------------
# one iteration (maps.py:839-1047)
# ---------------------------------------------------------------------------
def run_iteration(state, params, iteration_key, n_particles=None, particles_per_cell=None, rng_stage="erosion",
                  diag=None, nthreads=0, spawn=None, diag_lists=None):
    """``spawn``: optionally a precomputed ``(face, x, y, volume0)`` to use in
    place of the rain spawn (the same value ``spawn_particles`` returns).
    ``diag_lists(cl_cell, cl_delta, cl_vol, cl_mom, cl_count, cap)``: a test
    hook that sees each chunk's whole change lists before they are applied."""
    ep = _eparams(params)
    _check_plain(ep, state)
    key = iteration_key if isinstance(iteration_key, (tuple, list)) else (int(iteration_key),)
    rng = rng_of(params, rng_stage, *key)
    t0 = time.time()
    n_cells = int(np.count_nonzero(state.mask[state.interior] == MASK_ACTIVE))
    if n_particles is None:
        ppc = ep.particles_per_cell if particles_per_cell is None else particles_per_cell
        n_particles = max(1, int(round(ppc * n_cells)))
    sp = spawn_particles(state, n_particles, rng) if spawn is None else spawn
    stats = {"particles": 0, "entries": 0, "steps_mean": 0.0}
    if sp is None:
        return stats
    sp_face, sp_x, sp_y, volume0 = sp
    sp_sed = np.zeros(sp_face.shape[0])
    sp_dir = np.zeros(sp_face.shape[0])
    pend_cells = np.flatnonzero(state.pending.reshape(-1) > 0.0)
    released = 0.0
    if pend_cells.size:
        NE = state.NE
        pf = pend_cells // (NE * NE)
        rem = pend_cells - pf * NE * NE
        pei = rem // NE
        pej = rem - pei * NE
        loads = state.pending.reshape(-1)[pend_cells]
        released = float(loads.sum())
        sp_face = np.concatenate([sp_face, pf.astype(np.int64)])
        sp_x = np.concatenate([sp_x, (pei - state.H) + 0.5])
        sp_y = np.concatenate([sp_y, (pej - state.H) + 0.5])
        sp_sed = np.concatenate([sp_sed, loads])
        sp_dir = np.concatenate([sp_dir, rng.random(pend_cells.size) * (2.0 * np.pi)])
        state.pending.reshape(-1)[pend_cells] = 0.0
    state.pack()
    state.acc[...] = 0.0
    disc_sat = float(ep.disc_saturation)
    mom_scale = 1.0
    if float(ep.disc_saturation_cells) > 0.0 or float(ep.momentum_saturation_cells) > 0.0:
        total = float(volume0) * float(n_particles)
        n_rain = max(int(spawn_weights(state)[0].size), 1)
        rain_per_cell = max(total - float(state.inflow_volume), 0.0) / n_rain
        rain_per_cell = rain_per_cell if rain_per_cell > 0.0 else total / n_rain
        if float(ep.disc_saturation_cells) > 0.0:
            disc_sat = float(ep.disc_saturation_cells) * rain_per_cell
        if float(ep.momentum_saturation_cells) > 0.0:
            mom_scale = float(volume0) / (float(ep.momentum_saturation_cells) * rain_per_cell)
    unit = state.height_unit_m
    use_ecap = float(ep.slope_limit_erode) > 0.0
    if use_ecap:
        ecap = np.empty_like(state.height)
        lib().or_slope_erode_cap(state.F, state.NE, _p(state.height), _p(state.sediment), _p(state.mask),
                                 float(ep.slope_limit_erode), float(ep.slope_limit_exit), _p(ecap))
    else:
        ecap = np.zeros(1, np.float64)
    iter_deposit = cell_units(ep, "iter_deposit", unit)
    if float(ep.slope_limit_deposit) > 0.0:
        iter_deposit = min(iter_deposit, float(ep.slope_limit_deposit) * math.sqrt(2.0) * 0.3)
    max_steps = max_steps_of(ep, state.N)
    lateral = float(getattr(ep, "lateral_rate", 0.0)) > 0.0
    cap = change_list_cap(max_steps, lateral)
    P = sp_face.shape[0]
    chunk = chunk_size(ep, P)
    lake_park = 1.0 if (state.spherical and bool(getattr(ep, "lake_fill", False)) and bool(getattr(ep, "lake_balance", False))) else 0.0
    prm = TraceParams(dt=float(ep.dt), density=float(ep.density), friction=float(ep.friction),
                      deposition_rate=float(ep.deposition_rate), evap_rate=float(ep.evap_rate), k_mom=float(ep.k_mom),
                      mom_scale=mom_scale, k_disc=float(ep.k_disc), disc_saturation=disc_sat,
                      disc_exponent=float(ep.disc_exponent), slope_gain=float(ep.slope_gain),
                      slope_sat=float(ep.slope_saturation), erodibility=float(ep.erodibility),
                      cover_depth=cell_units(ep, "cover_depth", unit), min_volume=float(ep.min_volume_frac) * float(volume0),
                      max_steps=int(max_steps), max_erode=cell_units(ep, "max_erode", unit),
                      conc_scale=float(chunk / (volume0 * P)), pit_steps=int(ep.pit_steps),
                      deposit_on_exit=int(bool(state.deposit_on_exit)), ocean_rate=float(ep.ocean_deposition_rate),
                      ocean_steps=int(ep.ocean_steps), fan_room=cell_units(ep, "fan_room", unit),
                      dep_floor=cell_units(ep, "dep_floor_m", unit), lake_trap=float(ep.lake_trap), lake_park=lake_park,
                      lateral_rate=float(getattr(ep, "lateral_rate", 0.0)))
    cl_cell = np.empty(chunk * cap, np.int32)
    cl_delta = np.empty(chunk * cap, np.float64)
    cl_vol = np.empty(chunk * cap, np.float32)
    cl_mom = np.empty((chunk * cap, 2), np.float32)
    cl_count = np.zeros(chunk, np.int64)
    cl_lo = np.zeros(chunk, np.int64)
    cl_hi = np.zeros(chunk, np.int64)
    sp_death = np.zeros(chunk, np.int8)
    deaths = np.zeros(len(DEATH_NAMES), np.int64)
    roots = state.roots if state.roots is not None else np.zeros((1, 1, 1), np.float32)
    out = np.zeros(4, np.float64)
    entries = n_clamp = 0
    to_pending = lost = lost_offshore = 0.0
    t_trace = t_apply = 0.0
    L = lib()
    iter_erode = cell_units(ep, "iter_erode", unit)
    dep_floor = cell_units(ep, "dep_floor_m", unit)
    for c0 in range(0, P, chunk):
        c1 = min(P, c0 + chunk)
        m = c1 - c0
        t1 = time.time()
        sl = [np.ascontiguousarray(a[c0:c1]) for a in (sp_face, sp_x, sp_y, sp_sed, sp_dir)]
        L.or_trace_particles(m, *[_p(a) for a in sl], float(volume0), _p(state.samp), int(state.route is not None),
                             _p(state.mask), int(state.N), int(state.H), int(bool(state.spherical)), ctypes.byref(prm),
                             _p(roots), int(state.roots is not None), _p(cl_cell), _p(cl_delta), _p(cl_vol), _p(cl_mom),
                             _p(cl_count), _p(cl_lo), _p(cl_hi), _p(sp_death), int(nthreads))
        t2 = time.time()
        if diag is not None:
            diag(state, cl_cell, cl_vol, cl_count[:m], cap, sp_death[:m])
        if diag_lists is not None:
            diag_lists(cl_cell, cl_delta, cl_vol, cl_mom, cl_count[:m], cap)
        L.or_apply_changes(_p(cl_cell), _p(cl_delta), _p(cl_vol), _p(cl_mom), _p(cl_count), m, cap, _p(state.height),
                           _p(state.sediment), _p(state.acc), _p(state.pending), _p(state.samp), _p(state.disch_track),
                           _p(state.mom_track), _p(state.mask), iter_erode, iter_deposit, float(ep.fan_slope),
                           int(state.route is not None), dep_floor, float(ep.offshore_writeoff), _p(ecap), int(use_ecap),
                           _p(state.lake_load), _p(state.lake_id), _p(state.lake_room), _p(out))
        n_clamp += int(out[0])
        to_pending += out[1]
        lost += out[2]
        lost_offshore += out[3]
        t_trace += t2 - t1
        t_apply += time.time() - t2
        entries += int(cl_count[:m].sum())
        deaths += np.bincount(sp_death[:m], minlength=deaths.size)
    n = state.height.size
    L.or_fold_changes(n, _p(state.height), _p(state.sediment), _p(state.acc), _p(state.mask))
    L.or_ema_update(n, _p(state.discharge), _p(state.momentum), _p(state.disch_track), _p(state.mom_track), _p(state.mask),
                    float(ep.ema))
    stats.update({"particles": int(P), "entries": entries, "steps_mean": entries / max(P, 1), "volume": float(volume0),
                  "seconds": time.time() - t0})
    stats["deaths"] = {nm: int(c) for nm, c in zip(DEATH_NAMES, deaths)}
    stats["clamped"] = n_clamp
    stats["to_pending"] = to_pending
    stats["released"] = float(released)
    stats["pending_total"] = float(state.pending[state.interior].sum())
    stats["deficit_out"] = lost
    stats["lost_offshore"] = lost_offshore
    stats["seconds_trace"] = t_trace
    stats["seconds_apply"] = t_apply
    stats["chunk"] = int(chunk)
    stats["disc_saturation"] = disc_sat
    stats["mom_scale"] = mom_scale
    return stats


# ---------------------------------------------------------------------------
# routing surface (route.py:30-117)
# ---------------------------------------------------------------------------
def priority_flood_eps(surface, mask, owner, H, N, eps, base):
    """route.py:30-117: the epsilon-filled surface (heap-ordered flood in C)."""
    surface = np.ascontiguousarray(surface, np.float64)
    F, NE = surface.shape[0], surface.shape[1]
    filled = np.empty_like(surface)
    visited = np.empty(surface.shape, np.uint8)
    heap = np.empty(2 * surface.size, np.float64)  # (value, index) pairs, 16 B per cell
    lib().or_priority_flood_eps(F, NE, int(H), int(N), _p(surface), _p(np.ascontiguousarray(mask, np.uint8)),
                                _p(np.ascontiguousarray(owner, np.int64)), float(eps),
                                _p(np.ascontiguousarray(base, np.float64)), _p(filled), _p(visited), _p(heap))
    return filled


# ---------------------------------------------------------------------------
# the sea mask (maps.py:360-421, hydro/run.py:64-83)
# ---------------------------------------------------------------------------
def label_components(mask, owner, N, H):
    """hydro/lakes.py:19-46 (level all equal): (labels int32 (M,), n)."""
    mask = np.ascontiguousarray(np.asarray(mask, bool).reshape(-1), np.uint8)
    labels = np.empty(mask.size, np.int32)
    stack = np.empty(mask.size, np.int64)
    n = lib().or_label_components(_p(mask), _p(np.ascontiguousarray(owner, np.int64)), int(N), int(H), _p(labels),
                                  _p(stack))
    return labels, int(n)


def interior_cell_area(grid):
    """cubesphere.py:509-517: cell areas in m^2, float32, over the interior."""
    H, N = grid.H, grid.N
    area = (np.asarray(grid.sqrt_det_metric).astype(np.float64) * float(grid.cell_size_m) ** 2).astype(np.float32)
    return area[:, H:H + N, H:H + N]


def below_sea_components(surface, grid, min_fraction=0.02):
    """hydro/run.py:64-83: (below, labels, keep) flat over the interior."""
    N, H = grid.N, grid.H
    below = np.ascontiguousarray(np.asarray(surface) < 0.0).reshape(-1)
    if not below.any():
        return below, None, None
    labels, n = label_components(below, grid.owner, N, H)
    area = interior_cell_area(grid).reshape(-1).astype(np.float64)
    per = np.bincount(labels[below], weights=area[below], minlength=max(n, 1))
    keep = per >= float(min_fraction) * float(area.sum())
    keep[int(np.argmax(per))] = True
    return below, labels, keep


def ocean_min_fraction(params) -> float:
    """maps.py:770-774."""
    hydro = getattr(params, "hydro", None)
    return float(hydro.ocean_min_fraction) if hydro is not None and hasattr(params, "world") else 0.02


def refresh_base(state, min_fraction):
    """maps.py:360-421: base level 0 on the open ocean and on land, a closed
    basin's deepest surface inside it; halos copied from their owners."""
    base = np.zeros_like(state.height)
    info = {"closed_cells": 0, "closed_basins": 0, "sea_cells": 0, "deepest_basin_floor": 0.0}
    if not state.spherical or state.grid is None:
        state.base = base
        state.base_at = state.iteration
        return info
    inter = state.interior
    surf = (state.height[inter] + state.sediment[inter]).astype(np.float64)
    below, labels, keep = below_sea_components(surf.astype(np.float32), state.grid, float(min_fraction))
    if labels is not None:
        sea = below & keep[np.maximum(labels, 0)]
        closed = below & ~sea
        info["sea_cells"] = int(sea.sum())
        info["closed_cells"] = int(closed.sum())
        if closed.any():
            lab = labels[closed]
            floor = np.zeros(keep.size, dtype=np.float64)
            np.minimum.at(floor, lab, surf.reshape(-1)[closed])
            inner = np.zeros(surf.size, dtype=np.float64)
            inner[closed] = floor[lab]
            base[inter] = inner.reshape(surf.shape)
            info["closed_basins"] = int(np.unique(lab).size)
            info["deepest_basin_floor"] = float(floor.min())
    own = np.asarray(state.grid.owner).reshape(-1)
    flat = base.reshape(-1)
    flat[:] = flat[own]
    state.base = base
    state.base_at = state.iteration
    return info


# ---------------------------------------------------------------------------
# mass wasting (maps.py:1050-1085, particle.py:1104-1227)
# ---------------------------------------------------------------------------
DEP_FLOOR = 0.02  # particle.py:99


def talus_of(ep, hardness):
    """maps.py:1062: the talus slope, numpy's float32 arithmetic on the
    float32 hardness, widened to float64."""
    return (ep.talus_slope_soft + (ep.talus_slope_hard - ep.talus_slope_soft) * hardness).astype(np.float64)


def mass_wasting_pass(state, talus, rate, cap, mask=None, dep_floor=DEP_FLOOR):
    """maps.py:1073-1085: the three thermal passes over the whole state."""
    mask = state.mask if mask is None else np.ascontiguousarray(mask, np.uint8)
    F, NE = state.F, state.NE
    talus = np.ascontiguousarray(talus, np.float64)
    out_total = np.zeros_like(state.height)
    scale = np.zeros_like(state.height)
    L = lib()
    L.or_thermal_pass_a(F, NE, _p(state.height), _p(state.sediment), _p(state.metric), _p(talus), _p(mask), float(rate),
                        float(cap), _p(out_total), _p(scale))
    rscale = np.ones_like(state.height)
    L.or_thermal_pass_r(F, NE, _p(state.height), _p(state.sediment), _p(state.metric), _p(talus), _p(mask), float(rate),
                        _p(scale), _p(rscale), float(dep_floor))
    new_h = state.height.copy()
    new_s = state.sediment.copy()
    L.or_thermal_pass_b(F, NE, _p(state.height), _p(state.sediment), _p(state.metric), _p(talus), _p(mask), float(rate),
                        _p(scale), _p(rscale), _p(new_h), _p(new_s))
    state.height[...] = new_h
    state.sediment[...] = new_s


def thermal_erosion(state, params):
    """maps.py:1050-1070: the talus pass, then (creep_rate > 0) the creep
    pass with talus 0 and submerged cells frozen."""
    ep = _eparams(params)
    unit = state.height_unit_m
    talus = talus_of(ep, state.hardness)
    dep_floor = cell_units(ep, "dep_floor_m", unit)
    mass_wasting_pass(state, talus, float(ep.thermal_rate), cell_units(ep, "thermal_max", unit), dep_floor=dep_floor)
    if ep.creep_rate > 0.0:
        cmask = np.where(state.height + state.sediment < state.base, np.uint8(MASK_FROZEN), state.mask).astype(np.uint8)
        mass_wasting_pass(state, np.zeros_like(talus), float(ep.creep_rate), cell_units(ep, "thermal_max", unit), cmask,
                          dep_floor=dep_floor)


def slope_erode_cap(state, coef, exit_slope):
    """particle.py:1045-1076 over the state's surface."""
    out = np.empty_like(state.height)
    lib().or_slope_erode_cap(state.F, state.NE, _p(state.height), _p(state.sediment), _p(state.mask), float(coef),
                             float(exit_slope), _p(out))
    return out


# ---------------------------------------------------------------------------
# isostasy (maps.py:1325-1410) and the smoother (field.py:104-124, 179-188)
# ---------------------------------------------------------------------------
def exchange(grid, data, linear=False):
    """field.py:104-124 for a float64 scalar field, in place: each halo cell
    the weighted sum of its 16 interior sources.  The reference writes it as
    np.einsum, which leaves the summation order to numpy's SIMD kernels (so
    the reference's own result depends on the machine); the oracle fixes the
    order -- source after source, one rounding per product and per sum -- and
    the einsum agrees with it to the forward error bound of a 16-term dot
    product (tests/test_gpu_sidefuncs.py)."""
    hm = grid.halo
    flat = data.reshape(6 * grid.NE * grid.NE)
    w = (hm.w_lin if linear else hm.w).astype(flat.dtype)
    src = flat[hm.src]
    acc = np.zeros(hm.dst.size, flat.dtype)
    for s in range(w.shape[1]):
        acc = acc + w[:, s] * src[:, s]
    flat[hm.dst] = acc
    return data


def exchange_vector(grid, data):
    """field.py:104-124 for a (..., 2) tangent vector field, in place: each
    source rotated into the destination face's frame (R x, the two products
    summed), weighted and summed source after source (the einsum's order
    fixed as in `exchange`)."""
    hm = grid.halo
    flat = data.reshape(6 * grid.NE * grid.NE, 2)
    x = flat[hm.src]
    R = hm.rot
    a0 = np.zeros(hm.dst.size)
    a1 = np.zeros(hm.dst.size)
    for s in range(hm.w.shape[1]):
        ws = hm.w[:, s]
        a0 = a0 + ws * (R[:, s, 0, 0] * x[:, s, 0] + R[:, s, 0, 1] * x[:, s, 1])
        a1 = a1 + ws * (R[:, s, 1, 0] * x[:, s, 0] + R[:, s, 1, 1] * x[:, s, 1])
    flat[hm.dst, 0] = a0
    flat[hm.dst, 1] = a1
    return data


def laplacian(grid, data):
    """field.py:179-188: the flux-form Laplace-Beltrami operator per m^2,
    returned in float32 as FaceField.laplacian returns it."""
    d = np.ascontiguousarray(data, np.float64)
    ginv = np.ascontiguousarray(grid.metric_inv, np.float64)
    sq = np.ascontiguousarray(grid.sqrt_det_metric, np.float64)
    out = np.empty_like(d)
    lib().or_laplacian(6, grid.NE, _p(d), _p(ginv), _p(sq), _p(out))
    out /= grid.cell_size_m ** 2
    return out.astype(np.float32)


_KAPPA = {}


def smoothing_kappa(grid):
    """maps.py:1325-1363: 0.9 / (largest eigenvalue of -L in cell units), by
    40 power iterations from seeded white noise; cached per grid."""
    key = (grid.N, grid.H, float(grid.cell_size_m))
    if key not in _KAPPA:
        rng = np.random.default_rng(0)
        data = np.zeros((6, grid.NE, grid.NE))
        H, N = grid.H, grid.N
        data[:, H:H + N, H:H + N] = rng.standard_normal((6, N, N))
        exchange(grid, data)
        c2, lam = float(grid.cell_size_m) ** 2, 8.0
        for _ in range(40):
            exchange(grid, data, linear=True)
            Lv = laplacian(grid, data).astype(np.float64) * c2
            num = float(np.sqrt((Lv[:, H:-H, H:-H] ** 2).sum()))
            den = float(np.sqrt((data[:, H:H + N, H:H + N] ** 2).sum()))
            if num <= 0.0 or den <= 0.0:
                break
            lam = num / den
            data = Lv / num
        _KAPPA[key] = 0.9 / max(lam, 1e-6)
    return _KAPPA[key]


def isostasy_steps(state, ep):
    """maps.py:1392-1397: (sigma in cells, kappa, diffusion steps)."""
    cell = float(state.height_unit_m)
    sig = float(ep.flexure_km) * 1000.0 / cell
    sig = min(sig, state.N / 8.0)
    kappa = smoothing_kappa(state.grid)
    n = int(np.ceil(sig * sig / (2.0 * kappa))) if sig >= 0.5 else 0
    return sig, kappa, n


def apply_isostasy(state, params):
    """maps.py:1366-1410."""
    ep = _eparams(params)
    grid = state.grid
    d = state.iso_acc
    k = float(ep.isostasy)
    data = (-k * d).astype(np.float64)
    cell = float(state.height_unit_m)
    sig, kappa, n = isostasy_steps(state, ep)
    for _ in range(n):
        exchange(grid, data, linear=True)
        data += kappa * laplacian(grid, data).astype(np.float64) * (cell * cell)
    exchange(grid, data, linear=True)
    act = state.mask == MASK_ACTIVE
    iact = state.mask[state.interior] == MASK_ACTIVE
    mean = float(data[state.interior][iact].mean()) if iact.any() else 0.0
    state.height[act] += (data[act] - mean).astype(state.height.dtype)
    moved = float(np.abs(d[state.interior]).sum())
    state.iso_acc[...] = 0.0
    return {"sigma_cells": sig, "diffusion_steps": n, "kappa": kappa, "load_abs_cells": moved,
            "rebound_max": float((data[state.interior] - mean).max()),
            "rebound_min": float((data[state.interior] - mean).min())}


# ---------------------------------------------------------------------------
# uplift, replay, datum (maps.py:1088-1213, 1413-1508)
# ---------------------------------------------------------------------------
def _applied_uplift(state, cap):
    u = state.uplift if cap is None else np.minimum(state.uplift, float(cap))
    mean = 0.0
    if state.spherical:
        iact = state.mask[state.interior] == MASK_ACTIVE
        mean = float(u[state.interior][iact].mean()) if iact.any() else 0.0
    return u, mean


def apply_uplift(state, cap=None):
    act = state.mask == MASK_ACTIVE
    u, mean = _applied_uplift(state, cap)
    state.height[act] += u[act] - mean


def uplift_cap(state, ep):
    if not state.spherical or float(getattr(ep, "uplift_max_m", 0.0)) <= 0.0:
        return None
    return cell_units(ep, "uplift_max_m", state.height_unit_m)


def uplift_replay(state, ep):
    if not state.spherical or str(getattr(ep, "uplift_mode", "stack")) != "replay":
        return None
    u, mean = _applied_uplift(state, uplift_cap(state, ep))
    act = state.mask == MASK_ACTIVE
    total = np.zeros_like(state.height)
    total[act] = float(int(ep.iterations)) * (u[act] - mean)
    return total


def held_land_fraction(state, params):
    if state.land_target is not None:
        return float(state.land_target)
    if float(getattr(getattr(params, "tectonics", None), "shelf_fraction", 0.0)) > 0.0:
        raise ValueError("shelf mode holds the bedrock's land fraction: set state.land_target")
    return float(params.world.land_fraction)


def hold_datum(state, land_fraction, level=0.0):
    inter = state.interior
    surf = (state.height[inter] + state.sediment[inter]).ravel()
    M = surf.size
    n_land = min(max(int(round(float(land_fraction) * M)), 0), M)
    if n_land == 0:
        q = float(surf.max()) + 1.0
    elif n_land == M:
        q = float(surf.min())
    else:
        k = M - n_land
        q = float(np.partition(surf, k)[k])
    q -= float(level)
    if q != 0.0:
        state.height -= q
        if state.route is not None:
            state.route -= q
        if state.base is not None:
            np.subtract(state.base, q, out=state.base, where=state.base < 0.0)
    return q


def start_replay(state, params):
    ep = _eparams(params)
    total = uplift_replay(state, ep)
    if total is None:
        return None
    state.height -= total
    state.exchange_halos()
    inter = state.interior
    info = {"replay_max": float(total[inter].max()), "replay_min": float(total[inter].min()),
            "land_fraction": float(np.mean((state.height + state.sediment)[inter] >= 0)), "datum_shift": 0.0}
    if hasattr(params, "world"):
        info["datum_shift"] = hold_datum(state, held_land_fraction(state, params))
    return info


def step(state, params, iteration_key, **kw):
    """maps.py:1516-1617: particles, mass wasting, uplift, isostasy (every
    isostasy_every iterations), datum, halos."""
    ep = _eparams(params)
    if state.spherical and bool(getattr(ep, "lake_balance", False)) and int(ep.flood_every) > 0:
        raise NotImplementedError("lake refreshes are outside the oracle's scope")
    if state.spherical and int(ep.glacial_every) > 0:
        raise NotImplementedError("the glacial pass is outside the oracle's scope")
    t0 = time.time()
    sea_every = int(getattr(ep, "sea_mask_every", 0))
    if state.spherical and sea_every > 0 and (getattr(state, "base_at", None) is None or state.iteration % sea_every == 0):
        # before the route: the flood seeds on the sea
        st_base = refresh_base(state, ocean_min_fraction(params))
    else:
        st_base = None
    flood = int(ep.flood_every)
    if flood > 0 and (state.route is None or state.iteration % flood == 0):
        state.refresh_route(cell_units(ep, "route_eps", state.height_unit_m))
    if not state.spherical and bool(getattr(ep, "window_lakes", False)) and flood > 0 and \
            (state.lake_flag is None or state.iteration % flood == 0):
        hydro = getattr(params, "hydro", None)
        min_depth = float(hydro.lake_min_depth) if hydro is not None else 0.5
        state.refresh_lakes_window(min_depth / float(state.height_unit_m), float(getattr(ep, "window_lake_evap", 0.0)))
    iso = state.spherical and float(getattr(ep, "isostasy", 0.0)) > 0.0
    if iso:
        if state.iso_acc is None:
            state.iso_acc = np.zeros_like(state.height)
        s0 = state.height + state.sediment
    st = run_iteration(state, params, iteration_key, **kw)
    t1 = time.time()
    thermal_erosion(state, params)
    if iso:
        state.iso_acc += (state.height + state.sediment) - s0
    apply_uplift(state, uplift_cap(state, ep))
    if iso and (state.iteration + 1) % max(int(getattr(ep, "isostasy_every", 10)), 1) == 0:
        st["isostasy"] = apply_isostasy(state, params)
    if state.spherical and hasattr(params, "world"):
        st["datum_shift"] = hold_datum(state, held_land_fraction(state, params))
    state.exchange_halos()
    state.iteration += 1
    if st_base is not None:
        st["sea"] = st_base
    st["seconds_particles"] = t1 - t0
    st["seconds_total"] = time.time() - t0
    return st


# ---------------------------------------------------------------------------
# the erosion brush (no counterpart in the reference)
# ---------------------------------------------------------------------------
def brush_weights(radius: int) -> np.ndarray:
    """Integer shares (units of 2^-16) of the offsets |o| < radius, weighted
    max(0, radius - |o|) and normalised to 2^16, on the (2A+1)^2 block with
    A = radius - 1 (A = 0 for radius 0 and 1); the centre's share is 0: it
    keeps the remainder.  The weight total is summed row-major, one term at
    a time."""
    A = int(radius) - 1 if int(radius) > 1 else 0
    K = 2 * A + 1
    wt = np.zeros((K, K))
    total = 0.0
    for dy in range(-A, A + 1):
        for dx in range(-A, A + 1):
            w = float(radius) - math.sqrt(float(dx * dx + dy * dy))
            wt[dy + A, dx + A] = w if w > 0.0 else 0.0
            total += w if w > 0.0 else 0.0
    out = np.zeros((K, K), np.int64)
    for dy in range(K):
        for dx in range(K):
            if (dy, dx) != (A, A) and total > 0.0:
                out[dy, dx] = int(math.floor(65536.0 * wt[dy, dx] / total + 0.5))
    return out


def brush_share(d: int, w: int) -> int:
    """sign(d) floor(|d| w / 2^16) in exact integer arithmetic."""
    m = -d if d < 0 else d
    s = (m * w) >> 16
    return -s if d < 0 else s


def brush_spread(src: np.ndarray, mask: np.ndarray, owner: np.ndarray | None, radius: int) -> np.ndarray:
    """Each non-zero src[s] (int64, (F, NE, NE)) gives brush_share to every
    valid target s + o, o != 0, inside its face's extended array and keeps
    the rest; a target is valid when it is active (mask 1) or its owner is;
    a target that is not its own owner passes its share to the owner (so
    does a source's remainder).  A scatter over the sources, one at a
    time."""
    W = brush_weights(radius)
    A = (W.shape[0] - 1) // 2
    F, NE = src.shape[0], src.shape[1]
    own = np.arange(src.size).reshape(src.shape) if owner is None else np.asarray(owner)
    flat_mask = mask.reshape(-1)
    out = np.zeros(src.size, dtype=object)
    for f, i, j in zip(*np.nonzero(src)):
        d = int(src[f, i, j])
        given = 0
        for dy in range(-A, A + 1):
            for dx in range(-A, A + 1):
                if dy == 0 and dx == 0:
                    continue
                ti, tj = i + dy, j + dx
                if not (0 <= ti < NE and 0 <= tj < NE):
                    continue
                o = int(own[f, ti, tj])
                if flat_mask[o] != MASK_ACTIVE:
                    continue
                sh = brush_share(d, int(W[dy + A, dx + A]))
                out[o] += sh
                given += sh
        keep = int(own[f, i, j])
        out[keep if flat_mask[keep] == MASK_ACTIVE else (f * NE + i) * NE + j] += d - given
    return np.array([int(v) for v in out], dtype=np.int64).reshape(src.shape)
