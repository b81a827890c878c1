"""Erosion state and iteration on the GPU: the drop-in for the reference's
particle pass (bake/globe/erosion/maps.py: ErosionState, run_iteration, step).

Typical use (the reference's own, bake/globe/erosion/maps.py:10-15)::

    state = ErosionState.from_grid(grid, bedrock_m, hardness, precip, evap, uplift_m, params.erosion)
    for it in range(n):
        step(state, params, it)
    fields = state.fields()          # FaceFields in metres, on the host

``params`` is anything with the reference's ``erosion`` group and ``rng``
(the reference's own ``WorldParams`` works), so the spawn stream is the
reference's: ``params.rng(stage, *iteration_key)``.

Two modes (``ErosionState.mode``):

* ``"check"`` -- the reference's algorithm bit for bit: the spawn drawn from
  numpy's generator on the host exactly as the reference draws it, then per
  chunk of ``erosion.chunk`` droplets a float64 trace on the frozen packed
  record (``pg_trace_particles``) and a serial apply in droplet order
  (``pg_apply_changes``), then fold and EMA -- all on the device.
* ``"fast"`` -- the production path: spawn, trace, fixed-point accumulation
  and every per-cell pass run on the device with no host round trip
  (``globe_hip.erosion.fast``).

Everything after the particles (mass wasting, uplift, isostasy, datum
hold, halo exchange) and the slope limits run on the device in both modes;
the routing flood (and a window's lakes) in the check mode.  The remaining
side passes (sea mask, the planet's lake balance, glacial carving) raise
NotImplementedError.
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .. import _lib
from .._lib import TraceParams, check, lib
from ..config import cell_units
from ..field import FaceField
from . import particle as pk

#: an iteration is split into at least this many chunks (maps.py:781-787)
MIN_CHUNKS = 16


def _eparams(params):
    return params.erosion if hasattr(params, "erosion") else params


def height_unit(grid, eparams) -> float:
    """maps.py:755-763: the kernel works in cell units."""
    u = float(eparams.height_unit_m) if eparams is not None else 0.0
    if u in (0.0, 1.0) or u == grid.cell_size_m:
        return float(grid.cell_size_m)
    raise ValueError(f"erosion.height_unit_m={u} is not supported: the kernel works in cell units ({grid.cell_size_m} m); set 0")


def max_steps_of(ep, N: int) -> int:
    return int(ep.max_steps) if ep.max_steps > 0 else 2 * int(N)


def chunk_size(ep, n_particles: int) -> int:
    return max(1, min(int(ep.chunk), -(-int(n_particles) // MIN_CHUNKS)))


def _host(x, dtype):
    d = x.data if isinstance(x, FaceField) else x
    if isinstance(d, torch.Tensor):
        d = d.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(d), dtype=dtype)


def _dev(a: np.ndarray, device) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


_TORCH_DTYPE = {np.float64: torch.float64, np.float32: torch.float32, np.uint8: torch.uint8}


def _to_dev(x, dtype, device, divide_by=None) -> torch.Tensor:
    """A field (numpy array, FaceField or tensor) on the device as ``dtype``,
    optionally divided by a scalar; a tensor already on the device stays
    there, a host array is converted on the host first (the same values)."""
    d = x.data if isinstance(x, FaceField) else x
    if isinstance(d, torch.Tensor) and d.device.type == "cuda":
        t = d.detach().to(device=device, dtype=_TORCH_DTYPE[dtype])
        return (t / divide_by if divide_by is not None else t).contiguous()
    a = _host(d, dtype)
    return _dev(a / divide_by if divide_by is not None else a, device)


class HaloTables:
    """A grid's halo gather tables resident on the device."""

    def __init__(self, grid, device):
        hm = grid.halo
        self.K, self.S = hm.src.shape
        self.dst = _dev(hm.dst.astype(np.int64), device)
        self.src = _dev(hm.src.astype(np.int64), device)
        self.w = _dev(hm.w.astype(np.float64), device)
        self.w_lin = _dev(hm.w_lin.astype(np.float64), device)
        self.rot = _dev(hm.rot.astype(np.float64), device)
        # the transpose's gather table: per source cell the entries k * S + s that read it, in order
        src = hm.src.astype(np.int64).reshape(-1)
        if np.isin(hm.dst, src).any():
            raise ValueError("halo table: a destination is also a source")
        order = np.argsort(src, kind="stable")
        cells, start = np.unique(src[order], return_index=True)
        self.n_src = int(cells.size)
        self.src_cell = _dev(cells.astype(np.int64), device)
        self.src_start = _dev(np.append(start, src.size).astype(np.int64), device)
        self.src_entry = _dev(order.astype(np.int64), device)

    def adjoint(self, adj: torch.Tensor) -> None:
        """The scalar exchange's transpose in place (pg_halo_scalar_adjoint):
        halo entries folded onto their sources, halo cells cleared."""
        pk._need(adj, torch.float64, None, "halo adjoint")
        check(lib().pg_halo_scalar_adjoint(pk._ptr(adj), self.K, self.S, pk._ptr(self.dst), pk._ptr(self.w), self.n_src,
                                           pk._ptr(self.src_cell), pk._ptr(self.src_start), pk._ptr(self.src_entry),
                                           pk._stream()), "halo_scalar_adjoint")

    def exchange(self, data: torch.Tensor, vector: bool = False, linear: bool = False) -> None:
        pk._need(data, torch.float64, None, "halo field")
        w = self.w_lin if linear else self.w
        check(lib().pg_exchange_halos(pk._ptr(data), 2 if vector else 1, self.K, self.S, pk._ptr(self.dst),
                                      pk._ptr(self.src), pk._ptr(w), pk._ptr(self.rot if vector else None), pk._stream()),
              "exchange_halos")


@dataclass
class ErosionState:
    """Device-resident erosion state; fields and dtypes as the reference's
    (bake/globe/erosion/maps.py:56-115), as torch tensors on one GPU."""

    height: torch.Tensor  # (F, NE, NE) float64, cell units
    sediment: torch.Tensor
    discharge: torch.Tensor
    momentum: torch.Tensor  # (F, NE, NE, 2) float64
    hardness: torch.Tensor  # float32
    precip: torch.Tensor  # float32
    evap: torch.Tensor  # float32
    uplift: torch.Tensor  # float64, cell units per iteration
    mask: torch.Tensor  # uint8
    metric: torch.Tensor  # (F, NE, NE, 3) float32
    metric_inv: torch.Tensor
    N: int
    H: int
    spherical: bool
    height_unit_m: float
    grid: object = None
    mode: str = "check"
    #: the check mode's apply: "ordered" (the serial apply's results computed in
    #: parallel, bit for bit) or "serial" (one lane walks the lists)
    check_apply: str = "ordered"
    check_max_sweeps: int = 1024
    #: sweeps enqueued as plain launches per chunk (the rest run in one cooperative launch)
    check_launched_sweeps: int = 40
    #: record HIP events around each check-mode trace launch (particle.check_trace_time reads them)
    check_time_trace: bool = False
    #: droplets per wave of the check-mode trace (a power of two in [1, 64]: the same results); 8 measured
    #: fastest on MI355X (fewer divergent branches per wave), CheckAutotuner re-measures it
    check_lanes: int = 8
    route: torch.Tensor | None = None
    lake_flag: torch.Tensor | None = None
    disch_track: torch.Tensor = field(default=None, repr=False)
    mom_track: torch.Tensor = field(default=None, repr=False)
    samp: torch.Tensor = field(default=None, repr=False)
    acc: torch.Tensor = field(default=None, repr=False)
    pending: torch.Tensor = field(default=None, repr=False)
    lake_load: torch.Tensor = field(default=None, repr=False)
    lake_id: torch.Tensor = field(default=None, repr=False)
    lake_room: torch.Tensor = field(default=None, repr=False)
    base: torch.Tensor = field(default=None, repr=False)
    iteration: int = 0
    land_target: float | None = None
    deposit_on_exit: bool = False
    roots: torch.Tensor | None = None
    inflow_volume: float = 0.0
    iso_acc: torch.Tensor | None = None  # maps.py:1369-1371: surface change not yet compensated
    base_at: int | None = None  # the iteration of the last sea-mask refresh (maps.py:420)
    _halo: HaloTables | None = field(default=None, repr=False)
    _side: dict = field(default_factory=dict, repr=False)
    _owner: np.ndarray | None = field(default=None, repr=False)
    _uplift_mean: dict = field(default_factory=dict, repr=False)
    _mask_host: np.ndarray | None = field(default=None, repr=False)
    _select_work: torch.Tensor | None = field(default=None, repr=False)
    _fast: object = field(default=None, repr=False)
    _ordered: dict = field(default_factory=dict, repr=False)

    def __post_init__(self):
        F, NE = self.height.shape[0], self.height.shape[1]
        dev = self.height.device
        if dev.type != "cuda":
            raise ValueError("ErosionState lives on the GPU: pass device tensors (or use from_grid / window)")
        z = lambda: torch.zeros((F, NE, NE), dtype=torch.float64, device=dev)
        if self.mode == "check":
            self._check_buffers()
        if self.pending is None:
            self.pending = z()
        if self.lake_load is None:
            self.lake_load = z()
        if self.lake_id is None:
            self.lake_id = torch.full((F, NE, NE), -1, dtype=torch.int32, device=dev)
        if self.lake_room is None:
            self.lake_room = torch.zeros(1, dtype=torch.float64, device=dev)
        if self.base is None:
            self.base = z()
        if NE != self.N + 2 * self.H:
            raise ValueError("array width must be N + 2H")
        if self.mode not in ("check", "fast"):
            raise ValueError(f"mode must be 'check' or 'fast', got {self.mode!r}")
        if self.check_apply not in ("ordered", "serial"):
            raise ValueError(f"check_apply must be 'ordered' or 'serial', got {self.check_apply!r}")
        for a, shape, n in ((self.sediment, (F, NE, NE), "sediment"), (self.discharge, (F, NE, NE), "discharge"),
                            (self.momentum, (F, NE, NE, 2), "momentum"), (self.hardness, (F, NE, NE), "hardness"),
                            (self.precip, (F, NE, NE), "precip"), (self.evap, (F, NE, NE), "evap"),
                            (self.uplift, (F, NE, NE), "uplift"), (self.mask, (F, NE, NE), "mask"),
                            (self.metric, (F, NE, NE, 3), "metric"), (self.metric_inv, (F, NE, NE, 3), "metric_inv")):
            if tuple(a.shape) != shape:
                raise ValueError(f"{n}: shape {tuple(a.shape)}, expected {shape}")
        self._mask_host = self.mask.cpu().numpy()

    # -- constructors --------------------------------------------------------
    @classmethod
    def from_grid(cls, grid, bedrock_m, hardness, precip, evap, uplift_m, eparams=None, mask=None, device="cuda",
                  mode: str = "check", metric=None, metric_inv=None) -> "ErosionState":
        """Global 6-face state from extended arrays / FaceFields in metres
        (bake/globe/erosion/maps.py:152-191).  Inputs may also be tensors
        already on the device (taken without a host round trip), and the
        metric tensors may be given (default: the grid's)."""
        unit = height_unit(grid, eparams)
        F, NE = 6, grid.NE
        dev = torch.device(device)
        m = np.ones((F, NE, NE), np.uint8) if mask is None else _host(mask, np.uint8)
        zeros = lambda *extra: torch.zeros((F, NE, NE) + extra, dtype=torch.float64, device=dev)
        return cls(height=_to_dev(bedrock_m, np.float64, dev, unit), sediment=zeros(), discharge=zeros(),
                   momentum=zeros(2), hardness=_to_dev(hardness, np.float32, dev),
                   precip=_to_dev(precip, np.float32, dev), evap=_to_dev(evap, np.float32, dev),
                   uplift=_to_dev(uplift_m, np.float64, dev, unit), mask=_dev(m, dev),
                   metric=_to_dev(grid.metric if metric is None else metric, np.float32, dev),
                   metric_inv=_to_dev(grid.metric_inv if metric_inv is None else metric_inv, np.float32, dev),
                   N=grid.N, H=grid.H, spherical=True, height_unit_m=unit, grid=grid, mode=mode)

    @classmethod
    def window(cls, height_m, sediment_m, discharge, momentum, hardness, precip, evap, uplift_m, mask, metric,
               metric_inv, H, height_unit_m, deposit_on_exit=False, route_m=None, device="cuda", mode: str = "check"):
        """Single-face window (F = 1, spherical False) from (NE, NE[, C])
        arrays in metres (bake/globe/erosion/maps.py:193-240)."""
        NE = np.asarray(height_m).shape[0]
        a = lambda x, dt, extra=(): _dev(_host(x, dt).reshape((1, NE, NE) + tuple(extra)), device)
        hu = float(height_unit_m)
        return cls(height=a(_host(height_m, np.float64) / hu, np.float64),
                   sediment=a(_host(sediment_m, np.float64) / hu, np.float64), discharge=a(discharge, np.float64),
                   momentum=a(momentum, np.float64, (2,)), hardness=a(hardness, np.float32), precip=a(precip, np.float32),
                   evap=a(evap, np.float32), uplift=a(_host(uplift_m, np.float64) / hu, np.float64),
                   mask=a(mask, np.uint8), metric=a(metric, np.float32, (3,)), metric_inv=a(metric_inv, np.float32, (3,)),
                   N=NE - 2 * H, H=H, spherical=False, height_unit_m=hu, deposit_on_exit=deposit_on_exit,
                   route=None if route_m is None else a(_host(route_m, np.float64) / hu, np.float64), mode=mode)

    # -- views ---------------------------------------------------------------
    @property
    def F(self) -> int:
        return self.height.shape[0]

    @property
    def NE(self) -> int:
        return self.height.shape[1]

    @property
    def device(self):
        return self.height.device

    @property
    def interior(self):
        return (slice(None), slice(self.H, self.H + self.N), slice(self.H, self.H + self.N))

    def surface(self) -> torch.Tensor:
        return self.height + self.sediment

    def height_m(self) -> np.ndarray:
        return self.height.cpu().numpy() * self.height_unit_m

    def sediment_m(self) -> np.ndarray:
        return self.sediment.cpu().numpy() * self.height_unit_m

    def total_mass(self) -> float:
        """Sum of height + sediment + pending + lake_load over active
        interior cells (maps.py:322-326), summed on the host in float64."""
        act = self._mask_host[self.interior] == pk.MASK_ACTIVE
        tot = (self.height + self.sediment + self.pending + self.lake_load)[self.interior].cpu().numpy()
        return float(np.sum(tot[act]))

    def fields(self, names=("height", "sediment", "discharge", "momentum")) -> dict:
        """Output FaceFields in metres, on the host (maps.py:738-752)."""
        if self.grid is None:
            raise ValueError("fields() needs a Grid")
        out = {}
        for n in names:
            if n == "height":
                out[n] = FaceField(self.grid, self.height_m().astype(np.float32), name="height")
            elif n == "sediment":
                out[n] = FaceField(self.grid, self.sediment_m().astype(np.float32), name="sediment")
            elif n == "discharge":
                out[n] = FaceField(self.grid, self.discharge.cpu().numpy().astype(np.float32), name="discharge")
            elif n == "momentum":
                out[n] = FaceField(self.grid, self.momentum.cpu().numpy().astype(np.float32), is_vector=True, name="momentum")
        return out

    def _check_buffers(self) -> None:
        """The check mode's working arrays -- the packed record, the live acc,
        the discharge / momentum tracks -- allocated on first use (the fast
        mode keeps its own in its workspace: at N_c=8192 these are 50 GB)."""
        F, NE, dev = self.height.shape[0], self.height.shape[1], self.height.device
        if self.disch_track is None:
            self.disch_track = torch.zeros((F, NE, NE), dtype=torch.float64, device=dev)
        if self.mom_track is None:
            self.mom_track = torch.zeros((F, NE, NE, 2), dtype=torch.float64, device=dev)
        if self.samp is None:
            self.samp = torch.zeros((F, NE, NE, pk.NS), dtype=torch.float32, device=dev)
        if self.acc is None:
            self.acc = torch.zeros((F, NE, NE), dtype=torch.float64, device=dev)

    def pack(self) -> None:
        self._check_buffers()
        pk.pack_samples(self.samp, self.height, self.sediment, self.route, self.route is not None, self.discharge,
                        self.momentum, self.evap, self.hardness, self.metric, self.metric_inv, self.base, self.lake_flag)

    # -- halos ---------------------------------------------------------------
    def halo_tables(self) -> HaloTables:
        if self._halo is None:
            self._halo = HaloTables(self.grid, self.device)
        return self._halo

    def exchange_halos(self) -> None:
        """Refresh the halos of the evolving fields on the device
        (maps.py:329-340): height, sediment, discharge by the cubic weights,
        momentum rotated into each destination face, route bilinear."""
        if not self.spherical:
            return
        if self.grid is None:
            raise ValueError("spherical ErosionState needs its Grid for halo exchange")
        ht = self.halo_tables()
        for arr in (self.height, self.sediment, self.discharge):
            ht.exchange(arr)
        ht.exchange(self.momentum, vector=True)
        if self.route is not None:
            ht.exchange(self.route, linear=True)

    def uplift_mean(self, cap) -> float:
        """The active-interior mean of the (capped) uplift field that
        apply_uplift removes (maps.py:1137-1147).  The field is an input that
        never changes during a run, so the mean is taken once per cap with
        numpy's pairwise mean -- the reference's own reduction."""
        key = None if cap is None else float(cap)
        if key not in self._uplift_mean:
            u = self.uplift.cpu().numpy()
            if cap is not None:
                u = np.minimum(u, float(cap))
            mean = 0.0
            if self.spherical:
                iact = self._mask_host[self.interior] == pk.MASK_ACTIVE
                mean = float(u[self.interior][iact].mean()) if iact.any() else 0.0
            self._uplift_mean[key] = mean
        return self._uplift_mean[key]

    def select_work(self) -> torch.Tensor:
        if self._select_work is None:
            n = int(lib().pg_select_work_bytes())
            self._select_work = torch.zeros((n + 7) // 8, dtype=torch.int64, device=self.device)
        return self._select_work

    def side_buffer(self, name: str) -> torch.Tensor:
        """Device scratch of the side passes, allocated on first use and kept
        (so a captured graph sees the same buffers on every replay)."""
        t = self._side.get(name)
        if t is None:
            shape, dev = tuple(self.height.shape), self.device
            if name == "wasting":
                t = torch.empty((4,) + shape, dtype=torch.float64, device=dev)
            elif name == "cmask":
                t = torch.empty(shape, dtype=torch.uint8, device=dev)
            elif name == "sums":
                t = torch.zeros((3, 4), dtype=torch.float64, device=dev)
            elif name == "reduce":
                t = torch.empty(int(lib().pg_interior_sums_work_bytes()) // 8, dtype=torch.float64, device=dev)
            elif name == "sqrt_det":
                if self.grid is None:
                    raise ValueError("the isostasy smoother needs the state's Grid")
                t = _dev(np.ascontiguousarray(self.grid.sqrt_det_metric, np.float32), dev)
            elif name in ("talus", "s0", "load", "load2", "ecap"):
                t = torch.empty(shape, dtype=torch.float64, device=dev)
            else:
                raise KeyError(name)
            self._side[name] = t
        return t


    # -- the sea mask (maps.py:360-421) --------------------------------------
    def _sea_tables(self) -> dict:
        """Each interior cell's area (cubesphere.py:509-517: float32 m^2) as
        an exact integer multiple of 2^-shift m^2, so the device sums a
        component's area exactly; the owner table and the union-find's
        scratch; built once."""
        t = self._side.get("sea")
        if t is None:
            H, N = self.H, self.N
            area = (np.asarray(self.grid.sqrt_det_metric).astype(np.float64) * float(self.grid.cell_size_m) ** 2)
            area = area.astype(np.float32)[:, H:H + N, H:H + N].astype(np.float64)
            _, e = np.frexp(area)
            shift = 24 - int(e.min())
            fx = np.ldexp(area, shift)
            assert np.array_equal(fx, np.floor(fx)) and fx.sum() < 2.0 ** 62
            dev = self.device
            n = int(lib().pg_sea_work_bytes(N))
            t = {"area_fx": _dev(fx.astype(np.int64), dev), "shift": shift, "area_sum": float(area.sum()),
                 "owner": _dev(self.owner_table(), dev),
                 "work": torch.empty((n + 7) // 8, dtype=torch.int64, device=dev)}
            self._side["sea"] = t
        return t

    def refresh_base(self, min_fraction: float) -> dict:
        """Recompute the per-cell base level (maps.py:360-421): 0 on the open
        ocean and on land, a closed basin's deepest surface inside it; the
        halos copy their owners.  The components, areas and floors are
        found on the device (pg_refresh_base); a component's area is summed
        exactly, where the reference sums float64 in cell order."""
        info = {"closed_cells": 0, "closed_basins": 0, "sea_cells": 0, "deepest_basin_floor": 0.0}
        if not self.spherical or self.grid is None:
            self.base.zero_()
            self.base_at = self.iteration
            return info
        t = self._sea_tables()
        thr = math.ldexp(float(min_fraction) * t["area_sum"], t["shift"])
        got = pk.refresh_base(self.height, self.sediment, t["owner"], t["area_fx"], thr, self.N, self.H, self.base,
                              t["work"])
        got.pop("rounds")
        info.update(got)
        self.base_at = self.iteration
        return info

    # -- routing surface (maps.py:342-358, route.py:30-117) ------------------
    def owner_table(self) -> np.ndarray:
        """Flat index of the interior cell owning every extended cell
        (``grid.owner`` in spherical mode, identity for a window)."""
        if self._owner is None:
            n = self.F * self.NE * self.NE
            own = self.grid.owner if self.spherical else np.arange(n, dtype=np.int64).reshape(self.height.shape)
            self._owner = np.ascontiguousarray(own, dtype=np.int64)
        return self._owner

    def owner_device(self) -> torch.Tensor | None:
        """The owner table on the device (int64, (F, NE, NE)) on the globe;
        None on a window, whose halo has no owners."""
        if not self.spherical:
            return None
        t = getattr(self, "_owner_dev", None)
        if t is None:
            t = _dev(self.owner_table(), self.device)
            self._owner_dev = t
        return t

    def _flood_tables(self) -> dict:
        """The flood's static cell classes and the extra in-edges of the cells
        a cross-face halo cell maps onto (see pg_flood_*), built once on the
        host from the mask and the owner table."""
        t = self._side.get("flood")
        if t is not None:
            return t
        F, NE, H, N = self.F, self.NE, self.H, self.N
        n = F * NE * NE
        own = self.owner_table().reshape(n)
        m = self._mask_host.reshape(n)
        idx = np.arange(n, dtype=np.int64)
        f, rem = np.divmod(idx, NE * NE)
        ei, ej = np.divmod(rem, NE)
        ring = (ei == 0) | (ej == 0) | (ei == NE - 1) | (ej == NE - 1)
        visited0 = (m == pk.MASK_OUTSIDE) | ring
        inter = (ei >= H) & (ei < H + N) & (ej >= H) & (ej < H + N)
        target = ~visited0 & (own == idx)
        offs = [(1, 0), (1, 1), (0, 1), (-1, 1), (-1, 0), (-1, -1), (0, -1), (1, -1)]  # particle.py:102
        cand = np.flatnonzero(inter & ~visited0)
        seed = np.zeros(n, bool)
        for di, dj in offs:
            nb = own[cand + di * NE + dj]
            nf, nrem = np.divmod(nb, NE * NE)
            ni, nj = np.divmod(nrem, NE)
            out = (m[nb] == pk.MASK_OUTSIDE) | (ni < H) | (ni >= H + N) | (nj < H) | (nj >= H + N)
            seed[cand[out]] = True
        cls = np.zeros(n, np.uint8)
        cls[target] = 2
        cls[seed] = 1
        # irregular edges p -> owner(p + d) != p + d into assignable cells
        src = np.flatnonzero(cls > 0)
        es, ed = [], []
        for di, dj in offs:
            q = src + di * NE + dj
            nb = own[q]
            irr = (nb != q) & (cls[nb] == 2)
            es.append(src[irr])
            ed.append(nb[irr])
        es, ed = np.concatenate(es), np.concatenate(ed)
        order = np.lexsort((es, ed))
        es, ed = es[order], ed[order]
        spec, start = np.unique(ed, return_index=True)
        cls[spec] = 3
        ptr = np.append(start, ed.size).astype(np.int32)
        dev = self.device
        t = {"cls": _dev(cls.reshape(self.height.shape), dev), "n_spec": int(spec.size),
             "spec_cell": _dev(spec.astype(np.int64), dev) if spec.size else None,
             "spec_ptr": _dev(ptr, dev) if spec.size else None,
             "spec_src": _dev(es.astype(np.int64), dev) if spec.size else None,
             "s": torch.empty_like(self.height), "W": torch.empty_like(self.height),
             "flag": torch.zeros(1, dtype=torch.int32, device=dev)}
        self._side["flood"] = t
        return t

    def refresh_route(self, eps: float, batch: int = 32) -> int:
        """Recompute the epsilon-filled routing surface from the current
        terrain on the device (maps.py:351-358 with route.py:30-117), then
        refresh its halos (linear).  Returns the relaxation sweeps taken."""
        t = self._flood_tables()
        pk.flood_init(self.height, self.sediment, self.base, t["cls"], self.H, self.N, t["s"], t["W"], t["flag"])
        sweeps, cap = 0, self.F * self.NE * self.NE + batch
        while True:
            pk.flood_sweeps(t["s"], t["cls"], t["W"], float(eps), t["spec_cell"], t["spec_ptr"], t["spec_src"], batch,
                            t["flag"])
            sweeps += batch
            if int(t["flag"].item()) == 0:
                break
            if sweeps > cap:
                raise RuntimeError("routing flood did not converge")
        if self.route is None:
            self.route = torch.empty_like(self.height)
        pk.flood_finish(t["s"], t["W"], self.route)
        if self.spherical:
            self.halo_tables().exchange(self.route, linear=True)
        return sweeps

    def refresh_lakes_window(self, min_depth: float, lake_evap: float = 0.0) -> dict:
        """maps.py:629-672 without the evaporation balance: every active
        interior cell the routing flood fills deeper than ``min_depth`` (cell
        units) is a flagged, overflowing lake."""
        if float(lake_evap) > 0.0:
            raise NotImplementedError("the window lakes' evaporation balance (erosion.window_lake_evap) is outside "
                                      "the HIP erosion path")
        # written in place: a captured fast-mode graph keeps reading the same buffer
        if self.lake_flag is None:
            self.lake_flag = torch.zeros(self.height.shape, dtype=torch.uint8, device=self.device)
        flag = self.lake_flag
        flag.zero_()
        if self.route is None:
            return {"lake_cells": 0}
        lake = ((self.route - (self.height + self.sediment)) > float(min_depth)) & (self.mask == pk.MASK_ACTIVE)
        inter = self.interior
        flag[inter] = lake[inter].to(torch.uint8)
        return {"lake_cells": int(lake[inter].sum().item())}


# ---------------------------------------------------------------------------
# spawning (maps.py:793-833)
# ---------------------------------------------------------------------------
def spawn_weights(state: ErosionState):
    """Flat indices and weights of the spawnable cells: interior, active,
    land (surface >= base), precip > 0 -- on the host, as the reference."""
    surf = (state.height + state.sediment).cpu().numpy()
    base = state.base.cpu().numpy()
    precip = state.precip.cpu().numpy()
    inter = np.zeros(surf.shape, bool)
    inter[state.interior] = True
    ok = inter & (state._mask_host == pk.MASK_ACTIVE) & (surf >= base) & (precip > 0)
    idx = np.flatnonzero(ok)
    return idx, precip.reshape(-1)[idx].astype(np.float64)


def spawn_particles(state: ErosionState, n_particles: int, rng: np.random.Generator):
    """Deterministic spawn, the reference's draw order: sorted uniforms for
    the inverse CDF, then the offsets (maps.py:809-833).  The inverse-CDF
    search runs on the device.  Returns device tensors (face, x, y) and the
    spawn volume, or None."""
    idx, w = spawn_weights(state)
    if idx.size == 0 or n_particles <= 0:
        return None
    total = float(w.sum())
    if total <= 0.0:
        return None
    cdf = np.cumsum(w)
    u = np.sort(rng.random(n_particles))
    off = rng.random((n_particles, 2))
    dev = state.device
    k = pk.spawn_cells(_dev(cdf, dev), _dev(u, dev)).cpu().numpy()
    flat = idx[k]
    NE = state.NE
    f = flat // (NE * NE)
    rem = flat - f * NE * NE
    ei = rem // NE
    ej = rem - ei * NE
    x = (ei - state.H).astype(np.float64) + off[:, 0]
    y = (ej - state.H).astype(np.float64) + off[:, 1]
    return _dev(f.astype(np.int64), dev), _dev(x, dev), _dev(y, dev), total / n_particles


def ocean_min_fraction(params) -> float:
    """``hydro.ocean_min_fraction`` (maps.py:770-774); 0.02 without a hydro
    group."""
    hydro = getattr(params, "hydro", None)
    return float(hydro.ocean_min_fraction) if hydro is not None and hasattr(params, "world") else 0.02


def _check_scope(ep, state):
    """The limits shared by the fast and differentiable modes (none beyond
    the ones each checks itself)."""
    return None


def require_plain_side_passes(ep, what: str) -> None:
    """Refuse the side passes that a caller with its own iteration loop
    (batched windows, the differentiable mode) does not run."""
    on = [n for n, v in (("thermal_rate", float(ep.thermal_rate) != 0.0), ("creep_rate", float(ep.creep_rate) > 0.0),
                         ("slope_limit_erode", float(getattr(ep, "slope_limit_erode", 0.0)) > 0.0),
                         ("slope_limit_deposit", float(getattr(ep, "slope_limit_deposit", 0.0)) > 0.0),
                         ("isostasy", float(getattr(ep, "isostasy", 0.0)) > 0.0),
                         ("flood_every", int(ep.flood_every) > 0)) if v]
    if on:
        raise NotImplementedError(f"{what} runs the plain profile: switch off {', '.join(on)}")


def trace_params(ep, state, volume0, P, chunk, disc_sat, mom_scale, max_steps, lake_park) -> TraceParams:
    """trace_particles' scalar arguments as run_iteration computes them
    (maps.py:971-997)."""
    unit = state.height_unit_m
    return TraceParams(
        dt=float(ep.dt), density=float(ep.density), friction=float(ep.friction), deposition_rate=float(ep.deposition_rate),
        evap_rate=float(ep.evap_rate), k_mom=float(ep.k_mom), mom_scale=float(mom_scale), k_disc=float(ep.k_disc),
        disc_saturation=float(disc_sat), disc_exponent=float(ep.disc_exponent), slope_gain=float(ep.slope_gain),
        slope_sat=float(ep.slope_saturation), erodibility=float(ep.erodibility),
        cover_depth=cell_units(ep, "cover_depth", unit), min_volume=float(ep.min_volume_frac) * float(volume0),
        max_steps=int(max_steps), max_erode=cell_units(ep, "max_erode", unit), conc_scale=float(chunk / (volume0 * P)),
        pit_steps=int(ep.pit_steps), deposit_on_exit=int(bool(state.deposit_on_exit)),
        ocean_rate=float(ep.ocean_deposition_rate), ocean_steps=int(ep.ocean_steps),
        fan_room=cell_units(ep, "fan_room", unit), dep_floor=cell_units(ep, "dep_floor_m", unit),
        lake_trap=float(ep.lake_trap), lake_park=float(lake_park), lateral_rate=float(getattr(ep, "lateral_rate", 0.0)))


def discharge_scales(ep, state, volume0, n_particles, n_rain=None):
    """disc_saturation and mom_scale in cells of upstream rain (maps.py:911-921)."""
    disc_sat = float(ep.disc_saturation)
    mom_scale = 1.0
    if float(ep.disc_saturation_cells) > 0.0 or float(ep.momentum_saturation_cells) > 0.0:
        total = float(volume0) * float(n_particles)
        if n_rain is None:
            n_rain = int(spawn_weights(state)[0].size)
        n_rain = max(int(n_rain), 1)
        rain_per_cell = max(total - float(state.inflow_volume), 0.0) / n_rain
        rain_per_cell = rain_per_cell if rain_per_cell > 0.0 else total / n_rain
        if float(ep.disc_saturation_cells) > 0.0:
            disc_sat = float(ep.disc_saturation_cells) * rain_per_cell
        if float(ep.momentum_saturation_cells) > 0.0:
            mom_scale = float(volume0) / (float(ep.momentum_saturation_cells) * rain_per_cell)
    return disc_sat, mom_scale


def slope_limit_cap(state: ErosionState, ep):
    """maps.py:923-927: soillib's per-cell erosion cap of this iteration on
    the surface it starts from (particle.py:1045-1076), or None when
    ``slope_limit_erode`` is off."""
    if not float(getattr(ep, "slope_limit_erode", 0.0)) > 0.0:
        return None
    ecap = state.side_buffer("ecap")
    pk.slope_erode_cap(state.height, state.sediment, state.mask, float(ep.slope_limit_erode),
                       float(ep.slope_limit_exit), ecap)
    return ecap


def slope_limited_deposit(ep, unit: float) -> float:
    """maps.py:928-931: iter_deposit, capped by soillib's deposition limit
    when ``slope_limit_deposit`` is on."""
    iter_deposit = cell_units(ep, "iter_deposit", unit)
    if float(getattr(ep, "slope_limit_deposit", 0.0)) > 0.0:
        iter_deposit = min(iter_deposit, float(ep.slope_limit_deposit) * math.sqrt(2.0) * 0.3)
    return iter_deposit


def lake_park_of(state, ep) -> float:
    """1 when lake loads are parked (maps.py:996, lake_fill_args)."""
    return 1.0 if (state.spherical and bool(getattr(ep, "lake_fill", False)) and bool(getattr(ep, "lake_balance", False))) else 0.0


# ---------------------------------------------------------------------------
# one iteration (maps.py:839-1047)
# ---------------------------------------------------------------------------
def _scratch(state: ErosionState, name: str, n: int, dtype) -> torch.Tensor:
    """A per-state device scratch tensor of at least n elements, reused across iterations."""
    t = state._ordered.get(name)
    if t is None or t.numel() < n or t.dtype != dtype:
        t = torch.empty(max(n, 1), dtype=dtype, device=state.device)
        state._ordered[name] = t
    return t


def run_iteration(state: ErosionState, params, iteration_key, n_particles=None, particles_per_cell=None,
                  rng_stage="erosion", log=None, diag=None, spawn=None) -> dict:
    """The particle part of an iteration: spawn (rain + re-injected
    stockpiles), trace and apply, fold, EMA.  Same arguments, stats keys and
    ``diag`` hook as the reference (maps.py:839-870); ``diag`` is called per
    chunk with device tensors.  ``spawn`` optionally replaces the rain draw
    with precomputed device tensors ``(face, x, y, volume0)``."""
    if state.mode == "fast":
        from . import fast

        return fast.run_iteration(state, params, iteration_key, n_particles=n_particles,
                                  particles_per_cell=particles_per_cell, rng_stage=rng_stage, log=log)
    ep = _eparams(params)
    if int(getattr(ep, "brush_radius", 0)) > 1:
        raise NotImplementedError("the brush is an extension of the fast mode; the check mode reproduces the "
                                  "reference, whose droplets change one cell per step")
    key = iteration_key if isinstance(iteration_key, (tuple, list)) else (int(iteration_key),)
    rng = params.rng(rng_stage, *key)
    t0 = time.time()
    dev = state.device
    n_cells = int(np.count_nonzero(state._mask_host[state.interior] == pk.MASK_ACTIVE))
    if n_particles is None:
        ppc = ep.particles_per_cell if particles_per_cell is None else particles_per_cell
        n_particles = max(1, int(round(ppc * n_cells)))
    sp = spawn_particles(state, n_particles, rng) if spawn is None else spawn
    stats = {"particles": 0, "entries": 0, "steps_mean": 0.0}
    if sp is None:
        return stats
    sp_face, sp_x, sp_y, volume0 = sp
    sp_sed = torch.zeros(sp_face.shape[0], dtype=torch.float64, device=dev)
    sp_dir = torch.zeros(sp_face.shape[0], dtype=torch.float64, device=dev)
    # re-inject the pending stockpiles after the rain, in cell order (maps.py:886-903)
    pend = state.pending.reshape(-1).cpu().numpy()
    pend_cells = np.flatnonzero(pend > 0.0)
    released = 0.0
    if pend_cells.size:
        NE = state.NE
        pf = pend_cells // (NE * NE)
        rem = pend_cells - pf * NE * NE
        pei = rem // NE
        pej = rem - pei * NE
        loads = pend[pend_cells]
        released = float(loads.sum())
        sp_face = torch.cat([sp_face, _dev(pf.astype(np.int64), dev)])
        sp_x = torch.cat([sp_x, _dev((pei - state.H) + 0.5, dev)])
        sp_y = torch.cat([sp_y, _dev((pej - state.H) + 0.5, dev)])
        sp_sed = torch.cat([sp_sed, _dev(loads, dev)])
        sp_dir = torch.cat([sp_dir, _dev(rng.random(pend_cells.size) * (2.0 * np.pi), dev)])
        state.pending.reshape(-1)[_dev(pend_cells, dev)] = 0.0
    state.pack()
    state.acc.zero_()
    disc_sat, mom_scale = discharge_scales(ep, state, volume0, n_particles)
    unit = state.height_unit_m
    ecap = slope_limit_cap(state, ep)
    iter_deposit = slope_limited_deposit(ep, unit)
    max_steps = max_steps_of(ep, state.N)
    lateral = float(getattr(ep, "lateral_rate", 0.0)) > 0.0
    cap = pk.change_list_cap(max_steps, lateral)
    P = sp_face.shape[0]
    chunk = chunk_size(ep, P)
    prm = trace_params(ep, state, volume0, P, chunk, disc_sat, mom_scale, max_steps, lake_park_of(state, ep))
    cl_cell = torch.empty(chunk * cap, dtype=torch.int32, device=dev)
    cl_delta = torch.empty(chunk * cap, dtype=torch.float64, device=dev)
    cl_vol = torch.empty(chunk * cap, dtype=torch.float32, device=dev)
    cl_mom = torch.empty((chunk * cap, 2), dtype=torch.float32, device=dev)
    cl_count = torch.zeros(chunk, dtype=torch.int64, device=dev)
    cl_lo = torch.zeros(chunk, dtype=torch.int64, device=dev)
    cl_hi = torch.zeros(chunk, dtype=torch.int64, device=dev)
    sp_death = torch.zeros(chunk, dtype=torch.int8, device=dev)
    n_chunks = -(-P // chunk)
    outs = torch.zeros((n_chunks, 4), dtype=torch.float64, device=dev)
    counts = torch.zeros(7, dtype=torch.int64, device=dev)
    iter_erode = cell_units(ep, "iter_erode", unit)
    dep_floor = cell_units(ep, "dep_floor_m", unit)
    use_roots = state.roots is not None
    t_trace = t_apply = 0.0
    lake_park = lake_park_of(state, ep)
    # the one-call chunk loop hands a chunk with lake parks (or one that does not converge) to the
    # serial apply on the device; the chunk-by-chunk loop (diag hook) takes the serial apply for lake parks
    one_call = state.check_apply == "ordered" and diag is None
    ordered = state.check_apply == "ordered" and (lake_park == 0.0 or one_call)
    if ordered:
        key = (chunk, cap, int(state.check_max_sweeps))
        if state._ordered.get("key") != key:
            nbytes = pk.ordered_workspace_bytes(chunk, cap, int(state.check_max_sweeps))
            state._ordered.update({"key": key, "work": torch.empty(nbytes, dtype=torch.uint8, device=dev),
                                   "status": torch.zeros(4, dtype=torch.int32, device=dev)})
        ord_work, ord_status = state._ordered["work"], state._ordered["status"]
        ord_status.zero_()
    t1 = time.time()
    pk.set_trace_lanes(int(state.check_lanes))
    moves_dev = torch.zeros(1, dtype=torch.int64, device=dev) if one_call else None
    if one_call:
        # the whole chunk loop in one library call (no host round trip per chunk)
        pk.check_chunks(sp_face, sp_x, sp_y, sp_sed, sp_dir, volume0, chunk, state.samp, state.route is not None,
                        state.mask, state.N, state.H, state.spherical, prm, state.roots, use_roots, cl_cell, cl_delta,
                        cl_vol, cl_mom, cl_count, cl_lo, cl_hi, sp_death, state.height, state.sediment, state.acc,
                        state.pending, state.disch_track, state.mom_track, iter_erode, iter_deposit, float(ep.fan_slope),
                        dep_floor, float(ep.offshore_writeoff), ecap, ecap is not None, outs, counts, ord_status,
                        ord_work, _scratch(state, "ordered_drops", 4 * P, torch.float64), int(state.check_max_sweeps),
                        int(state.check_launched_sweeps), moves=moves_dev, timing=state.check_time_trace,
                        lake_load=state.lake_load, lake_id=state.lake_id, lake_room=state.lake_room)
    for ci, c0 in enumerate(range(0, P, chunk) if not one_call else ()):
        c1 = min(P, c0 + chunk)
        m = c1 - c0
        t1 = time.time()
        pk.trace_particles(sp_face[c0:c1], sp_x[c0:c1], sp_y[c0:c1], sp_sed[c0:c1], sp_dir[c0:c1], volume0, state.samp,
                           state.route is not None, state.mask, state.N, state.H, state.spherical, prm, state.roots,
                           use_roots, cl_cell, cl_delta, cl_vol, cl_mom, cl_count[:m], cl_lo[:m], cl_hi[:m], sp_death[:m])
        t2 = time.time()
        if diag is not None:
            diag(state, cl_cell, cl_vol, cl_count[:m], cap, sp_death[:m])
        if ordered:
            pk.apply_changes_ordered(cl_cell, cl_delta, cl_vol, cl_mom, cl_count[:m], cap, state.height, state.sediment,
                                     state.acc, state.pending, state.samp, state.disch_track, state.mom_track, state.mask,
                                     iter_erode, iter_deposit, float(ep.fan_slope), state.route is not None, dep_floor,
                                     float(ep.offshore_writeoff), ecap, ecap is not None, outs[ci], ord_status, ord_work,
                                     int(state.check_max_sweeps), int(state.check_launched_sweeps))
        else:
            pk.apply_changes(cl_cell, cl_delta, cl_vol, cl_mom, cl_count[:m], cap, state.height, state.sediment, state.acc,
                             state.pending, state.samp, state.disch_track, state.mom_track, state.mask, iter_erode,
                             iter_deposit, float(ep.fan_slope), state.route is not None, dep_floor,
                             float(ep.offshore_writeoff), ecap, ecap is not None, state.lake_load, state.lake_id,
                             state.lake_room, outs[ci])
        pk.chunk_stats(cl_count[:m], sp_death[:m], m, counts)
        t_trace += t2 - t1
        t_apply += time.time() - t2
    pk.fold_changes(state.height, state.sediment, state.acc, state.mask)
    pk.ema_update(state.discharge, state.momentum, state.disch_track, state.mom_track, state.mask, float(ep.ema))
    o = outs.cpu().numpy()
    cnt = counts.cpu().numpy()
    if ordered:
        ost = ord_status.cpu().numpy()
        err = pk.ordered_status_error(ost)
        if err is not None:
            raise RuntimeError(f"run_iteration: {err}")
        stats["sweeps_max"] = int(ost[1])
        stats["sweeps_total"] = int(ost[2])
        stats["serial_chunks"] = int(ost[3])
        if moves_dev is not None:
            stats["moves"] = int(moves_dev.item())
    # per-chunk partial results summed in chunk order, as the reference's loop does
    n_clamp = 0
    to_pending = lost = lost_offshore = 0.0
    for r in o:
        n_clamp += int(r[0])
        to_pending += float(r[1])
        lost += float(r[2])
        lost_offshore += float(r[3])
    entries = int(cnt[0])
    stats.update({"particles": int(P), "entries": entries, "steps_mean": entries / max(P, 1), "volume": float(volume0),
                  "seconds": time.time() - t0})
    stats["deaths"] = {n: int(c) for n, c in zip(pk.DEATH_NAMES, cnt[1:])}
    stats["clamped"] = n_clamp
    stats["to_pending"] = to_pending
    stats["released"] = float(released)
    stats["pending_total"] = float(state.pending[state.interior].sum().item())
    stats["deficit_out"] = lost
    stats["lost_offshore"] = lost_offshore
    stats["seconds_trace"] = t_trace
    stats["seconds_apply"] = t_apply
    stats["chunk"] = int(chunk)
    stats["disc_saturation"] = disc_sat
    stats["mom_scale"] = mom_scale
    if log is not None:
        log(f"  particles {P} volume {volume0:.3f} mean steps {stats['steps_mean']:.1f} deaths {stats['deaths']} "
            f"clamped {n_clamp} pending {stats['pending_total']:.1f} ({stats['seconds']:.2f}s)")
    return stats


# ---------------------------------------------------------------------------
# mass wasting (maps.py:1050-1085) and isostasy (maps.py:1325-1410)
# ---------------------------------------------------------------------------
def thermal_erosion(state: ErosionState, params) -> None:
    """One pass of talus-limited mass wasting, then (``creep_rate > 0``) the
    creep pass with talus 0 and submerged cells frozen (maps.py:1050-1070),
    on the device.  A pass at rate 0 moves nothing (every outflow is
    0.5 x 0 x excess and the state is rewritten unchanged), so it is skipped."""
    ep = _eparams(params)
    unit = state.height_unit_m
    rate, creep = float(ep.thermal_rate), float(ep.creep_rate)
    if rate == 0.0 and not creep > 0.0:
        return
    talus = state.side_buffer("talus")
    cap = cell_units(ep, "thermal_max", unit)
    dep_floor = cell_units(ep, "dep_floor_m", unit)
    if rate != 0.0:
        pk.talus(state.hardness, float(ep.talus_slope_soft), float(ep.talus_slope_hard), talus)
        _mass_wasting_pass(state, talus, rate, cap, None, dep_floor)
    if creep > 0.0:
        cmask = state.side_buffer("cmask")
        pk.creep_mask(state.height, state.sediment, state.base, state.mask, cmask)
        pk.scale(None, 0.0, talus)
        _mass_wasting_pass(state, talus, creep, cap, cmask, dep_floor)


def _mass_wasting_pass(state: ErosionState, talus, rate: float, cap: float, mask=None, dep_floor: float = pk.DEP_FLOOR):
    """maps.py:1073-1085: thermal_pass_a, _r, _b and the write-back."""
    mask = state.mask if mask is None else mask
    pk.mass_wasting_pass(state.height, state.sediment, state.metric, talus, mask, rate, cap, dep_floor,
                         state.side_buffer("wasting"))


_KAPPA_CACHE: dict = {}


def smoothing_kappa(state: ErosionState) -> float:
    """The largest monotone explicit step of the Laplace-Beltrami smoother,
    0.9 / lam_max with lam_max by 40 power iterations from seeded white
    noise (maps.py:1325-1363), on the device; cached per grid.  The norms
    are device reductions (fixed order), so kappa agrees with numpy's to
    rounding."""
    grid = state.grid
    if grid is None:
        raise ValueError("the isostasy smoother needs the state's Grid")
    key = (grid.N, grid.H, float(grid.cell_size_m))
    if key not in _KAPPA_CACHE:
        rng = np.random.default_rng(0)
        N, H, NE = grid.N, grid.H, grid.NE
        host = np.zeros((6, NE, NE))
        host[:, H:H + N, H:H + N] = rng.standard_normal((6, N, N))
        dev = state.device
        data = _dev(host, dev)
        ht = state.halo_tables()
        ht.exchange(data)
        Lv = torch.empty_like(data)
        sq = state.side_buffer("sqrt_det")
        sums = state.side_buffer("sums")
        red = state.side_buffer("reduce")
        cs2 = float(grid.cell_size_m) ** 2
        c2, lam = float(grid.cell_size_m) ** 2, 8.0
        for _ in range(40):
            ht.exchange(data, linear=True)
            pk.laplacian(data, state.metric_inv, sq, cs2, 1, 0.0, c2, Lv)
            pk.interior_sums(Lv, None, N, H, 1, red, sums[0])
            pk.interior_sums(data, None, N, H, 1, red, sums[1])
            s = sums[:2, 0].cpu().numpy()
            num, den = math.sqrt(float(s[0])), math.sqrt(float(s[1]))
            if num <= 0.0 or den <= 0.0:
                break
            lam = num / den
            pk.divide(Lv, num, data)
        _KAPPA_CACHE[key] = 0.9 / max(lam, 1e-6)
    return _KAPPA_CACHE[key]


def isostasy_steps(state: ErosionState, ep):
    """maps.py:1392-1397: (sigma in cells, kappa, diffusion steps)."""
    cell = float(state.height_unit_m)
    sig = float(ep.flexure_km) * 1000.0 / cell
    sig = min(sig, state.N / 8.0)
    kappa = smoothing_kappa(state)
    n = int(np.ceil(sig * sig / (2.0 * kappa))) if sig >= 0.5 else 0
    return sig, kappa, n


def apply_isostasy(state: ErosionState, params) -> dict:
    """Flexural isostatic response to the surface change accumulated in
    ``state.iso_acc`` (maps.py:1366-1410), on the device: the load -k x
    iso_acc, n explicit steps of the seam-aware Laplace-Beltrami smoother
    (linear halo exchange before each), the mean over the active interior
    removed, added to ``height`` of the active cells; iso_acc is zeroed.
    The stats' rebound extremes and load are device scalars in fast mode
    (nothing waits) and floats in check mode."""
    ep = _eparams(params)
    if not state.spherical:
        raise ValueError("isostasy is a global pass (spherical states only)")
    if state.iso_acc is None:
        state.iso_acc = torch.zeros_like(state.height)
    cell = float(state.height_unit_m)
    sig, kappa, n = isostasy_steps(state, ep)
    a, b = state.side_buffer("load"), state.side_buffer("load2")
    pk.scale(state.iso_acc, -float(ep.isostasy), a)
    ht = state.halo_tables()
    sq = state.side_buffer("sqrt_det")
    cs2 = float(state.grid.cell_size_m) ** 2
    c2 = cell * cell
    for _ in range(n):
        ht.exchange(a, linear=True)
        pk.laplacian(a, state.metric_inv, sq, cs2, 2, kappa, c2, b)
        a, b = b, a
    ht.exchange(a, linear=True)
    sums, red = state.side_buffer("sums"), state.side_buffer("reduce")
    pk.interior_sums(a, state.mask, state.N, state.H, 0, red, sums[0])       # mean over the active interior
    pk.interior_sums(a, None, state.N, state.H, 0, red, sums[1])             # extremes over the interior
    pk.interior_sums(state.iso_acc, None, state.N, state.H, 2, red, sums[2])  # load moved, |d|
    pk.rebound(state.height, a, state.mask, sums[0])
    pk.scale(None, 0.0, state.iso_acc)
    st = {"sigma_cells": sig, "diffusion_steps": n, "kappa": kappa}
    if state.mode == "fast":
        st["device_sums"] = sums
        return st
    v = sums.cpu().numpy()
    mean = v[0, 0] / v[0, 1] if v[0, 1] > 0 else 0.0
    st.update({"load_abs_cells": float(v[2, 0]), "rebound_max": float(v[1, 2] - mean),
               "rebound_min": float(v[1, 3] - mean)})
    return st


def prepare_side_passes(state: ErosionState, params) -> None:
    """Allocate the side passes' buffers and compute the smoother's step
    before a graph capture (both wait on the device)."""
    ep = _eparams(params)
    if float(ep.thermal_rate) != 0.0 or float(ep.creep_rate) > 0.0:
        for n in ("talus", "wasting", "cmask"):
            state.side_buffer(n)
    if float(getattr(ep, "slope_limit_erode", 0.0)) > 0.0:
        state.side_buffer("ecap")
    if state.spherical and float(getattr(ep, "isostasy", 0.0)) > 0.0:
        if state.iso_acc is None:
            state.iso_acc = torch.zeros_like(state.height)
        for n in ("s0", "load", "load2", "sums", "reduce", "sqrt_det"):
            state.side_buffer(n)
        isostasy_steps(state, ep)
    torch.cuda.synchronize()


# ---------------------------------------------------------------------------
# uplift, replay, datum (maps.py:1088-1213, 1413-1508)
# ---------------------------------------------------------------------------
def uplift_cap(state: ErosionState, ep):
    if not state.spherical or float(getattr(ep, "uplift_max_m", 0.0)) <= 0.0:
        return None
    return cell_units(ep, "uplift_max_m", state.height_unit_m)


def apply_uplift(state: ErosionState, cap=None) -> None:
    """height += min(uplift, cap) - mean on active cells, mean-free on the
    globe (maps.py:1088-1134)."""
    pk.apply_uplift(state.height, state.uplift, state.mask, cap, state.uplift_mean(cap))


def uplift_replay(state: ErosionState, ep):
    """Total uplift the run will apply, or None (maps.py:1159-1175)."""
    if not state.spherical or str(getattr(ep, "uplift_mode", "stack")) != "replay":
        return None
    cap = uplift_cap(state, ep)
    u = state.uplift if cap is None else torch.clamp(state.uplift, max=float(cap))
    act = state.mask == pk.MASK_ACTIVE
    total = torch.zeros_like(state.height)
    total[act] = float(int(ep.iterations)) * (u[act] - state.uplift_mean(cap))
    return total


def held_land_fraction(state: ErosionState, params) -> float:
    """maps.py:1433-1443."""
    if state.land_target is not None:
        return float(state.land_target)
    if float(getattr(getattr(params, "tectonics", None), "shelf_fraction", 0.0)) > 0.0:
        raise ValueError("shelf mode holds the bedrock's land fraction: set state.land_target")
    return float(params.world.land_fraction)


def hold_datum_device(state: ErosionState, land_fraction: float, level: float = 0.0) -> torch.Tensor:
    """The datum hold on the device; returns the shift as a device scalar
    (no host synchronisation).  q is the exact (M - round(lf M))-th smallest
    interior surface (maps.py:1484-1495)."""
    M = state.F * state.N * state.N
    n_land = min(max(int(round(float(land_fraction) * M)), 0), M)
    q = torch.zeros(1, dtype=torch.float64, device=state.device)
    k = M - 1 if n_land == 0 else (0 if n_land == M else M - n_land)
    check(lib().pg_select_kth_surface(pk._ptr(state.height), pk._ptr(state.sediment), state.F, state.N, state.H, k,
                                      pk._ptr(q), pk._ptr(state.select_work()), pk._stream()), "hold_datum")
    if n_land == 0:
        q += 1.0
    if float(level) != 0.0:
        q -= float(level)
    check(lib().pg_datum_shift(state.height.numel(), pk._ptr(state.height), pk._ptr(state.route), pk._ptr(state.base),
                               pk._ptr(q), pk._stream()), "datum_shift")
    return q


def hold_datum(state: ErosionState, land_fraction: float, level: float = 0.0) -> float:
    """maps.py:1446-1508; returns the shift applied (cell units)."""
    return float(hold_datum_device(state, land_fraction, level).item())


def start_replay(state: ErosionState, params):
    """maps.py:1178-1213."""
    ep = _eparams(params)
    total = uplift_replay(state, ep)
    if total is None:
        return None
    state.height -= total
    state.exchange_halos()
    inter = state.interior
    surf = (state.height + state.sediment)[inter]
    info = {"replay_max": float(total[inter].max().item()), "replay_min": float(total[inter].min().item()),
            # an exact count over M, correctly rounded once (numpy's mean of the boolean mask)
            "land_fraction": int((surf >= 0).sum().item()) / surf.numel(), "datum_shift": 0.0}
    if hasattr(params, "world"):
        info["datum_shift"] = hold_datum(state, held_land_fraction(state, params))
    return info


def sea_refresh_due(state: ErosionState, ep) -> bool:
    """Whether this iteration refreshes the sea mask (maps.py:1541-1543)."""
    sea_every = int(getattr(ep, "sea_mask_every", 0))
    return bool(state.spherical and sea_every > 0 and (state.base_at is None or state.iteration % sea_every == 0))


def refresh_route_and_lakes(state: ErosionState, params, due=None) -> None:
    """The routing flood (every ``flood_every``) and then a window's lakes
    (maps.py:1548-1566); ``due`` overrides the strides (False: neither)."""
    ep = _eparams(params)
    if due is False:
        return
    lakes = lakes_refresh_due(state, ep) if due is None else (not state.spherical and bool(getattr(ep, "window_lakes", False)))
    if route_refresh_due(state, ep) if due is None else int(ep.flood_every) > 0:
        state.refresh_route(cell_units(ep, "route_eps", state.height_unit_m))
    if lakes:
        hydro = getattr(params, "hydro", None)
        min_depth = float(hydro.lake_min_depth) if hydro is not None else 0.5
        state.refresh_lakes_window(min_depth / float(state.height_unit_m), float(getattr(ep, "window_lake_evap", 0.0)))


def route_refresh_due(state: ErosionState, ep) -> bool:
    """Whether this iteration refreshes the routing surface (maps.py:1548-1549)."""
    flood = int(ep.flood_every)
    return bool(flood > 0 and (state.route is None or state.iteration % flood == 0))


def lakes_refresh_due(state: ErosionState, ep) -> bool:
    """Whether this iteration refreshes a window's lakes (maps.py:1557-1566)."""
    return bool(not state.spherical and bool(getattr(ep, "window_lakes", False)) and int(ep.flood_every) > 0
                and (state.lake_flag is None or state.iteration % int(ep.flood_every) == 0))


def step(state: ErosionState, params, iteration_key, log=None, after_refresh=None, isostasy_due=None, sea_due=None,
         route_due=None, **kw) -> dict:
    """One iteration in the reference's order (maps.py:1516-1617): the sea
    mask (every ``sea_mask_every``), the routing flood (every
    ``flood_every``) and a window's lakes, particles, mass wasting, uplift,
    isostasy (every ``isostasy_every`` iterations), datum hold (globe), halo
    exchange.  ``isostasy_due``, ``sea_due`` and ``route_due`` (the routing
    flood and a window's lakes) override the stride's choice (graph capture:
    the sea mask and the flood are host-synchronous and run outside the
    graph)."""
    ep = _eparams(params)
    if state.spherical and bool(getattr(ep, "lake_balance", False)) and int(ep.flood_every) > 0:
        raise NotImplementedError("lake refreshes (erosion.lake_balance) are outside the HIP erosion path")
    if state.spherical and int(ep.glacial_every) > 0:
        raise NotImplementedError("the glacial pass is outside the HIP erosion path (set glacial_every = 0)")
    t0 = time.time()
    st_base = None
    if sea_refresh_due(state, ep) if sea_due is None else sea_due:
        # before the route: the flood seeds on the sea (maps.py:1541-1547)
        st_base = state.refresh_base(ocean_min_fraction(params))
    refresh_route_and_lakes(state, params, route_due)
    if after_refresh is not None:
        after_refresh(state)
    tr = time.time() - t0
    iso = state.spherical and float(getattr(ep, "isostasy", 0.0)) > 0.0
    if iso:
        if state.iso_acc is None:
            state.iso_acc = torch.zeros_like(state.height)
        s0 = state.side_buffer("s0")
        pk.surface(state.height, state.sediment, s0)
    st = run_iteration(state, params, iteration_key, log=log, **kw)
    t1 = time.time()
    thermal_erosion(state, params)
    if iso:
        pk.iso_accumulate(state.iso_acc, state.height, state.sediment, s0)
    apply_uplift(state, uplift_cap(state, ep))
    if iso:
        due = isostasy_due if isostasy_due is not None else \
            (state.iteration + 1) % max(int(getattr(ep, "isostasy_every", 10)), 1) == 0
        if due:
            st["isostasy"] = apply_isostasy(state, params)
    if state.spherical and hasattr(params, "world"):
        q = hold_datum_device(state, held_land_fraction(state, params), float(getattr(state, "sea_drop", 0.0)))
        st["datum_shift"] = q if state.mode == "fast" else float(q.item())
    state.exchange_halos()
    state.iteration += 1
    if st_base is not None:
        st["sea"] = st_base
    st["seconds_route"] = tr
    st["seconds_particles"] = t1 - t0 - tr
    st["seconds_total"] = time.time() - t0
    return st


__all__ = ["ErosionState", "run_iteration", "thermal_erosion", "apply_uplift", "uplift_replay", "start_replay",
           "apply_isostasy", "hold_datum", "step", "spawn_particles", "height_unit", "max_steps_of"]


class CheckAutotuner:
    """Picks the exact mode's droplets per wave from real iterations: the
    first len(candidates) iterations each run one setting with the trace
    launches timed (HIP events on the launch stream), normalised by the
    iteration's droplet moves; the fastest per move is kept.  Results do not
    depend on the setting, so no iteration is wasted.  ``step(run, key)``
    runs ``run(key)`` (one iteration of the state, returning its stats) and
    returns what it returns."""

    def __init__(self, state: ErosionState, candidates=(64, 32, 16, 8, 4)):
        self.state = state
        self.candidates = [int(c) for c in candidates]
        self.times: dict = {}
        self.best = None

    def step(self, run, iteration_key):
        if self.best is not None:
            return run(iteration_key)
        lanes = self.candidates[len(self.times)]
        timed = self.state.check_time_trace
        self.state.check_lanes = lanes
        self.state.check_time_trace = True
        pk.check_trace_time()
        out = run(iteration_key)
        ms, _ = pk.check_trace_time()
        self.state.check_time_trace = timed
        self.times[lanes] = ms / max(int(out["moves"]), 1) * 1e6  # ns per move
        if len(self.times) == len(self.candidates):
            self.finish()
        return out

    def finish(self):
        """Keep the fastest setting measured so far."""
        if self.times:
            self.best = min(self.times, key=self.times.get)
            self.state.check_lanes = self.best
        return self.best

