"""Erosion parameters, the seeded generator stream, and the length conversion.

The erosion stage is parameterised exactly like the reference's
``ErosionParams`` (bake/globe/config.py:486-583): same field names, same
defaults, same meaning.  Any object with these attributes (including the
reference's own ``WorldParams``/``ErosionParams``) can be handed to
:mod:`globe_hip.erosion.maps`; the dataclasses here are a self-contained
stand-in so the package runs without the reference installed.

``WorldParams.rng`` reproduces bake/globe/config.py:830-842: the entropy key
``[seed + stage_offset, len(keys), *keys]`` (masked to 64 bits) fed to
``numpy.random.default_rng``, so a run draws the same spawn stream as the
reference for the same seed and iteration key.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field

import numpy as np

#: per-stage seed offsets (bake/globe/config.py:23-33)
STAGE_SEED_OFFSETS = {
    "tectonics": 1000,
    "climate": 2000,
    "erosion": 3000,
    "hydro": 4000,
    "watersheds": 5000,
    "refine": 6000,
    "derive": 7000,
    "tiles": 8000,
    "stub": 9000,
}


@dataclass
class WorldGroup:
    """bake/globe/config.py:67-83 (the fields the erosion stage reads)."""

    seed: int = 0
    N_c: int = 1024
    cell_size_m: float = 9773.0
    land_fraction: float = 0.30
    halo: int = 4


@dataclass
class TectonicsGroup:
    """Only ``shelf_fraction`` is read here (the datum hold's land fraction,
    bake/globe/erosion/maps.py:1413-1443).  0 holds ``world.land_fraction``."""

    shelf_fraction: float = 0.0


@dataclass
class HydroGroup:
    """The fields the erosion stage reads: ``lake_min_depth`` (a window's
    lakes, bake/globe/erosion/maps.py:1563-1566; bake/globe/config.py:615)
    and ``ocean_min_fraction`` (the sea mask, maps.py:770-774;
    config.py:622)."""

    lake_min_depth: float = 0.5  # m
    ocean_min_fraction: float = 0.02


@dataclass
class ErosionParams:
    """Field-for-field the reference's ``ErosionParams``
    (bake/globe/config.py:486-583).  Lengths listed in ``LENGTH_PARAMS_M`` are
    metres and are converted to cell units by :func:`cell_units`."""

    iterations: int = 800
    particles_per_cell: float = 0.25
    dt: float = 1.2
    density: float = 1.0
    friction: float = 0.25
    deposition_rate: float = 0.1
    evap_rate: float = 0.001
    k_mom: float = 1.0
    k_disc: float = 1.0
    ema: float = 0.1
    thermal_rate: float = 0.5
    creep_rate: float = 0.02
    thermal_max: float = 50.0
    talus_slope_soft: float = 0.6
    talus_slope_hard: float = 1.2
    min_volume: float = 0.5
    min_volume_frac: float = 5e-5
    max_steps: int = 0
    checkpoint_every: int = 50
    quicklook_every: int = 50
    slope_gain: float = 2.0
    slope_saturation: float = 0.0
    erodibility: float = 0.2
    cover_depth: float = 5.0
    height_unit_m: float = 0.0
    disc_saturation: float = 32.0
    disc_exponent: float = 0.5
    disc_saturation_cells: float = 0.0
    momentum_saturation_cells: float = 0.0
    slope_limit_erode: float = 0.0
    slope_limit_deposit: float = 0.0
    slope_limit_exit: float = 0.02
    max_erode: float = 12.5
    iter_erode: float = 25.0
    iter_deposit: float = 50.0
    ocean_deposition_rate: float = 0.3
    ocean_steps: int = 64
    fan_room: float = 50.0
    fan_slope: float = 0.05
    offshore_writeoff: float = 0.25
    glacial_every: int = 10
    ice_evap: float = 0.0
    glacial_from: float = 0.75
    glacial_rate: float = 6.0
    glacial_ramp: int = 4
    glacial_max: float = 20.0
    climate_at_surface: bool = True
    ice_age_c: float = 8.0
    ice_history: bool = True
    ice_warm_c: float = 4.0
    ice_full: float = 0.9
    ice_sheet: bool = True
    ice_sea_share: float = 0.0
    ice_carve_flux: bool = False
    ice_yield_m: float = 8.0
    ice_ddf_mm: float = 4.0
    ice_snow_c: float = 1.0
    ice_rounds: int = 15
    ice_aridity: float = 0.7
    glacial_sticky: bool = True
    moraine_frac: float = 1.0
    strata_amp: float = 0.0
    strata_bed_m: float = 200.0
    strata_warp_m: float = 800.0
    strata_warp_km: float = 600.0
    isostasy: float = 0.8
    flexure_km: float = 60.0
    isostasy_every: int = 20
    uplift_max_m: float = 0.0
    uplift_mode: str = "replay"
    resume: bool = True
    sea_mask_every: int = 10
    basin_fill_grade: float = 0.1
    basin_fill_min_m: float = 1.0
    lake_balance: bool = True
    lateral_rate: float = 0.0
    window_lake_evap: float = 0.0
    window_lakes: bool = False
    lake_fill: bool = True
    lake_fill_rise_m: float = 1.0
    lake_fill_load: float = 0.55
    lake_trap: float = 0.9
    dep_floor_m: float = 1.0
    flood_every: int = 10
    route_eps: float = 0.05
    pit_steps: int = 16
    chunk: int = 2048
    backend: str = "cpu"
    #: the fast mode's erosion brush (no field of the reference, whose droplets change one cell per step):
    #: each chunk's height changes spread over the offsets |o| < brush_radius cells, weighted
    #: max(0, brush_radius - |o|); 0 = off, 1 = the cell alone
    brush_radius: int = 0


#: parameters declared in metres (bake/globe/config.py:603-605)
LENGTH_PARAMS_M = ("cover_depth", "max_erode", "iter_erode", "iter_deposit", "thermal_max", "fan_room",
                   "route_eps", "glacial_rate", "glacial_max", "uplift_max_m")


def cell_units(ep, name: str, cell_size_m: float) -> float:
    """A length parameter (metres) in the kernel's cell units
    (bake/globe/config.py:608-610)."""
    return float(getattr(ep, name)) / float(cell_size_m)


#: side passes switched off in the plain profile: the particle pass, uplift,
#: datum hold and halo exchange are all that remains of an iteration
PLAIN_PROFILE = dict(
    k_mom=0.0, k_disc=0.0, thermal_rate=0.0, creep_rate=0.0, flood_every=0, sea_mask_every=0,
    lake_balance=False, lake_fill=False, glacial_every=0, isostasy=0.0, basin_fill_grade=0.0,
    ice_sheet=False, ice_history=False, lateral_rate=0.0, strata_amp=0.0, backend="hip",
)


def plain_erosion_params(**overrides) -> ErosionParams:
    """``ErosionParams`` with every side pass off (see ``PLAIN_PROFILE``)."""
    d = dict(PLAIN_PROFILE)
    d.update(overrides)
    return ErosionParams(**d)


@dataclass
class WorldParams:
    """The groups of the reference's ``WorldParams`` the erosion stage reads,
    with the same ``rng`` (bake/globe/config.py:731-842)."""

    world: WorldGroup = field(default_factory=WorldGroup)
    tectonics: TectonicsGroup = field(default_factory=TectonicsGroup)
    erosion: ErosionParams = field(default_factory=ErosionParams)
    hydro: HydroGroup = field(default_factory=HydroGroup)

    @property
    def cell_size_m(self) -> float:
        return float(self.world.cell_size_m)

    def rng(self, stage: str, *extra: int) -> np.random.Generator:
        """Seeded generator for a stage and integer sub-keys (same key layout
        as the reference, so the streams are identical)."""
        if stage not in STAGE_SEED_OFFSETS:
            raise KeyError(f"unknown stage {stage!r}")
        mask = (1 << 64) - 1
        key = [(int(self.world.seed) + STAGE_SEED_OFFSETS[stage]) & mask, len(extra), *[int(e) & mask for e in extra]]
        return np.random.default_rng(key)

    def with_erosion(self, **overrides) -> "WorldParams":
        return dataclasses.replace(self, erosion=dataclasses.replace(self.erosion, **overrides))
