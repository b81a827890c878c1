"""Synthetic single-face windows and small globes for the erosion tests.

The window recipes restate the reference's test helper ``make_window``
(bake/tests/test_erosion.py:47-126): flat metric, heights in cells,
``closed`` (no sea, halo outside), ``tilt`` (a plane rising along +i with a
deep sea strip at i = 0, frozen walls, an uplift ridge), ``shelf`` (tilt with
a sea strip filled to the deposition floor and walled off) and ``dome``.
They return the keyword arguments of ``ErosionState.window`` so the same
inputs build the oracle's state and the device state.
"""
from __future__ import annotations

import numpy as np

from globe_hip.config import ErosionParams, WorldGroup, WorldParams, cell_units


def small_world(seed: int = 0, halo: int = 4) -> WorldParams:
    """The reference's ``small_world`` preset as far as erosion reads it
    (bake/globe/config.py:911-941): 50 m cells, land fraction 0.3, the
    shipped erosion defaults with 60 iterations and no basin fill."""
    p = WorldParams(world=WorldGroup(seed=seed, N_c=128, cell_size_m=50.0, land_fraction=0.3, halo=halo))
    p.erosion = ErosionParams(iterations=60, checkpoint_every=30, quicklook_every=30, basin_fill_grade=0.0,
                              climate_at_surface=False, ice_aridity=0.0, ice_age_c=0.0, ice_history=False,
                              ice_sheet=False, glacial_rate=50.0)
    return p


def window_params(seed: int = 0, **erosion) -> WorldParams:
    """small_world with the side passes the HIP path leaves out switched off
    (thermal, creep, routing flood, sea mask, lakes, glacial, isostasy)."""
    p = small_world(seed)
    d = dict(thermal_rate=0.0, creep_rate=0.0, flood_every=0, sea_mask_every=0, lake_balance=False, lake_fill=False,
             glacial_every=0, isostasy=0.0)
    d.update(erosion)
    for k, v in d.items():
        setattr(p.erosion, k, v)
    return p


def noise2d(NE: int, rng: np.random.Generator, octaves: int = 5, base: float = 3.0) -> np.ndarray:
    out = np.zeros((NE, NE))
    amp, tot = 1.0, 0.0
    x = np.linspace(0, 1, NE, endpoint=False)
    X, Y = np.meshgrid(x, x, indexing="ij")
    for o in range(octaves):
        f = base * 2 ** o
        n = int(np.ceil(f)) + 2
        lat = rng.random((n + 1, n + 1))
        qx, qy = X * f, Y * f
        i0 = np.floor(qx).astype(int)
        j0 = np.floor(qy).astype(int)
        tx, ty = qx - i0, qy - j0
        tx = tx * tx * (3 - 2 * tx)
        ty = ty * ty * (3 - 2 * ty)
        v = (lat[i0, j0] * (1 - tx) * (1 - ty) + lat[i0 + 1, j0] * tx * (1 - ty) + lat[i0, j0 + 1] * (1 - tx) * ty
             + lat[i0 + 1, j0 + 1] * tx * ty)
        out += amp * (v * 2 - 1)
        tot += amp
        amp *= 0.5
    return out / tot


def window_inputs(N: int, mode: str, params: WorldParams, seed: int = 0, relief: float = 6.0, uplift_total: float = 8.0,
                  iters: int = 150, deposit_on_exit: bool = False) -> dict:
    H = params.world.halo
    NE = N + 2 * H
    rng = np.random.default_rng(seed)
    cs = params.cell_size_m
    x = (np.arange(NE) - H + 0.5) / N
    X, Y = np.meshgrid(x, x, indexing="ij")
    mask = np.zeros((NE, NE), np.uint8)
    mask[H:H + N, H:H + N] = 1
    precip = np.ones((NE, NE), np.float32)
    upl = np.zeros((NE, NE))
    if mode == "closed":
        h = relief * (0.5 + 0.5 * noise2d(NE, rng)) + 0.5
    elif mode == "dome":
        h = relief * (0.5 + 0.5 * noise2d(NE, rng)) + 0.5
        upl = uplift_total / iters * np.exp(-((X - 0.5) ** 2 + (Y - 0.5) ** 2) / (2 * 0.18 ** 2))
    elif mode in ("tilt", "shelf"):
        h = relief * (0.35 * noise2d(NE, rng) + X) + 0.5
        if mode == "tilt":
            h = np.where(X < 6.0 / N, -20.0, h)
            wall = ((Y < 2.0 / N) | (Y > 1 - 2.0 / N) | (X > 1 - 2.0 / N)) & (mask == 1)
        else:
            h = np.where(X < 12.0 / N, -cell_units(params.erosion, "dep_floor_m", cs), h)
            wall = ((Y < 2.0 / N) | (Y > 1 - 2.0 / N) | (X > 1 - 2.0 / N) | (X < 2.0 / N)) & (mask == 1)
        h = np.where(wall, relief + 5.0, h)
        mask[wall] = 2
        precip[wall] = 0
        h = np.where(mask == 0, relief + 10.0, h)
        upl = uplift_total / iters * np.exp(-((X - 0.85) ** 2) / (2 * 0.12 ** 2))
    else:
        raise ValueError(mode)
    metric = np.zeros((NE, NE, 3), np.float32)
    metric[..., 0] = 1.0
    metric[..., 2] = 1.0
    z = np.zeros((NE, NE))
    return dict(height_m=h * cs, sediment_m=z, discharge=z, momentum=np.zeros((NE, NE, 2)),
                hardness=np.full((NE, NE), 0.5, np.float32), precip=precip, evap=np.ones((NE, NE), np.float32),
                uplift_m=upl * cs, mask=mask, metric=metric, metric_inv=metric, H=H, height_unit_m=cs,
                deposit_on_exit=deposit_on_exit)


def relief_globe(N: int = 48, H: int = 2, seed: int = 0):
    """A globe of 50 m cells with ~60 cells of relief, noise and a sea (so
    slopes exceed the talus and depressions abound): (grid, dict of extended
    inputs in metres) for ``ErosionState.from_grid``."""
    from globe_hip.cubesphere import get_grid

    g = get_grid(N, H, 50.0)
    rng = np.random.default_rng(seed)
    p = g.centers
    bed = 1500.0 * (np.sin(4 * p[..., 0] + 1.0) * np.cos(3 * p[..., 1]) + 0.6 * np.sin(7 * p[..., 2]))
    bed = bed + 250.0 * rng.standard_normal(bed.shape)
    F, NE = 6, g.NE
    hard = rng.uniform(0, 1, (F, NE, NE)).astype(np.float32)
    ones = np.ones((F, NE, NE), np.float32)
    return g, dict(bedrock_m=bed, hardness=hard, precip=ones, evap=ones, uplift_m=np.zeros((F, NE, NE)))
