"""Synthetic upstream fields for the erosion stage's benchmark workloads.

The bedrock is 3-D lattice value noise summed over octaves (fBm) sampled at
the cell centres on the unit sphere, the recipe of the reference's test
stubs (bake/globe/stubs.py:14-44: a random lattice per octave, smoothstep
trilinear interpolation, base frequency 2, frequency x2 and amplitude x0.5
per octave, normalised by the amplitude sum).  Because the lattice lives in
3-D space the field is continuous across cube edges, halos included.

Workloads (sizes and parameters of the benchmark configurations):

* ``plain_globe(N=128, ...)``: bedrock fBm x 5000 m, shifted so 30 % of the
  interior is land, uniform rain 1, evap 1, hardness 0, no uplift;
* ``earth_globe(N=1024, ...)``: the same bedrock, rain
  ``2 cos(lat)^2 + 0.3 + 0.2 fBm``, uplift = 64 Gaussian blobs (sigma
  200 km of arc, peak 1 mm per iteration).
"""
from __future__ import annotations

import math

import numpy as np

from .config import ErosionParams, WorldGroup, WorldParams, plain_erosion_params
from .cubesphere import Grid, get_grid

EARTH_RADIUS_M = 6371.0e3


def fbm_at(p: np.ndarray, rng: np.random.Generator, octaves: int = 5, base_freq: float = 2.0) -> np.ndarray:
    """Value-noise fBm at unit vectors ``p`` (..., 3), roughly in [-1, 1]."""
    p = np.asarray(p, dtype=np.float64)
    shape = p.shape[:-1]
    q3 = p.reshape(-1, 3)
    out = np.zeros(q3.shape[0])
    amp, total = 1.0, 0.0
    for o in range(octaves):
        freq = base_freq * 2 ** o
        n = int(np.ceil(freq)) + 2
        m = n + 1
        lattice = rng.random((m, m, m)).reshape(-1)
        q = (q3 + 1.0) * (0.5 * freq)
        i0 = np.floor(q).astype(np.int64)
        t = q - i0
        t = t * t * (3.0 - 2.0 * t)
        np.clip(i0, 0, n - 1, out=i0)
        base = (i0[:, 0] * m + i0[:, 1]) * m + i0[:, 2]
        acc = np.zeros_like(out)
        for dx in (0, 1):
            wx = t[:, 0] if dx else 1.0 - t[:, 0]
            for dy in (0, 1):
                wxy = wx * (t[:, 1] if dy else 1.0 - t[:, 1])
                off = (dx * m + dy) * m
                acc += wxy * ((1.0 - t[:, 2]) * lattice[base + off] + t[:, 2] * lattice[base + off + 1])
        out += amp * (acc * 2.0 - 1.0)
        total += amp
        amp *= 0.5
    return (out / total).reshape(shape)


def earth_cell_size(N: int) -> float:
    """Cell size (m) at which N cells per face edge give Earth's radius."""
    return 2.0 * math.pi * EARTH_RADIUS_M / (4.0 * N)


def _bedrock(grid: Grid, rng, octaves: int, amp_m: float, land_fraction: float) -> np.ndarray:
    bed = fbm_at(grid.centers, rng, octaves) * amp_m
    q = np.quantile(bed[grid.interior], 1.0 - land_fraction)
    return bed - q


def gaussian_blobs(grid: Grid, rng, n: int, sigma_m: float, peak_m: float, radius_m: float) -> np.ndarray:
    """Sum of ``n`` Gaussian bumps of arc-length width ``sigma_m`` centred at
    uniform random points of the sphere, on every extended cell (each bump
    evaluated within 6 sigma of its centre, where it is above 1.5e-8 of its
    peak)."""
    c = rng.normal(size=(n, 3))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    P = grid.centers.reshape(-1, 3)
    out = np.zeros(P.shape[0])
    cos_cut = np.cos(min(6.0 * sigma_m / radius_m, np.pi))
    for k in range(n):
        d = P @ c[k]
        near = np.flatnonzero(d > cos_cut)
        arc = np.arccos(np.clip(d[near], -1.0, 1.0)) * radius_m
        out[near] += peak_m * np.exp(-0.5 * (arc / sigma_m) ** 2)
    return out.reshape(grid.centers.shape[:-1])


def workload_params(N: int, H: int = 2, seed: int = 0, max_steps: int = 64, **erosion) -> WorldParams:
    """World + plain-profile erosion parameters of the benchmark workloads.
    The generic droplet-model knobs map onto the reference's fields as:
    inertia 0.05 -> friction (1 - 0.05) / dt; capacity 4 -> erodibility 4;
    erode = deposit 0.3 -> deposition_rate 0.3; evaporation 0.01 ->
    evap_rate 0.01 / dt; gravity 4 -> slope_gain 4; one droplet per cell."""
    dt = 1.2
    ep = plain_erosion_params(dt=dt, friction=(1.0 - 0.05) / dt, erodibility=4.0, deposition_rate=0.3,
                              evap_rate=0.01 / dt, slope_gain=4.0, particles_per_cell=1.0, max_steps=max_steps,
                              uplift_mode="stack")
    for k, v in erosion.items():
        setattr(ep, k, v)
    return WorldParams(world=WorldGroup(seed=seed, N_c=N, cell_size_m=earth_cell_size(N), land_fraction=0.30, halo=H),
                       erosion=ep)


def plain_globe(N: int = 128, H: int = 2, seed: int = 0, octaves: int = 6):
    """Plain erosion inputs: (grid, dict of extended float arrays in metres)."""
    grid = get_grid(N, H, earth_cell_size(N))
    p = WorldParams(world=WorldGroup(seed=seed, N_c=N, cell_size_m=grid.cell_size_m, halo=H))
    bed = _bedrock(grid, p.rng("stub", 1), octaves, 5000.0, 0.30)
    shape = bed.shape
    return grid, {
        "bedrock_m": bed,
        "hardness": np.zeros(shape, np.float32),
        "precip": np.ones(shape, np.float32),
        "evap": np.ones(shape, np.float32),
        "uplift_m": np.zeros(shape),
    }


def earth_globe(N: int = 1024, H: int = 2, seed: int = 0, octaves: int = 6, n_blobs: int = 64):
    """Earth-scale inputs: latitudinal rain and Gaussian uplift blobs."""
    grid, f = plain_globe(N, H, seed, octaves)
    p = WorldParams(world=WorldGroup(seed=seed, N_c=N, cell_size_m=grid.cell_size_m, halo=H))
    lat = grid.latitude()
    rain = 2.0 * np.cos(lat) ** 2 + 0.3 + 0.2 * fbm_at(grid.centers, p.rng("stub", 2), 4)
    f["precip"] = rain.astype(np.float32)
    f["uplift_m"] = gaussian_blobs(grid, p.rng("stub", 3), n_blobs, 200.0e3, 1.0e-3, EARTH_RADIUS_M)
    return grid, f


def value_noise_2d(NE: int, rng: np.random.Generator, octaves: int = 6, base: float = 4.0) -> np.ndarray:
    """Normalised 2-D fBm of smoothstep-interpolated lattice value noise on
    an NE x NE patch (lacunarity 2, gain 0.5), values in about [-1, 1]."""
    out = np.zeros((NE, NE))
    x = np.arange(NE) / NE
    amp, tot = 1.0, 0.0
    for o in range(octaves):
        f = base * 2 ** o
        n = int(np.ceil(f)) + 2
        lat = rng.random((n + 1, n + 1)) * 2.0 - 1.0
        q = x * f
        i0 = np.floor(q).astype(np.int64)
        t = q - i0
        t = t * t * (3.0 - 2.0 * t)
        a = lat[i0][:, i0] * (1 - t)[None, :] + lat[i0][:, i0 + 1] * t[None, :]
        b = lat[i0 + 1][:, i0] * (1 - t)[None, :] + lat[i0 + 1][:, i0 + 1] * t[None, :]
        out += amp * (a * (1 - t)[:, None] + b * t[:, None])
        tot += amp
        amp *= 0.5
    return out / tot


#: the reference's zoom erosion profile at the 305 m level (bake/globe/refine/
#: zoom.py:31-84 ZOOM_EROSION with zoom_params' level rules, :196-219: the
#: discharge scales at their cell caps, the wider slope limit above 150 m,
#: the talus rate above 40 m), max_steps aside -- the parameters a zoom tile
#: is eroded with
ZOOM_PROFILE_305M = dict(
    erodibility=1.0, k_disc=10.0, disc_exponent=0.0, disc_saturation=32.0, min_volume_frac=0.01, friction=0.0,
    evap_rate=0.001 / 1.2, max_erode=1e9, iter_erode=1e9, iter_deposit=1e9, thermal_max=1e9, thermal_rate=0.8,
    talus_slope_soft=0.8, talus_slope_hard=0.8, creep_rate=0.0, window_lakes=False, flood_every=5,
    disc_saturation_cells=1280.0, momentum_saturation_cells=512.0, slope_gain=0.589, slope_limit_erode=0.5,
    slope_limit_deposit=0.25, k_mom=1.0,
)


def zoom_windows(n: int = 64, seed: int = 0, side_lo: int = 128, side_hi: int = 1024, life_lo: int = 32,
                 life_hi: int = 256, H: int = 2, cell_m: float = 50.0, relief_m: float = 400.0, **erosion):
    """Synthetic zoom windows: side log-uniform in [side_lo, side_hi] cells,
    each with its own fBm bedrock seed (relief_m of relief on a 100 m
    floor), halo H outside the mask, uniform rain, one droplet per cell and
    a lifetime log-uniform in [life_lo, life_hi].  Returns (window keyword
    arguments for ErosionState.window, WorldParams) per window: the plain
    profile of workload_params with the window's seed and lifetime, then the
    ``erosion`` overrides (e.g. ZOOM_PROFILE_305M)."""
    rng = np.random.default_rng(seed)
    sides = np.exp(rng.uniform(np.log(side_lo), np.log(side_hi), n)).astype(np.int64)
    lives = np.exp(rng.uniform(np.log(life_lo), np.log(life_hi), n)).astype(np.int64)
    out = []
    for w in range(n):
        N = int(sides[w])
        NE = N + 2 * H
        h = 100.0 + relief_m * (0.5 + 0.5 * value_noise_2d(NE, np.random.default_rng(seed * 1000 + w + 1)))
        mask = np.zeros((NE, NE), np.uint8)
        mask[H:H + N, H:H + N] = 1
        metric = np.zeros((NE, NE, 3), np.float32)
        metric[..., 0] = 1.0
        metric[..., 2] = 1.0
        z = np.zeros((NE, NE))
        win = dict(height_m=h, sediment_m=z, discharge=z, momentum=np.zeros((NE, NE, 2)),
                   hardness=np.zeros((NE, NE), np.float32), precip=np.ones((NE, NE), np.float32),
                   evap=np.ones((NE, NE), np.float32), uplift_m=z, mask=mask, metric=metric, metric_inv=metric,
                   H=H, height_unit_m=cell_m)
        ep = workload_params(N, H, **erosion).erosion
        ep.max_steps = int(lives[w])
        p = WorldParams(world=WorldGroup(seed=seed * 1000 + w, N_c=N, cell_size_m=cell_m, land_fraction=0.3, halo=H),
                        erosion=ep)
        out.append((win, p))
    return out


# ---------------------------------------------------------------------------
# the planet-wide workload, built on the device (N_c = 8192: 403 M cells,
# whose host construction would take tens of minutes and ~250 GB)
# ---------------------------------------------------------------------------
def _face_points_device(grid: Grid, f: int, device):
    """(s, t, p) of face f's extended cell centres on the device: the
    gnomonic coordinates tan((u - 1/2) pi / 2), tan((v - 1/2) pi / 2) and the
    unit vectors (the equi-angular mapping of cubesphere.to_sphere_v)."""
    import torch

    from .cubesphere import BASES, HALF_PI

    e = (torch.arange(grid.NE, dtype=torch.float64, device=device) - grid.H + 0.5) / grid.N
    U, V = torch.meshgrid(e, e, indexing="ij")
    s = torch.tan((U - 0.5) * HALF_PI)
    t = torch.tan((V - 0.5) * HALF_PI)
    b = torch.from_numpy(BASES[f]).to(device)
    q = b[2] + s[..., None] * b[0] + t[..., None] * b[1]
    return s, t, q / torch.linalg.vector_norm(q, dim=-1, keepdim=True)


def fbm_lattices(rng: np.random.Generator, octaves: int = 5, base_freq: float = 2.0):
    """fbm_at's random lattices, drawn from ``rng`` in fbm_at's order:
    (frequency, n, m, lattice values) per octave."""
    out = []
    for o in range(octaves):
        freq = base_freq * 2 ** o
        n = int(np.ceil(freq)) + 2
        m = n + 1
        out.append((freq, n, m, rng.random((m, m, m)).reshape(-1)))
    return out


def fbm_device(p, lattices):
    """fbm_at on the device at unit vectors ``p`` (..., 3) (a float64
    tensor) from fbm_lattices' draws: the same smoothstep trilinear value
    noise, float64 torch ops."""
    import torch

    q3 = p.reshape(-1, 3)
    out = torch.zeros(q3.shape[0], dtype=torch.float64, device=p.device)
    amp, total = 1.0, 0.0
    for freq, n, m, lat in lattices:
        lattice = torch.from_numpy(lat).to(p.device)
        q = (q3 + 1.0) * (0.5 * freq)
        i0 = torch.floor(q)
        t = q - i0
        t = t * t * (3.0 - 2.0 * t)
        i0 = i0.to(torch.int64).clamp_(0, n - 1)
        base = (i0[:, 0] * m + i0[:, 1]) * m + i0[:, 2]
        del q, i0
        acc = torch.zeros_like(out)
        for dx in (0, 1):
            wx = t[:, 0] if dx else 1.0 - t[:, 0]
            for dy in (0, 1):
                wxy = wx * (t[:, 1] if dy else 1.0 - t[:, 1])
                off = (dx * m + dy) * m
                acc += wxy * ((1.0 - t[:, 2]) * lattice[base + off] + t[:, 2] * lattice[base + off + 1])
        out += amp * (acc * 2.0 - 1.0)
        total += amp
        amp *= 0.5
    return (out / total).reshape(p.shape[:-1])


def metric_device(grid: Grid, device):
    """The cell metric and its inverse, (6, NE, NE, 3) float32 on the device.
    The equi-angular metric depends only on the gnomonic coordinates (s, t),
    the same on every face: with r^2 = 1 + s^2 + t^2,
    g_uu = (pi/2)^2 (1 + s^2)^2 (1 + t^2) / r^4, g_uv = -(pi/2)^2 (1 + s^2)(1 + t^2) s t / r^4,
    g_vv = (pi/2)^2 (1 + t^2)^2 (1 + s^2) / r^4 (Grid.metric_uv's Jacobian products in closed form),
    scaled to cell units as Grid.metric is."""
    import torch

    from .cubesphere import HALF_PI

    s, t, _ = _face_points_device(grid, 0, device)
    s2, t2 = s * s, t * t
    r4 = (1.0 + s2 + t2) ** 2
    k = HALF_PI * HALF_PI
    scale = (grid.R_planet / (grid.N * grid.cell_size_m)) ** 2
    g = torch.stack([k * (1.0 + s2) ** 2 * (1.0 + t2) / r4, -k * (1.0 + s2) * (1.0 + t2) * s * t / r4,
                     k * (1.0 + t2) ** 2 * (1.0 + s2) / r4], dim=-1) * scale
    g32 = g.to(torch.float32)
    g64 = g32.to(torch.float64)
    det = g64[..., 0] * g64[..., 2] - g64[..., 1] ** 2
    gi = torch.stack([g64[..., 2] / det, -g64[..., 1] / det, g64[..., 0] / det], dim=-1).to(torch.float32)
    return g32.expand(6, *g32.shape).contiguous(), gi.expand(6, *gi.shape).contiguous()


def planet_globe(N: int = 8192, H: int = 2, seed: int = 0, octaves: int = 10, device="cuda"):
    """The planet-wide workload's inputs built on the device: (grid,
    fields, metric, metric_inv), fields as device tensors -- bedrock =
    ``octaves``-octave fBm x 5000 m shifted so 30 % of the interior is land,
    rain 2 cos(lat)^2 + 0.3 + 0.2 fBm (4 octaves), no hardness, evaporation
    1, no uplift.  The same recipe as earth_globe (whose host arithmetic it
    follows up to the device's tan / floating-point order)."""
    import torch

    grid = get_grid(N, H, earth_cell_size(N))
    p = WorldParams(world=WorldGroup(seed=seed, N_c=N, cell_size_m=grid.cell_size_m, halo=H))
    F, NE = 6, grid.NE
    bed_lat = fbm_lattices(p.rng("stub", 1), octaves)
    rain_lat = fbm_lattices(p.rng("stub", 2), 4)
    bed = torch.empty((F, NE, NE), dtype=torch.float64, device=device)
    rain = torch.empty((F, NE, NE), dtype=torch.float32, device=device)
    for f in range(F):
        _, _, pts = _face_points_device(grid, f, device)
        bed[f] = fbm_device(pts, bed_lat) * 5000.0
        lat = torch.asin(pts[..., 2].clamp(-1.0, 1.0))
        rain[f] = (2.0 * torch.cos(lat) ** 2 + 0.3 + 0.2 * fbm_device(pts, rain_lat)).to(torch.float32)
        del pts, lat
    inter = bed[:, H:H + N, H:H + N].reshape(-1)
    srt = torch.sort(inter).values
    pos = 0.7 * (srt.numel() - 1)
    lo = int(math.floor(pos))
    q = float(srt[lo]) + (pos - lo) * (float(srt[min(lo + 1, srt.numel() - 1)]) - float(srt[lo]))
    del srt, inter
    bed -= q
    metric, metric_inv = metric_device(grid, device)
    fields = {"bedrock_m": bed, "hardness": torch.zeros((F, NE, NE), dtype=torch.float32, device=device),
              "precip": rain, "evap": torch.ones((F, NE, NE), dtype=torch.float32, device=device),
              "uplift_m": torch.zeros((F, NE, NE), dtype=torch.float64, device=device)}
    return grid, fields, metric, metric_inv
