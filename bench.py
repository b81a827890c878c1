"""Benchmark: droplet-steps/s of the erosion stage's particle pass on one GPU.

Workload (the Earth-scale forward configuration): cube-sphere N_c = 1024,
halo 2, 6 x 1024^2 = 6,291,456 cells, Earth radius (9773 m cells); bedrock =
6-octave fBm x 5000 m shifted to 30 % land; rain 2 cos(lat)^2 + 0.3 + 0.2 fBm;
uplift = 64 Gaussian blobs (sigma 200 km, peak 1 mm / iteration) replayed
each iteration; one droplet per active cell (6.29 M per iteration), lifetime
128; plain profile.  A "step" is one full erosion iteration (spawn, trace,
apply, fold, EMA, uplift, datum hold, halo exchange).

``value``: the exact mode -- the reference's semantics (float64 droplets,
its 2048-droplet chunks, the serial apply's results computed in parallel,
bit for bit) -- droplet moves executed / wall time of the timed iterations,
inputs resident in HBM.  It is the mode that meets the hypsometry / eroded
volume tolerance against the reference algorithm at this size.

``fast_mode``: the float32 fast mode (order-independent fixed-point apply,
16 chunks) on the same workload and iterations, with its deviation from the
exact mode measured live (hypsometry CDF gap, eroded volume).

Also reported: ``roofline`` of the exact mode's trace kernel (156
algorithmic bytes per move against 8 TB/s, durations from HIP events on its
launch stream); ``cpu_baseline`` = the oracle (a C restatement of the
reference's numba path: parallel trace over 2048-droplet chunks, serial
apply) timed on this host's cores on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "procedural-generation_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = "droplet-steps/s at N_c=1024 vs numba host; fwd+adjoint/fwd ratio"
#: algorithmic HBM bytes per move of the exact mode's trace (float64 droplets on the packed float32
#: record, plain profile, each value counted once per step): the 2x2 stencil's surface, discharge and two
#: momentum channels (4 x 4 x 4 = 64), the cell's 11 scalars (metric 3, inverse metric 3, evaporation,
#: hardness, sediment, base, mask: 44), the next stencil's surface + next base + mask (24), and the change-list
#: entry written (cell 4, delta 8, volume 4, momentum 2 x 4 = 24)
B_MOVE_EXACT = 156.0
#: the same for the fast mode's particle pass (float32 record, four int64 fixed-point writes instead of the entry)
B_STEP_FAST = 164.0
HBM_PEAK_GBS = 8000.0
#: algorithmic HBM bytes per extended cell of one brush launch pair: the chunk's int64 change in (8), the
#: target flag in (1), the spread change out (8), the fold's flag read (1); the apron is served from L2
B_CELL_BRUSH = 18.0
PROFILE_COUNTERS = os.path.join(ROOT, "profiles", "trace_check_kernel_earth_n1024_counters.md")


def profiled_traffic():
    """HBM bytes per trace launch of this workload, from the committed counter
    profile it cites (None when the profile is absent)."""
    try:
        with open(PROFILE_COUNTERS) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1].split()[0])
    except OSError:
        return None
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_inputs(N, H, max_steps, iterations):
    from globe_hip import stubs

    t = time.time()
    grid, f = stubs.earth_globe(N, H)
    p = stubs.workload_params(N, H, max_steps=max_steps, uplift_mode="replay", iterations=iterations)
    _ = grid.metric, grid.metric_inv, grid.halo
    log(f"inputs N={N}: {time.time() - t:.1f}s")
    return grid, f, p


def surface_stats(st):
    """(interior surface, unit) on the host, float64."""
    return (st.height + st.sediment)[st.interior].cpu().numpy(), st.height_unit_m


def deviation(s0, sa, shift_a, sb, shift_b, unit):
    """Land hypsometry (64-bin CDF) max gap and relative eroded-volume
    difference of run b against run a, from the same initial surface s0."""
    la, lb = sa[sa >= 0] * unit, sb[sb >= 0] * unit
    hi = max(la.max(), lb.max())
    ha, _ = np.histogram(la, 64, (0.0, hi))
    hb, _ = np.histogram(lb, 64, (0.0, hi))
    gap = float(np.abs(np.cumsum(ha) / ha.sum() - np.cumsum(hb) / hb.sum()).max())
    ea = float(np.maximum(s0 - (sa + shift_a), 0.0).sum())
    eb = float(np.maximum(s0 - (sb + shift_b), 0.0).sum())
    return gap, (eb - ea) / ea


def exact_run(grid, f, p, steps, warmup, lanes=64):
    """The headline: the exact mode (check mode with the parallel apply)."""
    import torch

    from globe_hip.erosion import maps
    from globe_hip.erosion import particle as pk

    dev = torch.device("cuda:0")
    args = (grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"], p.erosion)
    st = maps.ErosionState.from_grid(*args, device=dev, mode="check")
    s0, unit = surface_stats(st)
    maps.start_replay(st, p)
    shift = 0.0
    log(f"exact mode: warmup {warmup}")
    # the warm-up iterations pick the trace's droplets per wave (the same results at every setting)
    tuner = None
    if lanes:
        st.check_lanes = lanes
    else:
        # the measured best first: a short warm-up tries the likeliest settings
        tuner = maps.CheckAutotuner(st, candidates=(8, 16, 4, 64)[:max(1, min(warmup, 4))])
    for it in range(warmup):
        if tuner is not None:
            shift += float(tuner.step(lambda k: maps.step(st, p, k), it)["datum_shift"])
        else:
            shift += float(maps.step(st, p, it)["datum_shift"])
    if tuner is not None:
        if tuner.best is None:
            tuner.finish()
        log(f"exact mode: droplets per wave {st.check_lanes} (trace ns per move {tuner.times})")
    torch.cuda.synchronize()
    st.check_time_trace = True
    pk.check_trace_time()  # drop anything recorded before the timed region
    log(f"exact mode: timing {steps} iterations")
    stats = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for it in range(warmup, warmup + steps):
        s = maps.step(st, p, it)
        shift += float(s["datum_shift"])
        stats.append(s)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    trace_ms, launches = pk.check_trace_time()
    moves = sum(s["moves"] for s in stats)
    particles = sum(s["particles"] for s in stats)
    entries = sum(s["entries"] for s in stats)
    chunks = sum(-(-s["particles"] // s["chunk"]) for s in stats)
    sweeps = sum(s["sweeps_total"] for s in stats)
    log(f"exact mode: {steps} iterations in {dt:.2f}s: {moves} moves, {particles} droplets, {chunks} chunks, "
        f"sweeps/chunk {sweeps / max(chunks, 1):.1f} (max {max(s['sweeps_max'] for s in stats)}); trace "
        f"{trace_ms:.0f} ms over {launches} launches")
    sf, _ = surface_stats(st)
    return dict(seconds=dt, moves=moves, particles=particles, entries=entries, chunks=chunks, trace_ms=trace_ms,
                launches=launches, sweeps_per_chunk=sweeps / max(chunks, 1),
                sweeps_max=max(s["sweeps_max"] for s in stats), s0=s0, s=sf, shift=shift, unit=unit,
                iterations=warmup + steps, lanes=int(st.check_lanes),
                lane_times=None if tuner is None else {str(k): v for k, v in tuner.times.items()})


def fast_run(grid, f, p, iterations, steps, chunks=16, brush=0):
    """The fast mode on the same workload for the same iterations; the last
    `steps` timed.  With brush = r > 1 the erosion brush of radius r is on;
    its LDS tile is picked by the autotuner over the first warm-up
    iterations."""
    import copy

    import torch

    from globe_hip.erosion import batch as bt
    from globe_hip.erosion import fast, maps

    dev = torch.device("cuda:0")
    if brush:
        p = copy.deepcopy(p)
        p.erosion.brush_radius = int(brush)
    args = (grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"], p.erosion)
    st = maps.ErosionState.from_grid(*args, device=dev, mode="fast")
    st.fast_chunks = chunks
    st.fast_sync_stats = False
    maps.start_replay(st, p)
    shift = 0.0
    warm = iterations - steps
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    if brush:
        warm = max(warm, 4)  # the autotuner's grid candidate and three tiles
        iterations = warm + steps
        tuner = bt.Autotuner(st, candidates=[(4 * cus, 0)], tiles=(0, 1, 2))
    else:
        # the persistent grid and register budget over the warm-up iterations, the measured best first
        tuner = bt.Autotuner(st, candidates=[(10 * cus, 1), (4 * cus, 0), (5 * cus, 1), (12 * cus, 2)][:max(warm, 1)])
    for it in range(warm):
        if tuner is not None and tuner.best is None:
            box = {}
            tuner.step(lambda k: box.setdefault("s", maps.step(st, p, k)), it)
            shift += float(box["s"]["datum_shift"].item())
        else:
            shift += float(maps.step(st, p, it)["datum_shift"].item())
    tuner.finish()
    if brush:
        log(f"fast mode brush {brush}: LDS tile {tuner.best_tile} (ms per brush launch per tile {tuner.tile_times})")
        st.fast_grid_blocks, st.fast_variant = 0, 0
    else:
        log(f"fast mode: (grid, variant) {tuner.best} (trace ns per move {tuner.times})")
    torch.cuda.synchronize()
    fast.trace_time()
    fast.brush_time()
    st.fast_time_trace = True
    raw = []
    t0 = time.perf_counter()
    for it in range(warm, iterations):
        s = maps.step(st, p, it)
        shift += float(s["datum_shift"].item())
        raw.append(st._fast.stats.clone())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    trace_ms, launches = fast.trace_time()
    brush_ms, brush_launches = fast.brush_time()
    dec = [fast.decode_stats(r.cpu().numpy(), 2.0 ** fast.HEIGHT_BITS) for r in raw]
    moves = sum(d["moves"] for d in dec)
    sf, _ = surface_stats(st)
    log(f"fast mode ({chunks} chunks, brush {brush}): {steps} iterations in {dt:.3f}s, {moves} moves")
    return dict(seconds=dt, moves=moves, trace_ms=trace_ms, launches=launches, s=sf, shift=shift, brush_ms=brush_ms,
                brush_launches=brush_launches, tile=tuner.best_tile, tile_ms=tuner.tile_times,
                cells=int(st.height.numel()), budget_exact=all(_budget_exact(d) for d in dec),
                grid=tuner.best, grid_ns=tuner.times)


def _budget_exact(d):
    from globe_hip._lib import FS

    r = d["raw_fx"]
    lhs = (int(r[FS["DEPOSITED_FX"]]) - int(r[FS["ERODED_FX"]]) + int(r[FS["TO_PENDING_FX"]])
           + int(r[FS["LOST_OFFSHORE_FX"]]) + int(r[FS["EXIT_FX"]]) + int(r[FS["VANISHED_FX"]]))
    return lhs == int(r[FS["RELEASED_FX"]])


def cpu_baseline(grid, f, p, fraction=1):
    """The oracle on one iteration of the same workload (1/fraction of the
    droplets; the whole spawn by default: ~10-15 s of CPU work), reference
    chunking and serial apply, on this host's cores."""
    from oracle import host as oracle

    # every core this process may use (the host's share: OMP_NUM_THREADS where the host sets one)
    threads = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    args = (grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"], p.erosion)
    st = oracle.OracleState.from_grid(*args)
    n_cells = int((st.mask[st.interior] == 1).sum())
    P = n_cells // fraction
    moves = [0]

    def diag(state, cl_cell, cl_vol, cl_count, cap, sp_death):
        n = cl_count.shape[0]
        v = cl_vol[: n * cap].reshape(n, cap)
        live = np.arange(cap)[None, :] < cl_count[:, None]
        moves[0] += int(((v >= 0) & live).sum())

    t = time.perf_counter()
    s = oracle.run_iteration(st, p, 0, n_particles=P, diag=diag, nthreads=threads)
    dt = time.perf_counter() - t
    log(f"cpu baseline: {P} droplets, {moves[0]} moves in {dt:.2f}s on {threads} threads "
        f"(trace {s['seconds_trace']:.2f}s, apply {s['seconds_apply']:.2f}s)")
    return {"value": moves[0] / dt, "unit": "droplet-steps/s", "cores": threads, "kind": "port",
            "sample": f"one iteration of the N_c=1024 workload with {P} droplets"
                      f"{'' if fraction == 1 else f' (1/{fraction} of the spawn)'}, reference chunking ({s['chunk']}) and "
                      f"serial apply; {moves[0]} moves in {dt:.2f} s"}


def adjoint_run(N=256, iterations=50, checkpoint_every=10, max_steps=128, seed=0, feedback=False):
    """Wall time of the differentiable mode's forward pass alone and of
    forward + hand-written adjoint (gradient of the hypsometry misfit with
    respect to the erodibility field and the six scalars), float32, on the
    N_c=256 earth globe, 50 iterations, checkpoints every 10; frozen-terrain
    gradient unless feedback."""
    import torch

    from globe_hip import stubs
    from globe_hip.erosion import adjoint as adj
    from globe_hip.erosion import maps

    t = time.time()
    grid, f = stubs.earth_globe(N, 2)
    p = stubs.workload_params(N, 2, max_steps=max_steps, iterations=iterations)
    st = maps.ErosionState.from_grid(grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"],
                                     p.erosion, device="cuda", mode="fast")
    run = adj.GradRun(st, p, iterations, real="f32", checkpoint_every=checkpoint_every, feedback=feedback)
    g = torch.Generator().manual_seed(seed)
    erod = (1.0 + 0.2 * (2.0 * torch.rand(tuple(st.height.shape), generator=g, dtype=torch.float64) - 1.0)).float().cuda()
    theta = adj.theta_nominal(run)
    log(f"adjoint inputs N={N}: {time.time() - t:.1f}s")

    def fwd():
        out = run.forward(erod, theta, tapes="none")
        return float(adj.hypsometry_loss(out["height"] + out["sediment"], run))

    def fwd_adj():
        e, th = erod.clone().requires_grad_(True), theta.clone().requires_grad_(True)
        loss = adj.hypsometry_loss(adj.erosion(e, th, run), run)
        loss.backward()
        return float(loss.detach()), e.grad, th.grad

    fwd()
    fwd_adj()  # warm-up: tapes and checkpoints allocated
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss_f = fwd()
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss_a, ge, gt = fwd_adj()
    torch.cuda.synchronize()
    t_fa = time.perf_counter() - t0
    log(f"adjoint (feedback {feedback}): forward {t_fwd:.3f}s, forward+adjoint {t_fa:.3f}s, ratio {t_fa / t_fwd:.2f}; "
        f"loss {loss_f:.6g}")
    return {"workload": f"earth globe N_c={N} H=2, {iterations} iterations, lifetime {max_steps}, float32, "
                        f"checkpoint every {checkpoint_every}; loss = 64-bin land hypsometric CDF vs Beta(2,5) on "
                        f"[0, 4000] m + 0.1 (mean land - 800 m)^2; erodibility field 1 +- 0.2; "
                        f"{'surface feedback' if feedback else 'frozen terrain'} gradient",
            "forward_s": t_fwd, "forward_adjoint_s": t_fa, "ratio": t_fa / t_fwd, "loss": loss_a,
            "grad_erod_norm": float(ge.double().norm()), "grad_theta": [float(v) for v in gt]}


def batch_run(n=64, iterations=50, seed=0, per_window_iterations=3):
    """Batched zoom windows: n windows (side log-uniform in [128, 1024],
    own fBm seed, lifetime log-uniform in [32, 256], one droplet per cell),
    the batched iteration autotuned then graph-captured and replayed for
    `iterations` iterations, against the same windows run one fast-mode pass
    per window."""
    import torch

    from globe_hip import stubs
    from globe_hip.erosion import batch as bt
    from globe_hip.erosion import fast

    t = time.time()
    wins = stubs.zoom_windows(n, seed=seed)
    b = bt.WindowBatch([w for w, _ in wins], [p for _, p in wins], device="cuda")
    cells = int(sum((x - 4) ** 2 for x in b.NE))
    log(f"batch inputs: {n} windows, {cells} interior cells, {time.time() - t:.1f}s")
    tuner = bt.Autotuner(b)
    it = 0
    while tuner.best is None:
        tuner.step(b.step, it)
        it += 1
    log(f"autotuned (grid, variant): {tuner.best} (trace ns per move per candidate {tuner.times})")
    g = bt.BatchGraph(b)
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for k in range(iterations):
        g.step(it + k)
        stats.append(b.stats.clone())
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    moves = sum(int(s[fast.FS["MOVES"]]) for s in stats)
    b2 = bt.WindowBatch([w for w, _ in wins], [p for _, p in wins], device="cuda")
    # the plain form: each window a default fast-mode pass (occupancy grid, natural register budget)
    b2.step_per_window(0)  # warm-up: per-window workspaces
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pw_moves = 0
    for k in range(per_window_iterations):
        b2.step_per_window(1 + k)
        torch.cuda.synchronize()
        pw_moves += b2.moves(per_window=True)
    t_pw = time.perf_counter() - t0
    per_it_b, per_it_pw = t_batch / iterations, t_pw / per_window_iterations
    # the two forms run different iterations of the same windows (the per-window sample is short), and the
    # moves per iteration drift as the terrain evolves: the speed-up is the ratio of moves per second
    speedup = (moves / t_batch) / (pw_moves / t_pw)
    log(f"batched: {iterations} iterations in {t_batch:.2f}s ({1000 * per_it_b:.1f} ms/iteration, "
        f"{moves / t_batch:.3g} moves/s); per window: {1000 * per_it_pw:.1f} ms/iteration "
        f"({pw_moves / t_pw:.3g} moves/s); speed-up {speedup:.2f} in moves/s, {per_it_pw / per_it_b:.2f} in "
        f"ms/iteration")
    return {"workload": f"{n} zoom windows, side log-uniform [128, 1024] ({cells} interior cells), halo 2, own fBm "
                        f"seed each, lifetime log-uniform [32, 256], 1 droplet/cell, fast mode 16 chunks",
            "iterations": iterations, "batched_s": t_batch, "batched_ms_per_iteration": 1000 * per_it_b,
            "batched_moves_per_s": moves / t_batch, "per_window_ms_per_iteration": 1000 * per_it_pw,
            "per_window_moves_per_s": pw_moves / t_pw, "speedup": speedup,
            "speedup_ms_per_iteration": per_it_pw / per_it_b,
            "moves_per_iteration": {"batched": moves / iterations, "per_window": pw_moves / per_window_iterations},
            "grid_blocks": tuner.best[0], "trace_variant": tuner.best[1],
            "autotune_trace_ns_per_move": {f"{g}/{v}": ms for (g, v), ms in tuner.times.items()}}


#: algorithmic bytes per extended cell of the side passes (float64 state, float32 metric / sqrt g)
#: thermal a: h, s, metric, talus, mask in + scale out = 45; r: + scale in, rscale out = 53;
#: b: + scale, rscale in, new h, s out = 69; write-back copy 32  -> one mass-wasting pass 199 B;
#: thermal_erosion = talus (12) + pass + creep mask (26) + zero talus (8) + pass = 444 B
WASTING_BYTES_PER_CELL = 199.0
THERMAL_CREEP_BYTES_PER_CELL = 2 * WASTING_BYTES_PER_CELL + 12.0 + 26.0 + 8.0
#: one smoothing step: data, metric_inv, sqrt g in, data out
SMOOTH_BYTES_PER_CELL = 32.0


def side_run(grid, f, p, reps=5, cpu=True):
    """The side passes at the earth preset's settings on the same globe:
    thermal + creep (maps.thermal_erosion) and one isostatic rebound
    (maps.apply_isostasy), timed with HIP events on the launch stream."""
    import torch

    from globe_hip.erosion import maps

    dev = torch.device("cuda:0")
    ep = p.erosion
    for k, v in dict(thermal_rate=0.5, creep_rate=0.02, thermal_max=50.0, isostasy=0.8, flexure_km=60.0).items():
        setattr(ep, k, v)
    args = (grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"], ep)
    st = maps.ErosionState.from_grid(*args, device=dev, mode="fast")
    maps.prepare_side_passes(st, p)
    n = st.height.numel()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def timed(fn, setup=None):
        ms = []
        for r in range(reps + 1):
            if setup is not None:
                setup()
            torch.cuda.synchronize()
            ev[0].record()
            fn()
            ev[1].record()
            torch.cuda.synchronize()
            if r:
                ms.append(ev[0].elapsed_time(ev[1]))
        return sorted(ms)[len(ms) // 2]

    t_w = timed(lambda: maps.thermal_erosion(st, p))
    rng = np.random.default_rng(0)
    acc = rng.standard_normal(tuple(st.height.shape)) * 1e-3
    sig, kappa, steps = maps.isostasy_steps(st, ep)

    acc_dev = torch.from_numpy(acc).to(dev)
    t_i = timed(lambda: maps.apply_isostasy(st, p), setup=lambda: st.iso_acc.copy_(acc_dev))
    wasting_gbs = THERMAL_CREEP_BYTES_PER_CELL * n / (t_w * 1e-3) / 1e9
    smooth_gbs = steps * SMOOTH_BYTES_PER_CELL * n / (t_i * 1e-3) / 1e9
    out = {"workload": f"the same globe, thermal_rate 0.5 + creep_rate 0.02 (two mass-wasting passes), isostasy 0.8 "
                       f"with flexure 60 km = {sig:.2f} cells: {steps} smoothing steps of kappa {kappa:.4f}",
           "thermal_creep_ms": t_w, "thermal_creep_GBps": wasting_gbs, "thermal_frac_of_hbm": wasting_gbs / HBM_PEAK_GBS,
           "isostasy_ms": t_i, "smoothing_steps": steps, "smoothing_GBps": smooth_gbs,
           "smoothing_frac_of_hbm": smooth_gbs / HBM_PEAK_GBS,
           "bytes_per_cell": {"thermal_erosion": THERMAL_CREEP_BYTES_PER_CELL, "smoothing_step": SMOOTH_BYTES_PER_CELL}}
    log(f"side passes: thermal+creep {t_w:.2f} ms ({wasting_gbs:.0f} GB/s), isostasy {t_i:.1f} ms for {steps} steps "
        f"({smooth_gbs:.0f} GB/s)")
    eps = float(ep.route_eps) / float(grid.cell_size_m)
    sweeps = st.refresh_route(eps)
    t_r = timed(lambda: st.refresh_route(eps))
    filled = int(((st.route - (st.height + st.sediment)) > 0).sum().item())
    out.update({"route_ms": t_r, "route_sweeps": sweeps, "route_filled_cells": filled})
    log(f"side passes: routing flood {t_r:.1f} ms ({sweeps} relaxation sweeps, {filled} cells filled)")
    if cpu:
        from oracle import host

        ost = host.OracleState.from_grid(*args)
        t = time.perf_counter()
        host.mass_wasting_pass(ost, host.talus_of(ep, ost.hardness), 0.5, 50.0 / grid.cell_size_m,
                               dep_floor=float(ep.dep_floor_m) / grid.cell_size_m)
        out["cpu_thermal_pass_ms"] = 1000.0 * (time.perf_counter() - t)
        log(f"side passes: oracle thermal pass (1 thread) {out['cpu_thermal_pass_ms']:.0f} ms")
    return out


def copy_probe(mib):
    """A plain copy of known size: its FETCH_SIZE / WRITE_SIZE per dispatch
    calibrate the HBM counters (2 x mib MiB of traffic per copy)."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copy probe: 3 copies of {mib} MiB")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=1024)
    ap.add_argument("--chunks", type=int, default=16, help="the fast mode's chunks per iteration")
    ap.add_argument("--max-steps", type=int, default=128)
    ap.add_argument("--brush", type=int, default=3, help="the fast mode's brush leg: radius (0 = no leg)")
    ap.add_argument("--check-lanes", type=int, default=0,
                    help="droplets per wave of the exact mode's trace (0: autotuned over the warm-up iterations)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fast", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-side", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run 3 device-to-device copies of this many MiB (HBM counter calibration) and exit")
    a = ap.parse_args()
    if a.copy_probe_mib:
        copy_probe(a.copy_probe_mib)
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU: the particle pass shards by droplets with one int64 all-reduce per "
              f"iteration (designed, not built); --gpus {a.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    grid, f, p = build_inputs(a.N, 2, a.max_steps, 200)
    r = exact_run(grid, f, p, a.steps, a.warmup, a.check_lanes)
    value = r["moves"] / r["seconds"]
    trace_s = r["trace_ms"] / 1000.0
    achieved = B_MOVE_EXACT * r["moves"] / trace_s / 1e9 if trace_s > 0 else 0.0
    traffic = profiled_traffic() if a.N == 1024 else None
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "droplet-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1000.0 * r["seconds"] / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"earth globe N_c={a.N} H=2 (6x{a.N}^2 cells), fBm bedrock, latitudinal rain, "
                               f"64 uplift blobs replayed, 1 droplet/cell, lifetime {a.max_steps}; exact mode: the "
                               f"reference's float64 trace over its 2048-droplet chunks and its serial apply's results "
                               f"computed in parallel",
                   "droplets_per_iteration": r["particles"] // a.steps, "moves": r["moves"],
                   "steps_mean": r["entries"] / max(r["particles"], 1), "chunks_per_iteration": r["chunks"] // a.steps,
                   "apply_sweeps_per_chunk": r["sweeps_per_chunk"], "apply_sweeps_max": r["sweeps_max"],
                   "trace_droplets_per_wave": r["lanes"], "autotune_trace_ns_per_move": r["lane_times"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"trace_check_kernel (one lane per droplet, {r['lanes']} droplets per wave, float64, one "
                               f"launch per 2048-droplet chunk)",
                     "bytes_per_move": B_MOVE_EXACT, "launch_ms_mean": r["trace_ms"] / max(r["launches"], 1),
                     "launches": r["launches"],
                     "trace_share_of_wall": trace_s / r["seconds"]},
    }
    if not a.no_fast:
        import torch

        torch.cuda.empty_cache()
        fr = fast_run(grid, f, p, r["iterations"], a.steps, a.chunks)
        gap, ero = deviation(r["s0"], r["s"], r["shift"], fr["s"], fr["shift"], r["unit"])
        fach = B_STEP_FAST * fr["moves"] / (fr["trace_ms"] / 1000.0) / 1e9 if fr["trace_ms"] > 0 else 0.0
        out["fast_mode"] = {
            "value": fr["moves"] / fr["seconds"], "unit": "droplet-steps/s",
            "ms_per_step": 1000.0 * fr["seconds"] / a.steps, "dtype": "f32", "chunks_per_iteration": a.chunks,
            "vs_exact_after_iterations": r["iterations"], "hypsometry_cdf_gap": gap, "eroded_volume_rel": ero,
            "grid_blocks_variant": list(fr["grid"]) if fr["grid"] else None,
            "autotune_trace_ns_per_move": {f"{g}/{v}": t for (g, v), t in fr["grid_ns"].items()},
            "meets_tolerance": bool(gap <= 0.005 and abs(ero) <= 0.005),
            "roofline": {"achieved": fach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": fach / HBM_PEAK_GBS,
                         "bytes_per_move": B_STEP_FAST}}
        log(f"fast mode vs exact after {r['iterations']} iterations: hypsometry gap {100 * gap:.2f} %, eroded volume "
            f"{100 * ero:+.2f} %")
    if not a.no_fast and a.brush > 1:
        import torch

        torch.cuda.empty_cache()
        br = fast_run(grid, f, p, r["iterations"], a.steps, a.chunks, brush=a.brush)
        per_launch = br["brush_ms"] / max(br["brush_launches"], 1)
        bach = B_CELL_BRUSH * br["cells"] / (per_launch * 1e-3) / 1e9 if per_launch > 0 else 0.0
        out["fast_mode_brush"] = {
            "value": br["moves"] / br["seconds"], "unit": "droplet-steps/s", "brush_radius": a.brush,
            "ms_per_step": 1000.0 * br["seconds"] / a.steps, "dtype": "f32", "chunks_per_iteration": a.chunks,
            "mass_budget_exact": br["budget_exact"], "lds_tile": ["64x4", "32x8", "16x16"][br["tile"]],
            "autotune_ms_per_brush_launch": {["64x4", "32x8", "16x16"][k]: v for k, v in br["tile_ms"].items()},
            "brush_share_of_wall": br["brush_ms"] / 1000.0 / br["seconds"],
            "roofline": {"kernel": "brush_spread_kernel + brush_fold_kernel (one launch pair per chunk)",
                         "achieved": bach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": bach / HBM_PEAK_GBS,
                         "bytes_per_cell": B_CELL_BRUSH, "launch_ms_mean": per_launch,
                         "launches": br["brush_launches"]}}
        log(f"brush {a.brush}: {per_launch * 1000:.1f} us per launch pair ({bach:.0f} GB/s), "
            f"{br['moves'] / br['seconds']:.3g} moves/s")
    if not a.no_adjoint:
        import torch

        torch.cuda.empty_cache()
        out["adjoint"] = adjoint_run()
        out["fwd_adjoint_ratio"] = out["adjoint"]["ratio"]
        fb = adjoint_run(feedback=True)
        out["adjoint"]["ratio_with_surface_feedback"] = fb["ratio"]
    if not a.no_batch:
        import torch

        torch.cuda.empty_cache()
        out["batched_windows"] = batch_run()
    if not a.no_side:
        import torch

        torch.cuda.empty_cache()
        out["side_passes"] = side_run(grid, f, p, cpu=not a.no_cpu)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(grid, f, p)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
