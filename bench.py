"""Benchmark: droplet-steps/s of the erosion stage's particle pass on one GPU.

Workload (the Earth-scale forward configuration): cube-sphere N_c = 1024,
halo 2, 6 x 1024^2 = 6,291,456 cells, Earth radius (9773 m cells); bedrock =
6-octave fBm x 5000 m shifted to 30 % land; rain 2 cos(lat)^2 + 0.3 + 0.2 fBm;
uplift = 64 Gaussian blobs (sigma 200 km, peak 1 mm / iteration) replayed
each iteration; one droplet per active cell (6.29 M per iteration), lifetime
128; plain profile; fast mode (float32 droplets, int64 fixed-point
deposition).  A "step" is one full erosion iteration (spawn, trace, apply,
fold, EMA, uplift, datum hold, halo exchange).  ``value`` = droplet moves
executed / wall time of the timed iterations, inputs resident in HBM.

Also reported: ``roofline`` of the dominant kernel (the persistent trace),
164 algorithmic bytes per droplet-step against 8 TB/s, its duration taken
live with HIP events on its launch stream; ``cpu_baseline`` = the oracle (a C
restatement of the reference's numba path: parallel trace over 2048-droplet
chunks, serial apply) timed on this host's cores on a bounded sample.

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
#: algorithmic HBM bytes per droplet-step (FP32 fast mode, plain profile, layout-independent, each value
#: counted once per step): the current 2x2 stencil x 4 channels (64), 11 cell scalars (44), the next stencil
#: + base + mask (24) and four int64 fixed-point writes -- height change, discharge, two momentum tracks (32)
B_STEP = 164.0
#: what this implementation's two launches per chunk move per step on top of that: the census pass
#: (trace_kernel) reads the 132 B, adds the erosion request and floor key (16) and the three tracks (24) and
#: writes the move's 16-B tape row; the applying pass (replay_kernel) reads the row (16) and the cell's
#: erosion scale (4) and adds the height change (8)
BYTES_PER_STEP_CENSUS = 188.0
BYTES_PER_STEP_APPLY = 28.0
HBM_PEAK_GBS = 8000.0
#: HBM bytes per chunk (census launch, applying launch) of this workload (N_c=1024, 16 chunks) by
#: (persistent grid, register-budget variant), from rocprofv3 FETCH_SIZE (x 2.000, calibrated on a 1-GiB
#: copy) + WRITE_SIZE (x 1.000); profiles/trace_kernel_earth_n1024_counters.md holds the runs
TRAFFIC_PER_CHUNK = {(1024, 0): (5.454e9, 1.851e9), (1280, 1): (10.181e9, 2.187e9), (1536, 2): (21.021e9, 2.282e9),
                     (2560, 1): (14.789e9, 2.486e9)}
#: relaxed 64-bit atomicAdd / atomicMin per droplet move of this workload by pass (census: the erosion
#: request and floor key on eroding steps and the three tracks on land; applying pass: the height change
#: and a land death's final deposits),
#: from rocprofv3 TCC_ATOMIC_sum (one per lane-atomic: 268.43 M for the probe's 268.44 M adds); every one is
#: performed past L2 (TCC_EA0_ATOMIC_sum equal); profiles/trace_kernel_earth_n1024_atomics.md
ATOMICS_PER_MOVE = {"census": 3.02, "applying": 0.37}


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


def gpu_run(grid, f, p, steps, warmup, chunks):
    import torch

    from globe_hip.erosion import fast, maps

    dev = torch.device("cuda:0")
    args = (grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"], p.erosion)
    st = maps.ErosionState.from_grid(*args, device=dev, mode="fast")
    st.fast_chunks = chunks
    st.fast_sync_stats = False
    maps.start_replay(st, p)
    # warm-up iterations double as the autotuner's trials of the persistent
    # trace grid (the results do not depend on it)
    from globe_hip.erosion import batch as bt

    cus = torch.cuda.get_device_properties(0).multi_processor_count
    tuner = bt.Autotuner(st, candidates=[(4 * cus, 0), (5 * cus, 1), (6 * cus, 2), (10 * cus, 1)][:max(warmup, 1)])
    log("warmup")
    for it in range(warmup):
        tuner.step(lambda k: maps.step(st, p, k), it)
    if tuner.best is None:
        tuner.finish()
    log(f"autotuned (grid, variant): {tuner.best} (trace ns per move per candidate {tuner.times})")
    torch.cuda.synchronize()
    fast.trace_time()  # drop warmup launches
    st.fast_time_trace = True
    stats = []
    log(f"timing {steps} iterations")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for it in range(warmup, warmup + steps):
        maps.step(st, p, it)
        stats.append(st._fast.stats.clone())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    (cen_ms, cen_n), (app_ms, app_n) = fast.trace_time_passes()
    trace_ms, launches = cen_ms + app_ms, cen_n + app_n
    raw = [s.cpu().numpy() for s in stats]
    dec = [fast.decode_stats(r, 2.0 ** fast.HEIGHT_BITS) for r in raw]
    moves = sum(d["moves"] for d in dec)
    particles = sum(d["particles"] for d in dec)
    entries = sum(d["entries"] for d in dec)
    log(f"{steps} iterations in {dt:.3f}s: {moves} moves, {particles} droplets, steps_mean {entries / max(particles, 1):.1f}; "
        f"trace {trace_ms:.1f} ms over {launches} launches (census {cen_ms:.1f} ms / {cen_n}, applying {app_ms:.1f} ms "
        f"/ {app_n})")
    return dict(seconds=dt, moves=moves, particles=particles, entries=entries, trace_ms=trace_ms, launches=launches,
                census_ms=cen_ms, census_launches=cen_n, apply_ms=app_ms, apply_launches=app_n,
                last=dec[-1], grid_blocks=st.fast_grid_blocks, variant=st.fast_variant,
                tune_ms={f"{g}/{v}": ms for (g, v), ms in tuner.times.items()})


def cpu_baseline(grid, f, p, fraction=8):
    """The oracle on one iteration of the same workload with 1/fraction of
    the droplets (reference chunking, serial apply), on this host's cores."""
    from oracle import host as oracle

    threads = min(16, len(os.sched_getaffinity(0)))
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
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
            "sample": f"one iteration of the N_c=1024 workload with {P} droplets (1/{fraction} of the spawn), "
                      f"reference chunking ({s['chunk']}) and serial apply; {moves[0]} moves in {dt:.2f} s"}


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
    log(f"batched: {iterations} iterations in {t_batch:.2f}s ({1000 * per_it_b:.1f} ms/iteration, "
        f"{moves / t_batch:.3g} moves/s); per window: {1000 * per_it_pw:.1f} ms/iteration "
        f"({pw_moves / t_pw:.3g} moves/s); speed-up {per_it_pw / per_it_b:.2f}")
    return {"workload": f"{n} zoom windows, side log-uniform [128, 1024] ({cells} interior cells), halo 2, own fBm "
                        f"seed each, lifetime log-uniform [32, 256], 1 droplet/cell, fast mode 16 chunks",
            "iterations": iterations, "batched_s": t_batch, "batched_ms_per_iteration": 1000 * per_it_b,
            "batched_moves_per_s": moves / t_batch, "per_window_ms_per_iteration": 1000 * per_it_pw,
            "per_window_moves_per_s": pw_moves / t_pw, "speedup": per_it_pw / per_it_b,
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


def atomic_probe(NE=1028, F=6, lanes=1 << 20, steps=64, hot=False):
    """The device's 64-bit atomicAdd rate on the trace's footprint (6 faces of
    NE^2 cells): a droplet-like walk adding to 4 planes per step (the applying
    pass's pattern), the same walk with 1 plane, and uniformly random cells."""
    import ctypes

    import torch

    from globe_hip import _lib

    n = F * NE * NE
    t = torch.zeros(4 * n, dtype=torch.int64, device="cuda:0")
    out = {}
    cases = (("walk_4_planes", 4, 0), ("walk_1_plane", 1, 0), ("random_1_plane", 1, 1))
    if hot:
        cases = cases + tuple((f"hot_{h}_cells", 1, h) for h in (1 << 20, 1 << 16, 1 << 12))
    for name, planes, pattern in cases:
        def run(seed):
            _lib.check(_lib.lib().pg_atomic_probe(ctypes.c_void_p(t.data_ptr()), n, NE, planes, lanes, steps, pattern,
                                                  seed, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
                       "pg_atomic_probe")
        run(1)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ms = []
        for r in range(3):
            ev[0].record()
            run(2 + r)
            ev[1].record()
            torch.cuda.synchronize()
            ms.append(ev[0].elapsed_time(ev[1]))
        m = sorted(ms)[1]
        out[name] = {"ms": m, "gatomics_per_s": lanes * steps * planes / (m * 1e-3) / 1e9}
    tot = int(t.sum().item())
    assert tot == 4 * lanes * steps * sum(c[1] for c in cases), "atomic probe lost updates"
    log("atomic probe (G atomics/s): " + ", ".join(f"{k} {v['gatomics_per_s']:.1f}" for k, v in out.items()))
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
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--N", type=int, default=1024)
    ap.add_argument("--chunks", type=int, default=16)
    ap.add_argument("--max-steps", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-side", action="store_true")
    ap.add_argument("--atomic-probe", action="store_true", help="only run the atomic-rate probe and exit")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run 3 device-to-device copies of this many MiB (HBM counter calibration) and exit")
    a = ap.parse_args()
    if a.copy_probe_mib:
        copy_probe(a.copy_probe_mib)
        return
    if a.atomic_probe:
        print(json.dumps({"atomic_probe": atomic_probe(hot=True)}), flush=True)
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU: the particle pass shards by droplets with one int64 all-reduce per "
              f"iteration (designed, not built); --gpus {a.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    grid, f, p = build_inputs(a.N, 2, a.max_steps, 200)
    r = gpu_run(grid, f, p, a.steps, a.warmup, a.chunks)
    value = r["moves"] / r["seconds"]
    trace_s = r["trace_ms"] / 1000.0
    achieved = B_STEP * r["moves"] / trace_s / 1e9 if trace_s > 0 else 0.0
    known = (a.N, a.chunks) == (1024, 16)
    traffic = TRAFFIC_PER_CHUNK.get((r["grid_blocks"], r["variant"])) if known else None
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
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"earth globe N_c={a.N} H=2 (6x{a.N}^2 cells), fBm bedrock, latitudinal rain, "
                               f"64 uplift blobs replayed, 1 droplet/cell, lifetime {a.max_steps}, fast mode "
                               f"{a.chunks} chunks/iteration",
                   "droplets_per_iteration": r["particles"] // a.steps, "moves": r["moves"],
                   "steps_mean": r["entries"] / max(r["particles"], 1), "grid_blocks": r["grid_blocks"],
                   "trace_variant": r["variant"],
                   "autotune_trace_ns_per_move": r["tune_ms"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": sum(traffic) if traffic else None,
                     "kernel": "one chunk's particle pass: trace_kernel<false, W, false> (census) + replay_kernel "
                               "(applying pass)",
                     "bytes_per_step": B_STEP,
                     "launch_ms_mean": r["trace_ms"] / max(r["census_launches"], 1),
                     "launches": r["census_launches"],
                     "passes": {
                         "census": {"kernel": f"trace_kernel<false, {4 + r['variant']}, false>",
                                    "ms_mean": r["census_ms"] / max(r["census_launches"], 1),
                                    "launches": r["census_launches"], "bytes_per_step": BYTES_PER_STEP_CENSUS,
                                    "traffic": traffic[0] if traffic else None},
                         "applying": {"kernel": "replay_kernel",
                                      "ms_mean": r["apply_ms"] / max(r["apply_launches"], 1),
                                      "launches": r["apply_launches"], "bytes_per_step": BYTES_PER_STEP_APPLY,
                                      "traffic": traffic[1] if traffic else None}}},
    }
    probe = atomic_probe()
    peak_atomics = probe["walk_4_planes"]["gatomics_per_s"]
    per_move = sum(ATOMICS_PER_MOVE.values())
    ach_atomics = per_move * r["moves"] / trace_s / 1e9 if trace_s > 0 else 0.0
    out["roofline"]["atomics"] = {"achieved": ach_atomics, "peak": peak_atomics, "unit": "G atomics/s",
                                  "frac": ach_atomics / peak_atomics, "per_move": ATOMICS_PER_MOVE,
                                  "peak_source": "measured: pg_atomic_probe, droplet-like walk over the same "
                                                 "footprint, 4 int64 planes", "probe": probe}
    for name, ms in (("census", r["census_ms"]), ("applying", r["apply_ms"])):
        g = ATOMICS_PER_MOVE[name] * r["moves"] / (ms / 1000.0) / 1e9 if ms > 0 else 0.0
        out["roofline"]["passes"][name]["atomics"] = {"achieved": g, "frac": g / peak_atomics}
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
