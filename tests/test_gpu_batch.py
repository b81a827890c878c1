"""Batched windows on the GPU: one persistent launch per chunk and pass over
windows of different sizes, seeds and lifetimes must give bit for bit what
each window gives alone in fast mode; the captured graph must replay the
eager batched iteration bit for bit; the autotuner must not change results."""
from __future__ import annotations

import pytest
import torch

from globe_hip import stubs
from globe_hip.erosion import batch as bt

pytestmark = pytest.mark.gpu

FIELDS = ("height", "sediment", "discharge", "momentum", "pending")


def _batch(cuda, n=6, seed=3, lo=24, hi=96, chunks=4, **erosion):
    wins = stubs.zoom_windows(n, seed=seed, side_lo=lo, side_hi=hi, life_lo=16, life_hi=96, **erosion)
    return bt.WindowBatch([w for w, _ in wins], [p for _, p in wins], device=cuda, chunks=chunks)


#: the reference's zoom profile (flood every 2 here, to refresh within a short run)
ZOOM = dict(stubs.ZOOM_PROFILE_305M, flood_every=2)


def test_zoom_profile_batched_equals_per_window(cuda):
    """The reference's zoom-tile erosion profile -- routing flood, talus
    pass, slope limits, discharge scales in cells of rain, the erf law --
    batched in one launch per pass and one fast-mode pass per window: the
    same bits."""
    a, b = _batch(cuda, seed=4, **ZOOM), _batch(cuda, seed=4, **ZOOM)
    for it in range(4):
        a.step(it)
        b.step_per_window(it)
        torch.cuda.synchronize()
        _same(a, b)
    assert all(st.route is not None for st in a.states)
    assert float((a.height - _batch(cuda, seed=4, **ZOOM).height).abs().max()) > 0


def test_zoom_profile_graph_replays_match_eager(cuda):
    a, b = _batch(cuda, seed=6, **ZOOM), _batch(cuda, seed=6, **ZOOM)
    g = bt.BatchGraph(b)
    for it in range(4):
        a.step(it)
        g.step(it)
        torch.cuda.synchronize()
        _same(a, b)


def _same(a, b):
    for n in FIELDS:
        x, y = getattr(a, n), getattr(b, n)
        assert torch.equal(x, y), (n, float((x - y).abs().max()))


def test_batched_equals_per_window(cuda):
    a, b = _batch(cuda), _batch(cuda)
    print("windows", a.NE, "lifetimes", a.max_steps, flush=True)
    for it in range(3):
        a.step(it)
        b.step_per_window(it)
        torch.cuda.synchronize()
        print(f"iteration {it}: moves batched {a.moves()} per window {b.moves(per_window=True)}", flush=True)
        _same(a, b)
    assert a.moves() > 0
    assert float((a.height - _batch(cuda).height).abs().max()) > 0


def test_batch_tape_replay_equals_the_traced_applying_pass(cuda):
    """The batch's census records every move and the applying pass replays
    it; with the tape off the applying pass traces again: bitwise the same,
    and no droplet is ever left untraced for want of tape rows."""
    from globe_hip import _lib

    a, b = _batch(cuda, seed=3), _batch(cuda, seed=3)
    b.use_tape = False
    for it in range(3):
        a.step(it)
        b.step(it)
        torch.cuda.synchronize()
        _same(a, b)
        assert int(a.stats[_lib.FS["TAPE_SHORT"]]) == 0
    assert a.tape is not None and b.tape is None


def test_batch_graph_replays_match_eager(cuda):
    a, b = _batch(cuda, seed=5), _batch(cuda, seed=5)
    g = bt.BatchGraph(b)
    for it in range(3):
        a.step(it)
        g.step(it)
        torch.cuda.synchronize()
        _same(a, b)


def test_autotuner_keeps_results(cuda):
    a, b = _batch(cuda, seed=7), _batch(cuda, seed=7)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    tuner = bt.Autotuner(b, candidates=[(cus, 0), (2 * cus, 1), (2 * cus, 2)])
    for it in range(5):
        a.step(it)
        tuner.step(b.step, it)
        torch.cuda.synchronize()
        _same(a, b)
    print("autotuner times (ms)", tuner.times, "best", tuner.best, flush=True)
    assert tuner.best in tuner.candidates


def _budget_closes(raw):
    """The fixed-point mass ledger: deposited - eroded + new stockpile + lost
    offshore + exits + vanished = released.  The totals are int64 sums of
    2^-48-cell amounts, accumulated with wrapping int64 atomics; over a
    15.6 M-cell batch they exceed 2^63, so the ledger is checked modulo
    2^64, where it is exact (each per-cell amount is far inside the range)."""
    from globe_hip import _lib

    FS = _lib.FS
    r = [int(x) for x in raw]
    lhs = (r[FS["DEPOSITED_FX"]] - r[FS["ERODED_FX"]] + r[FS["TO_PENDING_FX"]] + r[FS["LOST_OFFSHORE_FX"]]
           + r[FS["EXIT_FX"]] + r[FS["VANISHED_FX"]])
    return (lhs - r[FS["RELEASED_FX"]]) % (1 << 64) == 0, lhs, r[FS["RELEASED_FX"]]


def test_full_size_batch_equals_per_window(cuda):
    """The batched-windows configuration at its own size: 64 windows of side
    log-uniform in [128, 1024] (15.6 M interior cells), each its own fBm
    seed, lifetime log-uniform in [32, 256].  Two iterations batched (one
    launch per chunk and pass over all windows) and per window: every state
    array bit for bit, and the batch's fixed-point mass ledger closes exactly."""
    wins = stubs.zoom_windows(64, seed=0)
    a = bt.WindowBatch([w for w, _ in wins], [p for _, p in wins], device=cuda)
    b = bt.WindowBatch([w for w, _ in wins], [p for _, p in wins], device=cuda)
    sides = sorted(int(x) - 4 for x in a.NE)
    lives = sorted(int(x) for x in a.max_steps)
    print(f"64 windows: sides {sides[0]}..{sides[-1]}, lifetimes {lives[0]}..{lives[-1]}", flush=True)
    assert len(a.NE) == 64 and sides[0] >= 128 and sides[-1] <= 1024 and lives[0] >= 32 and lives[-1] <= 256
    for it in range(2):
        a.step(it)
        b.step_per_window(it)
        torch.cuda.synchronize()
        print(f"iteration {it}: moves batched {a.moves()} per window {b.moves(per_window=True)}", flush=True)
        assert a.moves() == b.moves(per_window=True)
        # the batch's passes count into its own statistics, each window's spawn (the stockpiles it
        # releases) into the window's: the ledger is over both
        ok, lhs, rel = _budget_closes(a.stats.cpu().numpy() + a.win_stats.sum(0).cpu().numpy())
        assert ok, (it, lhs, rel)
        _same(a, b)


def test_lateral_cuts_batched_equals_per_window(cuda):
    """Bank cuts (erosion.lateral_rate) in the batched launch and one
    fast-mode pass per window: the same bits, with the tape and without."""
    a, b, c = (_batch(cuda, seed=8, lateral_rate=20.0) for _ in range(3))
    c.use_tape = False
    for it in range(3):
        a.step(it)
        b.step_per_window(it)
        c.step(it)
        torch.cuda.synchronize()
        _same(a, b)
        _same(a, c)
    assert float((a.height - _batch(cuda, seed=8).height).abs().max()) > 0
