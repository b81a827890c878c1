"""The differentiable mode on the GPU: frozen-trajectory replay, the halo
exchange's transpose, and the hand-written adjoint against central finite
differences of the same frozen map and against the float64 adjoint.

Bars: replay at the nominal parameters reproduces the nominal run bit for
bit; <H^T a, x> = <a, H x> to 1e-12 relative; adjoint vs central
differences of the frozen map (float64, relative step 1e-6; 1e-3 and 1e-5
are printed beside them: larger steps cross the kinks of the capped
exchange) within 1e-4 relative on every checked component (plus 1e-5 of
the largest for the field's small components);
float32 adjoint vs float64 adjoint on the same trajectories within 1e-2.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from globe_hip import stubs
from globe_hip.erosion import adjoint as adj
from globe_hip.erosion import maps
from globe_hip.erosion import particle as pk

pytestmark = pytest.mark.gpu


def _state(N, cuda, n_blobs=8, **erosion):
    grid, f = stubs.earth_globe(N, 2, n_blobs=n_blobs)
    p = stubs.workload_params(N, 2, **erosion)
    st = maps.ErosionState.from_grid(grid, f["bedrock_m"], f["hardness"], f["precip"], f["evap"], f["uplift_m"],
                                     p.erosion, device=cuda, mode="fast")
    return st, p


def _erod(st, seed=0):
    g = torch.Generator().manual_seed(seed)
    e = 1.0 + 0.2 * (2.0 * torch.rand(tuple(st.height.shape), generator=g, dtype=torch.float64) - 1.0)
    return e.to(torch.float32).to(st.device)


def test_replay_at_nominal_is_the_nominal_run(cuda):
    st, p = _state(32, cuda)
    run = adj.GradRun(st, p, iterations=4, real="f64", checkpoint_every=2)
    erod, theta = _erod(st), adj.theta_nominal(run)
    print("grad mode N=32 T=4 f64: nominal then replay", flush=True)
    a = run.forward(erod, theta, record=True)
    b = run.replay(erod, theta)
    for n in ("height", "sediment"):
        assert torch.equal(a[n], b[n]), n
    # something happened
    assert float((a["height"] + a["sediment"] - run.s0["height"] - run.s0["sediment"]).abs().max()) > 0


def test_halo_adjoint_is_the_transpose(cuda):
    st, p = _state(32, cuda)
    ht = st.halo_tables()
    g = torch.Generator(device="cpu").manual_seed(3)
    x = torch.randn(tuple(st.height.shape), generator=g, dtype=torch.float64).to(cuda)
    a = torch.randn(tuple(st.height.shape), generator=g, dtype=torch.float64).to(cuda)
    print("halo transpose N=32", flush=True)
    hx = x.clone()
    ht.exchange(hx)  # halo cells <- H x_interior
    hta = a.clone()
    ht.adjoint(hta)
    # the exchange as a linear map y = E x (interior kept, halo replaced);
    # its transpose applied to a must satisfy <E^T a, x> = <a, E x>
    lhs = float((hta * x).sum())
    rhs = float((a * hx).sum())
    assert lhs == pytest.approx(rhs, rel=1e-12, abs=1e-12)
    # destinations are never sources
    assert not np.isin(ht.dst.cpu().numpy(), ht.src.cpu().numpy()).any()


def _fd_check(run, erod, theta, loss_of, cells, h_rel=1e-3):
    """Central differences of the frozen map, relative step h_rel (the
    erodibility field is float32: its steps are the ones float32 took)."""
    out = run.replay(erod, theta)
    base = loss_of(out["height"] + out["sediment"])
    fd_e, fd_t = [], []
    for c in cells:
        e1, e2 = erod.clone(), erod.clone()
        h = h_rel * float(erod.view(-1)[c])
        e1.view(-1)[c] += h
        e2.view(-1)[c] -= h
        h = float(e1.view(-1)[c].double() - e2.view(-1)[c].double()) / 2.0
        lp = loss_of(sum(run.replay(e1, theta)[n] for n in ("height", "sediment")))
        lm = loss_of(sum(run.replay(e2, theta)[n] for n in ("height", "sediment")))
        fd_e.append((lp - lm) / (2 * h))
    for i in range(6):
        t1, t2 = theta.clone(), theta.clone()
        h = h_rel * (abs(float(theta[i])) if float(theta[i]) != 0.0 else 1.0)
        t1[i] += h
        t2[i] -= h
        lp = loss_of(sum(run.replay(erod, t1)[n] for n in ("height", "sediment")))
        lm = loss_of(sum(run.replay(erod, t2)[n] for n in ("height", "sediment")))
        fd_t.append((lp - lm) / (2 * h))
    return base, np.array(fd_e), np.array(fd_t)


@pytest.mark.parametrize("ce,feedback", [(2, True), (6, True), (2, False)])
def test_adjoint_matches_central_differences(cuda, ce, feedback):
    T = 6
    st, p = _state(32, cuda)
    run = adj.GradRun(st, p, iterations=T, real="f64", checkpoint_every=ce, feedback=feedback)
    erod = _erod(st).requires_grad_(True)
    theta = adj.theta_nominal(run).requires_grad_(True)
    print(f"adjoint vs FD: N=32 T={T} checkpoint every {ce} feedback {feedback} f64", flush=True)
    surf = adj.erosion(erod, theta, run)
    land = (surf[st.interior].reshape(-1) > 0).detach()
    loss = adj.hypsometry_loss(surf, run, land=land)
    loss.backward()
    ge = erod.grad.detach().cpu().numpy().reshape(-1)
    gt = theta.grad.detach().numpy()
    # the frozen map: record the nominal trajectories, replay them
    run.forward(erod.detach(), theta.detach(), record=True, tapes="none")
    loss_of = lambda s: float(adj.hypsometry_loss(s, run, land=land))
    inter = np.zeros(st.height.shape, bool)
    inter[st.interior] = True
    cand = np.flatnonzero(inter.reshape(-1) & (np.abs(ge) > 0))
    rng = np.random.default_rng(0)
    big = cand[np.argsort(-np.abs(ge[cand]))[:8]]
    cells = np.concatenate([big, rng.choice(cand, 8, replace=False)])
    ad_e = ge[cells]
    print("erod  adjoint", ad_e, flush=True)
    print("theta adjoint", gt, flush=True)
    # the frozen map is piecewise smooth (caps, the c_eq > sed switch): steps
    # of 1e-3 and 1e-5 cross the kinks of the droplets nearest a switch, so
    # the bar is on the converged step; the others are printed beside it
    res = {}
    for h_rel in (1e-3, 1e-5, 1e-6):
        base, fd_e, fd_t = _fd_check(run, erod.detach(), theta.detach(), loss_of, cells, h_rel=h_rel)
        assert base == pytest.approx(float(loss), rel=1e-12)
        res[h_rel] = (fd_e, fd_t)
        print(f"h_rel {h_rel:g}: erod fd", fd_e, "\n  theta fd", fd_t, flush=True)
    fd_e, fd_t = res[1e-6]
    scale_e = np.abs(ad_e).max()
    assert np.all(np.abs(ad_e - fd_e) <= 1e-4 * np.abs(ad_e) + 1e-5 * scale_e), (ad_e, fd_e)
    for i in range(6):
        assert abs(gt[i] - fd_t[i]) <= 1e-4 * max(abs(gt[i]), abs(fd_t[i])) + 1e-9, (adj.THETA_NAMES[i], gt[i], fd_t[i])


def test_float32_adjoint_tracks_float64(cuda):
    T = 6
    st, p = _state(48, cuda)
    run32 = adj.GradRun(st, p, iterations=T, real="f32", checkpoint_every=3)
    erod = _erod(st)
    theta = adj.theta_nominal(run32)
    print(f"f32 vs f64 adjoint: N=48 T={T}", flush=True)
    run32.forward(erod, theta, record=True, tapes="none")
    e32, t32 = erod.clone().requires_grad_(True), theta.clone().requires_grad_(True)
    s32 = adj.erosion(e32, t32, run32)
    land = (s32[st.interior].reshape(-1) > 0).detach()
    adj.hypsometry_loss(s32, run32, land=land).backward()
    run64 = adj.GradRun(st, p, iterations=T, real="f64", checkpoint_every=3, replay_of=run32)
    e64, t64 = erod.clone().requires_grad_(True), theta.clone().requires_grad_(True)
    s64 = adj.erosion(e64, t64, run64)
    adj.hypsometry_loss(s64, run64, land=land).backward()
    a, b = e32.grad.double(), e64.grad.double()
    rel = float((a - b).norm() / b.norm())
    print("erod grad rel diff", rel, "theta", t32.grad.numpy(), t64.grad.numpy(), flush=True)
    assert rel < 1e-2
    tg32, tg64 = t32.grad.numpy(), t64.grad.numpy()
    assert np.all(np.abs(tg32 - tg64) <= 1e-2 * np.abs(tg64).max())


@pytest.mark.parametrize("real", ["f32", "f64"])
def test_adjoint_is_bitwise_reproducible(cuda, real):
    """The reverse sweep accumulates in 128-bit fixed point and the halo
    transpose gathers in a fixed order: two backward passes over the same
    run give the same bits (double atomics did not), with a different
    persistent grid for the forward recomputation."""
    T = 4
    st, p = _state(48, cuda)
    run = adj.GradRun(st, p, iterations=T, real=real, checkpoint_every=2, feedback=True)
    erod = _erod(st)
    theta = adj.theta_nominal(run)
    print(f"adjoint reproducibility N=48 T={T} {real}", flush=True)
    grads = []
    for rep in range(3):
        e, t = erod.clone().requires_grad_(True), theta.clone().requires_grad_(True)
        s = adj.erosion(e, t, run)
        adj.hypsometry_loss(s, run).backward()
        grads.append((e.grad.clone(), t.grad.clone()))
    for ge, gt in grads[1:]:
        assert torch.equal(ge, grads[0][0])
        assert torch.equal(gt, grads[0][1])
    assert float(grads[0][0].abs().max()) > 0 and float(grads[0][1].abs().max()) > 0


@pytest.mark.parametrize("real", ["f32", "f64"])
def test_tape_in_pieces_gives_the_whole_tapes_gradients(cuda, real):
    """The backward pass with the tape in pieces (each iteration re-traced
    piece by piece from its starting state, the pieces' reverse sweeps
    summed in the fixed-point accumulators, one flush) gives the one-tape
    gradients bit for bit; the piece size need not divide the spawn."""
    T = 4
    st, p = _state(48, cuda)
    whole = adj.GradRun(st, p, iterations=T, real=real, checkpoint_every=2, feedback=True)
    batch = max(1, whole.P // 3 + 17)
    pieces = adj.GradRun(st, p, iterations=T, real=real, checkpoint_every=2, feedback=True, tape_batch=batch)
    assert pieces.tape_batch == batch and pieces.tape_bytes < whole.tape_bytes
    erod = _erod(st)
    theta = adj.theta_nominal(whole)
    print(f"tape in pieces N=48 T={T} {real} P={whole.P} batch={batch}", flush=True)
    out = []
    for run in (whole, pieces):
        e, t = erod.clone().requires_grad_(True), theta.clone().requires_grad_(True)
        s = adj.erosion(e, t, run)
        adj.hypsometry_loss(s, run).backward()
        out.append((s.detach().clone(), e.grad.clone(), t.grad.clone()))
    assert torch.equal(out[0][0], out[1][0])
    assert torch.equal(out[0][1], out[1][1])
    assert torch.equal(out[0][2], out[1][2])
    assert float(out[0][1].abs().max()) > 0 and float(out[0][2].abs().max()) > 0
    # a piece records its tape only: the starting state it steers on (the run's own) is not written
    for name in whole.s0:
        assert torch.equal(whole.s0[name], pieces.s0[name]), name


def test_states_in_host_memory_give_the_same_gradients(cuda):
    """store="host": the starting state, checkpoints, a segment's starting
    states and a recorded run's nominal states in pinned host memory --
    the forward result, the gradients and a replay are the device store's
    bits (with the tape whole and in pieces)."""
    T = 5
    st, p = _state(48, cuda)
    erod = _erod(st)
    out = []
    for store, batch in (("device", None), ("host", None), ("host", 2000)):
        run = adj.GradRun(st, p, iterations=T, real="f32", checkpoint_every=2, tape_batch=batch, store=store)
        theta = adj.theta_nominal(run)
        print(f"state store {store}, tape_batch {batch}, N=48 T={T}", flush=True)
        e, t = erod.clone().requires_grad_(True), theta.clone().requires_grad_(True)
        s = adj.erosion(e, t, run)
        adj.hypsometry_loss(s, run).backward()
        run.forward(erod, theta, record=True, tapes="none")
        r = run.replay(erod, theta)
        out.append((s.detach().clone(), e.grad.clone(), t.grad.clone(), r["height"].clone()))
        if store == "host":
            assert not run.checkpoints[2]["height"].is_cuda and run.checkpoints[2]["height"].is_pinned()
    for o in out[1:]:
        for a, b in zip(out[0], o):
            assert torch.equal(a, b)
