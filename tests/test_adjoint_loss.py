"""The differentiable mode's loss on the CPU: the Beta(2, 5) target against
scipy's, the smoothed hypsometric CDF against the empirical one in the
sharp limit, and the loss's autograd gradient against finite differences."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.stats
import torch

from globe_hip.erosion import adjoint as adj


def test_beta25_cdf_is_scipys():
    x = torch.linspace(0.0, 1.0, 257, dtype=torch.float64)
    ref = scipy.stats.beta(2, 5).cdf(x.numpy())
    assert np.allclose(adj.beta25_cdf(x).numpy(), ref, rtol=0, atol=1e-14)


def test_soft_cdf_sharp_limit_is_the_empirical_cdf():
    rng = np.random.default_rng(1)
    h = torch.tensor(rng.uniform(-500, 4500, 5000))
    land = h > 0
    bins, zmax = 64, 4000.0
    z = np.arange(1, bins + 1) * zmax / bins
    hl = np.sort(h[land].numpy())
    emp = np.searchsorted(hl, z, side="right") / hl.size
    # loss with a vanishing width: sum (cdf - target)^2 + weight * (mean - m)^2
    tgt = scipy.stats.beta(2, 5).cdf(z / zmax)
    want = np.sum((emp - tgt) ** 2) + 0.1 * (hl.mean() - 800.0) ** 2
    got = float(adj.land_hypsometry_loss(h, land, tau_m=1e-6))
    assert got == pytest.approx(want, rel=1e-9)


def test_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    h = torch.tensor(rng.uniform(-300, 3000, 64), requires_grad=True)
    land = (h > 0).detach()
    assert torch.autograd.gradcheck(lambda x: adj.land_hypsometry_loss(x, land, tau_m=200.0), (h,), eps=1e-4,
                                    atol=1e-6, rtol=1e-4)
