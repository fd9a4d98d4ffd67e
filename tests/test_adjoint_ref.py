"""The differentiable CPU reference of the adjoint tests (tests/torch_ref.py)
against the oracle, and torch.autograd against finite differences on it.

The reference is pinned here the way the oracle is: its forward must give
the oracle's fields (clean rounding) with the extensions off and on, so
that its autograd gradient is the gradient of the same model the GPU
adjoint differentiates."""
import numpy as np
import pytest
import torch

from bgk_pcm_amd.params import ROUND_CLEAN, Params
from oracle_py import Oracle
from torch_ref import Cavity, probe_record


def run_both(p, t_hot, k_ratio, nsteps):
    theta = torch.tensor([p.stefan_number, k_ratio, t_hot], dtype=torch.float64)
    cav = Cavity(p, theta)
    for _ in range(nsteps):
        Trep = cav.step()
    o = Oracle(p, t_hot=t_hot, k_ratio=k_ratio)
    o.step(nsteps)
    return cav, Trep, o


@pytest.mark.parametrize("t_hot,k_ratio,tau_f", [(283.0, 1.0, 1.0), (283.5, 1.5, 1.0),
                                                 (284.0, 0.5, 0.8)])
def test_reference_forward_equals_oracle(t_hot, k_ratio, tau_f):
    p = Params(nx=41, ny=23, length=41.0, grav=1e-3, rounding=ROUND_CLEAN, tau_f=tau_f)
    cav, Trep, o = run_both(p, t_hot, k_ratio, 80)
    T = o.field("T")[1:-1, 1:-1]
    phi = o.field("phi")[1:-1, 1:-1]
    np.testing.assert_allclose(Trep.numpy(), T, rtol=1e-13, atol=0)
    np.testing.assert_allclose(cav.phi[1:-1, 1:-1].numpy(), phi, rtol=0, atol=1e-12)
    for k in (0, 3, 6):
        np.testing.assert_allclose(cav.f[k][1:-1, 1:-1].numpy(), o.field(f"f{k}")[1:-1, 1:-1],
                                   rtol=1e-13, atol=0)
        np.testing.assert_allclose(cav.g[k][1:-1, 1:-1].numpy(), o.field(f"g{k}")[1:-1, 1:-1],
                                   rtol=1e-13, atol=0)


def test_reference_autograd_matches_central_differences():
    p = Params(nx=14, ny=11, length=14.0, grav=2e-3, rounding=ROUND_CLEAN, stefan_number=0.1)
    pi, pj = [3, 7, 12], [2, 6, 10]
    theta = torch.tensor([0.1, 1.5, 283.5], dtype=torch.float64, requires_grad=True)
    target = probe_record(p, torch.tensor([0.12, 1.2, 283.2], dtype=torch.float64), pi, pj, 30)
    loss = ((probe_record(p, theta, pi, pj, 30) - target) ** 2).mean()
    loss.backward()
    g = theta.grad.numpy()
    for c, h in enumerate((1e-6, 1e-6, 1e-5)):
        tp = theta.detach().clone()
        tm = theta.detach().clone()
        tp[c] += h
        tm[c] -= h
        lp = ((probe_record(p, tp, pi, pj, 30) - target) ** 2).mean().item()
        lm = ((probe_record(p, tm, pi, pj, 30) - target) ** 2).mean().item()
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[c]) <= 1e-6 * abs(g[c]) + 1e-14, (c, fd, g[c])
