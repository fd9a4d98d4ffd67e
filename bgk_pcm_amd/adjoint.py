"""Gradients of a probe-temperature loss through the GPU time step.

`ProbeTemperature` is a torch.autograd.Function over the C ABI's adjoint
(bgk_adj_setup / bgk_adj_forward / bgk_adj_backward of include/bgk_pcm.h):

    theta = torch.tensor([stefan_number, k_ratio, t_hot], dtype=torch.float64,
                         requires_grad=True)
    T = probe_temperature(theta, base_params, probes_i, probes_j, nsteps, seg_len)
    loss = ((T - T_target) ** 2).mean()
    loss.backward()          # theta.grad = dL/d(St, k_ratio, t_hot)

Forward model: the FP64 moment-storage step (precision BGK_FP64_MOM, or
BGK_FP32_MOM through `precision`: the adjoint is then that of the step at
the FP32-stored states; clean rounding of STEP_7's collision) of the no-slip
cavity described by
`base_params` (the input.dat records), with the library's extensions
t_hot (hot-wall and initial liquid temperature in place of T0) and k_ratio
(solid/liquid conductivity ratio).  T[n, p] is the averaged temperature of
STEP_7/moments.f90:11-25 at interior node (probes_i[p], probes_j[p]) after
step n+1.  The state is checkpointed every seg_len steps on the GPU (and
every seg_outer steps first, with two levels) and recomputed segment by
segment in the backward pass; with n_states >= 0 the binomial (revolve)
schedule for that many state slots is used instead (0: as many as the
device's free memory holds).  The tensors live on
the CPU (the probe record is small); the lattice stays on the GPU.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from .lattice import Lattice
from .params import FP64_MOM, ROUND_CLEAN, Params


class ProbeTemperature(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta: torch.Tensor, base: Params, pi: Sequence[int], pj: Sequence[int],
                nsteps: int, seg_len: int, seg_outer: int = 0,
                precision: int = FP64_MOM, n_states: int = -1) -> torch.Tensor:
        st, kr, th = (float(x) for x in theta.detach().cpu().tolist())
        p = base.replace(precision=precision, rounding=ROUND_CLEAN, stefan_number=st)
        lat = Lattice(p, initialize=False, t_hot=th, k_ratio=kr)
        lat.initialization()
        if n_states >= 0:
            lat.adj_setup_binomial(nsteps, n_states, pi, pj)
        else:
            lat.adj_setup(nsteps, seg_len, pi, pj, seg_outer=seg_outer)
        rec = lat.adj_forward()
        ctx.lat = lat
        return torch.from_numpy(rec)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        lat = ctx.lat
        g = lat.adj_backward(grad_out.detach().cpu().numpy().astype(np.float64))
        lat.close()
        ctx.lat = None
        return torch.from_numpy(g), None, None, None, None, None, None, None, None


def probe_temperature(theta: torch.Tensor, base: Params, pi: Sequence[int], pj: Sequence[int],
                      nsteps: int, seg_len: int = 100, seg_outer: int = 0,
                      precision: int = FP64_MOM, n_states: int = -1) -> torch.Tensor:
    """T at the probes after each of nsteps steps, differentiable in
    theta = (stefan_number, k_ratio, t_hot)."""
    return ProbeTemperature.apply(theta, base, list(pi), list(pj), int(nsteps), int(seg_len),
                                  int(seg_outer), int(precision), int(n_states))


def seeded_probes(nx: int, ny: int, n: int, seed: int) -> tuple[list[int], list[int]]:
    """n distinct interior nodes drawn with numpy's PCG64 from `seed`."""
    rng = np.random.default_rng(seed)
    flat = rng.choice(nx * ny, size=n, replace=False)
    return [int(q % nx) + 1 for q in flat], [int(q // nx) + 1 for q in flat]
