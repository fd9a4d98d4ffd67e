"""Differentiable CPU restatement of the no-slip cavity step (torch, FP64).

Test infrastructure: the independent reference for the GPU adjoint.  It
restates the clean-rounding STEP_7 step in population form, the way the
oracle does (BC_f type 4, BC_g, streaming, momentsPre, phase_field, forces,
regularized collision, moments: STEP_7/main.f90:60-69), with the library's
two extensions (hot-wall / initial liquid temperature t_hot, conductivity
ratio k_ratio), written with torch operations so that torch.autograd gives
the gradient of a probe-temperature loss with respect to (stefan_number,
k_ratio, t_hot).  Its forward is pinned to the oracle in
tests/test_adjoint_ref.py; the GPU adjoint is checked against its gradient
in tests/test_gpu_adjoint.py.  Arrays are (Ny+2, Nx+2), indexed [j, i].
"""
from __future__ import annotations

import torch

from bgk_pcm_amd.params import Params

W = [4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36]
CX = [0, 1, 0, -1, 0, 1, -1, -1, 1]
CY = [0, 0, 1, 0, -1, 1, 1, -1, -1]
OPP = [0, 3, 4, 1, 2, 7, 8, 5, 6]
CS2 = 1.0 / 3.0
# phase-field taps C, E, N, W, S, NE, NW, SW, SE (STEP_7/phase_field.f90:27-45)
TAPS = [(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)]


def equilibria(m, u, v):
    """STEP_7/collision.f90:17-41 (and initialization.f90:104-128)."""
    uu, vv = u * u, v * v
    uv = 1.5 * (uu + vv)
    upv, umv = u + v, u - v
    return [W[0] * m * (1.0 - uv),
            W[1] * m * (1.0 + 3.0 * u + 4.5 * uu - uv),
            W[2] * m * (1.0 + 3.0 * v + 4.5 * vv - uv),
            W[3] * m * (1.0 - 3.0 * u + 4.5 * uu - uv),
            W[4] * m * (1.0 - 3.0 * v + 4.5 * vv - uv),
            W[5] * m * (1.0 + 3.0 * upv + 4.5 * upv * upv - uv),
            W[6] * m * (1.0 - 3.0 * umv + 4.5 * umv * umv - uv),
            W[7] * m * (1.0 - 3.0 * upv + 4.5 * upv * upv - uv),
            W[8] * m * (1.0 + 3.0 * umv + 4.5 * umv * umv - uv)]


def regularised(pops, m, u, v, omega):
    """STEP_7/collision.f90:44-150, clean rounding (temporaries in FP64)."""
    eq = equilibria(m, u, v)
    d = [pops[k] - eq[k] for k in range(9)]
    pxx = sum((CX[k] ** 2 - CS2) * d[k] for k in range(9))
    pyy = sum((CY[k] ** 2 - CS2) * d[k] for k in range(9))
    pxy = d[5] - d[6] + d[7] - d[8]
    out = []
    for k in range(9):
        h = (CX[k] ** 2 - CS2) * pxx + (CY[k] ** 2 - CS2) * pyy + 2.0 * CX[k] * CY[k] * pxy
        t = eq[k] + (0.5 * W[k]) / (CS2 * CS2) * h
        out.append(t - omega * (t - eq[k]))
    return out


class Cavity:
    """State and step of the no-slip cavity; theta = (St, k_ratio, t_hot)."""

    def __init__(self, p: Params, theta: torch.Tensor):
        self.p, self.nx, self.ny = p, p.nx, p.ny
        nx, ny = p.nx, p.ny
        st, kr, th = theta[0], theta[1], theta[2]
        self.st, self.kr, self.th = st, kr, th
        self.omf = 1.0 / p.tau_f
        self.alpha_l = CS2 * (p.tau_g - 0.5)
        self.omg_l = 1.0 / p.tau_g
        self.dT = p.t0 - p.t_in_sol
        self.beta = (1.0 - p.rho1 / p.rho0) / self.dT
        interf = nx // 40
        dt = torch.float64
        i = torch.arange(nx + 2).view(1, -1).expand(ny + 2, nx + 2)
        j = torch.arange(ny + 2).view(-1, 1).expand(ny + 2, nx + 2)
        liquid = i <= interf
        # STEP_7/initialization.f90:15-152 with T0 -> t_hot for the hot wall
        # and the liquid strip
        T = torch.where(liquid, th * torch.ones((), dtype=dt),
                        torch.full((ny + 2, nx + 2), p.t_in_sol, dtype=dt))
        rho = torch.where(liquid, torch.full((), p.rho1, dtype=dt),
                          torch.full((ny + 2, nx + 2), p.rho0, dtype=dt))
        u0 = p.re * CS2 * (p.tau_f - 0.5) / p.length
        u = torch.full((ny + 2, nx + 2), u0, dtype=dt)
        u = torch.where((j == 0) | (j == ny + 1), torch.zeros((), dtype=dt), u)
        v = torch.zeros((ny + 2, nx + 2), dtype=dt)
        self.uh, self.vh = u.detach().clone(), v.detach().clone()  # ghost velocities
        self.f = equilibria(rho, u, v)
        self.g = equilibria(T, u, v)
        self.phi = torch.where(liquid, torch.ones((), dtype=dt),
                               -torch.ones((ny + 2, nx + 2), dtype=dt))
        self.teast = T[1:ny + 1, nx]
        self.frozen6 = float(W[6] * p.rho0)
        self.frozen7 = float(W[7] * p.rho0)
        # which pulls of an interior node come from a ghost node
        ii = torch.arange(1, nx + 1).view(1, -1).expand(ny, nx)
        jj = torch.arange(1, ny + 1).view(-1, 1).expand(ny, nx)
        self.ghost = [(ii - CX[k] < 1) | (ii - CX[k] > nx) | (jj - CY[k] < 1) | (jj - CY[k] > ny)
                      for k in range(9)]
        self.frozen = {6: (ii == nx) & (jj == 1), 7: (ii == nx) & (jj == ny)}

    def _pull(self, a, k):
        nx, ny = self.nx, self.ny
        return a[1 - CY[k]:ny + 1 - CY[k], 1 - CX[k]:nx + 1 - CX[k]]

    def step(self):
        nx, ny, p = self.nx, self.ny, self.p
        Jn, In = slice(1, ny + 1), slice(1, nx + 1)
        # BC_f type 4: bounce-back of the own opposite population, except the
        # two ghosts BC_f never writes (their initial equilibria)
        fp = []
        for k in range(9):
            own = self.f[OPP[k]][Jn, In]
            val = torch.where(self.ghost[k], own, self._pull(self.f[k], k))
            if k in self.frozen:
                val = torch.where(self.frozen[k], torch.full((), getattr(self, f"frozen{k}"),
                                                              dtype=val.dtype), val)
            fp.append(val)
        # BC_g: W/E ghost columns at Tin and at the previous averaged T(Nx, j),
        # then rows 0 / Ny+1 copied from rows 1 / Ny over the whole width
        weq = equilibria(self.th, self.uh[Jn, 0], self.vh[Jn, 0])
        eeq = equilibria(self.teast, self.uh[Jn, nx + 1], self.vh[Jn, nx + 1])
        gp = []
        for k in range(9):
            mid = torch.cat([weq[k].view(-1, 1), self.g[k][Jn, In], eeq[k].view(-1, 1)], dim=1)
            ext = torch.cat([mid[:1], mid, mid[-1:]], dim=0)
            gp.append(self._pull(ext, k))
        # phi ghost rows (BC_g), taps
        ph = torch.cat([self.phi[1:2], self.phi[1:ny + 1], self.phi[ny:ny + 1]], dim=0)
        taps = [ph[1 + dy:ny + 1 + dy, 1 + dx:nx + 1 + dx] for dx, dy in TAPS]
        # momentsPre
        rho = sum(fp)
        T2 = sum(gp)
        u2 = (fp[1] - fp[3] + fp[5] - fp[6] - fp[7] + fp[8]) / rho
        v2 = (fp[2] - fp[4] + fp[5] + fp[6] - fp[7] - fp[8]) / rho
        # phase_field
        th_ = torch.tanh((T2 - 273.0) / p.t_width)
        rp, rm = 0.5 * (1.0 + th_), 0.5 * (1.0 - th_)
        kp = sum((0.5 * (1.0 + t)) ** 2 for t in taps) / 8.0 * rp
        km = sum((0.5 * (1.0 - t)) ** 2 for t in taps) / 8.0 * rm
        pc = taps[0]
        phi = pc + (0.5 * kp * (1.0 - pc) - 0.5 * km * (1.0 + pc))
        # forces
        T = T2 - (self.dT / self.st * (phi - pc) / 2.0)
        fy = rho * p.grav * self.beta * T * (1.0 + phi / 2.0)
        fx = -(u2 * (1.0 - phi) * 0.5)
        fy = fy - v2 * (1.0 - phi) * 0.5
        u = u2 + fx / (self.omf * rho)
        v = v2 + fy / (self.omf * rho)
        # collision, conductivity ratio through omega_g(phi)
        fl = 0.5 * (1.0 + phi)
        alpha = self.alpha_l * (fl + self.kr * (1.0 - fl))
        omg = 1.0 / (0.5 + alpha / CS2)
        fo = regularised(fp, rho, u, v, self.omf)
        go = regularised(gp, T, u, v, omg)
        # moments (averaged T)
        Trep = (sum(go) + T2) * 0.5
        pad = lambda a: torch.nn.functional.pad(a, (1, 1, 1, 1))
        self.f = [pad(x) for x in fo]
        self.g = [pad(x) for x in go]
        self.phi = torch.cat([self.phi[:, :1], torch.cat([self.phi[:1, 1:nx + 1], phi,
                                                          self.phi[-1:, 1:nx + 1]], dim=0),
                              self.phi[:, -1:]], dim=1)
        self.teast = Trep[:, nx - 1]
        return Trep


def probe_record(p: Params, theta: torch.Tensor, pi, pj, nsteps: int) -> torch.Tensor:
    """T (averaged) at probes (pi, pj) after each step, (nsteps, n_probes)."""
    cav = Cavity(p, theta)
    ii = torch.tensor([i - 1 for i in pi])
    jj = torch.tensor([j - 1 for j in pj])
    out = []
    for _ in range(nsteps):
        Trep = cav.step()
        out.append(Trep[jj, ii])
    return torch.stack(out)
