"""The five benchmark configurations of the BGK_PCM time step, as lattice
parameters, physics extensions and run lengths.  bench.py measures them and
tests/test_gpu_configs.py checks them; both take the definitions from here.

Relations used (STEP_7/diag.f90:31-40, setup.f90:102-105):
  Pr = (2 tau_f - 1)/(2 tau_g - 1),  nu = (tau_f - 1/2)/3,
  Ra = Pr grav (rho0/rho1 - 1) Ny^3 / nu^2,
  theta = (T - 273)/(T0 - T_in_sol): "solid at theta = 0, hot wall at 1" is
  STEP_7's T_in_sol = 273, T0 = 283 (UTIL/input.dat.0), which every
  configuration keeps.

Pr = 0.02 is tau_f = 0.51 with tau_g = 1.0 wherever tau_f is free: at the
same Rayleigh number this is the grav of the reference's own relaxation times
(tau_f 1, tau_g 0.51, Pr 50), because Ra depends on the relaxation times
through (2 tau_f - 1)(2 tau_g - 1) only.  The 4096^2 (Ra 1e7) and 2048^2
(Ra 1e6) cavities run stably for their 20000 steps with it (measured on
MI355X, FP64: max |v| 5.4e-3 and 2.6e-3 lattice units).
"""
from __future__ import annotations

from dataclasses import dataclass, field

from .params import FP32_MOM, FP64_MOM, ROUND_CLEAN, Params, grav_for_rayleigh

# tau_f, tau_g of Pr = 0.02 with tau_f free (see the module docstring)
PR002 = dict(tau_f=0.51, tau_g=1.0)
# the configurations' cold right wall (theta = 0, i.e. T_in_sol) and seeded
# initial noise of amplitude 1e-3 in theta units
LITERAL = dict(t_cold=273.0, noise_amp=1e-3, noise_seed=2024)


@dataclass
class Workload:
    name: str
    params: Params
    nsteps: int
    ext: dict = field(default_factory=dict)      # extensions of the literal configuration
    note: str = ""


def cavity(n: int, ra: float | None = None, **kw) -> Params:
    """UTIL/input.dat.0's no-slip cavity on an n x n lattice; grav set for
    Rayleigh number `ra` by the reference's definition when given."""
    p = Params(nx=n, ny=n, length=float(n), **kw)
    if ra is not None:
        p = p.replace(grav=grav_for_rayleigh(p, ra))
    return p


def config1() -> Workload:
    """256^2 melting cavity, Pr 0.02 at tau_f 0.8 (tau_g 15.5), Stefan 0.1,
    conductivity ratio 1, cold right wall and seeded noise, 10000 steps.
    Its Ra = 1e5 needs grav = 0.124 at these relaxation times: a free-fall
    velocity sqrt(grav beta dT H) of 0.9 lattice units (Mach 1.5), and
    STEP_7 itself reaches NaN before step 1000.  Measuring buoyancy from a
    reference temperature (t_ref 273 K or 278 K instead of STEP_7's absolute
    T) does not cure it: the run is non-finite within 500 steps with FP64
    populations and with FP32 moments alike (MI355X).  grav = 1e-4 is the
    stable run of the same cavity (Ra 81)."""
    return Workload("config1", cavity(256, tau_f=0.8, tau_g=15.5, grav=1e-4, stefan_number=0.1,
                                      iter=10000),
                    10000, dict(LITERAL),
                    "256^2, tau_f 0.8, tau_g 15.5 (Pr 0.02), grav 1e-4, Stefan 0.1")


def config2() -> Workload:
    """2048^2 FP64 throughput run: Ra 1e6, Stefan 0.1, Pr 0.02, cold right
    wall and seeded noise, 20000 steps."""
    return Workload("config2", cavity(2048, 1e6, stefan_number=0.1, iter=20000, **PR002),
                    20000, dict(LITERAL), "2048^2, Ra 1e6, Pr 0.02, Stefan 0.1")


def config3() -> Workload:
    """4096^2 FP32-storage vs FP64: Ra 1e7, Pr 0.02, Stefan 0.05,
    solid/liquid conductivity ratio 2, 20000 steps.  The FP32 forms follow
    the clean rounding of the collision (STEP_7's FP32 temporaries are not
    representable by them); their error is taken against the clean FP64
    run."""
    return Workload("config3", cavity(4096, 1e7, stefan_number=0.05, iter=20000,
                                      rounding=ROUND_CLEAN, **PR002),
                    20000, dict(k_ratio=2.0),
                    "4096^2, Ra 1e7, Pr 0.02, Stefan 0.05, conductivity ratio 2")


def config4() -> Workload:
    """512^2 adjoint: FP64 moments, 2000 steps, checkpoint every 100 steps,
    16 seeded probes; grav for Ra 1e6 at the reference's relaxation times.
    The gradient is taken at (Stefan, k_ratio, T_h) = (0.1, 1, 1) against the
    trajectory at (0.12, 1.5, 1.05)."""
    return Workload("config4", cavity(512, 1e6, stefan_number=0.1, rounding=ROUND_CLEAN,
                                      precision=FP64_MOM),
                    2000, {}, "512^2 adjoint, FP64 moments, Ra 1e6, 2000 steps, K = 100")


def config5(n: int = 8192) -> Workload:
    """8192^2 FP32-storage forward + adjoint: FP32 moments, Ra 1e7, Stefan
    0.1, 5000 steps with binomial checkpointing; the gradient check is the
    same workload at 1024^2 against the FP64 adjoint (two checkpoint levels)."""
    return Workload("config5", cavity(n, 1e7, stefan_number=0.1, rounding=ROUND_CLEAN,
                                      precision=FP32_MOM),
                    5000, {}, f"{n}^2 adjoint, FP32 moments, Ra 1e7, Stefan 0.1")


# the adjoint configurations' parameters (Stefan, k_ratio, T_h in theta units)
ADJ_THETA0 = (0.1, 1.0, 1.0)
ADJ_TARGET = (0.12, 1.5, 1.05)


def t_hot_of(p: Params, th: float) -> float:
    """Hot-wall temperature in kelvin of theta = th (T_crit 273)."""
    return 273.0 + th * (p.t0 - p.t_in_sol)
