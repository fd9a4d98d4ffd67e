"""Scalar diagnostics of the reference, computed on the host from the
device reductions, and the text records the reference writes.

``diag_values`` restates the formulas of STEP_7/diag.f90:15-114 (Grashof,
Prandtl and Rayleigh numbers, Fourier number, the correlation Nusselt
number and the melt-front estimates) from the four field sums that
``bgk_diag_sums`` returns.  Integer arithmetic follows the reference's
default INTEGER: ``Ny**3`` and ``Ny**2`` are 32-bit products (``Ny**3``
wraps for Ny > 1290) and ``Nx/Ny`` is an integer quotient.
"""
from __future__ import annotations

import math
from typing import Sequence, Tuple

from .params import Params


def _i32(x: int) -> int:
    x &= 0xFFFFFFFF
    return x - (1 << 32) if x >= (1 << 31) else x


def _pow(x: float, y: float) -> float:
    try:
        return math.pow(x, y)
    except (ZeroDivisionError, ValueError):
        if x == 0.0 and y < 0:
            return math.inf
        return math.nan


def diag_values(p: Params, it: int, sums: Sequence[float]) -> Tuple[list, bool]:
    """The 12 values of format 640 (diag.f90:112-116): it, theta, Nus,
    Nus_comp, Nus_comp_2, melt, melt_front, melt_front_expected, Gras, Pr,
    Ray, Fou; and whether the melt-front pair is defined (mod(it,100)==0 and
    theta in one of the two regimes)."""
    nx, ny = p.nx, p.ny
    dT = p.t0 - p.t_in_sol
    kvisc = (1.0 / 3.0) * (p.tau_f - 0.5)
    alpha = (1.0 / 3.0) * (p.tau_g - 0.5)
    beta_gras = (p.rho0 / p.rho1 - 1.0) / dT
    gras = p.grav * beta_gras * dT * float(_i32(ny * ny * ny)) / (kvisc * kvisc)
    pr = (2.0 * p.tau_f - 1.0) / (2.0 * p.tau_g - 1.0)
    ray = pr * gras
    fou = alpha * float(it) / float(_i32(ny * ny))
    theta = p.stefan_number * fou
    theta1 = _pow(ray, -0.5)
    theta2 = float(nx // ny) * _pow(ray, -0.25)
    a = _pow(2.0 * theta, -0.5)
    nus = a + (0.35 * _pow(ray, 0.25) - a) * _pow(
        1.0 + _pow(0.0175 * _pow(ray, 0.75) * _pow(theta, 1.5), -2.0), -0.5)
    nus_comp = sums[0] / float(ny)
    nus_comp_2 = sums[1] / float(ny)
    melt = front = front_exp = 0.0
    valid = False
    if it % 100 == 0:
        melt = sums[2] / float(_i32(nx * ny))
        if theta <= theta1:
            front = float(ny) / a
            front_exp = float(ny) / nus
            valid = True
        elif theta1 < theta < theta2:
            front = sums[3] / float(ny)
            front_exp = float(ny) * theta * _pow(ray, 0.25)
            valid = True
    return [float(it), theta, nus, nus_comp, nus_comp_2, melt, front, front_exp, gras, pr, ray,
            fou], valid


def fortran_e(x: float, w: int, d: int) -> str:
    """Fortran Ew.d edit descriptor (0.ddddE+xx form, right-justified)."""
    if x != x:
        return "NaN".rjust(w)
    if x in (math.inf, -math.inf):
        return ("Infinity" if x > 0 else "-Infinity").rjust(w)
    if x == 0.0:
        s = "0." + "0" * d + "E+00"
    else:
        m = f"{abs(x):.{d - 1}E}"
        mant, ex = m.split("E")
        digits = mant.replace(".", "")
        e = int(ex) + 1
        s = f"0.{digits}E{'+' if e >= 0 else '-'}{abs(e):02d}"
        if x < 0:
            s = "-" + s
    return s.rjust(w)


def field_check_record(it: int, vals: Sequence[float]) -> str:
    """One line of fort.64, format i8,1x,3(e16.8,1x) (field_check.f90:29)."""
    return (f"{it:8d} " + "".join(fortran_e(v, 16, 8) + " " for v in vals)).rstrip()


def probe_record(it: int, vals: Sequence[float]) -> str:
    """One line of fort.666, format i8,5(e14.6,1x) (probe.f90:15)."""
    return (f"{it:8d}" + "".join(fortran_e(v, 14, 6) + " " for v in vals)).rstrip()


def bulk_record(vals: Sequence[float]) -> str:
    """One line of non-dim_bulk_numbers.dat, format 1I8,11(1x,e13.6)
    (diag.f90:116)."""
    return f"{int(vals[0]):8d}" + "".join(" " + fortran_e(v, 13, 6) for v in vals[1:])
