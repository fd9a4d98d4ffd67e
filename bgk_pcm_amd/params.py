"""Simulation parameters: the 21-record ``input.dat`` of the reference and
the ctypes mirror of ``bgk_params`` (include/bgk_pcm.h).

The reference reads input.dat with 21 list-directed READs, one per record,
each reading a one-character label and one value (STEP_7/setup.f90:7-31,
``character :: id(50)`` at STEP_7/shared.f90:83): the first token of a line
is a label that is ignored, the second is the value, and the rest of the
line (comments) is skipped; lines after the 21st are never read.  Values use
Fortran real syntax (``1.0240d0``).  ``read_input_dat`` follows those rules
and raises ``ValueError`` where the Fortran runtime would stop with an error
(missing file or record, an integer field that is not an integer).
"""
from __future__ import annotations

import ctypes
import os
import re
from dataclasses import dataclass, field, fields, asdict
from typing import Dict

FP64 = 0
FP32_DEV = 1
FP32_MOM = 2
FP64_MOM = 3
ROUND_STEP7 = 0
ROUND_CLEAN = 1

# (struct member, label written by write_input_dat, Fortran type) in file order
INPUT_DAT_RECORDS = [
    ("iter", "iter", int),
    ("noutput", "noutput", int),
    ("dump", "dump", int),
    ("nx", "Nx", int),
    ("ny", "Ny", int),
    ("re", "Re", float),
    ("rho0", "rho0", float),
    ("rho1", "rho1", float),
    ("tau_f", "tau_f", float),
    ("tau_g", "tau_g", float),
    ("t0", "T0", float),
    ("t_in_sol", "T_in_sol", float),
    ("grav", "grav", float),
    ("eps", "eps", float),
    ("stefan_number", "stefan_number", float),
    ("t_width", "T_width", float),
    ("length", "length", float),
    ("west_bc", "westBC_st", int),
    ("east_bc", "eastBC_st", int),
    ("north_bc", "northBC_st", int),
    ("south_bc", "southBC_st", int),
]

VALID_BC = {
    "west_bc": (0, 1, 2, 3, 4),
    "east_bc": (0, 3, 4),
    "north_bc": (0, 1, 4, 5, 6),
    "south_bc": (0, 4, 6),
}


class BgkParamsC(ctypes.Structure):
    """Byte-for-byte mirror of ``struct bgk_params`` (include/bgk_pcm.h)."""

    _fields_ = [
        ("iter", ctypes.c_int32),
        ("noutput", ctypes.c_int32),
        ("dump", ctypes.c_int32),
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("reserved0", ctypes.c_int32),
        ("re", ctypes.c_double),
        ("rho0", ctypes.c_double),
        ("rho1", ctypes.c_double),
        ("tau_f", ctypes.c_double),
        ("tau_g", ctypes.c_double),
        ("t0", ctypes.c_double),
        ("t_in_sol", ctypes.c_double),
        ("grav", ctypes.c_double),
        ("eps", ctypes.c_double),
        ("stefan_number", ctypes.c_double),
        ("t_width", ctypes.c_double),
        ("length", ctypes.c_double),
        ("west_bc", ctypes.c_int32),
        ("east_bc", ctypes.c_int32),
        ("north_bc", ctypes.c_int32),
        ("south_bc", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("rounding", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("fp32_collision", ctypes.c_int32),
    ]


@dataclass
class Params:
    """input.dat records plus this library's options.

    Defaults are the values of the reference benchmark input
    UTIL/input.dat.0 (a no-slip cavity, Ra from grav=2.5e-4) except the
    grid, which defaults to 128 x 512 as there as well.
    """

    iter: int = 50000
    noutput: int = 2
    dump: int = 0
    nx: int = 128
    ny: int = 512
    re: float = 0.0
    rho0: float = 1.024
    rho1: float = 1.0
    tau_f: float = 1.0
    tau_g: float = 0.51
    t0: float = 283.0
    t_in_sol: float = 273.0
    grav: float = 0.00025
    eps: float = 0.0
    stefan_number: float = 0.1
    t_width: float = 10.0
    length: float = 512.0
    west_bc: int = 4
    east_bc: int = 4
    north_bc: int = 4
    south_bc: int = 4
    precision: int = FP64
    rounding: int = ROUND_STEP7
    device: int = 0
    fp32_collision: int = 0

    def to_c(self) -> BgkParamsC:
        c = BgkParamsC()
        for f in fields(self):
            setattr(c, f.name, getattr(self, f.name))
        c.reserved0 = 0
        return c

    def replace(self, **kw) -> "Params":
        d = asdict(self)
        d.update(kw)
        return Params(**d)

    def validate(self) -> None:
        """Checks the library makes before allocating (bgk_create)."""
        if self.nx < 1 or self.ny < 1:
            raise ValueError("Nx and Ny must be >= 1")
        for k, allowed in VALID_BC.items():
            if getattr(self, k) not in allowed:
                raise ValueError(f"{k}={getattr(self, k)} is not one of {allowed}")
        if self.tau_f == 0.0 or self.tau_g == 0.0:
            raise ValueError("tau_f and tau_g must be nonzero")
        if self.precision not in (FP64, FP32_DEV, FP32_MOM, FP64_MOM):
            raise ValueError("precision must be FP64, FP32_DEV, FP32_MOM or FP64_MOM")
        if self.rounding not in (ROUND_STEP7, ROUND_CLEAN):
            raise ValueError("rounding must be ROUND_STEP7 or ROUND_CLEAN")


def rayleigh(p: Params) -> float:
    """Rayleigh number as STEP_7/diag.f90:31-40 defines it,
    Ra = Pr * grav * beta_gras * Delta_T * Ny**3 / nu**2 with
    beta_gras = (rho0/rho1 - 1)/Delta_T, Pr = (2 tau_f - 1)/(2 tau_g - 1),
    nu = (tau_f - 1/2)/3, evaluated in floating point (the reference's
    32-bit Ny**3 wraps for Ny > 1290)."""
    nu = (p.tau_f - 0.5) / 3.0
    pr = (2.0 * p.tau_f - 1.0) / (2.0 * p.tau_g - 1.0)
    return pr * p.grav * (p.rho0 / p.rho1 - 1.0) * float(p.ny) ** 3 / nu ** 2


def grav_for_rayleigh(p: Params, ra: float) -> float:
    """grav that gives Rayleigh number `ra` (definition of `rayleigh`)."""
    return ra / rayleigh(p.replace(grav=1.0))


_FORTRAN_REAL = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([eEdDqQ][+-]?\d+)?$")
_FORTRAN_INT = re.compile(r"^[+-]?\d+$")


def _parse_value(tok: str, typ) -> float | int:
    tok = tok.rstrip(",")
    if typ is int:
        if not _FORTRAN_INT.match(tok):
            raise ValueError(f"integer expected, got {tok!r}")
        return int(tok)
    if not _FORTRAN_REAL.match(tok):
        raise ValueError(f"real expected, got {tok!r}")
    return float(re.sub(r"[dDqQ]", "e", tok))


def read_input_dat(path: str | os.PathLike, **options) -> Params:
    """Parse an input.dat as STEP_7/setup.f90:5-32 does."""
    with open(path, "r") as fh:
        lines = fh.read().splitlines()
    vals: Dict[str, object] = {}
    for rec, (name, label, typ) in enumerate(INPUT_DAT_RECORDS):
        if rec >= len(lines):
            raise ValueError(f"{path}: end of file before record {rec + 1} ({label})")
        toks = lines[rec].replace("\t", " ").split()
        if len(toks) < 2:
            raise ValueError(f"{path}: record {rec + 1} ({label}) has no value")
        try:
            vals[name] = _parse_value(toks[1], typ)
        except ValueError as e:
            raise ValueError(f"{path}: record {rec + 1} ({label}): {e}") from None
    vals.update(options)
    return Params(**vals)


def _fortran_real(x: float) -> str:
    r = repr(float(x))
    if "e" in r or "E" in r:
        m, e = r.lower().split("e")
        if "." not in m:
            m += ".0"
        return f"{m}d{e}"
    if "." not in r:
        r += ".0"
    return r + "d0"


def write_input_dat(p: Params, path: str | os.PathLike) -> None:
    """Write the 21 records in the reference layout (UTIL/input.dat.0)."""
    out = []
    for name, label, typ in INPUT_DAT_RECORDS:
        v = getattr(p, name)
        sval = str(int(v)) if typ is int else _fortran_real(v)
        out.append(f"  {label:<20s} {sval}")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")
