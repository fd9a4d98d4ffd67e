"""Python host mirror of the reference's time-step interface, over the C ABI.

``Lattice`` owns one lattice resident on the GPU (a ``bgk_handle`` of
include/bgk_pcm.h).  Its methods carry the reference's subroutine names
(STEP_7/main.f90:60-69: ``bc_f``, ``bc_g``, ``streaming``, ``moments_pre``,
``phase_field``, ``forces``, ``collision``, ``moments``; the diagnostics
``field_check``, ``mass_check``, ``diag``, ``probe``, ``prof_x``,
``prof_y``), so a script written against the reference's call sequence reads
the same.  Fields come back as numpy arrays shaped (Ny+2, Nx+2) and indexed
[j, i]: the Fortran (0:Nx+1, 0:Ny+1) column-major arrays read row by row.

There is no CPU fallback: if libbgk_pcm.so is missing or no GPU is visible,
construction raises.
"""
from __future__ import annotations

import ctypes
import os
import pathlib
from typing import Sequence

import numpy as np

from .params import Params, BgkParamsC
from . import diagnostics

LIB_PATH = pathlib.Path(__file__).resolve().parent / "lib" / "libbgk_pcm.so"

FIELD_NAMES = (
    [f"f{k}" for k in range(9)]
    + [f"fp{k}" for k in range(9)]
    + [f"g{k}" for k in range(9)]
    + [f"gp{k}" for k in range(9)]
    + ["rho", "u", "v", "T", "rho2", "u2", "v2", "T2", "phi", "phi_prec"]
)
FIELD_ID = {n: k for k, n in enumerate(FIELD_NAMES)}

ERRORS = {0: "BGK_OK", -1: "BGK_E_INVALID", -2: "BGK_E_HIP", -3: "BGK_E_NOMEM",
          -4: "BGK_E_UNSUPPORTED", -5: "BGK_E_NODEVICE"}


class BgkError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{ERRORS.get(code, code)}: {message}")
        self.code = code


class BgkInfoC(ctypes.Structure):
    _fields_ = [
        ("device_bytes", ctypes.c_int64),
        ("device_free", ctypes.c_int64),
        ("device_total", ctypes.c_int64),
        ("host_maxrss_mb", ctypes.c_double),
        ("pitch", ctypes.c_int32),
        ("fused_path", ctypes.c_int32),
    ]


class BgkExtC(ctypes.Structure):
    """struct bgk_ext of include/bgk_pcm.h (physics extensions)."""
    _fields_ = [("t_hot", ctypes.c_double), ("k_ratio", ctypes.c_double),
                ("east_cold", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("t_cold", ctypes.c_double), ("noise_amp", ctypes.c_double),
                ("noise_seed", ctypes.c_uint64), ("t_crit", ctypes.c_double),
                ("t_ref", ctypes.c_double)]


_lib = None

# name -> (argtypes) for every entry point of include/bgk_pcm.h
_H = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I32P = ctypes.POINTER(ctypes.c_int32)
SIGNATURES = {
    "bgk_abi_version": [],
    "bgk_last_error": [],
    "bgk_create": [ctypes.POINTER(BgkParamsC), ctypes.POINTER(ctypes.c_void_p)],
    "bgk_destroy": [_H],
    "bgk_set_stream": [_H, ctypes.c_void_p],
    "bgk_initialization": [_H],
    "bgk_set_extensions": [_H, ctypes.POINTER(BgkExtC)],
    "bgk_upload": [_H, ctypes.c_int32, _D, ctypes.c_int64],
    "bgk_download": [_H, ctypes.c_int32, _D, ctypes.c_int64],
    "bgk_bc_f": [_H],
    "bgk_bc_g": [_H],
    "bgk_streaming": [_H],
    "bgk_moments_pre": [_H],
    "bgk_phase_field": [_H],
    "bgk_forces": [_H],
    "bgk_collision": [_H],
    "bgk_moments": [_H],
    "bgk_step": [_H, ctypes.c_int32],
    "bgk_field_check": [_H, _D],
    "bgk_field_check_planes": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                               ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p, _D],
    "bgk_mass_check": [_H, _D],
    "bgk_diag_sums": [_H, _D],
    "bgk_gather": [_H, ctypes.c_int32, ctypes.c_int32, _I32P, _I32P, _D],
    "bgk_synchronize": [_H],
    "bgk_get_info": [_H, ctypes.POINTER(BgkInfoC)],
    "bgk_adj_setup": [_H, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _I32P, _I32P],
    "bgk_adj_setup_levels": [_H, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                             _I32P, _I32P],
    "bgk_adj_info": [_H, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32)],
    "bgk_adj_setup_binomial": [_H, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _I32P, _I32P],
    "bgk_adj_plan": [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)],
    "bgk_adj_forward": [_H, _D],
    "bgk_adj_backward": [_H, _D, _D],
}


def load_library(path: str | os.PathLike | None = None) -> ctypes.CDLL:
    """Load libbgk_pcm.so (built by __graft_entry__.build()); raises if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = pathlib.Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is not built: run `python -c 'import __graft_entry__ as g; "
                           f"g.build()'` (or make -C bgk_pcm_amd/csrc)")
    L = ctypes.CDLL(str(p))
    for name, args in SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_char_p if name == "bgk_last_error" else ctypes.c_int
    if path is None:
        _lib = L
    return L


def adj_plan(nsteps: int, n_states: int) -> int:
    """Forward steps the backward pass of a binomial schedule recomputes."""
    out = ctypes.c_int64()
    _check(load_library().bgk_adj_plan(nsteps, n_states, ctypes.byref(out)))
    return int(out.value)


def _check(rc: int) -> None:
    if rc != 0:
        raise BgkError(rc, load_library().bgk_last_error().decode(errors="replace"))


class Lattice:
    """One lattice on the GPU; see the module docstring.

    t_hot, k_ratio, t_cold, noise_amp, noise_seed, t_crit, t_ref: physics
    extensions of bgk_set_extensions (hot-wall and initial liquid
    temperature; solid/liquid conductivity ratio; east wall at t_cold when
    given; seeded initial temperature noise; the phase field's critical
    temperature; the buoyancy's reference temperature), applied before the
    initialization; the defaults give STEP_7."""

    def __init__(self, params: Params, initialize: bool = True, t_hot: float = 0.0,
                 k_ratio: float = 1.0, t_cold: float | None = None, noise_amp: float = 0.0,
                 noise_seed: int = 0, t_crit: float = 0.0, t_ref: float = 0.0):
        self.params = params
        self.nx, self.ny = params.nx, params.ny
        self._lib = load_library()
        self._c = params.to_c()
        h = ctypes.c_void_p()
        _check(self._lib.bgk_create(ctypes.byref(self._c), ctypes.byref(h)))
        self._h = h
        self.it = 0
        if (t_hot != 0.0 or k_ratio != 1.0 or t_cold is not None or noise_amp != 0.0
                or t_crit != 0.0 or t_ref != 0.0):
            self.set_extensions(t_hot, k_ratio, t_cold, noise_amp, noise_seed, t_crit, t_ref)
        if initialize:
            self.initialization()

    def set_extensions(self, t_hot: float = 0.0, k_ratio: float = 1.0,
                       t_cold: float | None = None, noise_amp: float = 0.0,
                       noise_seed: int = 0, t_crit: float = 0.0, t_ref: float = 0.0) -> None:
        e = BgkExtC(t_hot, k_ratio, 0 if t_cold is None else 1, 0,
                    0.0 if t_cold is None else t_cold, noise_amp, noise_seed, t_crit, t_ref)
        _check(self._lib.bgk_set_extensions(self._h, ctypes.byref(e)))

    # ---- adjoint (FP64 moment handles): probe-temperature record and its
    # gradient with respect to (stefan_number, k_ratio, t_hot)
    def adj_setup(self, nsteps: int, seg_len: int, pi: Sequence[int], pj: Sequence[int],
                  seg_outer: int = 0) -> None:
        """Checkpoints every seg_len steps; with seg_outer > 0 two levels
        (every seg_outer steps, then every seg_len steps of one outer segment)."""
        self._adj_n, self._adj_np = int(nsteps), len(pi)
        ai = np.ascontiguousarray(pi, dtype=np.int32)
        aj = np.ascontiguousarray(pj, dtype=np.int32)
        _check(self._lib.bgk_adj_setup_levels(self._h, nsteps, seg_outer, seg_len, len(pi),
                                              ai.ctypes.data_as(_I32P), aj.ctypes.data_as(_I32P)))

    def adj_setup_binomial(self, nsteps: int, n_states: int, pi: Sequence[int],
                           pj: Sequence[int]) -> None:
        """Binomial checkpointing with n_states snapshot slots (0: as many as
        the free device memory holds)."""
        self._adj_n, self._adj_np = int(nsteps), len(pi)
        ai = np.ascontiguousarray(pi, dtype=np.int32)
        aj = np.ascontiguousarray(pj, dtype=np.int32)
        _check(self._lib.bgk_adj_setup_binomial(self._h, nsteps, n_states, len(pi),
                                                ai.ctypes.data_as(_I32P),
                                                aj.ctypes.data_as(_I32P)))

    def adj_info(self) -> tuple[int, int]:
        """(device bytes of the adjoint's checkpoints and state, stored lattice states)."""
        b, n = ctypes.c_int64(), ctypes.c_int32()
        _check(self._lib.bgk_adj_info(self._h, ctypes.byref(b), ctypes.byref(n)))
        return int(b.value), int(n.value)

    def adj_forward(self) -> np.ndarray:
        """Run the recorded forward; returns T at the probes, (nsteps, n_probes)."""
        out = np.zeros((self._adj_n, max(self._adj_np, 1)))
        _check(self._lib.bgk_adj_forward(self._h, out.ctypes.data_as(_D)))
        self.it += self._adj_n
        return out[:, :self._adj_np].copy()

    def adj_backward(self, seeds: np.ndarray) -> np.ndarray:
        """dL/d(stefan_number, k_ratio, t_hot) from seeds = dL/dT(step, probe)."""
        sd = np.zeros((self._adj_n, max(self._adj_np, 1)))
        sd[:, :self._adj_np] = np.asarray(seeds, dtype=np.float64).reshape(self._adj_n, -1)
        g = np.zeros(3)
        _check(self._lib.bgk_adj_backward(self._h, sd.ctypes.data_as(_D), g.ctypes.data_as(_D)))
        return g

    # ---- lifetime
    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._lib.bgk_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- STEP_7/initialization.f90
    def initialization(self) -> None:
        _check(self._lib.bgk_initialization(self._h))
        self.it = 0

    # ---- the eight subroutines of the time step (STEP_7/main.f90:60-69)
    def bc_f(self):
        _check(self._lib.bgk_bc_f(self._h))

    def bc_g(self):
        _check(self._lib.bgk_bc_g(self._h))

    def streaming(self):
        _check(self._lib.bgk_streaming(self._h))

    def moments_pre(self):
        _check(self._lib.bgk_moments_pre(self._h))

    def phase_field(self):
        _check(self._lib.bgk_phase_field(self._h))

    def forces(self):
        _check(self._lib.bgk_forces(self._h))

    def collision(self):
        _check(self._lib.bgk_collision(self._h))

    def moments(self):
        _check(self._lib.bgk_moments(self._h))

    STAGES = ("bc_f", "bc_g", "streaming", "moments_pre", "phase_field", "forces", "collision",
              "moments")

    def stage(self, name: str) -> None:
        getattr(self, name)()

    def step(self, nsteps: int = 1) -> None:
        """nsteps time steps (fused kernel for the no-slip cavity)."""
        _check(self._lib.bgk_step(self._h, int(nsteps)))
        self.it += int(nsteps)

    # ---- host <-> device
    def download(self, field) -> np.ndarray:
        fid = FIELD_ID[field] if isinstance(field, str) else int(field)
        out = np.empty((self.ny + 2, self.nx + 2), dtype=np.float64)
        _check(self._lib.bgk_download(self._h, fid, out.ctypes.data_as(_D), self.nx + 2))
        return out

    def upload(self, field, arr: np.ndarray) -> None:
        fid = FIELD_ID[field] if isinstance(field, str) else int(field)
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != (self.ny + 2, self.nx + 2):
            raise ValueError(f"{field}: shape {a.shape}, expected {(self.ny + 2, self.nx + 2)}")
        _check(self._lib.bgk_upload(self._h, fid, a.ctypes.data_as(_D), self.nx + 2))

    def state(self, fields: Sequence[str] | None = None) -> dict:
        return {f: self.download(f) for f in (fields or FIELD_NAMES)}

    # ---- diagnostics (STEP_7/field_check.f90, mass_check.f90, diag.f90, probe.f90)
    def field_check(self) -> list:
        out = (ctypes.c_double * 3)()
        _check(self._lib.bgk_field_check(self._h, out))
        return list(out)

    def mass_check(self) -> list:
        out = (ctypes.c_double * 3)()
        _check(self._lib.bgk_mass_check(self._h, out))
        return list(out)

    def diag_sums(self) -> list:
        out = (ctypes.c_double * 4)()
        _check(self._lib.bgk_diag_sums(self._h, out))
        return list(out)

    def diag(self, it: int | None = None):
        """The 12 values diag.f90 writes with format 640, and whether the
        melt-front pair was set (only on steps divisible by 100)."""
        it = self.it if it is None else it
        return diagnostics.diag_values(self.params, it, self.diag_sums())

    def gather(self, field, i: Sequence[int], j: Sequence[int]) -> np.ndarray:
        fid = FIELD_ID[field] if isinstance(field, str) else int(field)
        ii = np.ascontiguousarray(i, dtype=np.int32)
        jj = np.ascontiguousarray(j, dtype=np.int32)
        if ii.shape != jj.shape:
            raise ValueError("i and j must have the same length")
        out = np.empty(ii.shape, dtype=np.float64)
        _check(self._lib.bgk_gather(self._h, fid, int(ii.size), ii.ctypes.data_as(_I32P),
                                    jj.ctypes.data_as(_I32P), out.ctypes.data_as(_D)))
        return out

    def probe(self) -> list:
        """u, v, rho, T, phi at (Nx/2, Ny/2) (STEP_7/probe.f90:7-15)."""
        i, j = [self.nx // 2], [self.ny // 2]
        return [float(self.gather(f, i, j)[0]) for f in ("u", "v", "rho", "T", "phi")]

    def prof_x(self) -> np.ndarray:
        """Rows i=1..Nx of u, v, rho, T, phi at j=Ny/2 (STEP_7/prof_x.f90:7-17)."""
        i = np.arange(1, self.nx + 1)
        j = np.full_like(i, self.ny // 2)
        return np.stack([self.gather(f, i, j) for f in ("u", "v", "rho", "T", "phi")], axis=1)

    def prof_y(self) -> np.ndarray:
        """Columns j=1..Ny of u, v, rho, T, phi at i=Nx/2 (STEP_7/prof_y.f90:7-17)."""
        j = np.arange(1, self.ny + 1)
        i = np.full_like(j, self.nx // 2)
        return np.stack([self.gather(f, i, j) for f in ("u", "v", "rho", "T", "phi")], axis=1)

    # ---- runtime
    def set_stream(self, stream_ptr: int | None) -> None:
        _check(self._lib.bgk_set_stream(self._h, ctypes.c_void_p(stream_ptr or None)))

    def synchronize(self) -> None:
        _check(self._lib.bgk_synchronize(self._h))

    def info(self) -> dict:
        inf = BgkInfoC()
        _check(self._lib.bgk_get_info(self._h, ctypes.byref(inf)))
        return {k: getattr(inf, k) for k, _ in BgkInfoC._fields_}
