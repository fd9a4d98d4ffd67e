"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) and runner of the
reference build (oracle/_ref).

Test infrastructure: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py only, as the checker.  Arrays are returned as
numpy views shaped (Ny+2, Nx+2) and indexed [j, i], i.e. the Fortran
(0:Nx+1, 0:Ny+1) column-major layout of the reference read row by row.
"""
from __future__ import annotations

import ctypes
import os
import pathlib
import subprocess
import tempfile

import numpy as np

from bgk_pcm_amd.params import Params, write_input_dat

REPO = pathlib.Path(__file__).resolve().parents[1]
ORACLE_DIR = REPO / "oracle"
ORACLE_SO = ORACLE_DIR / "liboracle.so"
REF_DIR = ORACLE_DIR / "_ref"

FIELD_NAMES = (
    [f"f{k}" for k in range(9)]
    + [f"fp{k}" for k in range(9)]
    + [f"g{k}" for k in range(9)]
    + [f"gp{k}" for k in range(9)]
    + ["rho", "u", "v", "T", "rho2", "u2", "v2", "T2", "phi", "phi_prec"]
)
FID = {n: k for k, n in enumerate(FIELD_NAMES)}
STAGES = ["bc_f", "bc_g", "streaming", "moments_pre", "phase_field", "forces", "collision", "moments"]

_lib = None


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR), "lib"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            build_oracle()
        L = ctypes.CDLL(str(ORACLE_SO))
        L.oracle_create.restype = ctypes.c_void_p
        L.oracle_create.argtypes = [ctypes.c_void_p]
        L.oracle_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_field.restype = ctypes.POINTER(ctypes.c_double)
        L.oracle_field.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.oracle_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        for name in ["oracle_initialization"] + [f"oracle_{s}" for s in STAGES]:
            getattr(L, name).argtypes = [ctypes.c_void_p]
        for name in ("oracle_field_check", "oracle_mass_check", "oracle_diag_sums"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.oracle_diag.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
        L.oracle_set_ext.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.oracle_set_ext2.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_uint64]
        L.oracle_set_ext3.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.oracle_tanh.restype = ctypes.c_double
        L.oracle_tanh.argtypes = [ctypes.c_double]
        _lib = L
    return _lib


class Oracle:
    """One CPU lattice of the reference STEP_7 (see oracle/bgk_oracle.h)."""

    def __init__(self, p: Params, initialize: bool = True, t_hot: float | None = None,
                 k_ratio: float = 1.0, t_cold: float | None = None, noise_amp: float = 0.0,
                 noise_seed: int = 0, t_crit: float = 273.0, t_ref: float = 0.0):
        self.p = p
        self.nx, self.ny = p.nx, p.ny
        self._c = p.to_c()
        self.h = lib().oracle_create(ctypes.byref(self._c))
        if not self.h:
            raise RuntimeError("oracle_create failed")
        if t_hot is not None or k_ratio != 1.0:
            lib().oracle_set_ext(self.h, p.t0 if t_hot is None else t_hot, k_ratio)
        if t_cold is not None or noise_amp != 0.0:
            lib().oracle_set_ext2(self.h, 0 if t_cold is None else 1,
                                  0.0 if t_cold is None else t_cold, noise_amp, noise_seed)
        if t_crit != 273.0 or t_ref != 0.0:
            lib().oracle_set_ext3(self.h, t_crit, t_ref)
        if initialize:
            lib().oracle_initialization(self.h)

    def close(self):
        if self.h:
            lib().oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def field(self, name) -> np.ndarray:
        fid = FID[name] if isinstance(name, str) else int(name)
        ptr = lib().oracle_field(self.h, fid)
        n = (self.nx + 2) * (self.ny + 2)
        return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(self.ny + 2, self.nx + 2)

    def state(self) -> np.ndarray:
        """All 46 fields, shape (46, Ny+2, Nx+2), a copy."""
        return np.stack([self.field(k).copy() for k in range(len(FIELD_NAMES))])

    def load(self, snap: np.ndarray) -> None:
        for k in range(len(FIELD_NAMES)):
            self.field(k)[...] = snap[k]

    def stage(self, name: str) -> None:
        getattr(lib(), f"oracle_{name}")(self.h)

    def step(self, n: int = 1) -> None:
        lib().oracle_step(self.h, int(n))

    def field_check(self):
        out = (ctypes.c_double * 3)()
        lib().oracle_field_check(self.h, out)
        return list(out)

    def mass_check(self):
        out = (ctypes.c_double * 3)()
        lib().oracle_mass_check(self.h, out)
        return list(out)

    def diag_sums(self):
        out = (ctypes.c_double * 4)()
        lib().oracle_diag_sums(self.h, out)
        return list(out)

    def diag(self, it: int):
        out = (ctypes.c_double * 12)()
        valid = ctypes.c_int()
        lib().oracle_diag(self.h, int(it), out, ctypes.byref(valid))
        return list(out), bool(valid.value)


# --------------------------------------------------------------- reference

REF_VARIANTS = {"step7": "step7/ref_harness", "step7_r8": "step7_r8/ref_harness"}


def have_ref(variant: str = "step7") -> bool:
    return (REF_DIR / REF_VARIANTS[variant]).exists()


def read_snapshot(path, nx: int, ny: int) -> np.ndarray:
    a = np.fromfile(path, dtype="<f8")
    n = (nx + 2) * (ny + 2)
    if a.size != len(FIELD_NAMES) * n:
        raise ValueError(f"{path}: {a.size} doubles, expected {len(FIELD_NAMES) * n}")
    return a.reshape(len(FIELD_NAMES), ny + 2, nx + 2)


def run_ref(p: Params, nsteps: int, stage_first: int = 1, stage_last: int = 0,
            variant: str = "step7", diag_every: int = 0, timeout: float = 600.0) -> dict:
    """Run the reference harness; returns snapshots and text outputs.

    keys: 'init', 'final', ('s', step, k) for the stage window, and
    'text:<file>' for each text file the reference wrote in its cwd.
    """
    exe = REF_DIR / REF_VARIANTS[variant]
    if not exe.exists():
        raise FileNotFoundError(f"{exe} not built (make -C oracle ref)")
    out = {}
    with tempfile.TemporaryDirectory(prefix="bgkref_") as wd:
        write_input_dat(p, os.path.join(wd, "input.dat"))
        snapdir = os.path.join(wd, "snap")
        os.mkdir(snapdir)
        r = subprocess.run([str(exe), str(nsteps), str(stage_first), str(stage_last), snapdir,
                            str(diag_every)], cwd=wd, capture_output=True, text=True,
                           timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"ref_harness failed: {r.stderr[-2000:]}")
        out["stdout"] = r.stdout
        out["init"] = read_snapshot(os.path.join(snapdir, "init.bin"), p.nx, p.ny)
        out["final"] = read_snapshot(os.path.join(snapdir, "final.bin"), p.nx, p.ny)
        for s in range(stage_first, stage_last + 1):
            for k in range(9):
                out[("s", s, k)] = read_snapshot(
                    os.path.join(snapdir, f"s{s:04d}_{k}.bin"), p.nx, p.ny)
        for fn in sorted(os.listdir(wd)):
            full = os.path.join(wd, fn)
            if os.path.isfile(full) and fn != "input.dat":
                with open(full, "r", errors="replace") as fh:
                    out["text:" + fn] = fh.read()
    return out
