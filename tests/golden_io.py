"""Loading of the golden fixtures written by tests/golden/make_golden.py."""
from __future__ import annotations

import json
import pathlib

import numpy as np

from bgk_pcm_amd.params import Params

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"

# fields whose halo cells the reference never writes (scratch of momentsPre /
# streaming): only their interior is defined
INTERIOR_ONLY = {f"fp{k}" for k in range(9)} | {f"gp{k}" for k in range(9)} | {
    "rho2", "u2", "v2", "T2"}


def fixtures():
    """State fixtures (driver_*.npz and formats_*.npz hold files only)."""
    return sorted(p.stem for p in GOLDEN.glob("*.npz")
                  if not p.stem.startswith(("driver_", "formats_")))


class Golden:
    def __init__(self, name: str):
        self.name = name
        self.d = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
        self.params = Params(**json.loads(str(self.d["params"])))
        self.nsteps = int(self.d["nsteps"])
        self.fields = [str(f) for f in self.d["fields"]]

    def snap(self, key: str) -> dict:
        a = self.d[key]
        return {f: a[k] for k, f in enumerate(self.fields)}

    def window(self):
        """[(step, [stage indices])] present in the fixture."""
        out = {}
        for k in self.d.files:
            if k.startswith("s") and "_" in k and k[1].isdigit():
                s, st = k[1:].split("_")
                out.setdefault(int(s), []).append(int(st))
        return sorted((s, sorted(v)) for s, v in out.items())

    def text(self, fn: str) -> str | None:
        key = "text_" + fn.replace(".", "_").replace("-", "_")
        return str(self.d[key]) if key in self.d.files else None


def bits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def mismatch_report(name: str, a: np.ndarray, b: np.ndarray) -> str:
    d = np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))
    nbad = int(np.count_nonzero(np.asarray(a).view(np.uint64) != np.asarray(b).view(np.uint64)))
    j, i = np.unravel_index(int(np.nanargmax(d)), d.shape) if d.size else (0, 0)
    return f"{name}: {nbad} nodes differ, max|d|={np.nanmax(d):.3e} at (i={i}, j={j})"


def compare_state(got: dict, want: dict, fields=None) -> list:
    """Bitwise comparison; returns the list of mismatch messages."""
    bad = []
    for f in fields or want.keys():
        a, b = got[f], want[f]
        if f in INTERIOR_ONLY:
            a, b = a[1:-1, 1:-1], b[1:-1, 1:-1]
        if not bits_equal(a, b):
            bad.append(mismatch_report(f, a, b))
    return bad
