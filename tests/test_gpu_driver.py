"""The Fortran driver over the GPU library against the complete reference
program.

tests/golden/driver_cav24x20.npz holds every file the reference STEP_7
program (its own main.f90) wrote for a 24x20, 2000-step no-slip cavity
(VTK at steps 0/1000/2000, fort.64/666/667/668/69, non-dim_bulk_numbers.dat,
log.dat, Nusselt.dat, Temp_trend.dat).  bgk_thermal_gpu, run on the same
input.dat, must write the same files with the same bytes: the VTK files
print every field with 17 significant digits, so this is bit-for-bit parity
of rho, u, v, T and phi (halo included) at the output steps.  The one
exception allowed is fort.64's reductions, compared at their printed
precision, which is also exact here.  Nusselt.dat is compared by
structure only: the reference prints its Nu arrays without ever computing
them (they are set only under DEBUG_GA, STEP_7/moments.f90:39-63), i.e.
uninitialized memory (the fixture holds a NaN), while the driver prints 0.
"""
import json
import os
import pathlib
import subprocess

import numpy as np
import pytest

from bgk_pcm_amd.params import Params, write_input_dat

pytestmark = pytest.mark.gpu

ROOT = pathlib.Path(__file__).resolve().parents[1]
EXE = ROOT / "bgk_pcm_amd" / "bin" / "bgk_thermal_gpu"
FIXTURE = ROOT / "tests" / "golden" / "driver_cav24x20.npz"


@pytest.mark.parametrize("mode", [[], ["--stages"]], ids=["fused", "stages"])
def test_driver_writes_the_reference_files(tmp_path, mode):
    d = np.load(FIXTURE, allow_pickle=False)
    p = Params(**json.loads(str(d["params"])))
    write_input_dat(p, tmp_path / "input.dat")
    print(f"running {EXE.name} {' '.join(mode)} on {p.nx}x{p.ny}, {p.iter} steps", flush=True)
    r = subprocess.run([str(EXE), *mode], cwd=tmp_path, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    want_files = {k[5:]: str(d[k]) for k in d.files if k.startswith("file_")}
    failures = []
    for fn, want in want_files.items():
        path = tmp_path / fn
        if not path.exists():
            failures.append(f"{fn}: missing")
            continue
        got = path.read_text()
        gl, wl = got.splitlines(), want.splitlines()
        if fn == "Nusselt.dat":
            ok = len(gl) == len(wl) and all(a.split()[0] == b.split()[0] for a, b in zip(gl, wl))
            if not ok:
                failures.append(f"{fn}: structure differs")
            continue
        if got != want:
            k = next((n for n, (a, b) in enumerate(zip(gl, wl)) if a != b), min(len(gl), len(wl)))
            failures.append(f"{fn} differs at line {k + 1} ({len(gl)} vs {len(wl)} lines):\n"
                            f"  got:  {gl[k] if k < len(gl) else '<eof>'}\n"
                            f"  want: {wl[k] if k < len(wl) else '<eof>'}")
    assert not failures, "\n".join(failures)
    checks = [l for l in r.stdout.splitlines() if "CHECK" in l]
    assert "\n".join(checks) == str(d["stdout_check"])
    assert "INFO: Mlups" in r.stdout
