"""Write the golden fixtures of tests/golden/ from the reference itself.

Runs the reference STEP_7 program, compiled from its own sources by
`make -C oracle ref` (needs the reference checkout at /root/reference, so it
runs only in the build container), through oracle/ref_harness.f90, and
stores its outputs as small .npz files (float64, no pickles) plus a
MANIFEST of sha256 sums.  The fixtures are data: inputs (the input.dat
records) and the reference's outputs (field snapshots and the text files
its diagnostics write).

    python tests/golden/make_golden.py

Each fixture holds:
  params      JSON of bgk_pcm_amd.params.Params for the run
  nsteps      steps run
  init        (F, Ny+2, Nx+2) state after initialization
  final       (F, Ny+2, Nx+2) state after nsteps
  s{step}_{k} (F, Ny+2, Nx+2) state before (k=0) / after stage k=1..8 of the
              window steps
  fields      names of the F stored fields (a subset of the 46)
  text_<fn>   text files written by the reference diagnostics

thot_pin_32x24.npz pins the hot-wall extension (t_hot) by a reduction to the
reference: at grav = 0 the reference run with T0' = 293, Stefan' = 0.25 must
equal the oracle at T0 = 283, Stefan 0.125, t_hot = 293 (`oracle_side`),
bit for bit (`python tests/golden/make_golden.py thot`).

formats_cav24x20.npz holds what the reference's output routines write
after 50 steps of a 24x20 cavity (oracle/_ref/vtkbin/ref_vtkbin): the split
binary VTK files of STEP_CUF as bytes (bin_<name>, uint8) and STEP_7's
dumpFile.dat as text (text_dumpFile), with `params` and `nsteps`
(`python tests/golden/make_golden.py formats` writes only this one).

long_runs.json holds sha256 digests of the final rho, u, v, T and phi (full
(Ny+2) x (Nx+2) float64 arrays, little-endian, row j after row j-1, halo
included) of the benchmark configurations at their full lengths, where a
stored array would be too large (`python tests/golden/make_golden.py long`,
about two and a half hours on 8 cores):
  cavity2048_ra1e6_pr002   the 2048^2 FP64 config (Ra 1e6, Stefan 0.1,
                           tau_f 0.51 / tau_g 1.0 for Pr 0.02) as STEP_7 can
                           express it (no noise, its averaged-T east wall):
                           written by the reference itself,
                           oracle/_ref/openmp_h (STEP_OPENMP, whose per-node
                           arithmetic is STEP_7's: the same bits)
  cavity2048_literal       the same with the config's cold east wall at
                           T_in_sol and seeded initial noise: written by the
                           C oracle (the extensions have no reference
                           counterpart)
  cavity256_literal        the 256^2 config (tau_f 0.8, tau_g 15.5, grav
                           1e-4, cold east wall, noise) after 10000 steps:
                           the C oracle
  cavity4096_ra1e7_pr002_st005  the 4096^2 benchmark cavity (Ra 1e7, Pr 0.02,
                           Stefan 0.05) as STEP_7 can express it (no
                           conductivity ratio) after 2000 steps: written
                           by the reference STEP_OPENMP build
  cavity8192_ra1e7_st01    the 8192^2 cavity of the large adjoint
                           configuration (Ra 1e7, Stefan 0.1, the reference's
                           relaxation times) after 500 steps: written by the
                           reference STEP_OPENMP build
Each entry also keeps min, max and mean of every field, for a readable
report when a digest differs.

driver_cav24x20.npz instead holds every file the complete reference program
(oracle/_ref/step7/bgk_thermal, its own main.f90) writes for a 24x20,
2000-step input, as text: file_<name> per file, the input.dat records in
`params`, and the deterministic stdout lines in `stdout_check`.
"""
from __future__ import annotations

import hashlib
import json
import os
import pathlib
import subprocess
import sys
import tempfile
from dataclasses import asdict

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parent))

from bgk_pcm_amd.params import Params  # noqa: E402
from bgk_pcm_amd.params import write_input_dat  # noqa: E402
from oracle_py import FIELD_NAMES, REF_DIR, run_ref  # noqa: E402

STATE = [f"f{k}" for k in range(9)] + [f"g{k}" for k in range(9)] + ["rho", "u", "v", "T", "phi"]
ALL = list(FIELD_NAMES)

# name: (params, nsteps, window steps, stages kept in the window, fields, diag_every)
CASES = {
    # no-slip cavity (UTIL/input.dat.0 physics), every stage of one step
    "stages_cav14x10": (Params(nx=14, ny=10), 40, [40], range(9), ALL, 0),
    # every boundary type of BC_f (BC_f.f90:9-196) at least once
    "bc_mixA_12x10": (Params(nx=12, ny=10, re=2.0, length=10.0, west_bc=1, east_bc=3,
                             north_bc=5, south_bc=6), 30, [30], range(3), ALL, 0),
    "bc_mixB_12x10": (Params(nx=12, ny=10, re=2.0, length=10.0, west_bc=2, east_bc=3,
                             north_bc=1, south_bc=4), 30, [30], range(3), ALL, 0),
    "bc_mixC_12x10": (Params(nx=12, ny=10, re=2.0, length=10.0, west_bc=3, east_bc=0,
                             north_bc=6, south_bc=0), 30, [30], range(3), ALL, 0),
    "bc_periodic_12x10": (Params(nx=12, ny=10, re=2.0, length=10.0, west_bc=0, east_bc=0,
                                 north_bc=0, south_bc=0), 30, [30], range(3), ALL, 0),
    # long runs: final state and the diagnostics text
    "run_cav32_1000": (Params(nx=32, ny=32), 1000, [], [], STATE, 1000),
    "run_cav60x16_u0_400": (Params(nx=60, ny=16, re=5.0, length=16.0), 400, [], [], STATE, 100),
    "run_clean_45x7_300": (Params(nx=45, ny=7, rounding=1), 300, [], [], STATE, 0),
}


DRIVER_CASE = ("driver_cav24x20", Params(nx=24, ny=20, iter=2000, noutput=2))


def driver_fixture():
    name, p = DRIVER_CASE
    exe = REF_DIR / "step7" / "bgk_thermal"
    with tempfile.TemporaryDirectory(prefix="bgkdrv_") as wd:
        write_input_dat(p, os.path.join(wd, "input.dat"))
        r = subprocess.run([str(exe)], cwd=wd, capture_output=True, text=True, timeout=600,
                           check=True)
        data = {"params": np.array(json.dumps(asdict(p)))}
        for fn in sorted(os.listdir(wd)):
            if fn == "input.dat":
                continue
            with open(os.path.join(wd, fn)) as fh:
                data["file_" + fn] = np.array(fh.read())
        data["stdout_check"] = np.array(
            "\n".join(l for l in r.stdout.splitlines() if "CHECK" in l))
    path = HERE / f"{name}.npz"
    np.savez_compressed(path, **data)
    print(f"{name}: {path.stat().st_size / 1024:.0f} KiB")
    return path


FORMATS_CASE = ("formats_cav24x20", Params(nx=24, ny=20, length=24.0, grav=1e-3), 50)


def formats_fixture():
    name, p, nsteps = FORMATS_CASE
    exe = REF_DIR / "vtkbin" / "ref_vtkbin"
    with tempfile.TemporaryDirectory(prefix="bgkfmt_") as wd:
        write_input_dat(p, os.path.join(wd, "input.dat"))
        subprocess.run([str(exe), str(nsteps)], cwd=wd, capture_output=True, text=True,
                       timeout=600, check=True)
        data = {"params": np.array(json.dumps(asdict(p))), "nsteps": np.array(nsteps)}
        for fn in sorted(os.listdir(wd)):
            if fn.endswith(".vtk"):
                data["bin_" + fn] = np.frombuffer(open(os.path.join(wd, fn), "rb").read(),
                                                  dtype=np.uint8)
        data["text_dumpFile"] = np.array(open(os.path.join(wd, "dumpFile.dat")).read())
    path = HERE / f"{name}.npz"
    np.savez_compressed(path, **data)
    print(f"{name}: {path.stat().st_size / 1024:.0f} KiB")
    return path


# The hot-wall extension pinned by a reduction to the reference: with grav
# = 0 the buoyancy vanishes, and the oracle at (T0, t_hot, Stefan) equals
# the reference at (T0' = t_hot, Stefan' = Stefan (t_hot - T_in_sol) / (T0 -
# T_in_sol)) - the same liquid and wall temperature, the same latent-heat
# factor (T0 - T_in_sol)/Stefan (80 here, exact in both) - bit for bit.
THOT_CASE = ("thot_pin_32x24", Params(nx=32, ny=24, length=24.0, grav=0.0, t0=293.0,
                                      stefan_number=0.25), 300,
             {"t0": 283.0, "stefan_number": 0.125, "t_hot": 293.0})


def thot_fixture():
    name, p, nsteps, oracle_side = THOT_CASE
    r = run_ref(p, nsteps, 1, 0)
    idx = [FIELD_NAMES.index(f) for f in STATE]
    data = {"params": np.array(json.dumps(asdict(p))), "nsteps": np.array(nsteps),
            "fields": np.array(STATE), "init": r["init"][idx], "final": r["final"][idx],
            "oracle_side": np.array(json.dumps(oracle_side))}
    path = HERE / f"{name}.npz"
    np.savez_compressed(path, **data)
    print(f"{name}: {path.stat().st_size / 1024:.0f} KiB")
    return path


LONG = HERE / "long_runs.json"


def field_digests(fields: dict) -> dict:
    """sha256 of each field's full array (float64, little-endian, C order of
    the (Ny+2, Nx+2) array) and its min / max / mean."""
    out = {}
    for f in ("rho", "u", "v", "T", "phi"):
        a = np.ascontiguousarray(fields[f], dtype="<f8")
        out[f] = {"sha256": hashlib.sha256(a.tobytes()).hexdigest(),
                  "min": float(a.min()), "max": float(a.max()), "mean": float(a.mean())}
    return out


def long_cases():
    """(name, params, nsteps, oracle extensions or None for the reference)."""
    from bgk_pcm_amd.params import grav_for_rayleigh
    p2 = Params(nx=2048, ny=2048, length=2048.0, tau_f=0.51, tau_g=1.0, stefan_number=0.1,
                iter=20000)
    p2 = p2.replace(grav=grav_for_rayleigh(p2, 1e6))
    p1 = Params(nx=256, ny=256, length=256.0, tau_f=0.8, tau_g=15.5, grav=1e-4,
                stefan_number=0.1, iter=10000)
    p3 = Params(nx=4096, ny=4096, length=4096.0, tau_f=0.51, tau_g=1.0, stefan_number=0.05,
                iter=2000)
    p3 = p3.replace(grav=grav_for_rayleigh(p3, 1e7))
    p5 = Params(nx=8192, ny=8192, length=8192.0, stefan_number=0.1, iter=500)
    p5 = p5.replace(grav=grav_for_rayleigh(p5, 1e7))
    literal = dict(t_cold=273.0, noise_amp=1e-3, noise_seed=2024)
    return [("cavity256_literal", p1, 10000, literal),
            ("cavity2048_ra1e6_pr002", p2, 20000, None),
            ("cavity2048_literal", p2, 20000, literal),
            ("cavity4096_ra1e7_pr002_st005", p3, 2000, None),
            ("cavity8192_ra1e7_st01", p5, 500, None)]


def long_fixture(only=None) -> None:
    import time
    from oracle_py import Oracle
    from oracle_py import read_snapshot
    data = json.loads(LONG.read_text()) if LONG.exists() else {}
    for name, p, nsteps, ext in long_cases():
        if only and name not in only:
            continue
        t0 = time.time()
        if ext is None:
            exe = REF_DIR / "openmp_h" / "ref_harness"
            with tempfile.TemporaryDirectory(prefix="bgklong_", dir="/tmp") as wd:
                write_input_dat(p, os.path.join(wd, "input.dat"))
                os.mkdir(os.path.join(wd, "snap"))
                subprocess.run([str(exe), str(nsteps), "1", "0", os.path.join(wd, "snap"), "0"],
                               cwd=wd, check=True, capture_output=True)
                snap = read_snapshot(os.path.join(wd, "snap", "final.bin"), p.nx, p.ny)
                fields = {f: snap[FIELD_NAMES.index(f)] for f in ("rho", "u", "v", "T", "phi")}
            source = "reference STEP_OPENMP (oracle/_ref/openmp_h/ref_harness)"
        else:
            o = Oracle(p, **ext)
            o.step(nsteps)
            fields = {f: o.field(f).copy() for f in ("rho", "u", "v", "T", "phi")}
            o.close()
            source = "C oracle (oracle/liboracle.so) with extensions " + json.dumps(ext)
        data[name] = {"params": asdict(p), "nsteps": nsteps, "extensions": ext,
                      "source": source, "digests": field_digests(fields)}
        LONG.write_text(json.dumps(data, indent=1) + "\n")
        print(f"{name}: {time.time() - t0:.0f} s", flush=True)


def update_manifest(path) -> None:
    mf = HERE / "MANIFEST.sha256"
    lines = [l for l in mf.read_text().splitlines() if not l.endswith("  " + path.name)]
    lines.append(f"{hashlib.sha256(path.read_bytes()).hexdigest()}  {path.name}")
    mf.write_text("\n".join(lines) + "\n")


def main() -> None:
    if len(sys.argv) > 1 and sys.argv[1] == "formats":
        update_manifest(formats_fixture())
        return
    if len(sys.argv) > 1 and sys.argv[1] == "thot":
        update_manifest(thot_fixture())
        return
    if len(sys.argv) > 1 and sys.argv[1] == "long":
        long_fixture(sys.argv[2:] or None)
        return
    manifest = []
    for name, (p, nsteps, window, stages, fields, diag_every) in CASES.items():
        variant = "step7_r8" if p.rounding == 1 else "step7"
        first = window[0] if window else 1
        last = window[-1] if window else 0
        r = run_ref(p, nsteps, first, last, variant=variant, diag_every=diag_every)
        idx = [FIELD_NAMES.index(f) for f in fields]
        data = {
            "params": np.array(json.dumps(asdict(p))),
            "nsteps": np.array(nsteps),
            "fields": np.array(fields),
            "init": r["init"][idx],
            "final": r["final"][idx],
        }
        for s in window:
            for k in stages:
                data[f"s{s}_{k}"] = r[("s", s, k)][idx]
        for key, val in r.items():
            if isinstance(key, str) and key.startswith("text:"):
                data["text_" + key[5:].replace(".", "_").replace("-", "_")] = np.array(val)
        if not np.all(np.isfinite(data["final"])):
            raise RuntimeError(f"{name}: non-finite reference output")
        path = HERE / f"{name}.npz"
        np.savez_compressed(path, **data)
        digest = hashlib.sha256(path.read_bytes()).hexdigest()
        manifest.append(f"{digest}  {path.name}")
        print(f"{name}: {path.stat().st_size / 1024:.0f} KiB")
    path = driver_fixture()
    manifest.append(f"{hashlib.sha256(path.read_bytes()).hexdigest()}  {path.name}")
    path = formats_fixture()
    manifest.append(f"{hashlib.sha256(path.read_bytes()).hexdigest()}  {path.name}")
    path = thot_fixture()
    manifest.append(f"{hashlib.sha256(path.read_bytes()).hexdigest()}  {path.name}")
    (HERE / "MANIFEST.sha256").write_text("\n".join(manifest) + "\n")


if __name__ == "__main__":
    main()
