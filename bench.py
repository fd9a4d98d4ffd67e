"""Throughput of the BGK_PCM time step on one MI355X.

Workload: the 4096^2 FP32-storage config, a melting cavity with Ra = 1e7
(grav set by the reference's definition), Stefan number 0.05 and a
solid/liquid conductivity ratio of 2 (the k_ratio extension), otherwise the reference benchmark input UTIL/input.dat.0 (no-slip
walls, tau_f=1, tau_g=0.51: the config's Pr 0.02 diverges at this Ra),
advanced by the fused step with FP32 moment storage (per node rho and
T - 273 in FP64; u, v and the thermal non-equilibrium moments in FP32; FP64
arithmetic).  One "step" is
one time step of the whole lattice (BC_f ... moments of the reference time
loop).  Inputs are resident in HBM before the timed region.

Printed JSON line (one):
  value          MLUPS of the moment-storage step = Nx*Ny*steps/t/1e6
  roofline       algorithmic HBM bytes per launch (88 B per lattice update
                 with tau_f = 1: 36 B of moments read + written, phi FP64
                 read + written; 112 B for tau_f != 1) over the fused
                 kernel's mean launch time, measured with HIP events on the
                 stream the kernel runs on; vs 8 TB/s
  fp32_population the fused step with FP32 population storage (f - w*rho0,
                 g - w*T_crit; 160 B per update), same workload and
                 measurements
  fp64           the FP64 fused kernel (304 B per update), likewise
  fp32_mom_error relative max-norm error (per field, liquid fraction for
                 phi) of the FP32 moment run against the FP64 moment run
                 with the same conductivity ratio after 1/4, 1/2 and all of
                 --error-steps (default 20000, the config's run length)
  fp32_error     the errors after --error-steps of both FP32 storages
                 against the FP64 population run with clean FP64
                 temporaries and against the STEP_7-rounding FP64 run, and
                 of the STEP_7 run against the clean one
  adjoint        gradient of a probe-temperature loss with respect to
                 (stefan_number, k_ratio, t_hot) by the GPU adjoint on the
                 adjoint workload (512^2, FP64 moments, 2000 steps,
                 checkpoint every 100, 16 seeded probes, target trajectory
                 at St 0.12, T_h 1.05, k_ratio 1.5), central differences
                 beside it, and the cost of the backward pass (recompute +
                 adjoint) per step over the forward step
  slab           the moment step on one GPU's share of an 8-GPU slab
                 decomposition of 8192^2: 8192 x 1026 nodes (1024 rows and
                 the two rows a neighbour would send), same measurements
  adjoint_large  the large adjoint workload (8192^2, FP32 moment storage,
                 Ra 1e7, St 0.1, 5000 steps, two levels of checkpoints every
                 289 and 17 steps, 16 seeded probes, the same kind of target
                 trajectory): device bytes of the handle and of the
                 adjoint's checkpoints, forward and backward time per step
                 and their ratio; and the same gradient by the FP32 and the
                 FP64 adjoint at 1024^2 with their relative difference
  cpu_baseline   the reference's own STEP_OPENMP time step (compiled from its
                 sources into oracle/_ref) timed on this host's cores on a
                 bounded sample of the same workload; the CPU oracle (a port)
                 if that build is absent

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N]
                       [--error-steps E] [--cpu-steps S] [--no-cpu-baseline]
                       [--no-fp64] [--no-adjoint] [--no-adjoint-large]
                       [--adjoint-large-steps N] [--adjoint-large-only]
                       [--calibrate-copy MiB]
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import re
import subprocess
import sys
import tempfile
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "MLUPS (D2Q9 f+g) on 1 MI355X: FP32-store vs FP64 at <=1e-5 error; % HBM peak"
# algorithmic HBM bytes per lattice update: state read + written, phi
# (FP64) read + written
BYTES_PER_LUP = {"mom_eq": 36 * 2 + 8 * 2, "mom": 48 * 2 + 8 * 2,
                 "fp32": 18 * 4 * 2 + 8 * 2, "fp64": 18 * 8 * 2 + 8 * 2}
FIELDS = ("rho", "u", "v", "T", "phi")


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--error-steps", type=int, default=20000)
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-adjoint-large", action="store_true")
    ap.add_argument("--adjoint-large-steps", type=int, default=5000)
    ap.add_argument("--adjoint-large-only", action="store_true",
                    help="only the large adjoint workload, without the 1024^2 check "
                         "(for profiling it)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB",
                    help="after the timed region, run a device-to-device copy of this size "
                         "(reference kernel for HBM counter scaling in profiles)")
    return ap.parse_args()


RAYLEIGH = 1e7
STEFAN = 0.05  # the 4096^2 FP32-storage config
K_RATIO = 2.0


def cavity(n: int):
    """The bench workload: UTIL/input.dat.0's cavity on an n x n lattice with
    grav set for Ra = 1e7 (the Rayleigh number of the 4096^2 FP32 config) by
    the reference's own definition.  Pr = 50 stays that of the reference
    relaxation times; grav = 2.5e-4 itself would put a hydrostatic density
    drop of ~3 across 4096 rows (STEP_7's buoyancy uses the absolute T) and
    the run diverges after a few thousand steps."""
    from bgk_pcm_amd.params import Params, grav_for_rayleigh
    p = Params(nx=n, ny=n, iter=20000, length=float(n))
    return p.replace(grav=grav_for_rayleigh(p, RAYLEIGH))


def timed(lat, stream, steps: int, warmup: int):
    """(kernel ms per step from HIP events, wall ms per step)."""
    import torch
    lat.step(warmup)
    lat.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    lat.step(steps)
    ev1.record(stream)
    lat.synchronize()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    return ev0.elapsed_time(ev1) / steps, wall


def gpu_run(p, stream, steps, warmup, ext=None):
    from bgk_pcm_amd.lattice import Lattice
    lat = Lattice(p, **(ext or {}))
    lat.set_stream(stream.cuda_stream)
    kms, wms = timed(lat, stream, steps, warmup)
    fc = lat.field_check()
    lat.close()
    return kms, wms, fc


def rel_errors(a, b):
    """Relative max-norm error max|a-b|/max|b| per field over the interior,
    with the liquid fraction (1+phi)/2 in place of phi."""
    import numpy as np
    out = {}
    for f in FIELDS:
        x, y = a[f][1:-1, 1:-1], b[f][1:-1, 1:-1]
        if f == "phi":
            x, y = (1 + x) / 2, (1 + y) / 2
        out[f] = float(np.abs(x - y).max() / np.abs(y).max())
    return out


def fp32_error(p64, storages: dict, steps: int, ext=None):
    """Errors after `steps` steps from the same initial state: each FP32
    storage run against the FP64 run with clean FP64 temporaries and against
    the STEP_7-rounding FP64 run, and the STEP_7 run against the clean one
    (the reference's own FP32-temporary noise floor); ext: physics
    extensions of every run."""
    from bgk_pcm_amd.lattice import Lattice
    from bgk_pcm_amd.params import FP64, ROUND_CLEAN, ROUND_STEP7
    assert p64.precision == FP64
    runs = {}
    todo = [("fp64_step7", p64.replace(rounding=ROUND_STEP7)),
            ("fp64_clean", p64.replace(rounding=ROUND_CLEAN))] + list(storages.items())
    for name, p in todo:
        with Lattice(p, **(ext or {})) as lat:
            lat.step(steps)
            runs[name] = {f: lat.download(f) for f in FIELDS}
    out = {"steps": steps}
    for name in storages:
        out[f"{name}_vs_fp64_clean"] = rel_errors(runs[name], runs["fp64_clean"])
        out[f"{name}_vs_fp64_step7"] = rel_errors(runs[name], runs["fp64_step7"])
    out["fp64_step7_vs_fp64_clean"] = rel_errors(runs["fp64_step7"], runs["fp64_clean"])
    return out


def mom_error_vs_fp64_mom(p, ext: dict, error_steps: int):
    """The FP32 moment run against the FP64 moment run with the same
    physics extensions (the same model in FP64; the clean FP64 population
    kernel agrees with it to 1e-10 at small sizes), both from the same
    initial state, after 1/4, 1/2 and all of error_steps."""
    from bgk_pcm_amd.lattice import Lattice
    from bgk_pcm_amd.params import FP32_MOM, FP64_MOM
    marks = sorted({max(error_steps // 4, 1), max(error_steps // 2, 1), error_steps})
    errs = []
    with Lattice(p.replace(precision=FP64_MOM), **ext) as l64, \
            Lattice(p.replace(precision=FP32_MOM), **ext) as l32:
        done = 0
        for m in marks:
            l64.step(m - done)
            l32.step(m - done)
            done = m
            e = rel_errors({f: l32.download(f) for f in FIELDS},
                           {f: l64.download(f) for f in FIELDS})
            errs.append(dict(steps=m, **e))
    return errs


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(p, steps: int):
    from bgk_pcm_amd.params import write_input_dat
    exe = ROOT / "oracle" / "_ref" / "openmp" / "ref_bench"
    sample = (f"{p.nx}x{p.ny} cavity, {steps} steps (the GPU run's input.dat; the reference "
              f"has no conductivity ratio), {cpu_model()}")
    if exe.exists():
        env = dict(os.environ)
        env.setdefault("OMP_PROC_BIND", "spread")
        env.setdefault("OMP_PLACES", "cores")
        with tempfile.TemporaryDirectory(prefix="bgkcpu_") as wd:
            write_input_dat(p, os.path.join(wd, "input.dat"))
            r = subprocess.run([str(exe), str(steps)], cwd=wd, capture_output=True, text=True,
                               timeout=900, env=env)
            if r.returncode != 0:
                raise RuntimeError(f"ref_bench failed: {r.stderr[-500:]}")
            line = [l for l in r.stdout.splitlines() if l.startswith("nx=")][-1]
            kv = dict(re.findall(r"(\w+)=\s*(\S+)", line))
            return {"value": float(kv["mlups"]), "unit": "MLUPS", "cores": int(kv["threads"]),
                    "kind": "reference",
                    "sample": sample + "; reference STEP_OPENMP subroutines (oracle/_ref/openmp), "
                              f"{float(kv['seconds']):.2f} s"}
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_py import Oracle
    o = Oracle(p)
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    return {"value": p.nx * p.ny * steps / dt / 1e6, "unit": "MLUPS",
            "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
            "kind": "port", "sample": sample + f"; C oracle with OpenMP, {dt:.2f} s"}


def traffic_from_profile(size: int, kind: str):
    """HBM bytes per launch measured with rocprofv3 counters, when a
    summary for this kernel and size is committed under profiles/."""
    f = ROOT / "profiles" / f"fused_{kind}_{size}x{size}_hbm_traffic.json"
    if not f.exists():
        return None
    try:
        return float(json.loads(f.read_text())["bytes_per_launch"])
    except Exception:
        return None


def adjoint_run(n: int = 512, nsteps: int = 2000, seg: int = 100, n_probes: int = 16):
    """The adjoint workload: gradient at the reference's parameters (St 0.1,
    T_h = T0, k_ratio 1) of the mean squared error of T at 16 seeded probes
    against the trajectory at St 0.12, T_h 1.05 (theta units, T0 - T_in_sol
    = 10 K per unit above T_crit = 273), k_ratio 1.5; grav for Ra = 1e6 (the
    2048^2 config's Rayleigh number)."""
    import numpy as np
    from bgk_pcm_amd.adjoint import seeded_probes
    from bgk_pcm_amd.lattice import Lattice
    from bgk_pcm_amd.params import FP64_MOM, ROUND_CLEAN, Params, grav_for_rayleigh
    base = Params(nx=n, ny=n, length=float(n), rounding=ROUND_CLEAN, precision=FP64_MOM,
                  stefan_number=0.1)
    base = base.replace(grav=grav_for_rayleigh(base, 1e6))
    dT = base.t0 - base.t_in_sol
    pi, pj = seeded_probes(n, n, n_probes, seed=0)

    def forward(theta, keep=False):
        st, kr, th = theta
        lat = Lattice(base.replace(stefan_number=st), initialize=False, t_hot=th, k_ratio=kr)
        lat.initialization()
        lat.adj_setup(nsteps, seg, pi, pj)
        t0 = time.perf_counter()
        rec = lat.adj_forward()
        tf = time.perf_counter() - t0
        if keep:
            return lat, rec, tf
        lat.close()
        return rec

    target = forward((0.12, 1.5, 273.0 + 1.05 * dT))
    theta0 = np.array([0.1, 1.0, base.t0])
    lat, rec, t_fwd = forward(theta0, keep=True)
    seeds = 2.0 * (rec - target) / rec.size
    t0 = time.perf_counter()
    grad = lat.adj_backward(seeds)
    t_bwd = time.perf_counter() - t0
    lat.close()
    # plain forward steps of the same handle type, for the cost ratio
    with Lattice(base) as lat:
        lat.step(50)
        lat.synchronize()
        t0 = time.perf_counter()
        lat.step(nsteps)
        lat.synchronize()
        t_plain = time.perf_counter() - t0
    loss = lambda r: float(((r - target) ** 2).mean())
    fd = []
    for c, h in enumerate((1e-6, 1e-6, 1e-6 * dT)):
        tp, tm = theta0.copy(), theta0.copy()
        tp[c] += h
        tm[c] -= h
        fd.append((loss(forward(tp)) - loss(forward(tm))) / (2 * h))
    rel = [abs(g - f) / abs(f) for g, f in zip(grad, fd)]
    return {"workload": f"{n}x{n} cavity, FP64 moments, {nsteps} steps, checkpoint every {seg}, "
                        f"{n_probes} probes; gradient at (St, k_ratio, t_hot) = "
                        f"({theta0[0]}, {theta0[1]}, {theta0[2]})",
            "loss": loss(rec), "grad": [float(g) for g in grad], "central_difference": fd,
            "fd_steps": [1e-6, 1e-6, 1e-6 * dT], "rel_error": rel,
            "forward_ms_per_step": round(t_plain / nsteps * 1e3, 5),
            "recorded_forward_ms_per_step": round(t_fwd / nsteps * 1e3, 5),
            "backward_ms_per_step": round(t_bwd / nsteps * 1e3, 5),
            "backward_over_forward": round(t_bwd / t_plain, 3)}


def progress(msg: str) -> None:
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def adjoint_large(n: int = 8192, nsteps: int = 5000, seg_outer: int = 289, seg: int = 17,
                  n_probes: int = 16, check_n: int | None = 1024):
    """The large adjoint workload: FP32 moment storage at n^2 (cavity of the
    bench, Ra 1e7, St 0.1) over nsteps steps with two levels of checkpoints
    (about nsteps^(1/3) steps apart in the inner level); the loss is the
    mean squared error of T at seeded probes against the trajectory at St
    0.12, T_h 1.05, k_ratio 1.5.  Then the FP32 and the FP64 adjoint of the
    same loss at check_n^2, for the gradient check."""
    import numpy as np
    from bgk_pcm_amd.adjoint import seeded_probes
    from bgk_pcm_amd.lattice import Lattice
    from bgk_pcm_amd.params import FP32_MOM, FP64_MOM, ROUND_CLEAN

    def record(base, theta, pi, pj, keep=False):
        st, kr, th = theta
        lat = Lattice(base.replace(stefan_number=st), initialize=False, t_hot=th, k_ratio=kr)
        lat.initialization()
        lat.adj_setup(nsteps, seg, pi, pj, seg_outer=seg_outer)
        t0 = time.perf_counter()
        rec = lat.adj_forward()
        tf = time.perf_counter() - t0
        if keep:
            return lat, rec, tf
        lat.close()
        return rec

    def gradient(base, pi, pj, target, theta0):
        lat, rec, tf = record(base, theta0, pi, pj, keep=True)
        t0 = time.perf_counter()
        g = lat.adj_backward(2.0 * (rec - target) / rec.size)
        tb = time.perf_counter() - t0
        adj_bytes, nstates = lat.adj_info()
        handle_bytes = lat.info()["device_bytes"]
        lat.close()
        return g, tf, tb, adj_bytes, nstates, handle_bytes

    out = {}
    runs = [(n, (FP32_MOM,))] + ([(check_n, (FP32_MOM, FP64_MOM))] if check_n else [])
    for size, prec_list in runs:
        base = cavity(size).replace(rounding=ROUND_CLEAN, stefan_number=0.1)
        dT = base.t0 - base.t_in_sol
        theta0 = (0.1, 1.0, base.t0)
        pi, pj = seeded_probes(size, size, n_probes, seed=0)
        for prec in prec_list:
            b = base.replace(precision=prec)
            progress(f"adjoint {size}^2 {'FP32' if prec == FP32_MOM else 'FP64'} moments: target")
            target = record(b, (0.12, 1.5, 273.0 + 1.05 * dT), pi, pj)
            progress(f"adjoint {size}^2: forward and backward")
            g, tf, tb, adj_bytes, nstates, hb = gradient(b, pi, pj, target, theta0)
            out[(size, prec)] = (g, tf, tb, adj_bytes, nstates, hb)
    g, tf, tb, adj_bytes, nstates, hb = out[(n, FP32_MOM)]
    # plain forward steps of the same handle type over the same steps (the
    # step gets slower as the melting front widens), for the cost ratio
    progress(f"adjoint {n}^2: plain forward steps")
    with Lattice(cavity(n).replace(precision=FP32_MOM, rounding=ROUND_CLEAN,
                                   stefan_number=0.1)) as lat:
        lat.synchronize()
        t0 = time.perf_counter()
        lat.step(nsteps)
        lat.synchronize()
        t_plain = (time.perf_counter() - t0) / nsteps
    res = {"workload": f"{n}x{n} cavity, FP32 moments, Ra 1e7, St 0.1, {nsteps} steps, "
                        f"checkpoints every {seg_outer} and {seg} steps, {n_probes} probes",
            "grad": [float(x) for x in g],
            "states_stored": nstates,
            "adjoint_device_bytes": adj_bytes, "handle_device_bytes": hb,
            "hbm_footprint_gb": round((adj_bytes + hb) / 1e9, 2),
            "forward_ms_per_step": round(t_plain * 1e3, 4),
            "recorded_forward_ms_per_step": round(tf / nsteps * 1e3, 4),
            "backward_ms_per_step": round(tb / nsteps * 1e3, 4),
            "backward_over_forward": round(tb / nsteps / t_plain, 3)}
    if check_n:
        g32, g64 = out[(check_n, FP32_MOM)][0], out[(check_n, FP64_MOM)][0]
        res["check"] = {"size": f"{check_n}x{check_n}", "steps": nsteps,
                        "grad_fp32_moments": [float(x) for x in g32],
                        "grad_fp64_moments": [float(x) for x in g64],
                        "rel_difference": [float(abs(a - b) / abs(b)) for a, b in zip(g32, g64)]}
    return res


def measure(p, stream, steps, warmup, kind, kernel, ext=None):
    """MLUPS from the wall clock and roofline figures of the fused kernel
    from HIP events on its stream (ext: physics extensions of the handle)."""
    n = p.nx
    kms, wms, _ = gpu_run(p, stream, steps, warmup, ext)
    lups = p.nx * p.ny
    nbytes = BYTES_PER_LUP[kind] * lups
    ach = nbytes / (kms * 1e-3) / 1e9
    return {"mlups": round(lups / (wms * 1e-3) / 1e6, 1), "ms_per_step": round(wms, 5),
            "kernel": kernel, "kernel_ms": round(kms, 5), "achieved_gbs": round(ach, 1),
            "frac": round(ach / HBM_PEAK_GBS, 4), "bytes_per_lup": BYTES_PER_LUP[kind],
            "algorithmic_bytes_per_launch": nbytes,
            "traffic": traffic_from_profile(n, kind)}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: the time step runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU slab decomposition is design-only)", file=sys.stderr)
        sys.exit(2)
    import torch
    from bgk_pcm_amd.params import FP32_DEV, FP32_MOM, ROUND_CLEAN

    torch.cuda.set_device(0)
    if args.adjoint_large_only:
        print(json.dumps({"adjoint_large": adjoint_large(nsteps=args.adjoint_large_steps,
                                                         check_n=None)}), flush=True)
        return
    stream = torch.cuda.Stream()
    n = args.size
    # the config's Stefan number and conductivity ratio (an extension of
    # every handle; the reference has none)
    p = cavity(n).replace(stefan_number=STEFAN)
    ext = {"k_ratio": K_RATIO}
    # the moment and FP32-population kernels follow the clean rounding of the
    # collision; beta (forces.f90:13) is then FP64 as well
    pm = p.replace(precision=FP32_MOM, rounding=ROUND_CLEAN)
    pp = p.replace(precision=FP32_DEV, rounding=ROUND_CLEAN)
    mom_kind = "mom_eq" if p.tau_f == 1.0 else "mom"

    mm = measure(pm, stream, args.steps, args.warmup, mom_kind, "k_fused_mom", ext)
    res = {
        "metric": METRIC,
        "value": mm["mlups"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mm["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"melting cavity {n}x{n}, Ra 1e7, Stefan {STEFAN}, solid/liquid "
                        f"conductivity ratio {K_RATIO:g} (UTIL/input.dat.0 otherwise), "
                        "fused time step, FP32 moment storage",
            "nx": n, "ny": n,
            "storage": "per node rho, T-273 FP64; u, v, (1-omega_g) Pi_g^neq FP32"
                       + ("" if mom_kind == "mom_eq" else ", (1-omega_f) Pi_f^neq FP32"),
            "arithmetic": "f64",
            "boundaries": "no-slip (type 4) on all walls",
            "tau_f": p.tau_f, "tau_g": p.tau_g, "grav": p.grav, "rayleigh": RAYLEIGH,
            "stefan_number": p.stefan_number, "k_ratio": K_RATIO,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": mm["achieved_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": mm["frac"],
            "traffic": mm["traffic"],
            "kernel": mm["kernel"],
            "kernel_ms": mm["kernel_ms"],
            "bytes_per_lup": mm["bytes_per_lup"],
            "algorithmic_bytes_per_launch": mm["algorithmic_bytes_per_launch"],
        },
    }
    res["fp32_population"] = measure(pp, stream, args.steps, args.warmup, "fp32", "k_fused32",
                                     ext)
    if not args.no_fp64:
        res["fp64"] = measure(p, stream, max(args.steps // 2, 1), args.warmup, "fp64",
                              "k_fused64", ext)
        res["mom_over_fp64"] = round(res["value"] / res["fp64"]["mlups"], 3)
        res["fp32_population_over_fp64"] = round(
            res["fp32_population"]["mlups"] / res["fp64"]["mlups"], 3)
    if args.error_steps > 0:
        progress(f"FP32 moment errors, conductivity ratio {K_RATIO:g}, {args.error_steps} steps")
        res["fp32_mom_error"] = {"k_ratio": K_RATIO,
                                 "vs_fp64_mom": mom_error_vs_fp64_mom(pm, ext, args.error_steps)}
        progress(f"FP32 storage errors, {args.error_steps} steps")
        res["fp32_error"] = fp32_error(p, {"mom": pm, "fp32_population": pp}, args.error_steps,
                                       ext)
    # one GPU's share of the 8192^2 lattice over 8 GPUs (design-only sharding)
    ps = cavity(8192).replace(ny=1026, precision=FP32_MOM, rounding=ROUND_CLEAN)
    res["slab"] = dict(measure(ps, stream, args.steps, args.warmup, mom_kind, "k_fused_mom"),
                       workload="8192 x 1026: one GPU's rows of 8192^2 on 8 GPUs + 2 halo rows")
    if not args.no_adjoint:
        progress("adjoint 512^2")
        res["adjoint"] = adjoint_run()
    if not args.no_adjoint_large:
        res["adjoint_large"] = adjoint_large(nsteps=args.adjoint_large_steps)
    if args.calibrate_copy > 0:
        # copies of known size for counter scaling: the runtime's device copy,
        # then 8-byte plain and non-temporal copies (libbgk_calib.so), the
        # access types of the fused kernels
        import ctypes
        nbytes = args.calibrate_copy * 1024 * 1024
        a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        calib = ctypes.CDLL(str(ROOT / "bgk_pcm_amd" / "lib" / "libbgk_calib.so"))
        calib.bgk_calib_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                         ctypes.c_int, ctypes.c_void_p]
        for nt in (0, 1):
            rc = calib.bgk_calib_copy(a.data_ptr(), b.data_ptr(), nbytes, nt,
                                      torch.cuda.current_stream().cuda_stream)
            assert rc == 0, rc
        torch.cuda.synchronize()
        res["calibration_copy_bytes"] = {"read": nbytes, "write": nbytes}
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(p, args.cpu_steps)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
