"""Throughput of the BGK_PCM time step on one MI355X.

Workload: the 4096^2 FP32-storage configuration (bgk_pcm_amd/configs.py,
config3): a melting cavity at Ra = 1e7 and Pr = 0.02 (tau_f 0.51, tau_g 1.0;
grav from the reference's definition), Stefan number 0.05 and a solid/liquid
conductivity ratio of 2, otherwise the reference benchmark input
UTIL/input.dat.0 (no-slip walls), advanced by the fused step with FP32
moment storage (per node rho and T - 273 in FP64; u, v and the flow and
thermal non-equilibrium moments in FP32; FP64 arithmetic).  One "step" is
one time step of the whole lattice (BC_f ... moments of the reference time
loop).  Inputs are resident in HBM before the timed region.

Order of the run: the error runs first (20000 steps of each storage), then
for each storage its per-launch timing (a few thousand steps, which bring
the GPU to its working clocks) followed by its W warm-up and K timed steps.

Printed JSON line (one):
  value          MLUPS of the moment-storage step = Nx*Ny*K/t/1e6 over the K
                 timed steps (one bgk_step call after W warm-up steps)
  roofline       algorithmic HBM bytes per launch (112 B per lattice update:
                 48 B of moments read + written, phi FP64 read + written)
                 over the fused kernel's time per launch, measured with HIP
                 events on the stream the kernel runs on as the difference of
                 a 400- and a 200-step call (which cancels the last launch of
                 a call, that also writes every macroscopic field), median of
                 three; vs 8 TB/s.  last_launch_extra_ms: a one-step call
                 less a plain launch (median of five)
  fp32_population the fused step with FP32 population storage (f -
                 w*rho_ref(row, phi), g - w*T_crit; 160 B per update), same
                 workload and measurements (the kernel's time per launch on
                 a handle that does not re-base its reference density;
                 ms_per_launch_with_rebase with the re-basing every 500
                 steps amortized)
  fp64           the FP64 fused kernel (304 B per update), likewise
  fp32_error     relative max-norm error (per field, liquid fraction for
                 phi) after 5000, 10000 and 20000 steps of the FP32 moment
                 and FP32 population storages, of the FP64 moment storage and
                 of the STEP_7-rounding FP64 run, all against the clean FP64
                 population run from the same initial state; within_1e-5 per
                 storage is the maximum over fields and marks against 1e-5
  slab           the moment step on one GPU's share of an 8-GPU slab
                 decomposition of 8192^2: 8192 x 1026 nodes (1024 rows and
                 the two rows a neighbour would send), same measurements
  adjoint        the adjoint configuration (config4: 512^2, FP64 moments,
                 2000 steps, checkpoint every 100, 16 seeded probes, target
                 trajectory at St 0.12, T_h 1.05, k_ratio 1.5): gradient by
                 the GPU adjoint, central differences beside it, and the cost
                 of the backward pass (recompute + adjoint) per step over the
                 forward step
  adjoint_large  the large adjoint configuration (config5: 8192^2, FP32
                 moments, Ra 1e7, St 0.1, 5000 steps, binomial
                 checkpointing with as many snapshots as the free device
                 memory holds, 16 seeded probes): device bytes of the handle
                 and of the adjoint's checkpoints, the forward steps the
                 backward pass recomputes, forward and backward time per
                 step and their ratio; and the same gradient by the FP32 and
                 the FP64 adjoint at 1024^2 with their relative difference
  cpu_baseline   the reference's own STEP_OPENMP time step (compiled from its
                 sources into oracle/_ref) timed on this host's cores on a
                 bounded sample of the same workload; the CPU oracle (a port)
                 if that build is absent

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N]
                       [--error-steps E] [--cpu-steps S] [--no-cpu-baseline]
                       [--no-fp64] [--no-adjoint] [--no-adjoint-large]
                       [--adjoint-large-steps N] [--adjoint-large-only]
                       [--headline-only] [--calibrate-copy MiB]
                       [--layout-probe NODES]
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import re
import statistics
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
BOUND = 1e-5


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
    ap.add_argument("--headline-only", action="store_true",
                    help="only the timed steps and the roofline of the three storages "
                         "(for profiling them)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB",
                    help="after the timed region, run a device-to-device copy of this size "
                         "(reference kernel for HBM counter scaling in profiles)")
    ap.add_argument("--layout-probe", type=int, default=0, metavar="NODES",
                    help="after the timed region, stream the FP32 moment step's 56 bytes in "
                         "and out per node over three planes and over 64-node tiles "
                         "(libbgk_calib.so) and report both rates")
    return ap.parse_args()


def layout_probe(nodes: int, launches: int = 50) -> dict:
    """Rate of a kernel that only streams the FP32 moment step's bytes (56 in
    and 56 out per node, non-temporal 16- and 8-byte accesses), with the
    three state parts as separate planes and as 64-node tiles
    (bgk_calib_stream56 in libbgk_calib.so): what each layout allows the
    step, apart from its arithmetic and its row-strip walk."""
    import ctypes
    import torch
    n = nodes // 64 * 64
    calib = ctypes.CDLL(str(ROOT / "bgk_pcm_amd" / "lib" / "libbgk_calib.so"))
    calib.bgk_calib_stream56.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                         ctypes.c_int, ctypes.c_void_p]
    src = torch.zeros(56 * n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream()
    out = {"nodes": n, "bytes_per_node": 112}
    for name, tiled in (("planes_tb_s", 0), ("tiles_tb_s", 1)):
        def launch():
            rc = calib.bgk_calib_stream56(src.data_ptr(), dst.data_ptr(), n, tiled,
                                          stream.cuda_stream)
            assert rc == 0, rc
        for _ in range(5):
            launch()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(launches):
            launch()
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / launches
        out[name] = round(112 * n / (ms * 1e-3) / 1e12, 3)
    del src, dst
    return out


def progress(msg: str) -> None:
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def workload(n: int):
    """config3 on an n x n lattice (grav for Ra 1e7 at that size)."""
    from bgk_pcm_amd import configs
    w = configs.config3()
    if n != w.params.nx:
        w.params = configs.cavity(n, 1e7, stefan_number=w.params.stefan_number,
                                  iter=w.params.iter, rounding=w.params.rounding,
                                  **configs.PR002)
    return w


def events(stream):
    import torch
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def call_ms(lat, stream, n: int) -> float:
    """HIP-event time of one bgk_step(n) call on the handle's stream."""
    e0, e1 = events(stream)
    e0.record(stream)
    lat.step(n)
    e1.record(stream)
    lat.synchronize()
    return e0.elapsed_time(e1)


def per_launch(lat, stream, m: int = 200, reps: int = 3):
    """(kernel ms per launch, extra ms of a call's last launch): a call of n
    steps costs a prologue, n - 1 plain launches and a last launch that also
    writes every macroscopic field, so T(2m) - T(m) = m plain launches
    (median over reps); the extra is the median time of one-step calls less
    one plain launch."""
    k = []
    for _ in range(reps):
        t1 = call_ms(lat, stream, m)
        t2 = call_ms(lat, stream, 2 * m)
        k.append((t2 - t1) / m)
    kms = statistics.median(k)
    one = statistics.median(call_ms(lat, stream, 1) for _ in range(5))
    return kms, one - kms


def measure(p, stream, steps, warmup, kind, kernel, ext=None):
    """MLUPS over `steps` timed steps (wall clock, one bgk_step call after
    `warmup` steps), and the roofline figures of the fused kernel from HIP
    events on its stream.  The per-launch timing runs first: its thousands
    of steps bring the GPU to its working clocks (measured: a 20-step call
    right after 5 warm-up steps on a GPU that idled takes 1.19x what the
    next calls take)."""
    from bgk_pcm_amd.lattice import Lattice
    lups = p.nx * p.ny
    more = {}
    if kind == "fp32":
        # FP32 populations re-base their reference density every 500 steps
        # with two other kernels; the fused kernel's own time per launch
        # comes from a handle that never re-bases (BGK_REBASE_EVERY=0 at
        # its creation)
        old = os.environ.get("BGK_REBASE_EVERY")
        os.environ["BGK_REBASE_EVERY"] = "0"
        try:
            with Lattice(p, **(ext or {})) as lat:
                lat.set_stream(stream.cuda_stream)
                kms, extra = per_launch(lat, stream)
        finally:
            if old is None:
                del os.environ["BGK_REBASE_EVERY"]
            else:
                os.environ["BGK_REBASE_EVERY"] = old
    with Lattice(p, **(ext or {})) as lat:
        lat.set_stream(stream.cuda_stream)
        if kind == "fp32":
            # calls of 500 and 1000 steps hold one and two re-basings: the
            # time per step with the re-basing amortized
            kr, _ = per_launch(lat, stream, m=500, reps=1)
            more["ms_per_launch_with_rebase"] = round(kr, 5)
        else:
            kms, extra = per_launch(lat, stream)
        lat.step(warmup)
        lat.synchronize()
        t0 = time.perf_counter()
        lat.step(steps)
        lat.synchronize()
        wall = (time.perf_counter() - t0) * 1e3 / steps
        lat.field_check()  # raises on a non-finite field
    nbytes = BYTES_PER_LUP[kind] * lups
    ach = nbytes / (kms * 1e-3) / 1e9
    return {"mlups": round(lups / (wall * 1e-3) / 1e6, 1), "ms_per_step": round(wall, 5),
            "kernel": kernel, "kernel_ms": round(kms, 5),
            "last_launch_extra_ms": round(extra, 5),
            "achieved_gbs": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4),
            "bytes_per_lup": BYTES_PER_LUP[kind], "algorithmic_bytes_per_launch": nbytes,
            "traffic": traffic_from_profile(p.nx, p.ny, kind), **more}


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


def errors(w, error_steps: int):
    """Every storage against the clean FP64 population run from the same
    initial state, after 1/4, 1/2 and all of error_steps."""
    from bgk_pcm_amd.lattice import Lattice
    from bgk_pcm_amd.params import FP32_DEV, FP32_MOM, FP64, FP64_MOM, ROUND_STEP7
    marks = sorted({max(error_steps // 4, 1), max(error_steps // 2, 1), error_steps})
    p = w.params

    def run(pp):
        out = {}
        with Lattice(pp, **w.ext) as lat:
            done = 0
            for m in marks:
                lat.step(m - done)
                done = m
                out[m] = {f: lat.download(f) for f in FIELDS}
        return out

    progress(f"errors: clean FP64 populations, {error_steps} steps")
    truth = run(p.replace(precision=FP64))
    res = {"vs": "clean FP64 populations (same initial state)", "marks": marks}
    for name, pp in (("fp32_moments", p.replace(precision=FP32_MOM)),
                     ("fp32_populations", p.replace(precision=FP32_DEV)),
                     ("fp64_moments", p.replace(precision=FP64_MOM)),
                     ("fp64_step7_rounding", p.replace(precision=FP64, rounding=ROUND_STEP7))):
        progress(f"errors: {name}")
        got = run(pp)
        res[name] = [dict(steps=m, **rel_errors(got[m], truth[m])) for m in marks]
        del got
    worst = lambda name: max(max(v for k, v in e.items() if k != "steps") for e in res[name])
    res["within_1e-5"] = {name: bool(worst(name) <= BOUND)
                          for name in ("fp32_moments", "fp32_populations", "fp64_moments")}
    return res


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


def traffic_from_profile(nx: int, ny: int, kind: str):
    """HBM bytes per launch measured with rocprofv3 counters, when a
    summary for this kernel and lattice is committed under profiles/."""
    f = ROOT / "profiles" / f"fused_{kind}_{nx}x{ny}_hbm_traffic.json"
    if not f.exists():
        return None
    try:
        return float(json.loads(f.read_text())["bytes_per_launch"])
    except Exception:
        return None


def adjoint_run():
    """config4: gradient at (St, k_ratio, T_h) = (0.1, 1, 1) of the mean
    squared error of T at 16 seeded probes against the trajectory at (0.12,
    1.5, 1.05); central differences with steps 1e-6 in St, k_ratio and T_h
    (theta units: 1e-5 K of t_hot)."""
    import numpy as np
    from bgk_pcm_amd import configs
    from bgk_pcm_amd.adjoint import seeded_probes
    from bgk_pcm_amd.lattice import Lattice
    w = configs.config4()
    base, nsteps, seg = w.params, w.nsteps, 100
    dT = base.t0 - base.t_in_sol
    pi, pj = seeded_probes(base.nx, base.ny, 16, seed=0)

    def forward(theta, keep=False):
        st, kr, th = theta
        lat = Lattice(base.replace(stefan_number=st), initialize=False,
                      t_hot=configs.t_hot_of(base, th), k_ratio=kr)
        lat.initialization()
        lat.adj_setup(nsteps, seg, pi, pj)
        t0 = time.perf_counter()
        rec = lat.adj_forward()
        tf = time.perf_counter() - t0
        if keep:
            return lat, rec, tf
        lat.close()
        return rec

    target = forward(configs.ADJ_TARGET)
    theta0 = np.array(configs.ADJ_THETA0)
    lat, rec, t_fwd = forward(theta0, keep=True)
    seeds = 2.0 * (rec - target) / rec.size
    t0 = time.perf_counter()
    grad = lat.adj_backward(seeds)
    t_bwd = time.perf_counter() - t0
    lat.close()
    grad[2] *= dT  # per theta unit, as the differences below
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
    for c in range(3):
        tp, tm = theta0.copy(), theta0.copy()
        tp[c] += 1e-6
        tm[c] -= 1e-6
        fd.append((loss(forward(tp)) - loss(forward(tm))) / 2e-6)
    rel = [abs(g - f) / abs(f) for g, f in zip(grad, fd)]
    return {"workload": f"{w.note}; 16 probes; gradient at (St, k_ratio, T_h) = "
                        f"{configs.ADJ_THETA0}, target {configs.ADJ_TARGET}",
            "loss": loss(rec), "grad": [float(g) for g in grad], "central_difference": fd,
            "fd_step": 1e-6, "rel_error": [float(x) for x in rel], "within_1e-4": bool(max(rel) <= 1e-4),
            "forward_ms_per_step": round(t_plain / nsteps * 1e3, 5),
            "recorded_forward_ms_per_step": round(t_fwd / nsteps * 1e3, 5),
            "backward_ms_per_step": round(t_bwd / nsteps * 1e3, 5),
            "backward_over_forward": round(t_bwd / t_plain, 3)}


def adjoint_large(nsteps: int = 5000, seg_outer: int = 289, seg: int = 17,
                  check_n: int | None = 1024):
    """config5: FP32 moment storage at 8192^2 over nsteps steps with
    binomial checkpointing in the memory the device has free (the 1024^2
    check with two levels of checkpoints every seg_outer and seg steps);
    the loss is the mean squared error of T at 16 seeded probes against the
    trajectory at (St, k_ratio, T_h) = (0.12, 1.5, 1.05).  Then the FP32 and
    the FP64 adjoint of the same loss at check_n^2, for the gradient
    check."""
    import numpy as np
    from bgk_pcm_amd import configs
    from bgk_pcm_amd.adjoint import seeded_probes
    from bgk_pcm_amd.lattice import Lattice
    from bgk_pcm_amd.lattice import adj_plan
    from bgk_pcm_amd.params import FP32_MOM, FP64_MOM
    w = configs.config5()

    def record(base, theta, pi, pj, keep=False):
        st, kr, th = theta
        lat = Lattice(base.replace(stefan_number=st), initialize=False,
                      t_hot=configs.t_hot_of(base, th), k_ratio=kr)
        lat.initialization()
        if base.nx == w.params.nx:
            lat.adj_setup_binomial(nsteps, 0, pi, pj)
        else:
            lat.adj_setup(nsteps, seg, pi, pj, seg_outer=seg_outer)
        t0 = time.perf_counter()
        rec = lat.adj_forward()
        tf = time.perf_counter() - t0
        if keep:
            return lat, rec, tf
        lat.close()
        return rec

    def gradient(base, pi, pj, target):
        lat, rec, tf = record(base, configs.ADJ_THETA0, pi, pj, keep=True)
        t0 = time.perf_counter()
        g = lat.adj_backward(2.0 * (rec - target) / rec.size)
        tb = time.perf_counter() - t0
        adj_bytes, nstates = lat.adj_info()
        handle_bytes = lat.info()["device_bytes"]
        lat.close()
        return g, tf, tb, adj_bytes, nstates, handle_bytes

    out = {}
    runs = [(w.params.nx, (FP32_MOM,))] + ([(check_n, (FP32_MOM, FP64_MOM))] if check_n else [])
    for size, prec_list in runs:
        base = configs.config5(size).params
        pi, pj = seeded_probes(size, size, 16, seed=0)
        for prec in prec_list:
            b = base.replace(precision=prec)
            progress(f"adjoint {size}^2 {'FP32' if prec == FP32_MOM else 'FP64'} moments: target")
            target = record(b, configs.ADJ_TARGET, pi, pj)
            progress(f"adjoint {size}^2: forward and backward")
            out[(size, prec)] = gradient(b, pi, pj, target)
    n = w.params.nx
    g, tf, tb, adj_bytes, nstates, hb = out[(n, FP32_MOM)]
    # plain forward steps of the same handle type over the same steps (the
    # step gets slower as the melting front widens), for the cost ratio
    progress(f"adjoint {n}^2: plain forward steps")
    with Lattice(w.params) as lat:
        lat.synchronize()
        t0 = time.perf_counter()
        lat.step(nsteps)
        lat.synchronize()
        t_plain = (time.perf_counter() - t0) / nsteps
    res = {"workload": f"{w.note}, {nsteps} steps, binomial checkpointing in the free "
                       "device memory, 16 probes",
           "grad": [float(x) for x in g],
           "states_stored": nstates,
           "recomputed_steps": adj_plan(nsteps, nstates - 2),
           "adjoint_device_bytes": adj_bytes, "handle_device_bytes": hb,
           "hbm_footprint_gb": round((adj_bytes + hb) / 1e9, 2),
           "forward_ms_per_step": round(t_plain * 1e3, 4),
           "recorded_forward_ms_per_step": round(tf / nsteps * 1e3, 4),
           "backward_ms_per_step": round(tb / nsteps * 1e3, 4),
           "backward_over_forward": round(tb / nsteps / t_plain, 3)}
    if check_n:
        g32, g64 = out[(check_n, FP32_MOM)][0], out[(check_n, FP64_MOM)][0]
        rel = [float(abs(a - b) / abs(b)) for a, b in zip(g32, g64)]
        res["check"] = {"size": f"{check_n}x{check_n}", "steps": nsteps,
                        "grad_fp32_moments": [float(x) for x in g32],
                        "grad_fp64_moments": [float(x) for x in g64],
                        "rel_difference": rel, "within_1e-4": bool(max(rel) <= 1e-4)}
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: the time step runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU slab decomposition is design-only)", file=sys.stderr)
        sys.exit(2)
    import torch
    from bgk_pcm_amd import configs
    from bgk_pcm_amd.params import FP32_DEV, FP32_MOM, FP64

    torch.cuda.set_device(0)
    if args.adjoint_large_only:
        print(json.dumps({"adjoint_large": adjoint_large(nsteps=args.adjoint_large_steps,
                                                         check_n=None)}), flush=True)
        return
    stream = torch.cuda.Stream()
    n = args.size
    w = workload(n)
    p, ext = w.params, w.ext
    pm = p.replace(precision=FP32_MOM)
    pp = p.replace(precision=FP32_DEV)
    mom_kind = "mom_eq" if p.tau_f == 1.0 else "mom"

    err = None
    if args.error_steps > 0 and not args.headline_only:
        err = errors(w, args.error_steps)
    progress("timed steps")
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
            "workload": f"melting cavity {n}x{n}, Ra 1e7, Pr 0.02, Stefan "
                        f"{p.stefan_number:g}, solid/liquid conductivity ratio "
                        f"{ext['k_ratio']:g} (UTIL/input.dat.0 otherwise), fused time step, "
                        "FP32 moment storage",
            "nx": n, "ny": n,
            "storage": "per node rho, T-273 FP64; u, v, (1-omega_f) Pi_f^neq, "
                       "(1-omega_g) Pi_g^neq FP32",
            "arithmetic": "f64",
            "boundaries": "no-slip (type 4) on all walls",
            "tau_f": p.tau_f, "tau_g": p.tau_g, "grav": p.grav, "rayleigh": 1e7,
            "prandtl": (2 * p.tau_f - 1) / (2 * p.tau_g - 1),
            "stefan_number": p.stefan_number, "k_ratio": ext["k_ratio"],
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
            "last_launch_extra_ms": mm["last_launch_extra_ms"],
            "bytes_per_lup": mm["bytes_per_lup"],
            "algorithmic_bytes_per_launch": mm["algorithmic_bytes_per_launch"],
        },
    }
    res["fp32_population"] = measure(pp, stream, args.steps, args.warmup, "fp32", "k_fused32",
                                     ext)
    if not args.no_fp64:
        res["fp64"] = measure(p.replace(precision=FP64), stream, max(args.steps // 2, 1),
                              args.warmup, "fp64", "k_fused64", ext)
        res["mom_over_fp64"] = round(res["value"] / res["fp64"]["mlups"], 3)
        res["fp32_population_over_fp64"] = round(
            res["fp32_population"]["mlups"] / res["fp64"]["mlups"], 3)
        res["mom_over_fp64_kernel"] = round(res["fp64"]["kernel_ms"] / mm["kernel_ms"], 3)
        res["fp32_population_over_fp64_kernel"] = round(
            res["fp64"]["kernel_ms"] / res["fp32_population"]["kernel_ms"], 3)
    if err is not None:
        res["fp32_error"] = err
    if args.layout_probe > 0:
        progress("layout probe")
        res["layout_probe"] = layout_probe(args.layout_probe)
    if args.headline_only:
        print(json.dumps(res), flush=True)
        return
    # one GPU's share of the 8192^2 lattice over 8 GPUs (design-only sharding)
    ps = configs.cavity(8192, 1e7, stefan_number=p.stefan_number, rounding=p.rounding,
                        precision=FP32_MOM, **configs.PR002).replace(ny=1026)
    res["slab"] = dict(measure(ps, stream, args.steps, args.warmup, mom_kind, "k_fused_mom", ext),
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
