"""The five benchmark configurations (bgk_pcm_amd/configs.py) at their full
sizes and lengths.

 - config 1, 256^2 x 10000 steps: the FP64 fused step against the oracle
   run live, bit for bit (STEP_7's physics); the literal configuration (cold
   right wall, seeded noise) on the FP64 population kernel against the
   oracle's digest (tests/golden/long_runs.json), bit for bit, and on FP64
   moments against the oracle within 1e-10.
 - config 2, 2048^2 x 20000 steps: the FP64 fused step against the digest
   of the reference program's own run (STEP_OPENMP built from its sources:
   STEP_7's arithmetic), bit for bit after the full run, which is how the
   reference validates a run (UTIL/RunMe_thread.sh:7); the literal
   configuration against the oracle's digest.
 - config 3, 4096^2: the FP64 fused step of the cavity as STEP_7 can express
   it (no conductivity ratio) against the digest of the reference's
   STEP_OPENMP run after 2000 steps, bit for bit; over 20000 steps (with
   the conductivity ratio 2) both FP32 storages against the clean FP64
   run from the same initial state, every field within 1e-5 (relative
   max-norm, liquid fraction for phi) after 5000, 10000 and 20000 steps.
 - config 4, 512^2 adjoint: the gradient of the probe loss against central
   differences of the GPU forward (step 1e-6), within 1e-4.
 - config 5, 8192^2: the FP64 fused step against the digest of the
   reference's STEP_OPENMP run after 500 steps, bit for bit; the FP32
   moment forward against the FP64 moment
   forward after its 5000 steps (rho, T, phi within 1e-5, u and v within
   2e-5); FP32 forward + adjoint with binomial checkpointing in
   the memory the device has free fits in HBM, gives a finite gradient and
   costs at most 3x the plain forward steps (recomputations included); at
   1024^2 the FP32 adjoint is within 1e-4 of the FP64 adjoint (two
   checkpoint levels there).
"""
import hashlib
import json
import pathlib
import time

import numpy as np
import pytest

from bgk_pcm_amd import configs
from bgk_pcm_amd.adjoint import seeded_probes
from bgk_pcm_amd.lattice import Lattice
from bgk_pcm_amd.params import FP32_DEV, FP32_MOM, FP64, FP64_MOM, ROUND_CLEAN, ROUND_STEP7
from golden_io import compare_state
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

STATE = ["rho", "u", "v", "T", "phi"]
LONG = json.loads((pathlib.Path(__file__).resolve().parent / "golden" /
                   "long_runs.json").read_text())


def rel_errors(got, want):
    """Relative max-norm per field over the interior, liquid fraction
    (1 + phi)/2 in place of phi."""
    out = {}
    for f in STATE:
        a, b = got[f][1:-1, 1:-1], want[f][1:-1, 1:-1]
        if f == "phi":
            a, b = (1.0 + a) / 2.0, (1.0 + b) / 2.0
        out[f] = float(np.abs(a - b).max() / np.abs(b).max())
    return out


def digest_mismatches(lat, entry):
    bad = []
    for f in STATE:
        a = np.ascontiguousarray(lat.download(f), dtype="<f8")
        want = entry["digests"][f]
        if hashlib.sha256(a.tobytes()).hexdigest() != want["sha256"]:
            bad.append(f"{f}: digest differs (min {a.min():.17g} / {want['min']:.17g}, "
                       f"max {a.max():.17g} / {want['max']:.17g}, "
                       f"mean {a.mean():.17g} / {want['mean']:.17g})")
    return bad


def check_params(p, entry):
    d = entry["params"]
    for k in ("nx", "ny", "tau_f", "tau_g", "grav", "stefan_number", "rounding"):
        assert d[k] == getattr(p, k), (k, d[k], getattr(p, k))


# ------------------------------------------------------------------ config 1

def test_config1_fp64_is_bit_exact_over_10000_steps():
    w = configs.config1()
    print("config 1, 256^2, STEP_7 physics: FP64 fused, 10000 steps", flush=True)
    o = Oracle(w.params)
    o.step(w.nsteps)
    with Lattice(w.params) as lat:
        lat.step(w.nsteps)
        got = {f: lat.download(f) for f in STATE}
    assert compare_state(got, {f: o.field(f) for f in STATE}, STATE) == []
    assert np.isfinite(got["T"]).all()


def test_config1_literal_fp64_populations_full_run():
    """Cold right wall and seeded noise on the FP64 population kernel:
    equal to the oracle's digest after the 10000 steps."""
    w = configs.config1()
    entry = LONG["cavity256_literal"]
    check_params(w.params, entry)
    assert entry["extensions"] == w.ext and entry["nsteps"] == w.nsteps
    print("config 1 literal (cold wall, noise): FP64 fused, 10000 steps", flush=True)
    with Lattice(w.params, **w.ext) as lat:
        lat.step(w.nsteps)
        assert digest_mismatches(lat, entry) == []


def test_config1_literal_fp64_moments():
    """The same on FP64 moment storage, which rebuilds the populations from
    moments: within 1e-10 of the oracle (clean rounding)."""
    w = configs.config1()
    p = w.params.replace(rounding=ROUND_CLEAN)
    print("config 1 literal: FP64 moments, 10000 steps", flush=True)
    o = Oracle(p, **w.ext)
    o.step(w.nsteps)
    with Lattice(p.replace(precision=FP64_MOM), **w.ext) as lat:
        lat.step(w.nsteps)
        err = rel_errors({f: lat.download(f) for f in STATE}, {f: o.field(f) for f in STATE})
    print("relative max-norm errors:", err)
    assert max(err.values()) <= 1e-10, err


# ------------------------------------------------------------------ config 2

def test_config2_fp64_equals_the_reference_after_20000_steps():
    w = configs.config2()
    entry = LONG["cavity2048_ra1e6_pr002"]
    check_params(w.params, entry)
    assert entry["extensions"] is None and entry["nsteps"] == w.nsteps
    print("config 2, 2048^2 Ra 1e6 Pr 0.02: FP64 fused 20000 steps vs the reference's "
          "digest", flush=True)
    with Lattice(w.params) as lat:
        lat.step(7000)
        lat.step(13000)
        assert digest_mismatches(lat, entry) == []


def test_config2_literal_fp64_full_run():
    w = configs.config2()
    entry = LONG["cavity2048_literal"]
    check_params(w.params, entry)
    assert entry["extensions"] == w.ext and entry["nsteps"] == w.nsteps
    print("config 2 literal (cold wall, noise): FP64 fused 20000 steps vs the oracle's "
          "digest", flush=True)
    with Lattice(w.params, **w.ext) as lat:
        lat.step(w.nsteps)
        assert digest_mismatches(lat, entry) == []


def test_config2_first_steps_against_the_oracle():
    """The first 20 steps against the oracle run live (a readable report
    should the digest ever differ)."""
    w = configs.config2()
    print("config 2: FP64 fused, 20 steps vs the oracle", flush=True)
    o = Oracle(w.params, **w.ext)
    o.step(20)
    with Lattice(w.params, **w.ext) as lat:
        lat.step(20)
        got = {f: lat.download(f) for f in STATE}
    assert compare_state(got, {f: o.field(f) for f in STATE}, STATE) == []


# ------------------------------------------------------------------ config 3

def test_config3_cavity_fp64_equals_the_reference_after_2000_steps():
    """Config 3's cavity as STEP_7 can express it (4096^2, Ra 1e7, Pr 0.02,
    Stefan 0.05; no conductivity ratio): the FP64 fused step against the
    digest of the reference's own STEP_OPENMP run after 2000 steps, bit for
    bit, at the benchmark's size."""
    w = configs.config3()
    entry = LONG["cavity4096_ra1e7_pr002_st005"]
    p = w.params.replace(precision=FP64, rounding=ROUND_STEP7)
    check_params(p, entry)
    assert entry["extensions"] is None
    print("config 3 cavity, 4096^2, STEP_7 physics: FP64 fused, "
          f"{entry['nsteps']} steps vs the reference's digest", flush=True)
    with Lattice(p) as lat:
        lat.step(entry["nsteps"])
        assert digest_mismatches(lat, entry) == []


MARKS = (5000, 10000, 20000)


@pytest.fixture(scope="module")
def config3_fp64():
    """The clean FP64 population run of config 3 at the marks."""
    w = configs.config3()
    p = w.params.replace(precision=FP64)
    out = {}
    print("config 3, 4096^2: clean FP64 populations, 20000 steps", flush=True)
    with Lattice(p, **w.ext) as lat:
        done = 0
        for m in MARKS:
            lat.step(m - done)
            done = m
            out[m] = {f: lat.download(f) for f in STATE}
    return out


@pytest.mark.parametrize("precision", [FP32_MOM, FP32_DEV], ids=["fp32_moments",
                                                                  "fp32_populations"])
def test_config3_fp32_storage_within_1e5_at_every_mark(config3_fp64, precision):
    w = configs.config3()
    p = w.params.replace(precision=precision)
    errs = {}
    print(f"config 3, 4096^2: precision {precision}, 20000 steps", flush=True)
    with Lattice(p, **w.ext) as lat:
        done = 0
        for m in MARKS:
            lat.step(m - done)
            done = m
            errs[m] = rel_errors({f: lat.download(f) for f in STATE}, config3_fp64[m])
            print(m, errs[m], flush=True)
    for m, e in errs.items():
        assert max(e.values()) <= 1e-5, (m, e)


# ------------------------------------------------------------------ config 4

def adjoint_problem(w, n_probes=16, seg_outer=0, seg=100, binomial=False):
    base = w.params
    pi, pj = seeded_probes(base.nx, base.ny, n_probes, seed=0)

    def record(theta, keep=False):
        st, kr, th = theta
        lat = Lattice(base.replace(stefan_number=st), initialize=False,
                      t_hot=configs.t_hot_of(base, th), k_ratio=kr)
        lat.initialization()
        if binomial:
            lat.adj_setup_binomial(w.nsteps, 0, pi, pj)
        else:
            lat.adj_setup(w.nsteps, seg, pi, pj, seg_outer=seg_outer)
        rec = lat.adj_forward()
        if keep:
            return lat, rec
        lat.close()
        return rec

    return record


def test_config4_adjoint_matches_central_differences():
    w = configs.config4()
    record = adjoint_problem(w)
    print("config 4, 512^2 adjoint: target trajectory", flush=True)
    target = record(configs.ADJ_TARGET)
    theta0 = np.array(configs.ADJ_THETA0)
    lat, rec = record(theta0, keep=True)
    print("config 4: backward", flush=True)
    grad = lat.adj_backward(2.0 * (rec - target) / rec.size)
    lat.close()
    # t_hot enters in kelvin: the gradient is per kelvin, the step 1e-6 in
    # theta units is 1e-5 K
    dT = w.params.t0 - w.params.t_in_sol
    loss = lambda r: float(((r - target) ** 2).mean())
    for c, h in enumerate((1e-6, 1e-6, 1e-6)):
        tp, tm = theta0.copy(), theta0.copy()
        tp[c] += h
        tm[c] -= h
        print(f"config 4: central difference {c}", flush=True)
        fd = (loss(record(tp)) - loss(record(tm))) / (2 * h)
        g = grad[c] * (dT if c == 2 else 1.0)
        print(f"parameter {c}: adjoint {g:.12e} central difference {fd:.12e}", flush=True)
        assert abs(g - fd) <= 1e-4 * abs(fd), (c, g, fd)


# ------------------------------------------------------------------ config 5

def test_config5_cavity_fp64_equals_the_reference_after_500_steps():
    """Config 5's cavity at its full size (8192^2, Ra 1e7, Stefan 0.1, the
    reference's relaxation times): the FP64 fused step against the digest of
    the reference's own STEP_OPENMP run after 500 steps, bit for bit."""
    w = configs.config5()
    entry = LONG["cavity8192_ra1e7_st01"]
    p = w.params.replace(precision=FP64, rounding=ROUND_STEP7)
    check_params(p, entry)
    assert entry["extensions"] is None
    print(f"config 5 cavity, 8192^2: FP64 fused, {entry['nsteps']} steps vs the "
          "reference's digest", flush=True)
    with Lattice(p) as lat:
        lat.step(entry["nsteps"])
        assert digest_mismatches(lat, entry) == []


def test_config5_fp32_forward_against_fp64_at_8192():
    """Config 5's forward model at its full size and length: the FP32 moment
    run against the FP64 moment run (the same model in FP64) from the same
    initial state after 5000 steps.  rho, T and the liquid fraction stay
    within 1e-5 (measured 1.2e-8, 1.2e-8, 3.2e-7); the velocities within
    2e-5 (measured u 3.1e-6, v 1.17e-5): the v difference grows as the
    convection develops (relative 1.2e-6, 1.7e-6, 3.6e-6, 6.9e-6, 1.17e-5 at
    500, 1000, 2000, 3000, 5000 steps; absolute ~ steps^2), the FP32
    rounding of u and v (3e-8 relative per step) amplified by the flow,
    so config 5 misses the 1e-5 of config 3's bound in v by 17%."""
    w = configs.config5()
    got = {}
    for prec in (FP32_MOM, FP64_MOM):
        print(f"config 5, 8192^2: precision {prec}, {w.nsteps} steps", flush=True)
        with Lattice(w.params.replace(precision=prec)) as lat:
            lat.step(w.nsteps)
            got[prec] = {f: lat.download(f) for f in STATE}
    err = rel_errors(got[FP32_MOM], got[FP64_MOM])
    print("relative max-norm errors:", err, flush=True)
    assert max(err[f] for f in ("rho", "T", "phi")) <= 1e-5, err
    assert max(err["u"], err["v"]) <= 2e-5, err


def test_config5_fp32_adjoint_fits_and_matches_fp64_at_1024():
    from bgk_pcm_amd.lattice import Lattice as L
    w = configs.config5()
    record = adjoint_problem(w, binomial=True)
    print("config 5, 8192^2: target trajectory (5000 steps)", flush=True)
    target = record(configs.ADJ_TARGET)
    lat, rec = record(configs.ADJ_THETA0, keep=True)
    adj_bytes, nstates = lat.adj_info()
    info = lat.info()
    print(f"config 5: {nstates} states, {(adj_bytes + info['device_bytes']) / 1e9:.1f} GB of "
          f"{info['device_total'] / 1e9:.1f}; backward", flush=True)
    assert adj_bytes + info["device_bytes"] <= info["device_total"]
    t0 = time.perf_counter()
    g = lat.adj_backward(2.0 * (rec - target) / rec.size)
    t_bwd = time.perf_counter() - t0
    lat.close()
    assert np.isfinite(g).all() and np.all(g != 0.0), g
    # cost of the backward pass (recomputations included) against plain
    # forward steps of the same lattice
    with Lattice(w.params) as lat:
        lat.synchronize()
        t0 = time.perf_counter()
        lat.step(w.nsteps)
        lat.synchronize()
        t_fwd = time.perf_counter() - t0
    ratio = t_bwd / t_fwd
    print(f"config 5: backward {t_bwd / w.nsteps * 1e3:.3f} ms/step, forward "
          f"{t_fwd / w.nsteps * 1e3:.3f} ms/step: {ratio:.3f}x", flush=True)
    assert ratio <= 3.0, ratio
    grads = {}
    for prec in (FP32_MOM, FP64_MOM):
        w2 = configs.config5(1024)
        w2.params = w2.params.replace(precision=prec)
        record = adjoint_problem(w2, seg_outer=289, seg=17)
        print(f"config 5 check, 1024^2 precision {prec}", flush=True)
        tgt = record(configs.ADJ_TARGET)
        lat, rec = record(configs.ADJ_THETA0, keep=True)
        grads[prec] = lat.adj_backward(2.0 * (rec - tgt) / rec.size)
        lat.close()
    rel = np.abs(grads[FP32_MOM] - grads[FP64_MOM]) / np.abs(grads[FP64_MOM])
    print("FP32 vs FP64 adjoint at 1024^2:", rel, flush=True)
    assert rel.max() <= 1e-4, rel
