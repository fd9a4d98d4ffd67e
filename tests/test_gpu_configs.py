"""The benchmark suite's smaller configurations against the oracle (the
4096^2 and 8192^2 ones are measured by bench.py).

 - 256^2 cavity, tau_f 0.8 with tau_g 15.5 (Pr 0.02), Stefan 0.1, 10000
   steps.  Its Rayleigh number 1e5 needs grav = 0.124, with which STEP_7
   itself reaches NaN before step 1000 (its buoyancy uses the absolute
   temperature), so grav is 1e-4 as in a stable run of the same cavity.
   The FP64 fused step against STEP_7's arithmetic (the oracle), bit for
   bit, over the whole run; the config's cold east wall (T_c = T_in_sol)
   and seeded initial noise (1e-3 in theta units) are extensions of the
   moment handles and run through FP64 moment storage against the oracle
   with the same extensions (clean rounding).
 - 2048^2 cavity with grav for Ra 1e6 (Stefan 0.1, the reference's
   relaxation times): the FP64 fused step bit for bit against the oracle
   for 100 of its 20000 steps (the oracle updates ~2e7 nodes per second on
   the host); its seeded noise is an extension, as above.
"""
import numpy as np
import pytest

from bgk_pcm_amd.lattice import Lattice
from bgk_pcm_amd.params import FP64_MOM, ROUND_CLEAN, Params, grav_for_rayleigh
from golden_io import compare_state
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

STATE = ["rho", "u", "v", "T", "phi"]


def cavity_256(**kw):
    return Params(nx=256, ny=256, length=256.0, tau_f=0.8, tau_g=15.5, grav=1e-4,
                  stefan_number=0.1, **kw)


def rel_errors(lat, o):
    out = {}
    for f in STATE:
        a, b = lat.download(f)[1:-1, 1:-1], o.field(f)[1:-1, 1:-1]
        if f == "phi":
            a, b = (1.0 + a) / 2.0, (1.0 + b) / 2.0
        out[f] = float(np.abs(a - b).max() / np.abs(b).max())
    return out


def test_cavity_256_fp64_is_bit_exact_over_10000_steps():
    p = cavity_256()
    print("config 256^2, tau_f 0.8, Pr 0.02: FP64 fused, 10000 steps", flush=True)
    o = Oracle(p)
    o.step(10000)
    with Lattice(p) as lat:
        lat.step(10000)
        got = {f: lat.download(f) for f in STATE}
    assert compare_state(got, {f: o.field(f) for f in STATE}, STATE) == []
    assert np.isfinite(got["T"]).all()


def test_cavity_256_cold_wall_and_noise_fp64_moments():
    """The literal config's east wall at T_c = T_in_sol and initial noise;
    FP64 moments rebuild the populations from moments, so they agree with
    the oracle to rounding (measured on MI355X: at most 2e-11, liquid
    fraction), not bit for bit."""
    p = cavity_256(rounding=ROUND_CLEAN)
    ext = dict(t_cold=p.t_in_sol, noise_amp=1e-3, noise_seed=2024)
    print("config 256^2 with cold east wall and noise: FP64 moments, 10000 steps", flush=True)
    o = Oracle(p, **ext)
    o.step(10000)
    with Lattice(p.replace(precision=FP64_MOM), **ext) as lat:
        lat.step(10000)
        err = rel_errors(lat, o)
    print("relative max-norm errors:", err)
    assert max(err.values()) <= 1e-10, err


def test_cavity_2048_fp64_is_bit_exact():
    p = Params(nx=2048, ny=2048, length=2048.0, stefan_number=0.1)
    p = p.replace(grav=grav_for_rayleigh(p, 1e6))
    print("config 2048^2, Ra 1e6: FP64 fused, 100 steps", flush=True)
    o = Oracle(p)
    o.step(100)
    with Lattice(p) as lat:
        lat.step(60)
        lat.step(40)
        got = {f: lat.download(f) for f in STATE}
    assert compare_state(got, {f: o.field(f) for f in STATE}, STATE) == []
