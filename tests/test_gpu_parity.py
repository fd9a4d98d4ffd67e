"""The HIP library against the reference's own outputs and the oracle.

Bar: bit-identical FP64 results (the reference STEP_7 path is deterministic
IEEE double arithmetic plus FP32 roundings, all reproduced on gfx950).

 - every per-subroutine kernel in isolation, from reference snapshots;
 - the initial state, whole runs through the per-subroutine kernels and
   through the fused kernel, against the reference fixtures;
 - the fused kernel against the oracle over more sizes, both rounding
   policies, any split of the steps into calls, and mixed with the
   per-subroutine path.
"""
import numpy as np
import pytest

from bgk_pcm_amd.lattice import FIELD_NAMES, Lattice
from bgk_pcm_amd.params import Params
from golden_io import INTERIOR_ONLY, Golden, compare_state, fixtures
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

CASES = fixtures()
STATE = ["rho", "u", "v", "T", "phi"]
POPS = [f"f{k}" for k in range(9)] + [f"g{k}" for k in range(9)]


def gpu_state(lat, fields):
    return {f: lat.download(f) for f in fields}


def compare_interior(got, want, fields):
    bad = []
    for f in fields:
        a, b = got[f][1:-1, 1:-1], want[f][1:-1, 1:-1]
        if not np.array_equal(np.ascontiguousarray(a).view(np.uint64),
                              np.ascontiguousarray(b).view(np.uint64)):
            d = np.abs(a - b)
            bad.append(f"{f}: {np.count_nonzero(a != b)} nodes differ, max|d|={d.max():.3e}")
    return bad


@pytest.mark.parametrize("name", CASES)
def test_initialization_matches_reference(name):
    g = Golden(name)
    print(f"case {name}: initialization", flush=True)
    with Lattice(g.params) as lat:
        fields = [f for f in g.fields if not f.startswith(("fp", "gp")) and not f.endswith("2")]
        assert compare_state(gpu_state(lat, fields), g.snap("init"), fields) == []


@pytest.mark.parametrize("name", [c for c in CASES if Golden(c).window()])
def test_each_subroutine_kernel_in_isolation(name):
    g = Golden(name)
    print(f"case {name}: stage kernels", flush=True)
    with Lattice(g.params, initialize=False) as lat:
        for step, stages in g.window():
            for k in stages:
                if k == 0 or (k - 1) not in stages:
                    continue
                for f, a in g.snap(f"s{step}_{k - 1}").items():
                    lat.upload(f, a)
                lat.stage(Lattice.STAGES[k - 1])
                bad = compare_state(gpu_state(lat, g.fields), g.snap(f"s{step}_{k}"), g.fields)
                assert bad == [], f"step {step} {Lattice.STAGES[k - 1]}: {bad[:5]}"


@pytest.mark.parametrize("name", CASES)
def test_subroutine_path_whole_run(name):
    g = Golden(name)
    print(f"case {name}: per-subroutine run of {g.nsteps} steps", flush=True)
    with Lattice(g.params) as lat:
        for _ in range(g.nsteps):
            for s in Lattice.STAGES:
                lat.stage(s)
        bad = compare_state(gpu_state(lat, g.fields), g.snap("final"), g.fields)
        assert bad == [], bad[:5]


@pytest.mark.parametrize("name", CASES)
def test_bgk_step_whole_run(name):
    """bgk_step: fused kernel for the no-slip cavity, the per-subroutine
    kernels otherwise.  Macroscopic fields are compared halo included;
    populations on the interior (their ghosts are scratch of the fused
    kernel)."""
    g = Golden(name)
    print(f"case {name}: bgk_step({g.nsteps})", flush=True)
    with Lattice(g.params) as lat:
        lat.step(g.nsteps)
        want = g.snap("final")
        got = gpu_state(lat, STATE + POPS)
        assert compare_state(got, want, STATE) == []
        assert compare_interior(got, want, POPS) == []


SIZES = [(67, 45, 150, 0), (130, 33, 100, 1), (1, 1, 40, 0), (2, 3, 40, 1), (3, 2, 40, 0),
         (97, 70, 60, 1)]


@pytest.mark.parametrize("nx,ny,nsteps,rounding", SIZES)
def test_fused_vs_oracle(nx, ny, nsteps, rounding):
    p = Params(nx=nx, ny=ny, rounding=rounding, re=3.0, length=float(ny))
    print(f"fused vs oracle {nx}x{ny} {nsteps} steps rounding={rounding}", flush=True)
    o = Oracle(p)
    o.step(nsteps)
    with Lattice(p) as lat:
        lat.step(nsteps)
        got = gpu_state(lat, STATE + POPS)
    want = {f: o.field(f) for f in STATE + POPS}
    assert compare_state(got, want, STATE) == []
    assert compare_interior(got, want, POPS) == []


@pytest.mark.parametrize("rounding", [0, 1])
def test_conductivity_ratio_is_bit_exact_on_both_fp64_paths(rounding):
    """The conductivity-ratio extension (a thermal relaxation rate per node
    from its phi after phase_field) on an FP64 handle: the fused kernel and
    the per-subroutine kernels against the oracle with the same ratio, bit
    for bit; the ratio does change the result."""
    p = Params(nx=67, ny=45, rounding=rounding, re=3.0, length=45.0)
    kr = 1.7
    print(f"k_ratio {kr}, 67x45, rounding={rounding}: fused 150 steps", flush=True)
    o = Oracle(p, k_ratio=kr)
    o.step(150)
    want = {f: o.field(f).copy() for f in STATE + POPS}
    with Lattice(p, k_ratio=kr) as lat:
        lat.step(150)
        got = gpu_state(lat, STATE + POPS)
    assert compare_state(got, want, STATE) == []
    assert compare_interior(got, want, POPS) == []
    o1 = Oracle(p)
    o1.step(150)
    assert not np.array_equal(o1.field("T"), want["T"])
    print(f"k_ratio {kr}: per-subroutine 40 steps", flush=True)
    o = Oracle(p, k_ratio=kr)
    o.step(40)
    with Lattice(p, k_ratio=kr) as lat:
        for _ in range(40):
            for s in Lattice.STAGES:
                lat.stage(s)
        got = gpu_state(lat, STATE + POPS)
    want = {f: o.field(f) for f in STATE + POPS}
    assert compare_state(got, want, STATE) == []
    assert compare_interior(got, want, POPS) == []


def test_any_split_of_steps_and_mixing_paths():
    p = Params(nx=40, ny=36, re=2.0, length=36.0)
    o = Oracle(p)
    o.step(40)
    with Lattice(p) as lat:
        for n in (1, 2, 3, 5, 7):
            lat.step(n)          # 18 fused steps in 5 calls
        for _ in range(4):       # 4 per-subroutine steps
            for s in Lattice.STAGES:
                lat.stage(s)
        lat.step(11)             # 11 fused
        for _ in range(7):
            lat.step(1)          # 7 single fused steps
        got = gpu_state(lat, STATE + POPS)
    want = {f: o.field(f) for f in STATE + POPS}
    assert compare_state(got, want, STATE) == []
    assert compare_interior(got, want, POPS) == []


def test_upload_download_roundtrip_all_fields():
    p = Params(nx=20, ny=14)
    rng = np.random.default_rng(7)
    with Lattice(p, initialize=False) as lat:
        for f in FIELD_NAMES:
            a = rng.standard_normal((p.ny + 2, p.nx + 2))
            lat.upload(f, a)
            assert np.array_equal(lat.download(f), a), f


EXT_ALL = dict(t_hot=284.5, k_ratio=1.6, t_cold=273.5, noise_amp=2e-3, noise_seed=77,
               t_crit=274.0, t_ref=278.0)


@pytest.mark.parametrize("rounding", [0, 1])
@pytest.mark.parametrize("ext_name", ["t_hot", "cold_noise", "t_crit", "t_ref", "all"])
def test_extensions_are_bit_exact_on_both_fp64_paths(rounding, ext_name):
    """Every physics extension on an FP64 population handle (hot-wall
    temperature, cold east wall, seeded initial noise, T_crit of the phase
    field, Boussinesq reference of the buoyancy, with the conductivity ratio
    in "all"): the fused kernel over 200 steps and the per-subroutine
    kernels over 30 steps against the oracle with the same extensions, bit
    for bit, and each extension changes the result."""
    ext = {"t_hot": dict(t_hot=284.5), "cold_noise": dict(t_cold=273.5, noise_amp=2e-3,
                                                          noise_seed=77),
           "t_crit": dict(t_crit=274.0), "t_ref": dict(t_ref=278.0), "all": EXT_ALL}[ext_name]
    p = Params(nx=67, ny=45, rounding=rounding, re=3.0, length=45.0, grav=1e-3)
    print(f"extensions {ext}, 67x45, rounding={rounding}: fused 200 steps", flush=True)
    o = Oracle(p, **ext)
    o.step(200)
    want = {f: o.field(f).copy() for f in STATE + POPS}
    with Lattice(p, **ext) as lat:
        lat.step(200)
        got = gpu_state(lat, STATE + POPS)
    assert compare_state(got, want, STATE) == []
    assert compare_interior(got, want, POPS) == []
    o1 = Oracle(p)
    o1.step(200)
    assert not np.array_equal(o1.field("T"), want["T"])
    print(f"extensions {ext_name}: per-subroutine 30 steps", flush=True)
    o = Oracle(p, **ext)
    o.step(30)
    with Lattice(p, **ext) as lat:
        for _ in range(30):
            for s in Lattice.STAGES:
                lat.stage(s)
        got = gpu_state(lat, STATE + POPS)
    want = {f: o.field(f) for f in STATE + POPS}
    assert compare_state(got, want, STATE) == []
    assert compare_interior(got, want, POPS) == []


def test_hot_wall_extension_reduces_to_the_reference_on_the_gpu():
    """The FP64 fused step with t_hot = 293 at T0 = 283, Stefan 0.125 and
    grav = 0 equals the reference run at T0' = 293, Stefan' = 0.25 bit for
    bit (tests/golden/thot_pin_32x24.npz; see the oracle's test of the same
    reduction)."""
    import json
    g = Golden("thot_pin_32x24")
    side = json.loads(str(g.d["oracle_side"]))
    p = g.params.replace(t0=side["t0"], stefan_number=side["stefan_number"])
    print("t_hot reduction 32x24, 300 steps", flush=True)
    with Lattice(p, t_hot=side["t_hot"]) as lat:
        lat.step(g.nsteps)
        got = gpu_state(lat, STATE + POPS)
    want = g.snap("final")
    assert compare_state(got, want, STATE) == []
    assert compare_interior(got, want, POPS) == []
