"""FP32 population storage (f - w*rho0, g - w*T_crit in FP32) against the FP64
oracle.

Error measure: relative max-norm max|a - b| / max|b| over the interior, per
field, with the liquid fraction (1+phi)/2 in place of phi.  Tolerance: 1e-5
for the default form (FP64 arithmetic on FP32 storage), against both the
STEP_7 rounding policy and the clean FP64 one; 1e-4 for the form with FP32
collision arithmetic (bgk_params.fp32_collision = 1).
"""
import numpy as np
import pytest

from bgk_pcm_amd.lattice import BgkError, Lattice
from bgk_pcm_amd.params import FP32_DEV, Params
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "u", "v", "T", "phi")


def rel_errors(lat, o):
    out = {}
    for f in FIELDS:
        a = lat.download(f)[1:-1, 1:-1]
        b = o.field(f)[1:-1, 1:-1]
        if f == "phi":
            a, b = (1.0 + a) / 2.0, (1.0 + b) / 2.0
        out[f] = float(np.abs(a - b).max() / np.abs(b).max())
    return out


@pytest.fixture(scope="module")
def oracles():
    base = Params(nx=64, ny=64, length=64.0)
    res = {}
    for rounding in (0, 1):
        o = Oracle(base.replace(rounding=rounding))
        o.step(4000)
        res[rounding] = o
    return base, res


@pytest.mark.parametrize("rounding", [0, 1])
def test_fp32_storage_error_bound(oracles, rounding):
    base, res = oracles
    print(f"fp32 storage, 64x64, 4000 steps vs FP64 rounding={rounding}", flush=True)
    with Lattice(base.replace(precision=FP32_DEV)) as lat:
        lat.step(4000)
        err = rel_errors(lat, res[rounding])
    print("relative max-norm errors:", err)
    assert all(np.isfinite(v) for v in err.values())
    assert max(err.values()) <= 1e-5, err


@pytest.mark.parametrize("rounding", [0, 1])
def test_fp32_storage_with_conductivity_ratio(rounding):
    """The conductivity-ratio extension on FP32 population storage, against
    the oracle with the same ratio (64x64, 4000 steps, bound 1e-5)."""
    base = Params(nx=64, ny=64, length=64.0, rounding=rounding)
    print(f"fp32 storage, k_ratio 2, 64x64, 4000 steps vs FP64 rounding={rounding}", flush=True)
    o = Oracle(base, k_ratio=2.0)
    o.step(4000)
    with Lattice(base.replace(precision=FP32_DEV), k_ratio=2.0) as lat:
        lat.step(4000)
        err = rel_errors(lat, o)
    print("relative max-norm errors:", err)
    assert max(err.values()) <= 1e-5, err


def test_fp32_collision_arithmetic_error_bound(oracles):
    base, res = oracles
    with Lattice(base.replace(precision=FP32_DEV, fp32_collision=1)) as lat:
        lat.step(4000)
        err = rel_errors(lat, res[1])
    print("relative max-norm errors (FP32 collision):", err)
    assert max(err.values()) <= 1e-4, err


def test_fp32_initial_populations():
    p = Params(nx=30, ny=20, re=4.0, length=20.0)
    o = Oracle(p)
    with Lattice(p.replace(precision=FP32_DEV)) as lat:
        for k in range(9):
            for name, m0 in ((f"f{k}", p.rho0), (f"g{k}", 273.0)):
                a, b = lat.download(name), o.field(name)
                # only the FP32 rounding of the deviation from w*m0
                w = (4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36)[k]
                dev = np.abs(b - w * m0)
                assert np.all(np.abs(a - b) <= dev * 2.0 ** -24 + 1e-15 * w * m0), name
        for f in FIELDS:
            assert np.array_equal(lat.download(f), o.field(f)), f


def test_fp32_handle_rejects_per_subroutine_calls_and_other_walls():
    with Lattice(Params(nx=8, ny=8, precision=FP32_DEV)) as lat:
        with pytest.raises(BgkError) as e:
            lat.streaming()
        assert e.value.code == -4
        with pytest.raises(BgkError) as e:
            lat.download("fp3")
        assert e.value.code == -4
    with pytest.raises(BgkError) as e:
        Lattice(Params(nx=8, ny=8, precision=FP32_DEV, west_bc=0, east_bc=0))
    assert e.value.code == -4


def test_row_reference_density_rebase(monkeypatch):
    """The FP32 f members are stored against a reference density per row,
    reset to the row's mean density every BGK_REBASE_EVERY steps (a tuning
    knob read at bgk_create).  Populations downloaded after re-basings
    decode against the current table: their sum is the density field; and
    the run re-based every 7 steps and the one never re-based both stay
    within the 1e-5 bound of the FP64 run."""
    p = Params(nx=48, ny=40, length=40.0, grav=2e-3)
    o = Oracle(p.replace(rounding=1))
    o.step(1500)
    for every in ("7", "0"):
        monkeypatch.setenv("BGK_REBASE_EVERY", every)
        print(f"fp32 storage, re-basing every {every} steps, 48x40, 1500 steps", flush=True)
        with Lattice(p.replace(precision=FP32_DEV, rounding=1)) as lat:
            lat.step(1500)
            err = rel_errors(lat, o)
            f = sum(lat.download(f"f{k}") for k in range(9))[1:-1, 1:-1]
            rho = lat.download("rho")[1:-1, 1:-1]
        print("relative max-norm errors:", err)
        assert max(err.values()) <= 1e-5, err
        # rho is moments.f90's sum of the stored f (its FP64 rounding apart)
        assert np.abs(f - rho).max() <= 1e-13, np.abs(f - rho).max()
