"""FP32 moment storage (bgk_params.precision = BGK_FP32_MOM) against the FP64
oracle.

The moment kernel rebuilds every pulled population from the upstream node's
(rho, T) in FP64 and (u, v, Pf, Pg) in FP32, so it agrees with the clean
FP64 policy up to the FP32 rounding of u, v and the non-equilibrium moments.
Error measure: relative max-norm max|a - b| / max|b| per field, with the
liquid fraction (1+phi)/2 in place of phi; on the small grids u and v are
both scaled by the largest velocity component (in a 1 x 1 cavity u is a
rounding residue ~1e-9 beside v ~1e-5).  Tolerances: 1e-6 on the whole
array (halo included) after 60 steps on small and ragged grids, where FP32
storage errors are ~1e-8; 1e-5 after 4000 steps on 64 x 64 against both
rounding policies (the bound the FP32 population storage is held to).
"""
import numpy as np
import pytest

from bgk_pcm_amd.lattice import BgkError, Lattice
from bgk_pcm_amd.params import FP32_DEV, FP32_MOM, FP64, FP64_MOM, ROUND_CLEAN, Params
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "u", "v", "T", "phi")


def rel_errors(lat, o, interior=True, speed_scale=False):
    out = {}
    vmax = max(np.abs(o.field("u")).max(), np.abs(o.field("v")).max())
    for f in FIELDS:
        a = lat.download(f)
        b = o.field(f)
        if interior:
            a, b = a[1:-1, 1:-1], b[1:-1, 1:-1]
        if f == "phi":
            a, b = (1.0 + a) / 2.0, (1.0 + b) / 2.0
        scale = vmax if (speed_scale and f in ("u", "v")) else np.abs(b).max()
        out[f] = float(np.abs(a - b).max() / scale) if scale > 0 else float(np.abs(a).max())
    return out


@pytest.fixture(scope="module")
def oracles():
    res = {}
    for tau_f, roundings in ((1.0, (0, 1)), (0.8, (1,))):
        base = Params(nx=64, ny=64, length=64.0, tau_f=tau_f)
        for rounding in roundings:
            o = Oracle(base.replace(rounding=rounding))
            o.step(4000)
            res[tau_f, rounding] = (base, o)
    return res


@pytest.mark.parametrize("tau_f,rounding", [(1.0, 0), (1.0, 1), (0.8, 1)])
def test_mom_storage_error_bound(oracles, rounding, tau_f):
    """tau_f = 1 stores no flow non-equilibrium moments (the collision
    returns the equilibrium); tau_f = 0.8 exercises the general layout.  At
    tau_f = 0.8 the STEP_7 policy itself is 2.3e-5 (u, v) away from clean
    FP64 after these 4000 steps (its FP32 ft no longer cancel as they do for
    omega_f = 1), so the bound is checked against clean FP64 there."""
    base, o = oracles[tau_f, rounding]
    print(f"moment storage, 64x64, 4000 steps, tau_f={tau_f} vs FP64 rounding={rounding}",
          flush=True)
    with Lattice(base.replace(precision=FP32_MOM, rounding=rounding)) as lat:
        lat.step(4000)
        err = rel_errors(lat, o)
    print("relative max-norm errors:", err)
    assert all(np.isfinite(v) for v in err.values())
    assert max(err.values()) <= 1e-5, err


SHAPES = [(1, 1), (2, 3), (3, 2), (37, 29), (130, 5), (129, 17), (300, 9), (530, 6)]


@pytest.mark.parametrize("nx,ny", SHAPES)
@pytest.mark.parametrize("rows,variant", [(0, 0), (1, 1), (2, 2), (7, 3), (3, 4), (5, 5),
                                          (2, 6), (4, 7), (1, 8)])
@pytest.mark.parametrize("tau_f", [1.0, 0.8])
def test_mom_shapes_and_strips(monkeypatch, nx, ny, rows, variant, tau_f):
    """Every grid shape, strip height (BGK_MOM_ROWS) and tiling variant
    (BGK_MOM_VARIANT: workgroup widths 64..512, 3- and 4-row rings; 0 =
    defaults) gives the oracle's fields, halo included, to FP32-storage
    precision; the run is split over two bgk_step calls."""
    monkeypatch.setenv("BGK_MOM_ROWS", str(rows))
    monkeypatch.setenv("BGK_MOM_VARIANT", str(variant))
    p = Params(nx=nx, ny=ny, length=float(max(nx, ny)), grav=1e-3, rounding=ROUND_CLEAN,
               tau_f=tau_f)
    print(f"moment storage {nx}x{ny} rows={rows} variant={variant} tau_f={tau_f}", flush=True)
    o = Oracle(p)
    o.step(60)
    with Lattice(p.replace(precision=FP32_MOM)) as lat:
        lat.step(25)
        lat.step(35)
        err = rel_errors(lat, o, interior=False, speed_scale=True)
    print("relative max-norm errors:", err)
    assert max(err.values()) <= 1e-6, err


def test_mom_initial_state_is_the_oracles():
    p = Params(nx=30, ny=20, re=4.0, length=20.0)
    o = Oracle(p)
    with Lattice(p.replace(precision=FP32_MOM)) as lat:
        for f in FIELDS:
            assert np.array_equal(lat.download(f), o.field(f)), f


def test_mom_reductions_follow_the_fields():
    p = Params(nx=48, ny=40, length=48.0, rounding=ROUND_CLEAN)
    o = Oracle(p)
    o.step(300)
    with Lattice(p.replace(precision=FP32_MOM)) as lat:
        lat.step(300)
        fc, fco = lat.field_check(), o.field_check()
        np.testing.assert_allclose(fc, fco, rtol=1e-9)
        np.testing.assert_allclose(lat.diag_sums(), o.diag_sums(), rtol=1e-6)


def test_mom_handle_rejects_population_access():
    with Lattice(Params(nx=8, ny=8, precision=FP32_MOM)) as lat:
        for name in ("f3", "g0", "fp1", "rho2"):
            with pytest.raises(BgkError) as e:
                lat.download(name)
            assert e.value.code == -4, name
        with pytest.raises(BgkError) as e:
            lat.upload("rho", np.ones((10, 10)))
        assert e.value.code == -4
        with pytest.raises(BgkError) as e:
            lat.streaming()
        assert e.value.code == -4
    with pytest.raises(BgkError) as e:
        Lattice(Params(nx=8, ny=8, precision=FP32_MOM, north_bc=0, south_bc=0))
    assert e.value.code == -4


# ---- FP64 moment storage (the adjoint's forward model) and the extensions

@pytest.mark.parametrize("nx,ny", [(1, 1), (3, 2), (37, 29), (130, 5), (300, 9)])
@pytest.mark.parametrize("tau_f", [1.0, 0.8])
def test_fp64_mom_matches_clean_oracle(nx, ny, tau_f):
    """FP64 moments: the clean FP64 step in moment form; only the order of
    the FP64 roundings differs from the population form of the oracle."""
    p = Params(nx=nx, ny=ny, length=float(max(nx, ny)), grav=1e-3, rounding=ROUND_CLEAN,
               tau_f=tau_f)
    print(f"FP64 moments {nx}x{ny} tau_f={tau_f}", flush=True)
    o = Oracle(p)
    o.step(60)
    with Lattice(p.replace(precision=FP64_MOM)) as lat:
        lat.step(60)
        err = rel_errors(lat, o, interior=False, speed_scale=True)
    print("relative max-norm errors:", err)
    assert max(err.values()) <= 1e-11, err


@pytest.mark.parametrize("precision,tol", [(FP64_MOM, 1e-10), (FP32_MOM, 1e-6)])
@pytest.mark.parametrize("t_hot,k_ratio", [(283.5, 1.0), (283.0, 1.5), (285.0, 0.5)])
def test_extensions_match_oracle(precision, tol, t_hot, k_ratio):
    """Hot-wall temperature and conductivity ratio against the oracle's
    restatement of the same extension, 300 steps on 48 x 40."""
    p = Params(nx=48, ny=40, length=48.0, grav=1e-3, rounding=ROUND_CLEAN)
    o = Oracle(p, t_hot=t_hot, k_ratio=k_ratio)
    o.step(300)
    with Lattice(p.replace(precision=precision), t_hot=t_hot, k_ratio=k_ratio) as lat:
        lat.step(300)
        err = rel_errors(lat, o, interior=False, speed_scale=True)
    print(f"extensions t_hot={t_hot} k_ratio={k_ratio}: {err}")
    assert max(err.values()) <= tol, err


@pytest.mark.parametrize("precision,tol", [(FP64_MOM, 1e-10), (FP32_MOM, 1e-6)])
def test_conductivity_ratio_general_layout_matches_oracle(precision, tol):
    """The conductivity ratio with tau_f = 0.8 (the 48-byte layout: flow
    non-equilibrium moments stored), 300 steps on 48 x 40."""
    p = Params(nx=48, ny=40, length=48.0, grav=1e-3, rounding=ROUND_CLEAN, tau_f=0.8)
    o = Oracle(p, k_ratio=2.0)
    o.step(300)
    with Lattice(p.replace(precision=precision), k_ratio=2.0) as lat:
        lat.step(300)
        err = rel_errors(lat, o, interior=False, speed_scale=True)
    print(f"k_ratio 2, tau_f 0.8: {err}")
    assert max(err.values()) <= tol, err


@pytest.mark.parametrize("precision,tol", [(FP64_MOM, 1e-10), (FP32_MOM, 1e-6)])
@pytest.mark.parametrize("t_crit,t_ref", [(275.0, 0.0), (273.0, 278.0), (271.5, 273.0)])
def test_critical_and_reference_temperatures_match_oracle(precision, tol, t_crit, t_ref):
    """The phase field's T_crit and the buoyancy's reference temperature
    (Boussinesq form) against the oracle, 300 steps on 48 x 40."""
    p = Params(nx=48, ny=40, length=48.0, grav=1e-3, rounding=ROUND_CLEAN)
    o = Oracle(p, t_crit=t_crit, t_ref=t_ref)
    o.step(300)
    with Lattice(p.replace(precision=precision), t_crit=t_crit, t_ref=t_ref) as lat:
        lat.step(300)
        err = rel_errors(lat, o, interior=False, speed_scale=True)
    print(f"t_crit={t_crit} t_ref={t_ref}: {err}")
    assert max(err.values()) <= tol, err


def test_extensions_are_accepted_by_every_handle_and_validated():
    """Every handle carries every extension (the FP64 and FP32 population
    kernels included); out-of-range values are refused with BGK_E_INVALID."""
    for prec in (FP64, FP32_DEV, FP32_MOM, FP64_MOM):
        with Lattice(Params(nx=8, ny=8, precision=prec)) as lat:
            lat.set_extensions(284.0, 2.0, t_cold=273.0, noise_amp=1e-3, noise_seed=3,
                               t_crit=275.0, t_ref=278.0)
            for bad in (dict(t_hot=284.0, k_ratio=0.0), dict(t_ref=-1.0), dict(t_crit=-2.0),
                        dict(noise_amp=-1e-3), dict(t_cold=-5.0)):
                with pytest.raises(BgkError) as e:
                    lat.set_extensions(**bad)
                assert e.value.code == -1, (prec, bad)


@pytest.mark.parametrize("precision,tol", [(FP64_MOM, 1e-10), (FP32_MOM, 1e-6)])
def test_cold_wall_and_noise_match_oracle(precision, tol):
    """Config 1's literal physics: cold east wall, seeded initial noise
    (amplitude 1e-3 in theta units), conductivity ratio 1."""
    p = Params(nx=48, ny=40, length=48.0, grav=1e-3, rounding=ROUND_CLEAN, tau_f=0.8)
    kw = dict(t_cold=273.0, noise_amp=1e-3, noise_seed=12345)
    o = Oracle(p, **kw)
    with Lattice(p.replace(precision=precision), **kw) as lat:
        for f in ("T", "rho", "phi"):
            assert np.array_equal(lat.download(f), o.field(f)), f
        o.step(300)
        lat.step(300)
        err = rel_errors(lat, o, interior=False, speed_scale=True)
    print(f"cold wall + noise: {err}")
    assert max(err.values()) <= tol, err
