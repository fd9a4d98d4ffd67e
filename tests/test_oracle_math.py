"""Checks of the oracle that the mathematics fixes, independent of the
reference build: each would fail for a plausible slip in the restatement
(a dropped term, a wrong sign or index, a swapped operand).

 - equilibria: at rest f_k = w_k*rho exactly; zeroth and first moments of
   the equilibrium are rho and rho*u;
 - streaming is the pull permutation fp_k(x) = f_k(x - c_k) (checked with
   numpy rolls on random data);
 - regularized collision conserves density exactly in exact arithmetic:
   sum_k f_k' = rho to rounding in the clean policy, to FP32 rounding of
   the nine temporaries in the STEP_7 policy;
 - phase field: phi = +1 and phi = -1 everywhere are fixed points for any
   temperature; with T far below the melting temperature (tanh = -1) a
   solid node cannot melt;
 - forces: in a solid node (phi = -1) with tau_f = 1 and rho = 1 the drag
   removes the velocity exactly; with phi = +1 and grav = 0 nothing changes;
 - moments are the averages of the pre- and post-collision moments;
 - field_check / mass_check / diag sums equal numpy's;
 - the 32-bit Ny**3 of diag.f90 wraps to 0 at Ny = 2048;
 - melting is monotone in a conduction-only run (physics sanity);
 - extensions: the phase field depends on T - T_crit only; the buoyancy
   vanishes at T = t_ref and scales as T - t_ref.
"""
import numpy as np
import pytest

from bgk_pcm_amd.params import Params
from oracle_py import Oracle

W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])


def random_state(o, rng, scale=1e-3):
    """Fill populations with a perturbed equilibrium-like random state."""
    ny, nx = o.ny + 2, o.nx + 2
    rho = 1.0 + 0.02 * rng.standard_normal((ny, nx))
    T = 278.0 + 3.0 * rng.standard_normal((ny, nx))
    for k in range(9):
        o.field(f"f{k}")[...] = W[k] * rho * (1 + scale * rng.standard_normal((ny, nx)))
        o.field(f"g{k}")[...] = W[k] * T * (1 + scale * rng.standard_normal((ny, nx)))


def test_rest_equilibrium_is_exact():
    o = Oracle(Params(nx=30, ny=20, re=0.0))
    rho = o.field("rho")
    for k in range(9):
        assert np.array_equal(o.field(f"f{k}"), W[k] * rho * 1.0)
        assert np.array_equal(o.field(f"g{k}"), W[k] * o.field("T") * 1.0)


def test_equilibrium_moments():
    p = Params(nx=30, ny=20, re=40.0, length=20.0)  # u0 = Re*nu/L = 0.33
    o = Oracle(p)
    f = np.stack([o.field(f"f{k}") for k in range(9)])
    rho, u = o.field("rho"), o.field("u")
    inner = (slice(1, -1), slice(1, -1))
    assert np.allclose(f.sum(0)[inner], rho[inner], rtol=1e-15, atol=0)
    jx = np.tensordot(CX, f, axes=1)
    assert np.allclose(jx[inner], (rho * u)[inner], rtol=1e-14, atol=1e-17)
    assert u[5, 5] != 0.0


def test_streaming_is_the_pull_permutation():
    rng = np.random.default_rng(3)
    o = Oracle(Params(nx=17, ny=13), initialize=False)
    random_state(o, rng)
    o.stage("streaming")
    for k in range(9):
        src = o.field(f"f{k}")
        # value at (j, i) after the pull is the source at (j - cy, i - cx)
        rolled = np.roll(np.roll(src, CY[k], axis=0), CX[k], axis=1)
        assert np.array_equal(o.field(f"fp{k}")[1:-1, 1:-1], rolled[1:-1, 1:-1]), k


@pytest.mark.parametrize("rounding,tol", [(1, 2e-15), (0, 1e-7)])
def test_collision_conserves_density(rounding, tol):
    rng = np.random.default_rng(5)
    o = Oracle(Params(nx=24, ny=18, rounding=rounding), initialize=False)
    random_state(o, rng)
    o.field("phi")[...] = rng.uniform(-1, 1, o.field("phi").shape)
    for s in ("streaming", "moments_pre", "phase_field", "forces", "collision"):
        o.stage(s)
    f = np.stack([o.field(f"f{k}") for k in range(9)])
    g = np.stack([o.field(f"g{k}") for k in range(9)])
    inner = (slice(1, -1), slice(1, -1))
    assert np.allclose(f.sum(0)[inner], o.field("rho")[inner], rtol=tol, atol=0)
    # the temperature populations conserve the (latent-heat corrected) T
    assert np.allclose(g.sum(0)[inner], o.field("T")[inner], rtol=tol, atol=0)


@pytest.mark.parametrize("phase", [1.0, -1.0])
def test_uniform_phase_is_a_fixed_point(phase):
    rng = np.random.default_rng(11)
    o = Oracle(Params(nx=20, ny=20), initialize=False)
    o.field("phi")[...] = phase
    o.field("T")[...] = rng.uniform(200.0, 350.0, o.field("T").shape)
    o.stage("phase_field")
    assert np.all(o.field("phi") == phase)


def test_cold_solid_cannot_melt():
    o = Oracle(Params(nx=20, ny=20), initialize=False)
    rng = np.random.default_rng(2)
    phi0 = rng.choice([-1.0, 1.0], size=o.field("phi").shape)
    o.field("phi")[...] = phi0
    o.field("T")[...] = 0.0  # tanh((0-273)/10) == -1 in double
    o.stage("phase_field")
    phi = o.field("phi")[1:-1, 1:-1]
    solid = phi0[1:-1, 1:-1] < 0
    assert np.all(phi[solid] == -1.0)
    assert np.all(phi <= phi0[1:-1, 1:-1])


def test_drag_stops_a_solid_node_and_spares_a_liquid_one():
    p = Params(nx=10, ny=8, tau_f=1.0, grav=0.0)
    o = Oracle(p, initialize=False)
    rng = np.random.default_rng(4)
    o.field("rho")[...] = 1.0
    o.field("T")[...] = 280.0
    u0 = rng.standard_normal(o.field("u").shape) * 1e-3
    v0 = rng.standard_normal(o.field("v").shape) * 1e-3
    o.field("u")[...] = u0
    o.field("v")[...] = v0
    phi = np.full(o.field("phi").shape, -1.0)
    phi[:, 5:] = 1.0
    o.field("phi")[...] = phi
    o.field("phi_prec")[...] = phi
    o.stage("forces")
    u, v = o.field("u")[1:-1, 1:-1], o.field("v")[1:-1, 1:-1]
    solid = phi[1:-1, 1:-1] < 0
    assert np.all(u[solid] == 0.0) and np.all(v[solid] == 0.0)
    assert np.array_equal(u[~solid], u0[1:-1, 1:-1][~solid])
    assert np.array_equal(v[~solid], v0[1:-1, 1:-1][~solid])
    assert np.array_equal(o.field("T"), np.full_like(u0, 280.0))  # phi == phi_prec


def test_moments_average_pre_and_post_collision():
    rng = np.random.default_rng(8)
    o = Oracle(Params(nx=16, ny=12, rounding=1), initialize=False)
    random_state(o, rng)
    for s in ("streaming", "moments_pre"):
        o.stage(s)
    T2, u2 = o.field("T2").copy(), o.field("u2").copy()
    for s in ("phase_field", "forces", "collision", "moments"):
        o.stage(s)
    f = np.stack([o.field(f"f{k}") for k in range(9)])
    g = np.stack([o.field(f"g{k}") for k in range(9)])
    inner = (slice(1, -1), slice(1, -1))
    rho = f.sum(0)
    assert np.allclose(o.field("rho")[inner], rho[inner], rtol=1e-15)
    assert np.allclose(o.field("T")[inner], ((g.sum(0) + T2) / 2)[inner], rtol=1e-15)
    ux = np.tensordot(CX, f, axes=1) / rho
    assert np.allclose(o.field("u")[inner], ((ux + u2) / 2)[inner], rtol=1e-12, atol=1e-18)


def test_reductions_equal_numpy():
    p = Params(nx=40, ny=30)
    o = Oracle(p)
    o.step(300)
    inner = (slice(1, -1), slice(1, -1))
    rho, T, phi = (o.field(f)[inner] for f in ("rho", "T", "phi"))
    np.testing.assert_allclose(o.field_check(), [rho.mean(), T.mean(), phi.mean()], rtol=1e-14)
    mass = (p.nx * p.ny * p.rho0 - rho.sum()) * 100 / (p.nx * p.ny * p.rho0)
    got = o.mass_check()
    assert abs(got[0] - mass) < 1e-9
    np.testing.assert_allclose(got[1:], [rho.mean(), T.mean()], rtol=1e-14)
    s = o.diag_sums()
    Tf = o.field("T")
    dT = p.t0 - p.t_in_sol
    want0 = np.sum(p.ny * ((p.t0 - Tf[1:-1, 2]) / 2) / dT)
    want1 = np.sum(p.ny * np.abs((-3 * Tf[1:-1, 1] + 4 * Tf[1:-1, 2] - Tf[1:-1, 3]) / 2) / dT)
    np.testing.assert_allclose(s[:2], [want0, want1], rtol=1e-13)
    assert s[2] == np.count_nonzero(phi > 0)
    ii = np.arange(1, p.nx + 1)[None, :].repeat(p.ny, 0)
    assert s[3] == ii[phi == 0].sum()


def test_grashof_wraps_like_default_integer():
    o = Oracle(Params(nx=8, ny=2048), initialize=True)
    out, _ = o.diag(1000)
    assert out[8] == 0.0  # Gras: 2048**3 = 2**33 wraps to 0 in 32 bits
    o2 = Oracle(Params(nx=8, ny=1024), initialize=True)
    assert o2.diag(1000)[0][8] > 0.0


def test_melting_is_monotone_without_convection():
    p = Params(nx=64, ny=8, grav=0.0)
    o = Oracle(p)
    melted = []
    for _ in range(6):
        o.step(400)
        melted.append(np.count_nonzero(o.field("phi")[1:-1, 1:-1] > 0))
    assert all(b >= a for a, b in zip(melted, melted[1:]))
    assert melted[-1] > melted[0] > 0


@pytest.mark.parametrize("k_ratio", [1.5, 0.4])
def test_conductivity_ratio_scales_the_thermal_relaxation(k_ratio):
    """Extension (no reference counterpart): the regularized collision
    scales the non-equilibrium second moment of g by 1 - omega_g, and the
    thermal diffusivity is alpha = cs2 (tau_g - 1/2).  With the conductivity
    ratio, the solid (phi = -1) must relax with alpha_s = k_ratio alpha_l,
    the liquid (phi = +1) with alpha_l, and phi = 0 with their mean."""
    tau_g = 0.8
    rng = np.random.default_rng(3)
    o = Oracle(Params(nx=12, ny=9, tau_g=tau_g, rounding=1), initialize=False,
               k_ratio=k_ratio)
    shape = o.field("T").shape
    phi = rng.choice([-1.0, 0.0, 1.0], size=shape)
    T = rng.uniform(270.0, 285.0, shape)
    u = rng.uniform(-1e-3, 1e-3, shape)
    v = rng.uniform(-1e-3, 1e-3, shape)
    o.field("phi")[...] = phi
    o.field("T")[...] = T
    o.field("u")[...] = u
    o.field("v")[...] = v
    o.field("rho")[...] = 1.0
    uu = 1.5 * (u * u + v * v)
    geq = [W[k] * T * (1 + 3 * (CX[k] * u + CY[k] * v) + 4.5 * (CX[k] * u + CY[k] * v) ** 2 - uu)
           for k in range(9)]
    for k in range(9):
        o.field(f"gp{k}")[...] = geq[k] * (1 + rng.uniform(-1e-3, 1e-3, shape))
        o.field(f"fp{k}")[...] = W[k]
    o.stage("collision")
    pxx = lambda pops: sum((CX[k] ** 2 - 1 / 3) * (pops[k] - geq[k]) for k in range(9))
    before = pxx([o.field(f"gp{k}") for k in range(9)])
    after = pxx([o.field(f"g{k}") for k in range(9)])
    alpha_l = (tau_g - 0.5) / 3
    alpha = np.where(phi < 0, k_ratio * alpha_l,
                     np.where(phi > 0, alpha_l, 0.5 * (1 + k_ratio) * alpha_l))
    expect = 1 - 1 / (0.5 + 3 * alpha)
    inner = (slice(1, -1), slice(1, -1))
    np.testing.assert_allclose((after / before)[inner], expect[inner], rtol=1e-9)


def test_initial_noise_is_splitmix64():
    """Extension: the noise generator is SplitMix64 (known answer: from state
    0 its first output is 0xE220A8397B1DCDAF); node q = i + (Nx+2) j draws
    the output of state seed + (q+1)*0x9E3779B97F4A7C15."""
    first = 0xE220A8397B1DCDAF
    p = Params(nx=6, ny=5, t0=283.0, t_in_sol=273.0)
    base = Oracle(p)
    amp = 1e-3
    o = Oracle(p, noise_amp=amp, noise_seed=0)
    dT = o.field("T") - base.field("T")
    # node q = 0 is the corner ghost (no noise); with seed = -0x9E37..15 (mod
    # 2^64) node q = 1 draws the first output of state 0
    o2 = Oracle(p, noise_amp=amp, noise_seed=(-0x9E3779B97F4A7C15) % 2 ** 64)
    u1 = (first >> 11) * 2.0 ** -53
    d2 = o2.field("T") - base.field("T")
    # q = 1 is (i=1, j=0): a ghost row, no noise; use q = (Nx+2)*1 + 1 instead
    q = (p.nx + 2) * 1 + 1
    o3 = Oracle(p, noise_amp=amp, noise_seed=(-(q) * 0x9E3779B97F4A7C15) % 2 ** 64)
    d3 = o3.field("T") - base.field("T")
    # T = 273 + noise, read back as a difference: exact to the rounding of 273 + x
    assert abs(d3[1, 1] - amp * (p.t0 - p.t_in_sol) * (2 * u1 - 1)) <= 1e-13
    inner = dT[1:-1, 1:-1]
    assert np.all(np.abs(inner) <= amp * 10.0) and np.abs(inner).max() > 0
    assert np.all(dT[0, :] == 0) and np.all(dT[:, 0] == 0) and np.all(d2[0, :] == 0)


def test_cold_east_wall_sets_the_east_ghost_equilibrium():
    """Extension: with a cold east wall, BC_g's east ghosts are equilibria
    at t_cold with the halo velocities (not at the averaged T(Nx, j))."""
    p = Params(nx=10, ny=8, re=2.0, length=10.0)
    o = Oracle(p, t_cold=270.0)
    o.step(5)
    o.stage("bc_f")
    o.stage("bc_g")
    u, v = o.field("u")[1:-1, -1], o.field("v")[1:-1, -1]
    uu = 1.5 * (u * u + v * v)
    for k in (3, 6, 7):
        cu = CX[k] * u + CY[k] * v
        eq = W[k] * 270.0 * (1 + 3 * cu + 4.5 * cu * cu - uu)
        np.testing.assert_allclose(o.field(f"g{k}")[1:-1, -1], eq, rtol=1e-15)


def test_critical_temperature_shifts_the_phase_field():
    """The phase field depends on T - T_crit only: T_crit = 283 on T equals the
    default 273 on T - 10 bit for bit (temperatures on a 1/64 grid, so the
    shift is exact)."""
    rng = np.random.default_rng(21)
    shape = (22, 22)
    T = 263.0 + rng.integers(0, 64 * 40, size=shape) / 64.0
    phi0 = rng.uniform(-1.0, 1.0, size=shape)
    out = []
    for t_crit, shift in ((283.0, 0.0), (273.0, 10.0)):
        o = Oracle(Params(nx=20, ny=20), initialize=False, t_crit=t_crit)
        o.field("phi")[...] = phi0
        o.field("T")[...] = T - shift
        o.stage("phase_field")
        out.append(o.field("phi").copy())
    assert np.array_equal(out[0], out[1])
    o = Oracle(Params(nx=20, ny=20), initialize=False)
    o.field("phi")[...] = phi0
    o.field("T")[...] = T
    o.stage("phase_field")
    assert not np.array_equal(o.field("phi"), out[0])  # the shift matters


def test_buoyancy_reference_temperature():
    """forces with t_ref: a liquid node at rest at T = t_ref feels no
    buoyancy, and the velocity increment scales as T - t_ref."""
    p = Params(nx=10, ny=8, tau_f=1.0, grav=1e-3)

    def dv(T, t_ref):
        o = Oracle(p, initialize=False, t_ref=t_ref)
        o.field("rho")[...] = 1.0
        o.field("T")[...] = T
        o.field("u")[...] = 0.0
        o.field("v")[...] = 0.0
        o.field("phi")[...] = 1.0
        o.field("phi_prec")[...] = 1.0
        o.stage("forces")
        return o.field("v")[1:-1, 1:-1].copy()

    assert np.all(dv(278.0, 278.0) == 0.0)
    a, b = dv(283.0, 278.0), dv(283.0, 273.0)
    assert np.all(a > 0.0)
    np.testing.assert_allclose(a / b, (283.0 - 278.0) / (283.0 - 273.0), rtol=1e-14)
    np.testing.assert_allclose(dv(283.0, 0.0) / b, 283.0 / 10.0, rtol=1e-14)
