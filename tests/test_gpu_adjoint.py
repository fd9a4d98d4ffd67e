"""The GPU adjoint (bgk_adj_*) against torch.autograd on the CPU reference
(tests/torch_ref.py, pinned to the oracle) and against central finite
differences of the GPU forward.

Loss: mean squared error of T at seeded-random interior probes after every
step against a target trajectory of the same model at other parameters,
gradient with respect to theta = (stefan_number, k_ratio, t_hot).
Tolerances: probe record vs the CPU reference 1e-11 relative (FP64 moment
form vs population form: only the order of roundings differs); gradient
vs autograd 1e-7 relative per component; gradient vs central differences
(steps 1e-6 St, 1e-6 k_ratio, 1e-4 t_hot) 1e-4 relative, the bar the
adjoint's gradient is held to.
"""
import numpy as np
import pytest
import torch

from bgk_pcm_amd.adjoint import probe_temperature, seeded_probes
from bgk_pcm_amd.lattice import BgkError, Lattice
from bgk_pcm_amd.params import FP32_DEV, FP32_MOM, FP64_MOM, ROUND_CLEAN, Params
from torch_ref import probe_record

pytestmark = pytest.mark.gpu


def gpu_record(p, theta, pi, pj, nsteps, seg):
    st, kr, th = theta
    with Lattice(p.replace(precision=FP64_MOM, rounding=ROUND_CLEAN, stefan_number=st),
                 initialize=False, t_hot=th, k_ratio=kr) as lat:
        lat.initialization()
        lat.adj_setup(nsteps, seg, pi, pj)
        return lat.adj_forward()


@pytest.mark.parametrize("tau_f,seg,k_ratio", [(1.0, 7, 1.5), (0.8, 50, 1.5), (1.0, 1, 1.5),
                                               (1.0, 16, 1.0)])
def test_adjoint_equals_autograd_of_the_cpu_reference(tau_f, seg, k_ratio):
    """k_ratio = 1 is the reference's physics: the forward takes the constant
    omega_g there, the gradient in k_ratio must not vanish."""
    p = Params(nx=30, ny=24, length=30.0, grav=2e-3, rounding=ROUND_CLEAN, tau_f=tau_f,
               stefan_number=0.1)
    nsteps = 120
    pi, pj = seeded_probes(p.nx, p.ny, 6, seed=4)
    # probes on the boundary-adjacent rows and columns too
    pi += [1, p.nx, 5]
    pj += [3, p.ny, 1]
    theta0 = (0.1, k_ratio, 283.5)
    target = probe_record(p, torch.tensor([0.12, 1.2, 283.2], dtype=torch.float64), pi, pj,
                          nsteps).numpy()
    print(f"adjoint vs autograd: {p.nx}x{p.ny}, {nsteps} steps, seg_len {seg}, tau_f {tau_f}, "
          f"k_ratio {k_ratio}", flush=True)
    # CPU reference gradient
    th = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    rec_ref = probe_record(p, th, pi, pj, nsteps)
    loss = ((rec_ref - torch.from_numpy(target)) ** 2).mean()
    loss.backward()
    g_ref = th.grad.numpy()
    # GPU record and adjoint gradient
    with Lattice(p.replace(precision=FP64_MOM, stefan_number=theta0[0]), initialize=False,
                 t_hot=theta0[2], k_ratio=theta0[1]) as lat:
        lat.initialization()
        lat.adj_setup(nsteps, seg, pi, pj)
        rec = lat.adj_forward()
        seeds = 2.0 * (rec - target) / rec.size
        g = lat.adj_backward(seeds)
    r = rec_ref.detach().numpy()
    assert np.abs(rec - r).max() <= 1e-11 * np.abs(r).max()
    print("gradient GPU", g, "autograd", g_ref)
    np.testing.assert_allclose(g, g_ref, rtol=1e-7, atol=0)


@pytest.mark.parametrize("nx,ny,re,seg", [(30, 24, 3.0, 9), (150, 20, 0.0, 11),
                                           (200, 12, 2.0, 25)],
                         ids=["re3", "three_tile_columns", "re2_four_tile_columns"])
def test_adjoint_with_inflow_velocity_and_tile_seams(nx, ny, re, seg):
    """u0 = Re nu / length != 0 puts constant ghost velocities into the
    thermal boundary equilibria (BC_g.f90:9-31) that the recomputed steps
    must take from the recorded run; Nx > 128 puts two or three seams of the
    adjoint kernel's 64-column tiles between probes (columns 64/65, 128/129,
    192/193 are probed), against autograd of the CPU reference."""
    p = Params(nx=nx, ny=ny, length=float(ny), grav=2e-3, rounding=ROUND_CLEAN, re=re,
               stefan_number=0.1)
    nsteps = 80
    # probes near the hot wall and the front (the Stefan number and t_hot act
    # there) and on both sides of every tile seam (the conductivity ratio
    # acts in the whole solid, so its gradient carries the seam adjoints)
    pi, pj = [2, 4, 3], [5, 7, 9]
    for col in (64, 65, 128, 129, 192, 193):
        if col <= nx and (col, 2 + col % 3) not in zip(pi, pj):
            pi.append(col)
            pj.append(2 + col % 3)
    theta0 = (0.1, 1.3, 283.5)
    target = probe_record(p, torch.tensor([0.12, 1.2, 283.2], dtype=torch.float64), pi, pj,
                          nsteps).numpy()
    print(f"adjoint vs autograd: {nx}x{ny}, Re {re}, {nsteps} steps, seg_len {seg}", flush=True)
    th = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    rec_ref = probe_record(p, th, pi, pj, nsteps)
    ((rec_ref - torch.from_numpy(target)) ** 2).mean().backward()
    g_ref = th.grad.numpy()
    with Lattice(p.replace(precision=FP64_MOM, stefan_number=theta0[0]), initialize=False,
                 t_hot=theta0[2], k_ratio=theta0[1]) as lat:
        lat.initialization()
        lat.adj_setup(nsteps, seg, pi, pj)
        rec = lat.adj_forward()
        g = lat.adj_backward(2.0 * (rec - target) / rec.size)
    r = rec_ref.detach().numpy()
    assert np.abs(rec - r).max() <= 1e-11 * np.abs(r).max()
    print("gradient GPU", g, "autograd", g_ref)
    np.testing.assert_allclose(g, g_ref, rtol=1e-7, atol=0)


def test_adjoint_misuse_is_reported():
    """A backward pass needs a recorded forward that matches the current
    state and constants; repeated probe nodes are refused."""
    p = Params(nx=20, ny=16, length=16.0, precision=FP64_MOM, rounding=ROUND_CLEAN)
    with Lattice(p) as lat:
        lat.adj_setup(10, 5, [3, 4], [5, 6])
        with pytest.raises(BgkError) as e:
            lat.adj_backward(np.zeros((10, 2)))
        assert e.value.code == -1 and "bgk_adj_forward" in str(e.value)
        lat.adj_forward()
        lat.step(1)  # the state moved on: the record no longer matches it
        with pytest.raises(BgkError) as e:
            lat.adj_backward(np.zeros((10, 2)))
        assert e.value.code == -1
        with pytest.raises(BgkError) as e:
            lat.adj_setup(10, 5, [3, 4, 3], [5, 6, 5])
        assert e.value.code == -1 and "share" in str(e.value)


def test_adjoint_matches_central_differences():
    p = Params(nx=64, ny=64, length=64.0, grav=1e-3, rounding=ROUND_CLEAN, stefan_number=0.1)
    nsteps, seg = 400, 40
    pi, pj = seeded_probes(p.nx, p.ny, 16, seed=1)
    theta0 = np.array([0.1, 1.5, 283.5])
    target = gpu_record(p, (0.12, 1.2, 283.2), pi, pj, nsteps, seg)
    loss = lambda rec: float(((rec - target) ** 2).mean())
    with Lattice(p.replace(precision=FP64_MOM, stefan_number=theta0[0]), initialize=False,
                 t_hot=theta0[2], k_ratio=theta0[1]) as lat:
        lat.initialization()
        lat.adj_setup(nsteps, seg, pi, pj)
        rec = lat.adj_forward()
        g = lat.adj_backward(2.0 * (rec - target) / rec.size)
    print("adjoint gradient", g, flush=True)
    for c, h in enumerate((1e-6, 1e-6, 1e-4)):
        tp, tm = theta0.copy(), theta0.copy()
        tp[c] += h
        tm[c] -= h
        fd = (loss(gpu_record(p, tp, pi, pj, nsteps, seg)) -
              loss(gpu_record(p, tm, pi, pj, nsteps, seg))) / (2 * h)
        print(f"component {c}: adjoint {g[c]:.10e} central difference {fd:.10e}")
        assert abs(fd - g[c]) <= 1e-4 * abs(g[c]), (c, fd, g[c])


@pytest.mark.parametrize("n_states", [-1, 5], ids=["segments", "binomial"])
def test_autograd_function_gives_the_adjoint_gradient(n_states):
    p = Params(nx=24, ny=20, length=24.0, grav=2e-3, rounding=ROUND_CLEAN)
    pi, pj = seeded_probes(p.nx, p.ny, 4, seed=9)
    nsteps = 60
    target = probe_record(p, torch.tensor([0.12, 1.2, 283.2], dtype=torch.float64), pi, pj,
                          nsteps)
    theta = torch.tensor([0.1, 1.5, 283.5], dtype=torch.float64, requires_grad=True)
    T = probe_temperature(theta, p, pi, pj, nsteps, seg_len=16, n_states=n_states)
    loss = ((T - target) ** 2).mean()
    loss.backward()
    th = torch.tensor([0.1, 1.5, 283.5], dtype=torch.float64, requires_grad=True)
    ((probe_record(p, th, pi, pj, nsteps) - target) ** 2).mean().backward()
    np.testing.assert_allclose(theta.grad.numpy(), th.grad.numpy(), rtol=1e-7)


def test_adjoint_needs_a_moment_handle():
    with Lattice(Params(nx=8, ny=8, precision=FP32_DEV)) as lat:
        with pytest.raises(BgkError) as e:
            lat.adj_setup(10, 5, [2], [3])
        assert e.value.code == -4
    with Lattice(Params(nx=8, ny=8, precision=FP64_MOM)) as lat:
        with pytest.raises(BgkError) as e:
            lat.adj_setup(10, 5, [0], [3])
        assert e.value.code == -1
        with pytest.raises(BgkError) as e:  # outer segment shorter than the inner one
            lat.adj_setup(10, 5, [2], [3], seg_outer=3)
        assert e.value.code == -1
    with Lattice(Params(nx=8, ny=8, precision=FP32_MOM)) as lat:
        lat.adj_setup(10, 5, [2], [3])


def adjoint_gradient(p, prec, nsteps, seg, seg_outer, pi, pj, target, theta=(0.1, 1.5, 283.5)):
    st, kr, th = theta
    with Lattice(p.replace(precision=prec, stefan_number=st), initialize=False, t_hot=th,
                 k_ratio=kr) as lat:
        lat.initialization()
        lat.adj_setup(nsteps, seg, pi, pj, seg_outer=seg_outer)
        info = lat.adj_info()
        rec = lat.adj_forward()
        return lat.adj_backward(2.0 * (rec - target) / rec.size), rec, info


@pytest.mark.parametrize("prec", [FP64_MOM, FP32_MOM])
@pytest.mark.parametrize("tau_f", [1.0, 0.8])
def test_two_level_checkpoints_give_the_same_gradient(prec, tau_f):
    """Two levels recompute the same states: the gradient is bit-identical to
    the one-level run, with nsteps/seg_outer + seg_outer/seg + seg + 1 - 1
    states stored (the first inner checkpoint is the outer one)."""
    p = Params(nx=40, ny=32, length=40.0, grav=2e-3, rounding=ROUND_CLEAN, tau_f=tau_f)
    nsteps = 230
    pi, pj = seeded_probes(p.nx, p.ny, 6, seed=5)
    target = gpu_record(p, (0.12, 1.2, 283.2), pi, pj, nsteps, 10)
    g1, r1, i1 = adjoint_gradient(p, prec, nsteps, 7, 0, pi, pj, target)
    g2, r2, i2 = adjoint_gradient(p, prec, nsteps, 7, 40, pi, pj, target)
    print("one level", g1, i1, "two levels", g2, i2, flush=True)
    assert np.array_equal(r1, r2)
    assert np.array_equal(g1, g2)
    assert i1[1] == (nsteps + 6) // 7 + 7 + 1
    assert i2[1] == (nsteps + 39) // 40 + (40 + 6) // 7 - 1 + 7 + 1
    assert i2[0] < i1[0]


def test_fp32_moment_adjoint_matches_the_fp64_adjoint():
    """The adjoint over FP32-stored states against the FP64 one on the same
    problem.  Tolerance 1e-4 relative per component (the gradient bar of the
    FD check); the FP32 storage's own forward error is ~1e-6."""
    p = Params(nx=96, ny=96, length=96.0, grav=1e-3, rounding=ROUND_CLEAN)
    nsteps = 600
    pi, pj = seeded_probes(p.nx, p.ny, 16, seed=3)
    target = gpu_record(p, (0.12, 1.2, 283.2), pi, pj, nsteps, 50)
    g64, r64, _ = adjoint_gradient(p, FP64_MOM, nsteps, 25, 150, pi, pj, target)
    g32, r32, _ = adjoint_gradient(p, FP32_MOM, nsteps, 25, 150, pi, pj, target)
    rel = np.abs(g32 - g64) / np.abs(g64)
    print("FP64 adjoint", g64, "FP32 adjoint", g32, "relative difference", rel,
          "record difference", np.abs(r32 - r64).max(), flush=True)
    assert np.all(rel <= 1e-4), rel


def test_adjoint_with_cold_wall_and_noise_matches_central_differences():
    """With the cold east wall the east ghosts no longer depend on the state;
    the adjoint drops that path (checked against central differences)."""
    p = Params(nx=40, ny=32, length=40.0, grav=2e-3, rounding=ROUND_CLEAN, stefan_number=0.1)
    nsteps, seg = 200, 30
    pi, pj = seeded_probes(p.nx, p.ny, 8, seed=2)
    pi += [p.nx, p.nx]
    pj += [3, 17]
    kw = dict(t_cold=273.0, noise_amp=1e-3, noise_seed=7)

    def rec_at(theta):
        st, kr, th = theta
        with Lattice(p.replace(precision=FP64_MOM, stefan_number=st), initialize=False,
                     t_hot=th, k_ratio=kr, **kw) as lat:
            lat.initialization()
            lat.adj_setup(nsteps, seg, pi, pj)
            return lat.adj_forward()

    target = rec_at((0.12, 1.2, 283.2))
    theta0 = np.array([0.1, 1.5, 283.5])
    with Lattice(p.replace(precision=FP64_MOM, stefan_number=theta0[0]), initialize=False,
                 t_hot=theta0[2], k_ratio=theta0[1], **kw) as lat:
        lat.initialization()
        lat.adj_setup(nsteps, seg, pi, pj)
        rec = lat.adj_forward()
        g = lat.adj_backward(2.0 * (rec - target) / rec.size)
    loss = lambda r: float(((r - target) ** 2).mean())
    for c, h in enumerate((1e-6, 1e-6, 1e-4)):
        tp, tm = theta0.copy(), theta0.copy()
        tp[c] += h
        tm[c] -= h
        fd = (loss(rec_at(tp)) - loss(rec_at(tm))) / (2 * h)
        print(f"component {c}: adjoint {g[c]:.10e} central difference {fd:.10e}")
        assert abs(fd - g[c]) <= 1e-4 * abs(g[c]), (c, fd, g[c])


def test_adjoint_with_critical_and_reference_temperatures_matches_central_differences():
    """T_crit and the Boussinesq buoyancy reference change the phase-field
    rates and the buoyancy term of the reversed update."""
    p = Params(nx=40, ny=32, length=40.0, grav=2e-3, rounding=ROUND_CLEAN, stefan_number=0.1)
    nsteps, seg = 200, 30
    pi, pj = seeded_probes(p.nx, p.ny, 8, seed=6)
    kw = dict(t_crit=274.0, t_ref=278.0)

    def rec_at(theta, keep=False):
        st, kr, th = theta
        lat = Lattice(p.replace(precision=FP64_MOM, stefan_number=st), initialize=False,
                      t_hot=th, k_ratio=kr, **kw)
        lat.initialization()
        lat.adj_setup(nsteps, seg, pi, pj)
        rec = lat.adj_forward()
        if keep:
            return lat, rec
        lat.close()
        return rec

    target = rec_at((0.12, 1.2, 283.2))
    theta0 = np.array([0.1, 1.5, 283.5])
    lat, rec = rec_at(theta0, keep=True)
    g = lat.adj_backward(2.0 * (rec - target) / rec.size)
    lat.close()
    loss = lambda r: float(((r - target) ** 2).mean())
    for c, h in enumerate((1e-6, 1e-6, 1e-4)):
        tp, tm = theta0.copy(), theta0.copy()
        tp[c] += h
        tm[c] -= h
        fd = (loss(rec_at(tp)) - loss(rec_at(tm))) / (2 * h)
        print(f"component {c}: adjoint {g[c]:.10e} central difference {fd:.10e}")
        assert abs(fd - g[c]) <= 1e-4 * abs(g[c]), (c, fd, g[c])


@pytest.mark.parametrize("n_states", [1, 2, 3, 5, 8, 0])
def test_binomial_schedule_gives_the_level_schedules_gradient(n_states):
    """Binomial checkpointing (bgk_adj_setup_binomial) recomputes the same
    states as the one-level schedule: the same gradient bit for bit, with 1
    snapshot (every state recomputed from the initial one) up to as many as
    memory holds (0: capped at the number of steps, every state stored)."""
    p = Params(nx=30, ny=24, length=30.0, grav=2e-3, rounding=ROUND_CLEAN, tau_f=0.8,
               stefan_number=0.1, precision=FP64_MOM)
    nsteps = 75
    pi, pj = seeded_probes(p.nx, p.ny, 6, seed=5)
    theta0 = (0.1, 1.4, 283.5)
    target = gpu_record(p, (0.12, 1.2, 283.2), pi, pj, nsteps, 10)
    grads = []
    for mode in ("levels", "binomial"):
        print(f"adjoint {mode}, {nsteps} steps, n_states {n_states}", flush=True)
        with Lattice(p.replace(stefan_number=theta0[0]), initialize=False, t_hot=theta0[2],
                     k_ratio=theta0[1]) as lat:
            lat.initialization()
            if mode == "levels":
                lat.adj_setup(nsteps, 9, pi, pj)
            else:
                lat.adj_setup_binomial(nsteps, n_states, pi, pj)
            rec = lat.adj_forward()
            grads.append(lat.adj_backward(2.0 * (rec - target) / rec.size))
    assert np.array_equal(grads[0], grads[1]), grads


def test_binomial_schedule_at_1024_gives_the_level_gradient():
    """At 1024^2 with FP32 moments, with few snapshot slots (leaf loops
    through the work states as well as split advances) and with every state
    stored, the binomial schedule's gradient equals the one-level schedule's
    bit for bit."""
    from bgk_pcm_amd import configs
    base = configs.config5(1024).params.replace(precision=FP32_MOM)
    nsteps = 240
    pi, pj = seeded_probes(base.nx, base.ny, 16, seed=0)
    theta0 = configs.ADJ_THETA0
    target = None
    grads = []
    for mode in ("levels", "binomial", "binomial_auto"):
        print(f"1024^2 adjoint {mode}, {nsteps} steps", flush=True)
        with Lattice(base.replace(stefan_number=theta0[0]), initialize=False,
                     t_hot=configs.t_hot_of(base, theta0[2]), k_ratio=theta0[1]) as lat:
            lat.initialization()
            if mode == "levels":
                lat.adj_setup(nsteps, 16, pi, pj)
            else:
                lat.adj_setup_binomial(nsteps, 6 if mode == "binomial" else 0, pi, pj)
            rec = lat.adj_forward()
            if target is None:
                target = rec * (1.0 + 1e-4)
            grads.append(lat.adj_backward(2.0 * (rec - target) / rec.size))
    assert np.isfinite(grads[0]).all() and np.all(grads[0] != 0.0), grads[0]
    assert np.array_equal(grads[0], grads[1]), grads
    assert np.array_equal(grads[0], grads[2]), grads

