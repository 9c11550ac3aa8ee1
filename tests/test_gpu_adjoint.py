"""The adjoint run (axs3d_adjoint_run) on the GPU:

* the per-step sensitivity densities satisfy Euler's identity for the
  stiffness and the mass, which are linear in (lambda, mu) and in rho:
  sum lambda K_lambda + mu K_mu = <psi, K u> and sum rho K_rho = <psi, M a>,
  with K u from the FP64 oracle (relative 1e-12);
* the gradient of a waveform misfit through the torch harness matches
  central finite differences of forward runs along 16 seeded model
  perturbations (relative 1e-3, the target; measured up to 5e-6, the
  truncation error of the central differences at h = 1e-4);
* the FP32+FP64 adjoint agrees with the FP64 adjoint (relative L2 1e-3);
* the checkpoint interval does not change the result (bitwise).
"""
import numpy as np
import pytest

import oracle
from axisem3d_amd import adjoint, preloop

pytestmark = pytest.mark.gpu

NSTEPS = 300


def _spec(nsteps=NSTEPS):
    return preloop.cartesian_homogeneous(
        nx=6, nz=7, width=6e3, depth=5.6e3, nsteps=nsteps, src_depth=2e3,
        n_receivers=4, seed=2,
        nodal_nr=lambda s, z: 3.0 + 9.0 * (s / 6e3) + 0.0 * z)


def _smooth(spec, rng, amp):
    """A smooth seeded relative perturbation per element GLL point."""
    sz = spec._gq["sz"]                     # (E, 2, 25)
    s, z = sz[:, 0, :], sz[:, 1, :]
    d = np.zeros_like(s)
    for _ in range(4):
        ks, kz = rng.uniform(-1.0, 1.0, 2) * 2 * np.pi / 3e3
        d += np.cos(ks * s + kz * z + rng.uniform(0, 2 * np.pi))
    return amp * d / np.abs(d).max()


def _wdot(x, y, nr):
    w = np.concatenate([np.r_[1.0, np.full(n // 2, 2.0)] for n in nr])
    return float(np.sum(w[:, None] * np.real(np.conj(x) * y)))


def test_sensitivity_densities_satisfy_euler_identity():
    from axisem3d_amd.solver import GpuSolver
    sp = _spec(10)
    rng = np.random.default_rng(5)
    rho, vp, vs = adjoint.model_of(sp)
    sp = adjoint.with_model(sp, rho * (1 + _smooth(sp, rng, 0.1)),
                            vp * (1 + _smooth(sp, rng, 0.1)),
                            vs * (1 + _smooth(sp, rng, 0.1)))
    o = oracle.Oracle(sp)
    n = o.nmodes
    psi, u, a = (rng.standard_normal((n, 3)) + 1j * rng.standard_normal(
        (n, 3)) for _ in range(3))
    g = GpuSolver(sp, "fp64")
    kern = g.adjoint_density(psi, u, a)
    g.close()
    lam, mu = sp.elem_iso[:, 0, :], sp.elem_iso[:, 1, :]
    rho_e = adjoint.model_of(sp)[0]
    got_k = np.sum(lam * kern[..., 0] + mu * kern[..., 1])
    ref_k = -_wdot(psi, o.stiffness(u), sp.solid_nr)
    assert abs(got_k - ref_k) < 1e-12 * abs(ref_k)
    m = np.repeat(sp.solid_mass, sp.solid_nr // 2 + 1)[:, None]
    got_m = np.sum(rho_e * kern[..., 2])
    ref_m = _wdot(psi, m * a, sp.solid_nr)
    assert abs(got_m - ref_m) < 1e-12 * abs(ref_m)


def _problem():
    sp = _spec()
    rho, vp, vs = adjoint.model_of(sp)
    rng = np.random.default_rng(11)
    # observed data: the model with a seeded +-2% vs perturbation
    d_obs = adjoint.forward_seismograms(
        sp, rho, vp, vs * (1 + _smooth(sp, rng, 0.02)), NSTEPS)
    return sp, (rho, vp, vs), d_obs


def _misfit(sp, model, d_obs):
    r = adjoint.forward_seismograms(sp, *model, NSTEPS)
    return 0.5 * np.sum((r - d_obs) ** 2)


def test_adjoint_gradient_matches_finite_differences():
    import torch
    sp, model, d_obs = _problem()
    S = adjoint.seismogram_function()
    t = [torch.tensor(x, requires_grad=True) for x in model]
    r = S.apply(*t, sp, NSTEPS, "fp64", None)
    J = 0.5 * ((r - torch.from_numpy(d_obs)) ** 2).sum()
    J.backward()
    grads = [x.grad.numpy() for x in t]
    assert float(J.detach()) > 0
    rng = np.random.default_rng(17)
    worst = 0.0
    for trial in range(16):
        dirs = [x * _smooth(sp, rng, 1.0) for x in model]
        h = 1e-4
        mp = [x + h * d for x, d in zip(model, dirs)]
        mm = [x - h * d for x, d in zip(model, dirs)]
        fd = (_misfit(sp, mp, d_obs) - _misfit(sp, mm, d_obs)) / (2 * h)
        ad = sum(float(np.sum(gx * d)) for gx, d in zip(grads, dirs))
        err = abs(fd - ad) / abs(fd)
        print(f"perturbation {trial}: FD {fd:.9e} adjoint {ad:.9e} "
              f"relative difference {err:.2e}")
        worst = max(worst, err)
    assert worst < 1e-3


def test_adjoint_on_general_material_spec_matches_fd():
    """An all-isotropic spec built through the general material arrays
    (make_media: mat1d and a flat prop3d, half the elements 3D) is
    uploaded from those arrays; the gradient must follow the model there
    too.  Four seeded central finite-difference directions, smooth in
    space as in the other gradient tests (a direction rough from slice to
    slice would load the Nyquist order of even Nr, where the adjoint is
    not exact).  Odd element Nr (5) everywhere, as the other 3D gradient
    tests: an even Nr's Nyquist row keeps the discrete stiffness from
    being exactly self-adjoint."""
    sp = preloop.cartesian_homogeneous(
        nx=6, nz=7, width=6e3, depth=5.6e3, nsteps=NSTEPS, src_depth=2e3,
        n_receivers=4, seed=2, nodal_nr=lambda s, z: np.full(s.shape, 5.0))
    E = sp.n_solid_elements
    preloop.make_media(sp, np.zeros(E, dtype=np.uint8),
                       field=lambda x, y, z: 0.02 * np.sin(x / 900.0),
                       mask3d=(np.arange(E) % 2) == 0)
    model = list(adjoint.model_of(sp)) + list(adjoint.model3d_of(sp))
    rng = np.random.default_rng(23)
    obs = list(model)
    obs[4] = model[4] * (1 + _smooth3d(sp, rng, 0.02))
    d_obs = adjoint.forward_seismograms3d(sp, obs, NSTEPS)
    r = adjoint.forward_seismograms3d(sp, model, NSTEPS)
    grads = adjoint.gradient3d(sp, model, r - d_obs, "fp64")

    def misfit(m):
        return 0.5 * np.sum(
            (adjoint.forward_seismograms3d(sp, m, NSTEPS) - d_obs) ** 2)
    worst = 0.0
    for trial in range(4):
        dirs = [model[i] * (_smooth(sp, rng, 1.0) if i < 3 else
                            _smooth3d(sp, rng, 1.0)) for i in range(5)]
        h = 1e-4
        fd = (misfit([x + h * d for x, d in zip(model, dirs)]) -
              misfit([x - h * d for x, d in zip(model, dirs)])) / (2 * h)
        ad = sum(float(np.sum(gx * d)) for gx, d in zip(grads, dirs))
        err = abs(fd - ad) / abs(fd)
        print(f"general-material spec, perturbation {trial}: FD {fd:.9e} "
              f"adjoint {ad:.9e} relative difference {err:.2e}")
        worst = max(worst, err)
    assert worst < 1e-3


def test_fp32_adjoint_matches_fp64_and_checkpoints_do_not_matter():
    sp, model, d_obs = _problem()
    r = adjoint.forward_seismograms(sp, *model, NSTEPS)
    djdr = r - d_obs
    g64 = adjoint.gradient(sp, *model, djdr, "fp64")
    g64b = adjoint.gradient(sp, *model, djdr, "fp64",
                            checkpoint_interval=7)
    g64c = adjoint.gradient(sp, *model, djdr, "fp64",
                            checkpoint_interval=NSTEPS)
    for a, b, c in zip(g64, g64b, g64c):
        assert np.array_equal(a, b) and np.array_equal(a, c)
    g32 = adjoint.gradient(sp, *model, djdr, "fp32")
    for a, b in zip(g32, g64):
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"FP32+FP64 vs FP64 adjoint: relative L2 {err:.2e}")
        assert err < 1e-3


def test_adjoint_rejects_unsupported_domains():
    # TI media are not supported by the adjoint run
    from axisem3d_amd.solver import GpuSolver
    from axisem3d_amd import capi
    sp = preloop.cartesian_homogeneous(
        nx=4, nz=4, width=4e3, depth=4e3, nsteps=10, src_depth=2e3,
        n_receivers=2, nodal_nr=lambda s, z: np.full(s.shape, 5.0))
    E = sp.n_solid_elements
    preloop.make_media(sp, np.ones(E, np.uint8), field=None,
                       mask3d=np.zeros(E, bool), att=None, seed=3)
    g = GpuSolver(sp, "fp64")
    with pytest.raises(capi.Axs3dError, match="adjoint"):
        g.adjoint_run(np.zeros((10, 2, 3)))
    g.close()


def _spec3d(nsteps=NSTEPS):
    # odd element Nr (5) everywhere, every other element 3D isotropic
    sp = preloop.cartesian_homogeneous(
        nx=6, nz=7, width=6e3, depth=5.6e3, nsteps=nsteps, src_depth=2e3,
        n_receivers=4, seed=2, nodal_nr=lambda s, z: np.full(s.shape, 5.0))
    E = sp.n_solid_elements
    preloop.make_3d_isotropic(sp, lambda x, y, z: 0.0 * x,
                              mask3d=(np.arange(E) % 2) == 0)
    return sp


def _smooth3d(spec, rng, amp):
    """A smooth seeded relative perturbation per slice of the 3D
    elements, [slice][25], from the slices' Cartesian coordinates."""
    sz = spec._gq["sz"]
    out = []
    for e in np.nonzero(np.asarray(spec.elem_kind) == 1)[0]:
        nr = int(spec.elem_nr[e])
        phi = 2 * np.pi * np.arange(nr) / nr
        s, z = sz[e, 0], sz[e, 1]
        out.append((s[None, :] * np.cos(phi)[:, None],
                    s[None, :] * np.sin(phi)[:, None],
                    np.broadcast_to(z[None, :], (nr, 25))))
    x = np.concatenate([o[0] for o in out])
    y = np.concatenate([o[1] for o in out])
    z = np.concatenate([o[2] for o in out])
    d = np.zeros_like(x)
    for _ in range(4):
        k = rng.uniform(-1.0, 1.0, 3) * 2 * np.pi / 3e3
        d += np.cos(k[0] * x + k[1] * y + k[2] * z + rng.uniform(0, 6.3))
    return amp * d / np.abs(d).max()


def test_adjoint3d_gradient_matches_finite_differences():
    """Mixed 1D/3D isotropic domain: the gradient with respect to rho, vp,
    vs per GLL point and vp, vs per slice of the 3D elements against
    central differences of forward runs along 8 seeded directions
    (relative 1e-3)."""
    sp = _spec3d()
    model = list(adjoint.model_of(sp)) + list(adjoint.model3d_of(sp))
    rng = np.random.default_rng(19)
    true = list(model)
    true[4] = model[4] * (1 + _smooth3d(sp, rng, 0.02))
    true[2] = model[2] * (1 + _smooth(sp, rng, 0.02))
    d_obs = adjoint.forward_seismograms3d(sp, true, NSTEPS)
    r = adjoint.forward_seismograms3d(sp, model, NSTEPS)
    grads = adjoint.gradient3d(sp, model, r - d_obs)
    assert np.abs(grads[3]).max() > 0 and np.abs(grads[4]).max() > 0

    def misfit(m):
        return 0.5 * np.sum((adjoint.forward_seismograms3d(sp, m, NSTEPS) -
                             d_obs) ** 2)
    worst = 0.0
    for trial in range(8):
        dirs = [model[i] * (_smooth(sp, rng, 1.0) if i < 3 else
                            _smooth3d(sp, rng, 1.0)) for i in range(5)]
        h = 1e-4
        fd = (misfit([m + h * d for m, d in zip(model, dirs)]) -
              misfit([m - h * d for m, d in zip(model, dirs)])) / (2 * h)
        ad = sum(float(np.sum(g * d)) for g, d in zip(grads, dirs))
        err = abs(fd - ad) / abs(fd)
        print(f"3D direction {trial}: FD {fd:.9e} adjoint {ad:.9e} "
              f"relative difference {err:.2e}")
        worst = max(worst, err)
    assert worst < 1e-3


def _spec_att(nsteps=NSTEPS, cg4=True):
    # mixed 1D/3D isotropic domain with 5-SLS attenuation (Q_mu 60):
    # the model parameters are the unrelaxed moduli, the anelastic
    # coefficients (dlambda, dmu) stay fixed
    sp = preloop.cartesian_homogeneous(
        nx=6, nz=7, width=6e3, depth=5.6e3, nsteps=nsteps, src_depth=2e3,
        n_receivers=4, seed=2, nodal_nr=lambda s, z: np.full(s.shape, 5.0))
    E = sp.n_solid_elements
    att = {"nsls": 5, "cg4": cg4, "f_min": 0.1, "f_max": 4.0, "f_ref": 1.0,
           "qkp": np.full((E, 25), 57823.0), "qmu": np.full((E, 25), 60.0)}
    preloop.make_3d_isotropic(sp, lambda x, y, z: 0.0 * x, att=att,
                              mask3d=(np.arange(E) % 2) == 0)
    return sp


@pytest.mark.parametrize("cg4", [True, False])
def test_adjoint_with_attenuation_matches_finite_differences(cg4):
    """The adjoint field runs the same SLS recursion backwards in time
    (its kernel depends on the lag only and the anelastic operator is
    symmetric), with its own memory variables; the forward memory
    variables are part of the checkpoints.  Gradient with respect to the
    unrelaxed moduli (as rho, vp, vs) against central differences along 8
    seeded directions (relative 1e-3), CG4 and FULL."""
    sp = _spec_att(cg4=cg4)
    model = list(adjoint.model_of(sp)) + list(adjoint.model3d_of(sp))
    rng = np.random.default_rng(23)
    true = list(model)
    true[4] = model[4] * (1 + _smooth3d(sp, rng, 0.02))
    true[2] = model[2] * (1 + _smooth(sp, rng, 0.02))
    d_obs = adjoint.forward_seismograms3d(sp, true, NSTEPS)
    r = adjoint.forward_seismograms3d(sp, model, NSTEPS)
    grads = adjoint.gradient3d(sp, model, r - d_obs)

    def misfit(m):
        return 0.5 * np.sum((adjoint.forward_seismograms3d(sp, m, NSTEPS) -
                             d_obs) ** 2)
    worst = 0.0
    for trial in range(8):
        dirs = [model[i] * (_smooth(sp, rng, 1.0) if i < 3 else
                            _smooth3d(sp, rng, 1.0)) for i in range(5)]
        h = 1e-4
        fd = (misfit([m + h * d for m, d in zip(model, dirs)]) -
              misfit([m - h * d for m, d in zip(model, dirs)])) / (2 * h)
        ad = sum(float(np.sum(g * d)) for g, d in zip(grads, dirs))
        err = abs(fd - ad) / abs(fd)
        print(f"attenuation cg4={cg4} direction {trial}: FD {fd:.9e} "
              f"adjoint {ad:.9e} relative difference {err:.2e}")
        worst = max(worst, err)
    assert worst < 1e-3
