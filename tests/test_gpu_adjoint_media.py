"""The adjoint for general solid media (axs3d_adjoint_run_media): dJ/d of
every material parameter (isotropic lambda, mu; transversely isotropic A,
C, F, L, N; anisotropic 21 Cij) per element GLL point of the 1D elements
and per slice of the 3D elements, and dJ/drho at fixed moduli.

Domains: a spherical sector (element types isotropic, TI and anisotropic
in turn, every other element 3D, strains rotated to RTZ, 5-SLS CG4
attenuation held fixed) and a Cartesian box with odd element Nr (5)
everywhere and no attenuation.  The misfit is the waveform L2 against the
seismograms of a seeded smooth perturbation of all parameters.

The FP64 gradient is checked against central differences of the FP64
forward misfit along seeded smooth directions, in all parameters at once
and in each of the three blocks alone (relative to the directional
derivative, or to a tenth of the all-parameter directions' mean where a
block's is smaller: 1e-3 on the sector, whose
element Nr varies from 10 to 25 and whose even-Nr Nyquist orders keep the
discrete stiffness from being exactly self-adjoint; 1e-5 on the odd-Nr
box), and the FP32+FP64 gradient against the FP64 one (relative 1e-3 per
array)."""

import numpy as np
import pytest

from axisem3d_amd import adjoint, preloop

pytestmark = pytest.mark.gpu

NSTEPS = 300


def _sector():
    sp = preloop.spherical_regional(
        ntheta=6, nrad=6, theta_max_deg=2.0, depth=60e3, nu_min=3,
        nu_max=12, nsteps=NSTEPS, n_receivers=3, attenuation=True, cg4=True,
        corr_len=10e3)
    E = sp.n_solid_elements
    preloop.make_media(sp, (np.arange(E) % 3).astype(np.uint8),
                       field=sp._field, mask3d=(np.arange(E) % 2) == 1,
                       att=sp._att_in, seed=2)
    preloop.finish_regional(sp, NSTEPS, src_depth=20e3, f_peak=0.5,
                            n_receivers=3, theta_max_deg=2.0)
    return sp


def _box():
    sp = preloop.cartesian_homogeneous(
        nx=6, nz=7, width=6e3, depth=5.6e3, nsteps=NSTEPS, src_depth=2e3,
        n_receivers=4, seed=2, nodal_nr=lambda s, z: np.full(s.shape, 5.0))
    E = sp.n_solid_elements
    preloop.make_media(sp, ((np.arange(E) + 1) % 3).astype(np.uint8),
                       field=lambda x, y, z: 0.02 * np.sin(x / 900.0),
                       mask3d=(np.arange(E) % 2) == 0, att=None, seed=3)
    return sp


CASES = {"sector": (_sector, 1e-3, 40e3), "box": (_box, 1e-5, 6e3)}


def _waves(rng, x, y, z, wavelength):
    d = np.zeros_like(x)
    for _ in range(4):
        k = rng.uniform(-1.0, 1.0, 3) * 2 * np.pi / wavelength
        d += np.cos(k[0] * x + k[1] * y + k[2] * z + rng.uniform(0, 6.3))
    return d / np.abs(d).max()


def _xyz(sp):
    """(s, 0, z) of every element GLL point (n_elements, 25), and (s, z,
    phi) of every slice GLL point of the 3D elements ([slice][25] each,
    element order)."""
    sz = sp._gq["sz"]
    s, z = sz[:, 0], sz[:, 1]
    ss, zs, ps = [], [], []
    for e in np.nonzero(np.asarray(sp.elem_kind) == 1)[0]:
        nr = int(sp.elem_nr[e])
        phi = 2 * np.pi * np.arange(nr) / nr
        ss.append(np.broadcast_to(s[e][None, :], (nr, 25)))
        zs.append(np.broadcast_to(z[e][None, :], (nr, 25)))
        ps.append(np.broadcast_to(phi[:, None], (nr, 25)))
    return (s, 0.0 * s, z), tuple(np.concatenate(a) for a in (ss, zs, ps))


def _direction(sp, rng, wl):
    """A seeded smooth direction in (mat1d, rho, prop3d): each parameter
    its own smooth field, scaled by the element's shear modulus (moduli)
    or by rho."""
    (x, y, z), (s3, z3, p3) = _xyz(sp)
    mtype = np.asarray(sp.elem_mtype)
    kind = np.asarray(sp.elem_kind)
    nr = np.asarray(sp.elem_nr)
    m_off = np.asarray(sp.mat1d_off)
    p_off = np.asarray(sp.prop_off)
    s_off = np.asarray(sp.elem_slice_off)
    mu = np.asarray(sp.elem_iso)[:, 1, :]
    rho = np.asarray(sp._rho1d)
    fields = [_waves(rng, x, y, z, wl) for _ in range(21)]
    # per slice: smooth in (s, z) with a low azimuthal order, so that the
    # axisymmetric wavefield of the axial source feels it
    fields3 = [_waves(rng, s3, 0.0 * s3, z3, wl) *
               (1.0 + 0.5 * np.cos(p3 + rng.uniform(0, 6.3)))
               for _ in range(21)]
    dm = np.zeros(np.asarray(sp.mat1d).shape)
    dp = np.zeros(np.asarray(sp.prop3d).size)
    s3 = 0
    for e in range(len(mtype)):
        npar = (2, 5, 21)[int(mtype[e])]
        if kind[e] == 0:
            for c in range(npar):
                dm[m_off[e] + 25 * c:m_off[e] + 25 * (c + 1)] = \
                    mu[e] * fields[c][e]
        else:
            n = int(nr[e])
            blk = np.stack([f[s3:s3 + n] for f in fields3[:npar]], axis=1)
            assert s_off[e] == s3
            dp[p_off[e]:p_off[e] + blk.size] = (mu[e][None, None, :] *
                                                blk).reshape(-1)
            s3 += n
    dr = rho * _waves(rng, x, y, z, wl)
    return dm, dr, dp


def _media(sp):
    return (np.array(sp.mat1d, dtype=np.float64), np.array(sp._rho1d),
            np.array(sp.prop3d, dtype=np.float64).reshape(-1))


def _dot(g, d):
    return sum(float(np.sum(np.asarray(a).reshape(-1) *
                            np.asarray(b).reshape(-1)))
               for a, b in zip(g, d) if a is not None)


@pytest.mark.parametrize("name", sorted(CASES))
def test_media_adjoint_matches_fd_and_fp64(name):
    make, tol, wl = CASES[name]
    sp = make()
    mtype = np.asarray(sp.elem_mtype)
    assert set(mtype.tolist()) == {0, 1, 2}
    assert np.any(np.asarray(sp.elem_kind) == 1)
    assert np.any(np.asarray(sp.elem_kind) == 0)
    media = _media(sp)
    rng = np.random.default_rng(47)
    true = [m + 0.02 * d for m, d in zip(media, _direction(sp, rng, wl))]
    d_obs = adjoint.forward_seismograms_media(sp, true, NSTEPS)
    r = adjoint.forward_seismograms_media(sp, media, NSTEPS)
    assert np.abs(r - d_obs).max() > 0
    g64 = adjoint.gradient_media(sp, media, r - d_obs, "fp64",
                                 checkpoint_interval=20)
    g32 = adjoint.gradient_media(sp, media, r - d_obs, "fp32",
                                 checkpoint_interval=20)
    for what, a, b in zip(("mat1d", "rho", "prop3d"), g32, g64):
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"media adjoint {name}: FP32+FP64 vs FP64 d/d{what} "
              f"{rel:.2e}")
        assert np.linalg.norm(b) > 0 and rel < 1e-3

    def misfit(m):
        return 0.5 * np.sum((adjoint.forward_seismograms_media(
            sp, m, NSTEPS) - d_obs) ** 2)
    worst = 0.0
    ref = []
    for trial in range(6):
        d = _direction(sp, rng, wl)
        # three directions in all parameters, then each block on its own
        if trial >= 3:
            d = tuple(x if i == trial - 3 else 0.0 * x
                      for i, x in enumerate(d))
        h = 1e-4
        fd = (misfit([m + h * x for m, x in zip(media, d)]) -
              misfit([m - h * x for m, x in zip(media, d)])) / (2 * h)
        ad = _dot(g64, d)
        if trial < 3:
            ref.append(abs(fd))
        # a block's directional derivative can be small by chance (random
        # signs of 21 parameter fields): measure it against a tenth of the
        # all-parameter directions' mean as well
        err = abs(fd - ad) / max(abs(fd), 0.1 * np.mean(ref))
        print(f"media adjoint {name} direction {trial}: FD {fd:.9e} "
              f"adjoint {ad:.9e} relative difference {err:.2e}")
        worst = max(worst, err)
    assert worst < tol
