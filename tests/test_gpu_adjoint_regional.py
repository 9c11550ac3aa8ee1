"""The adjoint on a reduced version of the spherical regional workload: a
3 x 120 km sector of the seeded 6-layer model with the +-3% 3D random
field on every element (cardinal path), per-point nu from 10 up to the
in-plane bound, 5-SLS CG4 attenuation, an axial moment tensor and 6
surface receivers.  The misfit is the waveform L2 against seismograms of
a seeded +-2% dvs perturbation of the 3D model; the forward run keeps
checkpoints every 20 of 400 steps.

The FP32+FP64 adjoint gradient (rho per GLL point, vp and vs per slice)
is checked against central differences of the FP64 forward misfit along
16 seeded directions (relative 1e-3), and against the FP64 adjoint
(relative 1e-3 per component)."""
import numpy as np
import pytest

from axisem3d_amd import adjoint, preloop

pytestmark = pytest.mark.gpu

NSTEPS = 400
INTERVAL = 20


def _spec():
    sp = preloop.spherical_regional(
        ntheta=12, nrad=8, theta_max_deg=3.0, depth=120e3, nu_min=10,
        nu_max=400, nsteps=NSTEPS, n_receivers=6, attenuation=True,
        cg4=True, corr_len=20e3)
    sp._build_3d()
    preloop.finish_regional(sp, NSTEPS, src_depth=30e3, f_peak=0.5,
                            n_receivers=6, theta_max_deg=3.0)
    return sp


def _waves(rng, x, y, z, amp, wavelength=40e3):
    d = np.zeros_like(x)
    for _ in range(4):
        k = rng.uniform(-1.0, 1.0, 3) * 2 * np.pi / wavelength
        d += np.cos(k[0] * x + k[1] * y + k[2] * z + rng.uniform(0, 6.3))
    return amp * d / np.abs(d).max()


def _smooth_point(sp, rng, amp):
    """Smooth seeded relative perturbation per element GLL point."""
    sz = sp._gq["sz"]
    return _waves(rng, sz[:, 0], 0.0 * sz[:, 0], sz[:, 1], amp)


def _smooth_slice(sp, rng, amp):
    """Smooth seeded relative perturbation per slice of the 3D elements,
    [slice][25], from the slices' Cartesian coordinates."""
    sz = sp._gq["sz"]
    xs, ys, zs = [], [], []
    for e in np.nonzero(np.asarray(sp.elem_kind) == 1)[0]:
        nr = int(sp.elem_nr[e])
        phi = 2 * np.pi * np.arange(nr) / nr
        s, z = sz[e, 0], sz[e, 1]
        xs.append(s[None, :] * np.cos(phi)[:, None])
        ys.append(s[None, :] * np.sin(phi)[:, None])
        zs.append(np.broadcast_to(z[None, :], (nr, 25)))
    return _waves(rng, np.concatenate(xs), np.concatenate(ys),
                  np.concatenate(zs), amp)


def test_regional_adjoint_matches_fd_and_fp64():
    sp = _spec()
    assert np.all(np.asarray(sp.elem_kind) == 1)
    assert sp.att["nsls"] == 5 and sp.att["cg4"]
    model = list(adjoint.model_of(sp)) + list(adjoint.model3d_of(sp))
    rng = np.random.default_rng(31)
    true = list(model)
    true[4] = model[4] * (1 + _smooth_slice(sp, rng, 0.02))
    d_obs = adjoint.forward_seismograms3d(sp, true, NSTEPS)
    r = adjoint.forward_seismograms3d(sp, model, NSTEPS)
    assert np.abs(r - d_obs).max() > 0
    g32 = adjoint.gradient3d(sp, model, r - d_obs, "fp32",
                             checkpoint_interval=INTERVAL)
    g64 = adjoint.gradient3d(sp, model, r - d_obs, "fp64",
                             checkpoint_interval=INTERVAL)
    for name, a, b in zip(("rho", "vp", "vs", "vp3d", "vs3d"), g32, g64):
        nb = np.linalg.norm(b)
        if nb == 0:
            continue
        rel = np.linalg.norm(a - b) / nb
        print(f"regional adjoint: FP32+FP64 vs FP64 d/d{name} {rel:.2e}")
        assert rel < 1e-3

    def misfit(m):
        return 0.5 * np.sum((adjoint.forward_seismograms3d(sp, m, NSTEPS) -
                             d_obs) ** 2)
    worst = 0.0
    for trial in range(16):
        dirs = [model[0] * _smooth_point(sp, rng, 1.0),
                0.0 * model[1], 0.0 * model[2],
                model[3] * _smooth_slice(sp, rng, 1.0),
                model[4] * _smooth_slice(sp, rng, 1.0)]
        h = 1e-4
        fd = (misfit([m + h * d for m, d in zip(model, dirs)]) -
              misfit([m - h * d for m, d in zip(model, dirs)])) / (2 * h)
        ad = sum(float(np.sum(g * d)) for g, d in zip(g32, dirs))
        err = abs(fd - ad) / abs(fd)
        print(f"regional adjoint direction {trial}: FD {fd:.9e} adjoint "
              f"{ad:.9e} relative difference {err:.2e}")
        worst = max(worst, err)
    assert worst < 1e-3
