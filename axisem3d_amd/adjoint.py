"""Gradient harness over the adjoint run.

The reference has no adjoint solver.  Here the sensitivity of any
seismogram misfit to the elastic model comes from the discrete adjoint of
the GPU time loop (``axs3d_adjoint_run``): the forward run from rest,
checkpoints on the device, the adjoint field backwards in time, and the
sensitivity densities accumulated on the fly.  This module wraps it as a
``torch.autograd.Function`` so that a misfit written in torch can be
differentiated end to end; torch is the harness only, every number comes
from the HIP kernels.

Model parameterisation: rho, vp, vs per GLL point of every element
(shape (n_elements, 25)), giving lambda = rho (vp^2 - 2 vs^2), mu = rho
vs^2 (Material::createIsotropic, Material.cpp:566-598) and the assembled
mass sum rho IF (GLLPoint.cpp:54-77).  The time step stays that of the
spec, so a perturbed model runs the same steps.
"""
import copy

import numpy as np

from .solver import GpuSolver


def with_model(spec, rho, vp, vs, vp3d=None, vs3d=None):
    """A copy of spec (isotropic solid) with lambda, mu and the point
    masses of the model (rho, vp, vs) at every element GLL point; 3D
    elements take their velocities per slice from vp3d, vs3d (the layout
    of spec.prop3d without the parameter axis, [slice][25]) with the
    density of their GLL point."""
    rho, vp, vs = (np.asarray(x, dtype=np.float64) for x in (rho, vp, vs))
    ne = len(spec.elem_points)
    for x in (rho, vp, vs):
        if x.shape != (ne, 25):
            raise ValueError("model arrays must be (n_elements, 25)")
    mtype = getattr(spec, "elem_mtype", None)
    if mtype is not None and np.any(np.asarray(mtype) != 0):
        raise ValueError("with_model: the adjoint covers isotropic media "
                         "only")
    sp = copy.copy(spec)
    lam = rho * (vp ** 2 - 2.0 * vs ** 2)
    mu = rho * vs ** 2
    sp.elem_iso = np.ascontiguousarray(np.stack([lam, mu], axis=1))
    if mtype is not None:
        # a spec with the general material arrays is uploaded from mat1d
        # (one (lambda, mu) block of 2 x 25 per element), not elem_iso
        off = np.asarray(spec.mat1d_off, dtype=np.int64)
        mat1d = np.array(spec.mat1d, dtype=np.float64, copy=True)
        kind = getattr(spec, "elem_kind", None)
        for e in range(ne):
            if kind is None or int(kind[e]) == 0:
                mat1d[off[e]:off[e] + 50] = sp.elem_iso[e].reshape(-1)
        sp.mat1d = mat1d
    ifac = np.asarray(spec.elem_gq)[:, 5, :]
    mass = np.zeros(len(spec.solid_nr))
    np.add.at(mass, np.asarray(spec.elem_points).reshape(-1),
              (rho * ifac).reshape(-1))
    sp.solid_mass = mass
    sp._rho1d = rho
    if vp3d is not None:
        r3 = _slice_rho(spec, rho)
        vp3d, vs3d = (np.asarray(x, dtype=np.float64) for x in (vp3d, vs3d))
        p3 = np.ascontiguousarray(np.stack(
            [r3 * (vp3d ** 2 - 2.0 * vs3d ** 2), r3 * vs3d ** 2], axis=1))
        # general-material specs keep prop3d flat, [slice][2][25] in
        # element order: the same values in the same order
        sp.prop3d = p3.reshape(-1) if mtype is not None else p3
    return sp


def with_media(spec, mat1d, rho, prop3d=None):
    """A copy of a general-media spec (make_media: spec.elem_mtype,
    spec.mat1d, spec.prop3d) with the material parameters mat1d (packed as
    spec.mat1d: per element its type's parameters x 25), prop3d (packed as
    spec.prop3d) and the point masses of the density rho (n_elements, 25)
    at every element GLL point (GLLPoint.cpp:54-77).  The attenuation
    state (dlambda, dmu) and the time step stay those of the spec."""
    sp = copy.copy(spec)
    sp.mat1d = np.ascontiguousarray(mat1d, dtype=np.float64)
    if prop3d is not None:
        sp.prop3d = np.ascontiguousarray(prop3d, dtype=np.float64)
    rho = np.asarray(rho, dtype=np.float64)
    ifac = np.asarray(spec.elem_gq)[:, 5, :]
    mass = np.zeros(len(spec.solid_nr))
    np.add.at(mass, np.asarray(spec.elem_points).reshape(-1),
              (rho * ifac).reshape(-1))
    sp.solid_mass = mass
    sp._rho1d = rho
    return sp


def forward_seismograms_media(spec, media, nsteps, precision="fp64"):
    """Seismograms of the general media (mat1d, rho, prop3d) from rest."""
    g = GpuSolver(with_media(spec, *media), precision, max_samples=nsteps)
    try:
        g.run(nsteps)
        return g.seismograms(nsteps)
    finally:
        g.close()


def gradient_media(spec, media, djdr, precision="fp64",
                   checkpoint_interval=None):
    """dJ/d(mat1d) (packed as spec.mat1d), dJ/drho (n_elements, 25) at
    fixed moduli and dJ/d(prop3d) (packed as spec.prop3d, or None) of the
    general media (mat1d, rho, prop3d) for the misfit derivative djdr
    (axs3d_adjoint_run_media)."""
    g = GpuSolver(with_media(spec, *media), precision, max_samples=len(djdr))
    try:
        return g.adjoint_run_media(djdr, checkpoint_interval)
    finally:
        g.close()


def _slice_rho(spec, rho):
    """rho of each slice's GLL points, [slice][25] (3D elements)."""
    kind = np.asarray(spec.elem_kind)
    nr = np.asarray(spec.elem_nr)
    return np.concatenate([np.repeat(rho[e][None, :], int(nr[e]), axis=0)
                           for e in range(len(kind)) if kind[e] == 1])


def model3d_of(spec):
    """(vp3d, vs3d) per slice of the 3D elements, [slice][25]."""
    r3 = _slice_rho(spec, np.asarray(spec._rho1d, dtype=np.float64))
    p3 = np.asarray(spec.prop3d).reshape(-1, 2, 25)
    lam, mu = p3[:, 0, :], p3[:, 1, :]
    return np.sqrt((lam + 2.0 * mu) / r3), np.sqrt(mu / r3)


def model_of(spec):
    """(rho, vp, vs) per element GLL point of a 1D isotropic spec, from its
    lambda, mu and the density the spec was built with."""
    rho = np.asarray(spec._rho1d, dtype=np.float64)
    lam, mu = spec.elem_iso[:, 0, :], spec.elem_iso[:, 1, :]
    vs = np.sqrt(mu / rho)
    vp = np.sqrt((lam + 2.0 * mu) / rho)
    return rho, vp, vs


def chain_rule(kern, rho, vp, vs):
    """dJ/d(rho, vp, vs) from dJ/d(lambda, mu, rho at fixed lambda, mu),
    kern of shape (n_elements, 25, 3)."""
    kl, km, kr = kern[..., 0], kern[..., 1], kern[..., 2]
    d_rho = kr + (vp ** 2 - 2.0 * vs ** 2) * kl + vs ** 2 * km
    d_vp = 2.0 * rho * vp * kl
    d_vs = -4.0 * rho * vs * kl + 2.0 * rho * vs * km
    return d_rho, d_vp, d_vs


def forward_seismograms(spec, rho, vp, vs, nsteps, precision="fp64"):
    """Seismograms (nsteps, n_receivers, 3) of the model from rest."""
    g = GpuSolver(with_model(spec, rho, vp, vs), precision,
                  max_samples=nsteps)
    try:
        g.run(nsteps)
        return g.seismograms(nsteps)
    finally:
        g.close()


def gradient(spec, rho, vp, vs, djdr, precision="fp64",
             checkpoint_interval=None):
    """dJ/d(rho, vp, vs) of a misfit J whose derivative with respect to
    the seismograms of the model is djdr (nsteps, n_receivers, 3)."""
    g = GpuSolver(with_model(spec, rho, vp, vs), precision,
                  max_samples=len(djdr))
    try:
        kern, _ = g.adjoint_run(djdr, checkpoint_interval)
    finally:
        g.close()
    return chain_rule(kern, np.asarray(rho), np.asarray(vp), np.asarray(vs))


def chain_rule3d(spec, kern, kern3d, rho, vp, vs, vp3d, vs3d):
    """dJ/d(rho, vp, vs) per element GLL point and dJ/d(vp3d, vs3d) per
    slice, from the adjoint's dJ/d(lambda, mu, rho) (1D elements and the
    mass) and dJ/d(lambda, mu) per slice (3D elements), with lambda = rho
    (vp^2 - 2 vs^2), mu = rho vs^2 slice by slice."""
    d_rho, d_vp, d_vs = chain_rule(kern, rho, vp, vs)
    r3 = _slice_rho(spec, rho)
    k3 = np.asarray(kern3d).reshape(-1, 2, 25)
    kl, km = k3[:, 0, :], k3[:, 1, :]
    d_vp3 = 2.0 * r3 * vp3d * kl
    d_vs3 = 2.0 * r3 * vs3d * (km - 2.0 * kl)
    per_slice = (vp3d ** 2 - 2.0 * vs3d ** 2) * kl + vs3d ** 2 * km
    kind = np.asarray(spec.elem_kind)
    nr = np.asarray(spec.elem_nr)
    off = 0
    for e in range(len(kind)):
        if kind[e] == 1:
            d_rho[e] += per_slice[off:off + int(nr[e])].sum(axis=0)
            off += int(nr[e])
    return d_rho, d_vp, d_vs, d_vp3, d_vs3


def forward_seismograms3d(spec, model, nsteps, precision="fp64"):
    """Seismograms of the model (rho, vp, vs, vp3d, vs3d) from rest."""
    g = GpuSolver(with_model(spec, *model), precision, max_samples=nsteps)
    try:
        g.run(nsteps)
        return g.seismograms(nsteps)
    finally:
        g.close()


def gradient3d(spec, model, djdr, precision="fp64", checkpoint_interval=None):
    """dJ/d of the model (rho, vp, vs, vp3d, vs3d) of a domain with 3D
    isotropic elements, from the adjoint run."""
    g = GpuSolver(with_model(spec, *model), precision, max_samples=len(djdr))
    try:
        kern, kern3d = g.adjoint_run(djdr, checkpoint_interval)
    finally:
        g.close()
    return chain_rule3d(spec, kern, kern3d, *model)


def seismogram_function():
    """torch.autograd.Function: (rho, vp, vs) float64 CPU tensors ->
    seismograms, backward through the adjoint run.  Built on first use so
    that importing this module does not import torch."""
    import torch

    class Seismograms(torch.autograd.Function):
        @staticmethod
        def forward(ctx, rho, vp, vs, spec, nsteps, precision, interval):
            m = [x.detach().cpu().numpy() for x in (rho, vp, vs)]
            ctx.model = m
            ctx.spec, ctx.precision, ctx.interval = spec, precision, interval
            seis = forward_seismograms(spec, *m, nsteps, precision)
            return torch.from_numpy(seis)

        @staticmethod
        def backward(ctx, grad):
            d = gradient(ctx.spec, *ctx.model, grad.detach().cpu().numpy(),
                         ctx.precision, ctx.interval)
            return (torch.from_numpy(d[0]), torch.from_numpy(d[1]),
                    torch.from_numpy(d[2]), None, None, None, None)

    return Seismograms
