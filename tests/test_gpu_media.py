"""GPU parity of transversely isotropic and anisotropic media (1D on the
Fourier path, 3D on the cardinal path), with the spherical RTZ rotation
and with SLS attenuation, against the FP64 oracle through the C ABI.

Tolerances: FP64 mode agrees with the oracle to rounding (relative 1e-13
per stage, 1e-12 seismogram L2); FP32+FP64 mode relative 1e-5 per stage
and 1e-4 seismogram L2."""
import numpy as np
import pytest

import oracle
from axisem3d_amd import preloop

pytestmark = pytest.mark.gpu


def _spec(types, mask3d, att, nsteps=300, spherical=True):
    if spherical:
        spec = preloop.spherical_regional(
            ntheta=6, nrad=6, theta_max_deg=2.0, depth=60e3, nu_min=3,
            nu_max=12, nsteps=nsteps, n_receivers=3, attenuation=att,
            corr_len=10e3)
        E = spec.n_solid_elements
        preloop.make_media(spec, types(E), field=spec._field,
                           mask3d=mask3d(E), att=spec._att_in, seed=2)
        preloop.finish_regional(spec, nsteps, src_depth=20e3, f_peak=0.5,
                                n_receivers=3, theta_max_deg=2.0)
    else:
        spec = preloop.cartesian_homogeneous(
            nx=6, nz=6, width=6e3, depth=6e3, nsteps=nsteps, src_depth=2e3,
            n_receivers=3,
            nodal_nr=lambda s, z: 3.0 + 9.0 * (s / 6e3) + 0.0 * z)
        E = spec.n_solid_elements
        preloop.make_media(spec, types(E), field=None, mask3d=mask3d(E),
                           att=None, seed=3)
    return spec


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _solver(spec, prec):
    from axisem3d_amd.solver import GpuSolver
    return GpuSolver(spec, prec)


MIXED_TYPES = lambda E: (np.arange(E) % 3).astype(np.uint8)
TI_ONLY = lambda E: np.ones(E, np.uint8)
ALL_1D = lambda E: np.zeros(E, bool)
ALL_3D = lambda E: np.ones(E, bool)
HALF_3D = lambda E: (np.arange(E) % 2) == 1


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 1e-5)])
@pytest.mark.parametrize("mask", [ALL_1D, ALL_3D, HALF_3D],
                         ids=["1d", "3d", "mixed"])
@pytest.mark.parametrize("att", [False, True])
def test_media_stiffness_matches_oracle(prec, tol, mask, att):
    spec = _spec(MIXED_TYPES, mask, att)
    assert spec.elem_theta is not None
    o = oracle.Oracle(spec)
    g = _solver(spec, prec)
    rng = np.random.default_rng(7)
    for call in range(2 if att else 1):
        u = rng.standard_normal((o.nmodes, 3)) + 1j * rng.standard_normal(
            (o.nmodes, 3))
        g.set_field("displ", u)
        g.set_field("stiff", np.zeros_like(u))
        g.computeStiffness()
        err = _rel(g.get_field("stiff"), o.stiffness(u))
        assert err < tol, (call, err)
    g.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 1e-5)])
def test_ti_cartesian_matches_oracle(prec, tol):
    spec = _spec(TI_ONLY, HALF_3D, False, spherical=False)
    assert getattr(spec, "elem_theta", None) is None
    o = oracle.Oracle(spec)
    g = _solver(spec, prec)
    rng = np.random.default_rng(8)
    u = rng.standard_normal((o.nmodes, 3)) + 1j * rng.standard_normal(
        (o.nmodes, 3))
    g.set_field("displ", u)
    g.set_field("stiff", np.zeros_like(u))
    g.computeStiffness()
    assert _rel(g.get_field("stiff"), o.stiffness(u)) < tol
    g.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-4)])
def test_media_seismograms_match_oracle(prec, tol):
    nsteps = 200
    spec = _spec(MIXED_TYPES, HALF_3D, True, nsteps=nsteps)
    ref = oracle.Oracle(spec).run(nsteps)
    g = _solver(spec, prec)
    g.run(nsteps)
    got = g.seismograms()
    g.close()
    assert np.abs(ref).max() > 0
    err = _rel(got, ref)
    print(f"TI/anisotropic {prec}: relative seismogram L2 = {err:.3e}")
    assert err < tol
