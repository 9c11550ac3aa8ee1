"""GPU parity of SLS attenuation on the Fourier path (1D-material
elements) and in domains that mix 1D and 3D elements, against the FP64
oracle through the C ABI.

Tolerances: FP64 mode agrees with the oracle to rounding (relative 1e-13
per stage, 1e-12 seismogram L2); FP32+FP64 mode relative 1e-5 per stage
and 1e-4 seismogram L2."""
import numpy as np
import pytest

import oracle
from axisem3d_amd import preloop

pytestmark = pytest.mark.gpu


def _spec(mask3d, nsteps=300, cg4=True):
    spec = preloop.spherical_regional(
        ntheta=6, nrad=6, theta_max_deg=2.0, depth=60e3, nu_min=3,
        nu_max=12, nsteps=nsteps, n_receivers=3, attenuation=True, cg4=cg4,
        corr_len=10e3)
    E = spec.n_solid_elements
    spec._build_3d(np.asarray(mask3d(E), dtype=bool))
    preloop.finish_regional(spec, nsteps, src_depth=20e3, f_peak=0.5,
                            n_receivers=3, theta_max_deg=2.0)
    return spec


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _solver(spec, prec):
    from axisem3d_amd.solver import GpuSolver
    return GpuSolver(spec, prec)


ALL_1D = lambda E: np.zeros(E, bool)
MIXED = lambda E: (np.arange(E) % 3) == 1


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 1e-5)])
@pytest.mark.parametrize("cg4", [True, False])
@pytest.mark.parametrize("mask", [ALL_1D, MIXED], ids=["1d", "mixed"])
def test_attenuation_stiffness_matches_oracle(prec, tol, cg4, mask):
    spec = _spec(mask, cg4=cg4)
    o = oracle.Oracle(spec)
    g = _solver(spec, prec)
    rng = np.random.default_rng(5)
    for call in range(3):
        # the memory variables carry state from call to call
        u = rng.standard_normal((o.nmodes, 3)) + 1j * rng.standard_normal(
            (o.nmodes, 3))
        g.set_field("displ", u)
        g.set_field("stiff", np.zeros_like(u))
        g.computeStiffness()
        err = _rel(g.get_field("stiff"), o.stiffness(u))
        assert err < tol, (call, err)
    g.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-4)])
@pytest.mark.parametrize("mask", [ALL_1D, MIXED], ids=["1d", "mixed"])
def test_attenuation_seismograms_match_oracle(prec, tol, mask):
    nsteps = 200
    spec = _spec(mask, nsteps=nsteps)
    ref = oracle.Oracle(spec).run(nsteps)
    g = _solver(spec, prec)
    g.run(nsteps)
    got = g.seismograms()
    g.close()
    assert np.abs(ref).max() > 0
    err = _rel(got, ref)
    print(f"attenuation {prec}: relative seismogram L2 = {err:.3e}")
    assert err < tol
