"""GPU parity of undulated geometry by particle relabelling (PRT branch of
SolidElement::displToStiff, SolidElement.cpp:115-159; axs3d_set_solid_prt)
against the FP64 oracle through the C ABI.

Domains: a Cartesian box whose top two element rows are undulated
(dz = a (z - z0)^2 cos(s / L), zero below z0) with 1D and 3D elements and
5-SLS attenuation, the same box undulated down past its moment source, and a spherical sector (isotropic, TI and anisotropic
elements in turn, half of them 3D, CG4 attenuation) whose top row is
undulated with an axisymmetric topography.  The GPU path applies the
operands to the Fourier coefficients before the transform (they do not
depend on phi); the oracle follows the reference's order (all nine series
through the FFT, the PRT per slice).

Tolerances: FP64 mode relative 1e-13 per stage and 1e-12 seismogram L2;
FP32+FP64 mode relative 1e-5 per stage and 1e-4 seismogram L2."""
import math

import numpy as np
import pytest

import oracle
from axisem3d_amd import preloop

pytestmark = pytest.mark.gpu

NSTEPS = 300


def _box(z0=4.0e3, a=0.08 / 1.6e3):
    mesh = preloop.StructuredMesh(np.linspace(0, 6e3, 7),
                                  np.linspace(0, 5.6e3, 8))
    sp = preloop.build_domain(
        mesh, lambda r: (2600.0 + 0 * r, 6000.0 + 0 * r, 3464.0 + 0 * r),
        lambda s, z: 3.0 + 9.0 * (s / 6e3) + 0.0 * z, courant=0.4)
    E = sp.n_solid_elements
    att = dict(nsls=5, cg4=False, f_min=0.05, f_max=5.0, f_ref=1.0,
               qkp=np.full((E, 25), 57823.0), qmu=np.full((E, 25), 80.0))
    preloop.make_3d_isotropic(
        sp, lambda x, y, z: 0.03 * np.sin(x / 900.0) * np.cos(y / 700.0),
        att, (np.arange(E) % 2) == 1)
    preloop.add_undulation(
        sp, lambda s, z: np.where(z > z0, a * (z - z0) ** 2, 0.0) *
        np.cos(s / 2.5e3))
    mt = np.array([1.0, 0.7, -0.4, 0.3, 0.2, -0.1]) * 1e15
    preloop.add_axial_moment_source(sp, 2.0e3, mt, 1.628 / (2 * math.pi),
                                    1.628, NSTEPS)
    preloop.add_receivers(sp, np.array([0.5e3, 3e3, 5.9e3, 5e3]),
                          np.array([5.6e3, 5.6e3, 5.2e3, 0.5e3]),
                          np.array([0.3, 1.1, 2.0, 4.0]))
    return sp


def _sector():
    sp = preloop.spherical_regional(
        ntheta=6, nrad=6, theta_max_deg=2.0, depth=60e3, nu_min=3,
        nu_max=12, nsteps=NSTEPS, n_receivers=3, attenuation=True, cg4=True,
        corr_len=10e3)
    E = sp.n_solid_elements
    preloop.make_media(sp, (np.arange(E) % 3).astype(np.uint8),
                       field=sp._field, mask3d=(np.arange(E) % 2) == 1,
                       att=sp._att_in, seed=2)
    r0 = 6371e3 - 10e3
    th_max = math.radians(2.0)

    def dz(s, z):
        r = np.hypot(s, z)
        th = np.arctan2(s, z)
        return np.where(r > r0, 400.0 * ((r - r0) / 10e3) ** 2, 0.0) * \
            np.cos(math.pi * th / th_max)
    preloop.add_undulation(sp, dz)
    preloop.finish_regional(sp, NSTEPS, src_depth=20e3, f_peak=0.5,
                            n_receivers=3, theta_max_deg=2.0)
    return sp


SPECS = {"box": _box, "sector": _sector,
         # the undulation reaches down past the source: the moment source
         # takes the PRT branch (SolidElement.cpp:293-331)
         "box_src": lambda: _box(1.6e3, 0.04 / 1.6e3)}


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 1e-5)])
@pytest.mark.parametrize("name", sorted(SPECS))
def test_prt_stiffness_matches_oracle(name, prec, tol):
    from axisem3d_amd.solver import GpuSolver
    sp = SPECS[name]()
    prt = sp.prt["elems"]
    assert 0 < len(prt) < sp.n_solid_elements
    assert np.any(np.asarray(sp.elem_kind)[prt] == 1)
    assert np.any(np.asarray(sp.elem_kind)[prt] == 0)
    assert np.any(sp.prt["x"][:, 1] != 0.0)
    o = oracle.Oracle(sp)
    g = GpuSolver(sp, prec)
    rng = np.random.default_rng(61)
    u = rng.standard_normal((o.nmodes, 3)) + 1j * rng.standard_normal(
        (o.nmodes, 3))
    g.set_field("displ", u)
    g.set_field("stiff", np.zeros_like(u))
    g.computeStiffness()
    err = _rel(g.get_field("stiff"), o.stiffness(u))
    print(f"PRT {name} {prec}: stiffness relative difference {err:.3e}")
    assert err < tol
    g.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-4)])
@pytest.mark.parametrize("name", sorted(SPECS))
def test_prt_seismograms_match_oracle(name, prec, tol):
    from axisem3d_amd.solver import GpuSolver
    sp = SPECS[name]()
    assert (int(sp.src_elem) in set(sp.prt["elems"].tolist())) == (
        name == "box_src")
    o = oracle.Oracle(sp)
    ref = o.run(NSTEPS)
    g = GpuSolver(sp, prec)
    g.run(NSTEPS)
    err = _rel(g.seismograms(), ref)
    uerr = _rel(g.get_field("displ"), o.displ)
    print(f"PRT {name} {prec}: seismogram L2 {err:.3e}, displacement L2 "
          f"{uerr:.3e}")
    assert np.abs(ref).max() > 0
    assert err < tol
    g.close()


def test_prt_adjoint_is_refused():
    """The adjoint run of an undulated domain is refused at the C ABI (the
    library raises before any kernel runs)."""
    from axisem3d_amd import capi
    from axisem3d_amd.solver import GpuSolver
    g = GpuSolver(_box(), "fp64")
    with pytest.raises(capi.Axs3dError, match="PRT"):
        g.adjoint_run(np.zeros((10, 4, 3)))
    g.close()
