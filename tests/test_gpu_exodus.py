"""Exodus meshes of the reference through the GPU path against the FP64
oracle.

The SEG salt local example (examples/03_salt_body_SEG_local/
input1D: its Exodus mesh, 1900 solid and 100 fluid elements, and its
inparam keys, from the committed fixture) set up by preloop.from_inparam
and run through the C ABI against the FP64 oracle: Clayton-Enquist and
Kosloff boundaries on the right and bottom, the 200 m water layer with its
free surface, the axial source with the inputs' Gaussian STF, Nr as the
inputs give it (1) and with Nr 9 everywhere, 5 receivers below the sea
floor.  The regional PREM mesh (spherical and semi-spherical elements,
transversely isotropic upper mantle, CG4 attenuation from the mesh's Q
and SLS, Clayton-Enquist and Kosloff boundaries on its right side and at
the core-mantle boundary, Nr 5, an axial source at 12 km depth, 4 surface
receivers).  The global PREM mesh of the anisotropy example (the fluid
outer core coupled at both core boundaries, linear elements at the
centre, transversely isotropic upper mantle, FULL attenuation, Nr 5, a
source at 12 km depth on the axis, 5 surface receivers from the pole to
the antipode).  Tolerances: seismograms FP64 relative L2 1e-12, FP32+FP64
1e-4."""
import numpy as np
import pytest

import oracle
from axisem3d_amd import preloop
from test_exodus import (fixture_inparam, fixture_mesh, prem_global_mesh,
                         prem_regional_mesh)

pytestmark = pytest.mark.gpu


def _spec(nr, nsteps):
    inp = fixture_inparam()
    inp["nr"] = dict(inp["nr"], constant=nr)
    ex = fixture_mesh()
    z = float(ex.coords[:, 1].max()) - 250.0
    return preloop.from_inparam(
        inp, ex, nsteps, receivers_s=[0.0, 1e3, 2e3, 3.5e3, 4.9e3],
        receivers_z=[z] * 5, receivers_phi=[0.0, 0.5, 1.0, 2.0, 3.0])


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-4)])
@pytest.mark.parametrize("nr,nsteps", [(1, 1200), (9, 500)])
def test_salt_example_matches_oracle(nr, nsteps, prec, tol):
    from axisem3d_amd.solver import GpuSolver
    sp = _spec(nr, nsteps)
    oracle.set_threads(8)
    try:
        ref = oracle.Oracle(sp).run(nsteps)
    finally:
        oracle.set_threads(1)
    g = GpuSolver(sp, prec)
    g.run(nsteps)
    got = g.seismograms()
    g.close()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"salt example Nr {nr} {prec}: {nsteps} steps, seismogram L2 "
          f"{err:.3e}")
    assert np.abs(ref).max() > 0
    assert err < tol


def _prem_spec(nsteps):
    ex = prem_regional_mesh()
    sp = preloop.from_exodus(ex, lambda s, z: np.full(np.shape(s), 5.0),
                             courant=0.6, attenuation="CG4")
    preloop.add_absorbing(sp, keys=("RIGHT", "BOTTOM"), sponge=True,
                          rel_spans=[0.05, 0.05],
                          t0=ex.global_variables["minimum_period"])
    R = float(np.hypot(ex.coords[:, 0], ex.coords[:, 1]).max())
    mt = np.array([4.71e24, 3.81e22, -4.74e24, 3.99e23, -8.05e23,
                   -1.23e24]) * 1e-7
    preloop.add_axial_moment_source(sp, R - 12e3, mt, 10.0, 1.628, nsteps,
                                    stf_order=0)
    th = np.radians([0.5, 1.0, 2.0, 4.0])
    preloop.add_receivers(sp, (R - 1e3) * np.sin(th), (R - 1e3) * np.cos(th),
                          np.array([0.0, 0.7, 1.5, 3.0]))
    return sp


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-4)])
def test_prem_regional_matches_oracle(prec, tol):
    from axisem3d_amd.solver import GpuSolver
    nsteps = 1200
    sp = _prem_spec(nsteps)
    oracle.set_threads(8)
    try:
        ref = oracle.Oracle(sp).run(nsteps)
    finally:
        oracle.set_threads(1)
    g = GpuSolver(sp, prec)
    g.run(nsteps)
    got = g.seismograms()
    g.close()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"PREM regional {prec}: {nsteps} steps, seismogram L2 {err:.3e}")
    assert np.abs(ref).max() > 0
    assert err < tol


def _prem_global_spec(nsteps):
    ex = prem_global_mesh()
    sp = preloop.from_exodus(ex, lambda s, z: np.full(np.shape(s), 5.0),
                             courant=0.6, attenuation="FULL")
    R = float(np.hypot(ex.coords[:, 0], ex.coords[:, 1]).max())
    mt = np.array([1.0, 0.3, -0.8, 0.4, -0.2, 0.5]) * 1e20
    preloop.add_axial_moment_source(sp, R - 12e3, mt, 30.0, 1.628, nsteps,
                                    stf_order=0)
    th = np.radians([1.0, 10.0, 45.0, 120.0, 179.0])
    preloop.add_receivers(sp, (R - 1e3) * np.sin(th), (R - 1e3) * np.cos(th),
                          np.array([0.0, 0.7, 1.5, 3.0, 4.0]))
    return sp


_GLOBAL = {}


def _prem_global_reference(nsteps):
    """The oracle's run of the global PREM spec, computed once for both
    precisions."""
    if nsteps not in _GLOBAL:
        sp = _prem_global_spec(nsteps)
        oracle.set_threads(8)
        try:
            o = oracle.Oracle(sp)
            seis = o.run(nsteps)
        finally:
            oracle.set_threads(1)
        _GLOBAL[nsteps] = (sp, seis, o.fdispl.copy())
    return _GLOBAL[nsteps]


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-4)])
def test_prem_global_matches_oracle(prec, tol):
    """1800 steps: the P wave has crossed the outer core by then, so the
    fluid potential is compared too (FP64, relative 1e-11: a field, not a
    seismogram, whose small parts weigh in its norm)."""
    from axisem3d_amd.solver import GpuSolver
    nsteps = 1800
    sp, ref, fref = _prem_global_reference(nsteps)
    assert len(sp.fluid_elem_nr) > 0 and sp.elem_mtype.sum() > 0
    g = GpuSolver(sp, prec)
    g.run(nsteps)
    got = g.seismograms()
    ef = np.linalg.norm(g.get_field("fdispl") - fref) / np.linalg.norm(fref)
    g.close()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"PREM global {prec}: {nsteps} steps, seismogram L2 {err:.3e}, "
          f"fluid field {ef:.3e}")
    assert np.abs(ref).max() > 0
    assert err < tol
    if prec == "fp64":
        assert ef < 1e-11
