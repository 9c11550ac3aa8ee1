"""GPU parity of fluid elements (displacement potential), the solid-fluid
coupling, the fluid axis and the fluid free surface against the FP64
oracle through the C ABI: a fluid layer inside a solid (two coupled
interfaces, fluid points on the axis) and an ocean on top of a solid (free
surface).

Tolerances: FP64 mode agrees with the oracle to rounding (relative 1e-13
per stage, 1e-12 seismogram L2); FP32+FP64 mode relative 1e-5 per stage
and 1e-4 seismogram L2."""
import math

import numpy as np
import pytest

import oracle
from axisem3d_amd import preloop

pytestmark = pytest.mark.gpu


def _spec(layer, nsteps=300):
    mesh = preloop.StructuredMesh(np.linspace(0, 4e3, 5),
                                  np.linspace(0, 6e3, 7))
    lo, hi = (2e3, 4e3) if layer == "middle" else (4e3, 6e3)

    def mat(r, rc):
        inl = (rc > lo) & (rc < hi)
        one = np.ones_like(r)
        return (np.where(inl, 1000.0, 2600.0) * one,
                np.where(inl, 1500.0, 6000.0) * one,
                np.where(inl, 0.0, 3464.0) * one)
    sp = preloop.build_domain(
        mesh, mat, lambda s, z: 3.0 + 8.0 * (s / 4e3) + 0.0 * z, courant=0.4)
    ix, iz = mesh.element_ix_iz()
    zc = (iz + 0.5) * 1e3
    preloop.make_fluid(sp, (zc > lo) & (zc < hi))
    mt = np.array([1.0, 0.7, 0.4, 0.3, 0.2, 0.1]) * 1e15
    preloop.add_axial_moment_source(sp, 1e3, mt, 1.628 / (2 * math.pi),
                                    1.628, nsteps)
    zr = 5.5e3 if layer == "middle" else 3.5e3
    preloop.add_receivers(sp, np.array([0.5e3, 2e3, 3.5e3]),
                          np.array([zr, zr, 1.5e3]),
                          np.array([0.3, 1.1, 2.0]))
    return sp


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _solver(spec, prec):
    from axisem3d_amd.solver import GpuSolver
    return GpuSolver(spec, prec)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 1e-5)])
@pytest.mark.parametrize("layer", ["middle", "ocean"])
def test_fluid_stiffness_and_coupling_match_oracle(prec, tol, layer):
    spec = _spec(layer)
    assert len(spec.sf_solid) > 0
    assert (len(spec.fluid_surface_points) > 0) == (layer == "ocean")
    o = oracle.Oracle(spec)
    g = _solver(spec, prec)
    assert g.fnmodes == o.fnmodes
    rng = np.random.default_rng(31)
    u = rng.standard_normal((o.nmodes, 3)) + 1j * rng.standard_normal(
        (o.nmodes, 3))
    chi = rng.standard_normal(o.fnmodes) + 1j * rng.standard_normal(
        o.fnmodes)
    g.set_field("displ", u)
    g.set_field("fdispl", chi)
    g.set_field("stiff", np.zeros_like(u))
    g.set_field("fstiff", np.zeros_like(chi))
    g.computeStiffness()
    ks, kf = o.stiffness(u), o.fluid_stiffness(chi)
    assert _rel(g.get_field("stiff"), ks) < tol
    assert _rel(g.get_field("fstiff"), kf) < tol
    # coupling on exact inputs
    g.set_field("stiff", ks)
    g.set_field("fstiff", kf)
    g.applyBC_BeforeAssemblingStiff()
    ws, wf = o.coupling(u, ks, kf)
    assert _rel(g.get_field("fstiff"), wf) < tol
    assert _rel(g.get_field("stiff"), ws) < tol
    g.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-4)])
@pytest.mark.parametrize("layer", ["middle", "ocean"])
def test_fluid_seismograms_match_oracle(prec, tol, layer):
    nsteps = 300
    spec = _spec(layer, nsteps)
    o = oracle.Oracle(spec)
    ref = o.run(nsteps)
    g = _solver(spec, prec)
    g.run(nsteps)
    got = g.seismograms()
    chi = g.get_field("fdispl")
    g.close()
    assert np.abs(ref).max() > 0 and np.abs(o.fdispl).max() > 0
    err = _rel(got, ref)
    print(f"fluid {layer} {prec}: relative seismogram L2 = {err:.3e}")
    assert err < tol
    assert _rel(chi, o.fdispl) < 10 * tol


def _spec3d(layer, nsteps=300):
    # element Nr 5 to 11 (point Nr 3 to 11), half the fluid elements 3D,
    # two of them on the axis
    sp = _spec(layer, nsteps)
    nfe = len(sp.fluid_elem_nr)
    return preloop.make_fluid3d(sp, np.arange(0, nfe, 2), amp=0.3, seed=5)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 1e-5)])
@pytest.mark.parametrize("layer", ["middle", "ocean"])
def test_fluid3d_stiffness_matches_oracle(prec, tol, layer):
    spec = _spec3d(layer)
    assert spec.fluid_elem_kind.any() and not spec.fluid_elem_kind.all()
    o = oracle.Oracle(spec)
    g = _solver(spec, prec)
    rng = np.random.default_rng(41)
    chi = rng.standard_normal(o.fnmodes) + 1j * rng.standard_normal(
        o.fnmodes)
    g.set_field("displ", np.zeros((o.nmodes, 3), complex))
    g.set_field("fdispl", chi)
    g.set_field("stiff", np.zeros((o.nmodes, 3), complex))
    g.set_field("fstiff", np.zeros_like(chi))
    g.computeStiffness()
    kf = o.fluid_stiffness(chi)
    assert _rel(g.get_field("fstiff"), kf) < tol
    g.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-4)])
def test_fluid3d_seismograms_match_oracle(prec, tol):
    nsteps = 300
    spec = _spec3d("middle", nsteps)
    o = oracle.Oracle(spec)
    ref = o.run(nsteps)
    g = _solver(spec, prec)
    g.run(nsteps)
    got = g.seismograms()
    chi = g.get_field("fdispl")
    g.close()
    err = _rel(got, ref)
    print(f"fluid 3D {prec}: relative seismogram L2 = {err:.3e}")
    assert err < tol
    assert _rel(chi, o.fdispl) < 10 * tol


def _spec_c3d(layer, nsteps=300):
    # fluid 3D mass on every third fluid point (coupled ones among them),
    # all pairs 3D (undulated interface normals with a phi component)
    sp = _spec(layer, nsteps)
    pts = np.unique(np.concatenate([np.arange(0, len(sp.fluid_nr), 3),
                                    sp.sf_fluid[::2]]))
    preloop.make_fluid_mass3d(sp, pts, amp=0.2, seed=7)
    return preloop.make_coupling3d(sp, amp=0.15, tilt=0.15, seed=8)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 1e-5)])
@pytest.mark.parametrize("layer", ["middle", "ocean"])
def test_coupling3d_and_fluid_mass3d_stages_match_oracle(prec, tol, layer):
    spec = _spec_c3d(layer)
    assert len(spec.sf3d_pairs) == len(spec.sf_solid)
    o = oracle.Oracle(spec)
    g = _solver(spec, prec)
    rng = np.random.default_rng(51)
    u = rng.standard_normal((o.nmodes, 3)) + 1j * rng.standard_normal(
        (o.nmodes, 3))
    ks = rng.standard_normal((o.nmodes, 3)) + 1j * rng.standard_normal(
        (o.nmodes, 3))
    kf = rng.standard_normal(o.fnmodes) + 1j * rng.standard_normal(o.fnmodes)
    # coupling (applyBC_BeforeAssemblingStiff)
    g.set_field("displ", u)
    g.set_field("stiff", ks)
    g.set_field("fstiff", kf)
    g.applyBC_BeforeAssemblingStiff()
    ws, wf = o.coupling(u, ks, kf)
    assert _rel(g.get_field("fstiff"), wf) < tol
    assert _rel(g.get_field("stiff"), ws) < tol
    # stiffness to acceleration (fluid Mass3D)
    g.set_field("stiff", ks)
    g.set_field("fstiff", kf)
    g.computeStiffToAccel()
    assert _rel(g.get_field("fstiff"), o.fluid_stiff_to_accel(kf)) < tol
    assert _rel(g.get_field("stiff"), o.stiff_to_accel(ks)) < tol
    g.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-4)])
@pytest.mark.parametrize("layer", ["middle", "ocean"])
def test_coupling3d_seismograms_match_oracle(prec, tol, layer):
    nsteps = 300
    spec = _spec_c3d(layer, nsteps)
    o = oracle.Oracle(spec)
    ref = o.run(nsteps)
    g = _solver(spec, prec)
    g.run(nsteps)
    got = g.seismograms()
    chi = g.get_field("fdispl")
    g.close()
    err = _rel(got, ref)
    print(f"3D coupling {layer} {prec}: relative seismogram L2 = {err:.3e}")
    assert err < tol
    assert _rel(chi, o.fdispl) < 10 * tol
