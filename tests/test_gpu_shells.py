"""A small version of the full-anisotropy shell workload: a spherical
sector of solid-fluid-solid shells (the middle third of the radius range
acoustic), every solid element 3D with a 21-Cij tensor per GLL point and
slice (seeded +-5% perturbations of the layered model), 5-SLS CG4
attenuation, element Nr from 273 up to 1050 (above 832), both
solid-fluid interfaces coupled, against the FP64 oracle through the C ABI.

The in-plane mesh is coarse next to the azimuthal resolution, so the
Courant number is divided by 60 to keep the step stable.

Tolerances: stage stiffness (solid and fluid, from a random state) FP64
relative 1e-13, FP32+FP64 1e-5; seismograms and fields after 150 steps
FP64 1e-12, FP32+FP64 1e-4."""
import numpy as np
import pytest

import oracle
from axisem3d_amd import preloop


def shells_spec(nsteps=150):
    sp = preloop.spherical_regional(
        ntheta=4, nrad=6, theta_max_deg=6.0, depth=600e3, nu_min=20,
        nu_max=520, nsteps=nsteps, n_receivers=3, attenuation=True,
        corr_len=100e3, bound_nr=False)
    ix, iz = sp._mesh.element_ix_iz()
    fluid = (iz == 2) | (iz == 3)
    att = dict(sp._att_in)
    att["qmu"], att["qkp"] = att["qmu"][~fluid], att["qkp"][~fluid]
    preloop.make_fluid(sp, fluid)
    sp._courant /= 60.0
    E = sp.n_solid_elements
    preloop.make_media(sp, np.full(E, 2, np.uint8), field=sp._field,
                       mask3d=np.ones(E, bool), att=att, seed=4)
    preloop.finish_regional(sp, nsteps, src_depth=50e3, f_peak=1.0,
                            n_receivers=3, theta_max_deg=6.0)
    return sp


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_shells_spec_is_what_it_claims():
    sp = shells_spec()
    assert sp.elem_nr.max() > 832
    assert np.all(np.asarray(sp.elem_mtype) == 2)
    assert np.all(np.asarray(sp.elem_kind) == 1)
    assert sp.att["nsls"] == 5 and sp.att["cg4"]
    assert len(sp.fluid_elem_nr) > 0
    # couplings on both the inner and the outer interface
    r = np.hypot(sp.solid_sz[sp.sf_solid, 0], sp.solid_sz[sp.sf_solid, 1])
    assert len(np.unique(np.round(r, -3))) == 2


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 1e-5)])
def test_shells_stiffness_matches_oracle(prec, tol):
    from axisem3d_amd.solver import GpuSolver
    sp = shells_spec()
    o = oracle.Oracle(sp)
    g = GpuSolver(sp, prec)
    rng = np.random.default_rng(21)
    u = rng.standard_normal((o.nmodes, 3)) + 1j * rng.standard_normal(
        (o.nmodes, 3))
    chi = rng.standard_normal(o.fnmodes) + 1j * rng.standard_normal(
        o.fnmodes)
    g.set_field("displ", u)
    g.set_field("fdispl", chi)
    g.set_field("stiff", np.zeros_like(u))
    g.set_field("fstiff", np.zeros_like(chi))
    g.computeStiffness()
    ref_s, ref_f = o.stiffness(u), o.fluid_stiffness(chi)
    es = _rel(g.get_field("stiff"), ref_s)
    ef = _rel(g.get_field("fstiff"), ref_f)
    print(f"shells {prec}: solid stiffness {es:.3e}, fluid {ef:.3e}")
    assert es < tol and ef < tol
    # the coupling on both interfaces, on exact inputs
    g.set_field("stiff", ref_s)
    g.set_field("fstiff", ref_f)
    g.applyBC_BeforeAssemblingStiff()
    ws, wf = o.coupling(u, ref_s, ref_f)
    assert _rel(g.get_field("fstiff"), wf) < tol
    assert _rel(g.get_field("stiff"), ws) < tol
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-4)])
def test_shells_seismograms_match_oracle(prec, tol):
    from axisem3d_amd.solver import GpuSolver
    nsteps = 150
    sp = shells_spec(nsteps)
    oracle.set_threads(8)
    try:
        o = oracle.Oracle(sp)
        ref = o.run(nsteps)
    finally:
        oracle.set_threads(1)
    g = GpuSolver(sp, prec)
    g.run(nsteps)
    err = _rel(g.seismograms(), ref)
    eu = _rel(g.get_field("displ"), o.displ)
    ef = _rel(g.get_field("fdispl"), o.fdispl)
    print(f"shells {prec}: seismogram L2 {err:.3e}, solid field {eu:.3e}, "
          f"fluid field {ef:.3e}")
    assert np.abs(ref).max() > 0
    assert err < tol and eu < tol and ef < tol
    g.close()
