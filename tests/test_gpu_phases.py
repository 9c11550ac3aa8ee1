"""The per-phase entry points, called in the order of
NewmarkTimeScheme::solve (NewmarkTimeScheme.cpp:41-108), against the bulk
step (axs3d_run) on the same domain: seismograms and final displacement,
for a solid Fourier-path domain, a fluid layer with 3D coupling and 3D
fluid mass, and a domain mixing 1D and 3D elements with attenuation.
The two forms sum the source and the element forces in a different order,
so they agree to rounding, not bitwise: relative 1e-12 in FP64, 1e-4 in
FP32+FP64 (the FP32 stiffness rounding of 60 steps)."""
import math

import numpy as np
import pytest

from axisem3d_amd import preloop

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 1e-4}
NSTEPS = 60


def _solid_spec():
    # Nr from 3 to 12 across the mesh, Fourier-path solid elements
    return preloop.cartesian_homogeneous(
        nx=6, nz=7, width=6e3, depth=5.6e3,
        nodal_nr=lambda s, z: 3.0 + 9.0 * (s / 6e3) + 0.0 * z,
        nsteps=300, src_depth=2.0e3, n_receivers=4, seed=3)


def _fluid_spec():
    # a fluid layer inside a solid; 3D fluid mass and 3D coupling normals
    mesh = preloop.StructuredMesh(np.linspace(0, 4e3, 5),
                                  np.linspace(0, 6e3, 7))
    lo, hi = 2e3, 4e3

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
                                    1.628, 300)
    preloop.add_receivers(sp, np.array([0.5e3, 2e3, 3.5e3]),
                          np.array([5.5e3, 5.5e3, 1.5e3]),
                          np.array([0.3, 1.1, 2.0]))
    pts = np.unique(np.concatenate([np.arange(0, len(sp.fluid_nr), 3),
                                    sp.sf_fluid[::2]]))
    preloop.make_fluid_mass3d(sp, pts, amp=0.2, seed=7)
    return preloop.make_coupling3d(sp, amp=0.15, tilt=0.15, seed=8)


def _mixed_att_spec():
    # spherical regional domain, every third element 3D (cardinal path),
    # CG4 attenuation on all of them
    spec = preloop.spherical_regional(
        ntheta=6, nrad=6, theta_max_deg=2.0, depth=60e3, nu_min=3,
        nu_max=12, nsteps=300, n_receivers=3, attenuation=True, cg4=True,
        corr_len=10e3)
    E = spec.n_solid_elements
    spec._build_3d((np.arange(E) % 3) == 1)
    preloop.finish_regional(spec, 300, src_depth=20e3, f_peak=0.5,
                            n_receivers=3, theta_max_deg=2.0)
    return spec


SPECS = {"solid": _solid_spec, "fluid": _fluid_spec,
         "mixed3d_att": _mixed_att_spec}


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("case", sorted(SPECS))
def test_phase_calls_match_bulk_step(case, prec):
    from axisem3d_amd.solver import GpuSolver
    spec = SPECS[case]()
    bulk = GpuSolver(spec, prec)
    phases = GpuSolver(spec, prec)
    try:
        bulk.run(NSTEPS)
        for t in range(NSTEPS):
            phases.applySources(t)
            phases.computeStiffness()
            phases.applyBC_BeforeAssemblingStiff()
            phases.recordOutput(t)
            phases.applyBC_AfterAssemblingStiff()
            phases.computeStiffToAccel()
            phases.newmark()
        phases.checkStability(NSTEPS)
        ref = bulk.seismograms(NSTEPS)
        got = phases.seismograms(NSTEPS)
        assert np.abs(ref).max() > 0
        err = _rel(got, ref)
        print(f"{case} {prec}: phase calls vs bulk step, seismogram "
              f"relative L2 = {err:.3e}")
        assert err < TOL[prec]
        assert _rel(phases.get_field("displ"), bulk.get_field("displ")) \
            < TOL[prec]
        if case == "fluid":
            assert _rel(phases.get_field("fdispl"),
                        bulk.get_field("fdispl")) < TOL[prec]
    finally:
        bulk.close()
        phases.close()
