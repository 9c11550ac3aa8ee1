"""The captured step graph (axs3d_run's default) against the same steps
launched kernel by kernel: both launch the same kernels on the same data,
so every field and every seismogram sample is bitwise equal.  The graph is
replayed in several runs with host reads of the fields between them, on a
domain of mixed 1D/3D attenuating elements and on a fluid layer with 3D
fluid elements, both with a source and receivers."""
import math

import numpy as np
import pytest

from axisem3d_amd import preloop

pytestmark = pytest.mark.gpu


def _regional(nsteps):
    spec = preloop.spherical_regional(
        ntheta=6, nrad=6, theta_max_deg=2.0, depth=60e3, nu_min=3,
        nu_max=12, nsteps=nsteps, n_receivers=3, attenuation=True, cg4=True,
        corr_len=10e3)
    spec._build_3d((np.arange(spec.n_solid_elements) % 3) == 1)
    preloop.finish_regional(spec, nsteps, src_depth=20e3, f_peak=0.5,
                            n_receivers=3, theta_max_deg=2.0)
    return spec


def _fluid(nsteps):
    mesh = preloop.StructuredMesh(np.linspace(0, 4e3, 5),
                                  np.linspace(0, 6e3, 7))

    def mat(r, rc):
        inl = (rc > 2e3) & (rc < 4e3)
        one = np.ones_like(r)
        return (np.where(inl, 1000.0, 2600.0) * one,
                np.where(inl, 1500.0, 6000.0) * one,
                np.where(inl, 0.0, 3464.0) * one)
    sp = preloop.build_domain(
        mesh, mat, lambda s, z: 3.0 + 8.0 * (s / 4e3) + 0.0 * z, courant=0.4)
    ix, iz = mesh.element_ix_iz()
    zc = (iz + 0.5) * 1e3
    preloop.make_fluid(sp, (zc > 2e3) & (zc < 4e3))
    preloop.make_fluid3d(sp, np.arange(0, len(sp.fluid_elem_nr), 2),
                         amp=0.3, seed=2)
    mt = np.array([1.0, 0.7, 0.4, 0.3, 0.2, 0.1]) * 1e15
    preloop.add_axial_moment_source(sp, 1e3, mt, 1.628 / (2 * math.pi),
                                    1.628, nsteps)
    preloop.add_receivers(sp, np.array([0.5e3, 2e3]), np.array([5.5e3, 1.5e3]),
                          np.array([0.3, 1.1]))
    return sp


FIELDS = ["displ", "veloc", "accel"]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("case", ["regional", "fluid"])
def test_graph_replay_equals_eager_launches(prec, case):
    from axisem3d_amd.solver import GpuSolver
    chunks = [40, 1, 59]
    spec = (_regional if case == "regional" else _fluid)(sum(chunks))
    fields = FIELDS + (["fdispl", "fveloc", "faccel"] if case == "fluid"
                       else [])
    gg = GpuSolver(spec, prec)
    ge = GpuSolver(spec, prec)
    ge.use_graph(False)
    for n in chunks:
        gg.run(n)
        ge.run(n)
        for f in fields:
            a, b = gg.get_field(f), ge.get_field(f)
            assert np.abs(b).max() > 0 or f != "displ"
            assert np.array_equal(a, b), (f, n)
    sg, se = gg.seismograms(), ge.seismograms()
    assert np.abs(se).max() > 0
    assert np.array_equal(sg, se)
    gg.close()
    ge.close()
