"""GPU parity of the Fourier-path solid stiffness loop against the FP64
oracle, through the C ABI (``include/axs3d.h``).

Tolerances: FP64 mode agrees with the oracle to rounding (relative
1e-12 per stage, relative seismogram L2 1e-12 after hundreds of steps);
FP32+FP64 mode is bounded by the north-star seismogram bar, relative L2
<= 1e-4 against FP64, and 1e-5 per stage.
"""
import numpy as np
import pytest

import oracle
from axisem3d_amd import preloop

pytestmark = pytest.mark.gpu

TOL_STAGE = {"fp64": 1e-12, "fp32": 1e-5}


def _solver(spec, prec, **kw):
    from axisem3d_amd.solver import GpuSolver
    return GpuSolver(spec, prec, **kw)


def _varnr_spec(nsteps=300, seed=3):
    # Nr from 3 to 12 across the mesh: odd and even Nr, zero padding
    def nodal_nr(s, z):
        return 3.0 + 9.0 * (s / 6e3) + 0.0 * z
    return preloop.cartesian_homogeneous(
        nx=6, nz=7, width=6e3, depth=5.6e3, nodal_nr=nodal_nr, nsteps=nsteps,
        src_depth=2.0e3, n_receivers=4, seed=seed)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _random_field(nmodes, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((nmodes, 3)) + 1j * rng.standard_normal(
        (nmodes, 3))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_stiffness_matches_oracle(prec):
    spec = _varnr_spec()
    assert len(set((spec.solid_nr % 2).tolist())) == 2   # odd and even Nr
    o = oracle.Oracle(spec)
    g = _solver(spec, prec)
    u = _random_field(o.nmodes, 11)
    g.set_field("displ", u)
    g.set_field("stiff", np.zeros_like(u))
    g.computeStiffness()
    got = g.get_field("stiff")
    ref = o.stiffness(u)
    assert _rel(got, ref) < TOL_STAGE[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_moment_source_matches_oracle(prec):
    spec = _varnr_spec()
    o = oracle.Oracle(spec)
    g = _solver(spec, prec)
    z = np.zeros((o.nmodes, 3), complex)
    tstep = int(np.argmax(np.abs(spec.stf)))
    g.set_field("stiff", z)
    g.applySources(tstep)
    got = g.get_field("stiff")
    ref = o.apply_source(z, tstep)
    assert np.linalg.norm(ref) > 0
    assert _rel(got, ref) < TOL_STAGE[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_point_phases_match_oracle(prec):
    spec = _varnr_spec()
    o = oracle.Oracle(spec)
    g = _solver(spec, prec)
    s0 = _random_field(o.nmodes, 5)
    # axial masking
    g.set_field("stiff", s0)
    g.applyBC_AfterAssemblingStiff()
    s1 = o.axial(s0)
    assert _rel(g.get_field("stiff"), s1) < TOL_STAGE[prec]
    # Nyquist + mass
    g.computeStiffToAccel()
    s2 = o.stiff_to_accel(s1)
    assert _rel(g.get_field("stiff"), s2) < TOL_STAGE[prec]
    # Newmark from random state
    u, v, a = (_random_field(o.nmodes, k) for k in (6, 7, 8))
    for name, f in (("displ", u), ("veloc", v), ("accel", a)):
        g.set_field(name, f)
    g.set_field("stiff", s2)
    g.newmark()
    s3, u3, v3, a3 = o.newmark(s2, u, v, a)
    tol = max(1e-12, TOL_STAGE[prec])
    assert _rel(g.get_field("displ"), u3) < tol
    assert _rel(g.get_field("veloc"), v3) < tol
    assert _rel(g.get_field("accel"), a3) < TOL_STAGE[prec]
    assert np.all(g.get_field("stiff") == 0)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-4)])
def test_seismograms_match_oracle(prec, tol):
    nsteps = 400
    spec = _varnr_spec(nsteps=nsteps)
    o = oracle.Oracle(spec)
    ref = o.run(nsteps)
    g = _solver(spec, prec)
    g.run(nsteps)
    got = g.seismograms()
    assert np.abs(ref).max() > 0
    err = _rel(got, ref)
    print(f"{prec}: relative seismogram L2 vs FP64 oracle = {err:.3e}")
    assert err < tol
    # wavefield too
    assert _rel(g.get_field("displ"), o.displ) < tol


def test_cartesian_8000_elements_fp32_vs_fp64():
    """The 8000-element nu=2 workload over 600 steps: FP32+FP64 against the
    GPU FP64 path (itself pinned to the oracle above), and the FP64 path
    against the oracle over the first 60 steps."""
    nsteps = 600
    spec = preloop.cartesian_homogeneous(nsteps=nsteps)
    g64 = _solver(spec, "fp64")
    g64.run(nsteps)
    g32 = _solver(spec, "fp32")
    g32.run(nsteps)
    s64, s32 = g64.seismograms(), g32.seismograms()
    assert np.abs(s64).max() > 0
    err = _rel(s32, s64)
    print(f"cartesian 8000: FP32 vs FP64 relative L2 = {err:.3e}")
    assert err < 1e-4
    o = oracle.Oracle(spec)
    ref = o.run(60)
    g = _solver(spec, "fp64")
    g.run(60)
    assert _rel(g.get_field("displ"), o.displ) < 1e-12


def test_stability_flag_raises():
    spec = _varnr_spec()
    g = _solver(spec, "fp64")
    u = np.zeros((g.nmodes, 3), complex)
    u[5, 1] = np.nan
    g.set_field("displ", u)
    from axisem3d_amd.capi import Axs3dError
    with pytest.raises(Axs3dError, match="checkStability"):
        g.run(1)
    g.close()
    # the message names the step at which the reference's per-step check
    # (NewmarkTimeScheme.cpp:72-74, at the start of step tstep, reported
    # as tstep + 1) sees the non-finite displacement, not the end of the
    # batch: started from a NaN at step 3, it reports 4
    for graph in (True, False):
        g = _solver(spec, "fp64")
        g.use_graph(graph)
        g.run(3)
        u = g.get_field("displ")
        u[5, 1] = np.nan
        g.set_field("displ", u)
        with pytest.raises(Axs3dError, match=r"current step  =  4\b"):
            g.run(10)
        g.close()
