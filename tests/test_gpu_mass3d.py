"""GPU parity of Mass3D (per-slice mass through C2R / R2C) against the
FP64 oracle, through the C ABI: the stage form (axs3d_stiff_to_accel) on a
random stiffness, then whole runs whose seismograms cross the 3D-mass
points.  Tolerances as for the other point phases: 1e-12 per stage in
FP64, 1e-5 in FP32+FP64; seismograms 1e-12 and 1e-4."""
import numpy as np
import pytest

import oracle
from axisem3d_amd import preloop

pytestmark = pytest.mark.gpu

TOL_STAGE = {"fp64": 1e-12, "fp32": 1e-5}


def _spec(nsteps=300, seed=4):
    # Nr from 3 to 12: odd and even Nr, axis points among the 3D masses
    def nodal_nr(s, z):
        return 3.0 + 9.0 * (s / 6e3) + 0.0 * z
    spec = preloop.cartesian_homogeneous(
        nx=6, nz=7, width=6e3, depth=5.6e3, nodal_nr=nodal_nr, nsteps=nsteps,
        src_depth=2.0e3, n_receivers=4, seed=seed)
    pts = np.arange(0, len(spec.solid_nr), 2)
    return preloop.make_mass3d(spec, pts, amp=0.2, seed=seed)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_mass3d_stage_matches_oracle(prec):
    from axisem3d_amd.solver import GpuSolver
    spec = _spec()
    o = oracle.Oracle(spec)
    g = GpuSolver(spec, prec)
    rng = np.random.default_rng(9)
    s0 = rng.standard_normal((o.nmodes, 3)) + 1j * rng.standard_normal(
        (o.nmodes, 3))
    g.set_field("stiff", s0)
    g.computeStiffToAccel()
    ref = o.stiff_to_accel(s0)
    assert _rel(g.get_field("stiff"), ref) < TOL_STAGE[prec]


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-4)])
def test_mass3d_seismograms_match_oracle(prec, tol):
    from axisem3d_amd.solver import GpuSolver
    nsteps = 300
    spec = _spec(nsteps=nsteps)
    o = oracle.Oracle(spec)
    ref = o.run(nsteps)
    g = GpuSolver(spec, prec)
    g.run(nsteps)
    got = g.seismograms()
    assert np.abs(ref).max() > 0
    err = _rel(got, ref)
    print(f"{prec}: Mass3D seismogram relative L2 vs FP64 oracle = {err:.3e}")
    assert err < tol
    assert _rel(g.get_field("displ"), o.displ) < tol
