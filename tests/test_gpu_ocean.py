"""GPU parity of the ocean-load mass (MassOceanLoad1D in the point
kernel, MassOceanLoad3D through the per-slice mass kernel) against the
FP64 oracle through the C ABI: the stage (axs3d_stiff_to_accel) on a
random stiffness, whole runs whose surface receivers sit on the loaded
points, and the per-phase form against the bulk step.

Tolerances: stage FP64 relative 1e-13, FP32+FP64 1e-5; seismograms FP64
1e-12, FP32+FP64 1e-4."""
import numpy as np
import pytest

import oracle
from axisem3d_amd import preloop

pytestmark = pytest.mark.gpu

SRD = 1.03e6
NSTEPS = 200


def _sector(kind):
    sp = preloop.spherical_regional(
        ntheta=6, nrad=4, theta_max_deg=2.0, depth=60e3, nu_min=3,
        nu_max=12, nsteps=NSTEPS, n_receivers=4, attenuation=False,
        corr_len=10e3)
    if kind == "3d-mass":
        # a 3D mass on every other top point: those get the 3D form
        top = np.nonzero(np.abs(np.hypot(sp.solid_sz[:, 0],
                                         sp.solid_sz[:, 1]) - 6371e3) < 1.0)[0]
        preloop.make_mass3d(sp, top[::2], amp=0.2, seed=3)
        srd = lambda s, z, phi: SRD
    elif kind == "3d-ocean":
        srd = lambda s, z, phi: SRD * (1.0 + 0.3 * np.cos(phi) +
                                       0.1 * np.sin(2 * phi))
    else:
        srd = lambda s, z, phi: SRD
    preloop.add_ocean_load(sp, srd)
    preloop.finish_regional(sp, NSTEPS, src_depth=15e3, f_peak=0.5,
                            n_receivers=4, theta_max_deg=2.0)
    return sp


KINDS = ["1d", "3d-mass", "3d-ocean"]


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 1e-5)])
@pytest.mark.parametrize("kind", KINDS)
def test_ocean_stage_matches_oracle(kind, prec, tol):
    from axisem3d_amd.solver import GpuSolver
    sp = _sector(kind)
    nv = sp.ocean_load["nval"]
    assert (np.any(nv > 1)) == (kind != "1d")
    o = oracle.Oracle(sp)
    g = GpuSolver(sp, prec)
    rng = np.random.default_rng(13)
    f = rng.standard_normal((o.nmodes, 3)) + 1j * rng.standard_normal(
        (o.nmodes, 3))
    g.set_field("stiff", f)
    g.computeStiffToAccel()
    err = _rel(g.get_field("stiff"), o.stiff_to_accel(f.copy()))
    print(f"ocean {kind} {prec}: stage {err:.3e}")
    assert err < tol
    g.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-4)])
@pytest.mark.parametrize("kind", KINDS)
def test_ocean_seismograms_match_oracle(kind, prec, tol):
    from axisem3d_amd.solver import GpuSolver
    sp = _sector(kind)
    o = oracle.Oracle(sp)
    ref = o.run(NSTEPS)
    g = GpuSolver(sp, prec)
    g.run(NSTEPS)
    err = _rel(g.seismograms(), ref)
    eu = _rel(g.get_field("displ"), o.displ)
    print(f"ocean {kind} {prec}: seismogram L2 {err:.3e}, field {eu:.3e}")
    assert np.abs(ref).max() > 0
    assert err < tol and eu < tol
    g.close()


@pytest.mark.parametrize("kind", KINDS)
def test_ocean_per_phase_matches_bulk(kind):
    from axisem3d_amd.solver import GpuSolver
    sp = _sector(kind)
    n = 80
    a = GpuSolver(sp, "fp64")
    a.run(n)
    b = GpuSolver(sp, "fp64")
    for t in range(n):
        b.applySources(t)
        b.computeStiffness()
        b.applyBC_BeforeAssemblingStiff()
        b.recordOutput(t)
        b.applyBC_AfterAssemblingStiff()
        b.computeStiffToAccel()
        b.applyBC_AfterComputingAccel()
        b.newmark()
    err = _rel(b.seismograms(n), a.seismograms(n))
    print(f"ocean {kind}: per-phase vs bulk {err:.3e}")
    assert err < 1e-12
    a.close()
    b.close()
