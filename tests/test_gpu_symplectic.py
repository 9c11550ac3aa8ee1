"""GPU parity of the symplectic time scheme (SymplecticTimeScheme,
src/core/time/SymplecticTimeScheme.cpp:18-142: per step the receivers, the
source, u += kappa_0 dt v, then four sub-steps of stiffness, boundaries,
mass and v += pi_k dt a, u += kappa_k dt v) against the FP64 oracle
through the C ABI, in the bulk form (axs3d_run) and the per-phase form.
Domains: a Cartesian box (element Nr 3 to 12) with absorbing boundaries,
and a spherical sector with every other element 3D and CG4 attenuation.
Tolerances: FP64 relative 1e-12, FP32+FP64 relative 1e-4."""
import math

import numpy as np
import pytest

import oracle
from axisem3d_amd import preloop

pytestmark = pytest.mark.gpu

NSTEPS = 150


def _box():
    mesh = preloop.StructuredMesh(np.linspace(0, 6e3, 7),
                                  np.linspace(0, 5.6e3, 8))
    sp = preloop.build_domain(
        mesh, lambda r: (2600.0 + 0 * r, 6000.0 + 0 * r, 3464.0 + 0 * r),
        lambda s, z: 3.0 + 9.0 * (s / 6e3) + 0.0 * z, courant=0.5)
    preloop.add_absorbing(sp, keys=("RIGHT", "BOTTOM"), sponge=True,
                          rel_spans=[0.25, 0.25], t0=0.3)
    mt = np.array([1.0, 0.7, -0.4, 0.3, 0.2, -0.1]) * 1e15
    preloop.add_axial_moment_source(sp, 3.6e3, mt, 1.628 / (2 * math.pi),
                                    1.628, NSTEPS)
    preloop.add_receivers(sp, np.array([0.5e3, 3e3, 5.9e3, 5e3]),
                          np.array([5.6e3, 5.6e3, 5.6e3, 0.5e3]),
                          np.array([0.3, 1.1, 2.0, 4.0]))
    sp.time_scheme = "symplectic"
    return sp


def _sector():
    sp = preloop.spherical_regional(
        ntheta=8, nrad=6, theta_max_deg=2.0, depth=60e3, nu_min=3,
        nu_max=12, nsteps=NSTEPS, n_receivers=4, attenuation=True, cg4=True,
        corr_len=10e3)
    E = sp.n_solid_elements
    sp._build_3d((np.arange(E) % 2) == 1)
    preloop.finish_regional(sp, NSTEPS, src_depth=20e3, f_peak=0.5,
                            n_receivers=4, theta_max_deg=2.0)
    sp.time_scheme = "symplectic"
    return sp


SPECS = {"box": _box, "sector3d": _sector}


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-4)])
@pytest.mark.parametrize("name", sorted(SPECS))
def test_symplectic_seismograms_match_oracle(name, prec, tol):
    from axisem3d_amd.solver import GpuSolver
    sp = SPECS[name]()
    o = oracle.Oracle(sp)
    ref = o.run(NSTEPS)
    g = GpuSolver(sp, prec)
    g.run(NSTEPS)
    err = _rel(g.seismograms(), ref)
    uerr = _rel(g.get_field("displ"), o.displ)
    print(f"symplectic {name} {prec}: seismogram L2 {err:.3e}, "
          f"displacement L2 {uerr:.3e}")
    assert np.abs(ref).max() > 0
    assert err < tol
    if prec == "fp64":
        assert uerr < 1e-12
    g.close()


@pytest.mark.parametrize("name", sorted(SPECS))
def test_symplectic_per_phase_matches_bulk(name):
    from axisem3d_amd.solver import GpuSolver
    sp = SPECS[name]()
    n = 60
    a = GpuSolver(sp, "fp64")
    a.run(n)
    b = GpuSolver(sp, "fp64")
    for t in range(n):
        b.recordOutput(t)
        b.applySources(t)
        b.symplecticLaunch()
        for isub in range(1, 5):
            b.computeStiffness()
            b.applyBC_BeforeAssemblingStiff()
            b.applyBC_AfterAssemblingStiff()
            b.computeStiffToAccel()
            b.applyBC_AfterComputingAccel()
            b.symplecticUpdate(isub)
    err = _rel(b.seismograms(n), a.seismograms(n))
    print(f"symplectic {name}: per-phase vs bulk seismogram L2 {err:.3e}")
    assert err < 1e-12
    a.close()
    b.close()
