"""GPU parity of the absorbing boundaries (Clayton-Enquist, solid and
fluid, 1D and per-slice forms; the Kosloff sponge, 1D and per slice, on
points with a 1D or a 3D mass and on deferred fluid points) and of the
station frames, against the FP64 oracle through the C ABI.

Domains: a Cartesian solid box (element Nr 3 to 12) with Clayton and
sponge on its right side and bottom; a spherical sector whose every
other element has a 3D medium (per-slice Clayton and sponge there), with
CG4 attenuation and receivers in RTZ, ENZ and xyz frames; a fluid layer
reaching the right side (fluid Clayton and sponge, half the fluid
elements 3D, some boundary points with a 3D mass).

Tolerances: FP64 relative 1e-12 (seismograms, fields, stage terms),
FP32+FP64 relative 1e-4 (seismograms) and 1e-5 (stage terms)."""
import math

import numpy as np
import pytest

import oracle
from axisem3d_amd import preloop

pytestmark = pytest.mark.gpu

NSTEPS = 200
TOL = {"fp64": 1e-12, "fp32": 1e-4}


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
    return sp


def _sector():
    sp = preloop.spherical_regional(
        ntheta=8, nrad=6, theta_max_deg=2.0, depth=60e3, nu_min=3,
        nu_max=12, nsteps=NSTEPS, n_receivers=4, attenuation=True, cg4=True,
        corr_len=10e3)
    E = sp.n_solid_elements
    sp._build_3d((np.arange(E) % 2) == 1)
    preloop.add_absorbing(sp, keys=("RIGHT", "BOTTOM"), sponge=True,
                          rel_spans=[0.25, 0.25], t0=2.0)
    preloop.finish_regional(sp, NSTEPS, src_depth=20e3, f_peak=0.5,
                            n_receivers=4, theta_max_deg=2.0)
    th = np.arccos(np.clip(sp.solid_sz[:, 1] / np.hypot(
        sp.solid_sz[:, 0], sp.solid_sz[:, 1]), -1, 1))
    n = len(sp.rec_elem)
    rng = np.random.default_rng(9)
    sp.rec_frames = {"wcs": np.array([1, 2, 3, 0][:n] * (n // 4 + 1))[:n],
                     "theta": rng.uniform(0.0, 0.04, n),
                     "baz": rng.uniform(0.0, 2 * math.pi, n),
                     "cartesian": 0}
    del th
    return sp


def _fluid():
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
    nfe = len(sp.fluid_elem_nr)
    preloop.make_fluid3d(sp, np.arange(0, nfe, 2), amp=0.3, seed=5)
    preloop.add_absorbing(sp, keys=("RIGHT", "BOTTOM"), sponge=True,
                          rel_spans=[0.25, 0.2], t0=0.3)
    mt = np.array([1.0, 0.7, 0.4, 0.3, 0.2, 0.1]) * 1e15
    preloop.add_axial_moment_source(sp, 1e3, mt, 1.628 / (2 * math.pi),
                                    1.628, NSTEPS)
    preloop.add_receivers(sp, np.array([0.5e3, 2e3, 3.5e3]),
                          np.array([5.5e3, 5.5e3, 1.5e3]),
                          np.array([0.3, 1.1, 2.0]))
    # 3D masses on some boundary points: the solid points' k_mass3d and
    # the deferred fluid points take the sponge after their mass
    right_s = np.nonzero(np.abs(sp.solid_sz[:, 0] - 4e3) < 1e-6)[0]
    right_f = np.nonzero(np.abs(sp.fluid_sz[:, 0] - 4e3) < 1e-6)[0]
    preloop.make_mass3d(sp, right_s[::2], amp=0.2, seed=3)
    preloop.make_fluid_mass3d(sp, right_f[::2], amp=0.2, seed=4)
    # a fluid point with a 3D mass couples through SolidFluidCoupling3D
    return preloop.make_coupling3d(sp, amp=0.15, tilt=0.15, seed=8)


SPECS = {"box": _box, "sector3d": _sector, "fluid": _fluid}


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _oracle_seismograms(sp, nsteps):
    o = oracle.Oracle(sp)
    seis = o.run(nsteps)
    fr = getattr(sp, "rec_frames", None)
    if fr is not None:
        seis = oracle.rotate_station(seis, fr["wcs"], fr["theta"],
                                     fr["baz"], sp.rec_phi,
                                     fr["cartesian"])
    return o, seis


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("name", sorted(SPECS))
def test_absorbing_seismograms_match_oracle(name, prec):
    from axisem3d_amd.solver import GpuSolver
    sp = SPECS[name]()
    assert len(sp.clayton_solid["point"]) > 0
    assert len(sp.sponge_solid["point"]) > 0
    o, ref = _oracle_seismograms(sp, NSTEPS)
    g = GpuSolver(sp, prec)
    g.run(NSTEPS)
    got = g.seismograms()
    err = _rel(got, ref)
    u_err = _rel(g.get_field("displ"), o.displ)
    print(f"absorbing {name} {prec}: seismogram L2 {err:.3e}, "
          f"displacement L2 {u_err:.3e}")
    assert np.abs(ref).max() > 0
    assert err < TOL[prec]
    if prec == "fp64":
        assert u_err < 1e-12
        if o.fnmodes:
            assert _rel(g.get_field("fdispl"), o.fdispl) < 1e-12
    g.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
@pytest.mark.parametrize("name", ["box", "sector3d"])
def test_absorbing_stage_terms_match_oracle(name, prec, tol):
    """axs3d_apply_bc_before on a zero stiffness leaves minus the Clayton
    force of the velocity; axs3d_apply_bc_after_accel then takes the sponge
    term of (u, v) from a given acceleration."""
    from axisem3d_amd.solver import GpuSolver
    sp = SPECS[name]()
    o = oracle.Oracle(sp)
    g = GpuSolver(sp, prec)
    rng = np.random.default_rng(17)
    n = o.nmodes
    u, v, a = ((rng.standard_normal((n, 3)) + 1j * rng.standard_normal(
        (n, 3))) * 1e-3 for _ in range(3))
    g.set_field("displ", u)
    g.set_field("veloc", v)
    g.set_field("stiff", np.zeros((n, 3), complex))
    g.applyBC_BeforeAssemblingStiff()
    ref, _ = o.clayton(v, np.zeros((n, 3), complex))
    got = g.get_field("stiff")
    assert np.abs(ref).max() > 0
    assert _rel(got, ref) < tol
    g.set_field("stiff", a)
    g.applyBC_AfterComputingAccel()
    ref, _ = o.sponge(u, v, a)
    assert _rel(g.get_field("stiff") - a, ref - a) < tol
    g.close()


@pytest.mark.parametrize("name", sorted(SPECS))
def test_absorbing_per_phase_matches_bulk(name):
    """The per-phase entry points in the order of NewmarkTimeScheme::solve
    (Clayton in apply_bc_before, the sponge in apply_bc_after_accel)
    against axs3d_run, FP64."""
    from axisem3d_amd.solver import GpuSolver
    sp = SPECS[name]()
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
    print(f"absorbing {name}: per-phase vs bulk seismogram L2 {err:.3e}")
    assert err < 1e-12
    a.close()
    b.close()
