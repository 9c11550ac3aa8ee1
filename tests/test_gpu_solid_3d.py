"""GPU parity of the cardinal (3D material) path against the FP64
oracle, through the C ABI: spherical sector with a 3D random field,
5-SLS attenuation (CG4 and FULL), odd and even element Nr (radices 2, 3,
4, 5, 11, 13 in the FFT).

Tolerances: FP64 mode agrees with the oracle to rounding (relative 1e-13
per stage, 1e-12 seismogram L2); FP32+FP64 mode relative 1e-5 per stage
and 1e-4 seismogram L2."""
import numpy as np
import pytest

import oracle
from axisem3d_amd import preloop

pytestmark = pytest.mark.gpu


def _spec(nsteps=300, att=True, cg4=True, nu_max=12):
    spec = preloop.spherical_regional(
        ntheta=6, nrad=6, theta_max_deg=2.0, depth=60e3, nu_min=3,
        nu_max=nu_max, nsteps=nsteps, n_receivers=3, attenuation=att,
        cg4=cg4, corr_len=10e3)
    spec._build_3d()
    preloop.finish_regional(spec, nsteps, src_depth=20e3, f_peak=0.5,
                            n_receivers=3, theta_max_deg=2.0)
    return spec


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _solver(spec, prec):
    from axisem3d_amd.solver import GpuSolver
    return GpuSolver(spec, prec)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 1e-5)])
@pytest.mark.parametrize("att,cg4", [(False, True), (True, True),
                                     (True, False)])
def test_cardinal_stiffness_matches_oracle(prec, tol, att, cg4):
    spec = _spec(att=att, cg4=cg4)
    assert len(set((spec.elem_nr % 2).tolist())) == 2
    o = oracle.Oracle(spec)
    g = _solver(spec, prec)
    rng = np.random.default_rng(4)
    u = rng.standard_normal((o.nmodes, 3)) + 1j * rng.standard_normal(
        (o.nmodes, 3))
    g.set_field("displ", u)
    g.set_field("stiff", np.zeros_like(u))
    g.computeStiffness()
    assert _rel(g.get_field("stiff"), o.stiffness(u)) < tol
    if att:
        # second call: memory variables carry state from the first
        g.set_field("stiff", np.zeros_like(u))
        g.computeStiffness()
        assert _rel(g.get_field("stiff"), o.stiffness(u)) < tol


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-4)])
def test_cardinal_seismograms_match_oracle(prec, tol):
    nsteps = 200
    spec = _spec(nsteps=nsteps)
    ref = oracle.Oracle(spec).run(nsteps)
    g = _solver(spec, prec)
    g.run(nsteps)
    got = g.seismograms()
    assert np.abs(ref).max() > 0
    err = _rel(got, ref)
    print(f"3D+attenuation {prec}: relative seismogram L2 = {err:.3e}")
    assert err < tol


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 2e-5)])
def test_cardinal_large_nr_matches_oracle(prec, tol):
    """Elements with Nr from ~600 to ~810 (FFT radix chains of 2, 3, 4, 5,
    7, 11, 13), 3D model and CG4 attenuation: stage-level stiffness."""
    spec = preloop.spherical_regional(
        ntheta=3, nrad=2, theta_max_deg=30.0, depth=12e3, nu_min=300,
        nu_max=400, nsteps=10, n_receivers=2, corr_len=50e3, bound_nr=False)
    spec._build_3d()
    preloop.finish_regional(spec, 10, src_depth=5e3, n_receivers=2)
    assert spec.elem_nr.max() > 600
    o = oracle.Oracle(spec)
    g = _solver(spec, prec)
    rng = np.random.default_rng(9)
    u = rng.standard_normal((o.nmodes, 3)) + 1j * rng.standard_normal(
        (o.nmodes, 3))
    g.set_field("displ", u)
    g.set_field("stiff", np.zeros_like(u))
    g.computeStiffness()
    err = _rel(g.get_field("stiff"), o.stiffness(u))
    print(f"large Nr {sorted(set(spec.elem_nr.tolist()))} {prec}: {err:.2e}")
    assert err < tol


def test_regional_reduced_fp32_vs_fp64_stable():
    """The regional workload at reduced width (nu 10..400 over the full
    30 degrees, 60 x 24 elements): 400 steps are stable and the FP32
    stiffness path stays within the north-star bar of the FP64 path at
    the near-source monitors and the surface receivers."""
    nsteps = 400
    spec = preloop.spherical_regional(ntheta=60, nrad=24, nsteps=nsteps,
                                      n_receivers=10, f_peak=0.05,
                                      courant=0.4)
    spec._build_3d()
    preloop.finish_regional(spec, nsteps, src_depth=50e3, f_peak=0.05,
                            n_receivers=10)
    s64 = _run(spec, "fp64", nsteps)
    s32 = _run(spec, "fp32", nsteps)
    assert np.isfinite(s64).all() and np.abs(s64).max() > 0
    err = _rel(s32, s64)
    print(f"regional 60x24: FP32 vs FP64 relative L2 = {err:.3e}")
    assert err < 1e-4


def _run(spec, prec, nsteps):
    g = _solver(spec, prec)
    g.run(nsteps)
    out = g.seismograms()
    g.close()
    return out


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batching_does_not_change_results(prec, monkeypatch):
    """A 50 KiB scratch splits the 3D elements into many batches (a few
    elements each); the kernels see the same operands, so the fields are
    bitwise those of the default (autotuned) batches."""
    nsteps = 60
    spec = _spec(nsteps)
    out = []
    for mib in (None, "0.05"):
        if mib is None:
            monkeypatch.delenv("AXS3D_CARD_SCRATCH_MB", raising=False)
        else:
            monkeypatch.setenv("AXS3D_CARD_SCRATCH_MB", mib)
        g = _solver(spec, prec)
        info = g.autotune_info()
        g.run(nsteps)
        out.append((g.get_field("displ"), g.seismograms(), info))
        g.close()
    assert abs(out[1][2]["cardinal_scratch_mib"] - 0.05) < 1e-6
    assert np.abs(out[0][0]).max() > 0
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])

