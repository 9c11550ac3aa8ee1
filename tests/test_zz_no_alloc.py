"""The allocation-free stiffness entry point (axs3d_compute_stiffness_static,
cardinal scratch in a file-scope device array) against
axs3d_compute_stiffness (the context's scratch), through the C ABI: two
contexts built from the same domain, the same displacement, one calls
each entry point, and the stiffness must agree bitwise (each element's
cardinal path does not depend on the batch it runs in).  Two calls in a
row, the second on a smaller domain (3D fluid elements), then the first
domain again, whose attenuation state has advanced by one call."""
import numpy as np
import pytest

from axisem3d_amd import preloop

pytestmark = pytest.mark.gpu


def _solid3d_att():
    # spherical sector, every element 3D with CG4 attenuation, element Nr
    # up to 25
    spec = preloop.spherical_regional(
        ntheta=12, nrad=8, theta_max_deg=3.0, depth=80e3, nu_min=3,
        nu_max=12, nsteps=50, n_receivers=3, attenuation=True, cg4=True,
        corr_len=10e3)
    spec._build_3d(np.ones(spec.n_solid_elements, bool))
    preloop.finish_regional(spec, 50, src_depth=20e3, f_peak=0.5,
                            n_receivers=3, theta_max_deg=3.0)
    return spec


def _fluid3d():
    # a fluid layer inside a solid, half the fluid elements 3D
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
    return preloop.make_fluid3d(sp, np.arange(0, nfe, 2), amp=0.3, seed=5)


def _pair(spec, prec, seed):
    from axisem3d_amd.solver import GpuSolver
    ga, gb = GpuSolver(spec, prec), GpuSolver(spec, prec)
    rng = np.random.default_rng(seed)
    u = (rng.standard_normal((ga.nmodes, 3)) +
         1j * rng.standard_normal((ga.nmodes, 3))) * 1e-3
    chi = None
    if ga.fnmodes > 0:
        chi = (rng.standard_normal(ga.fnmodes) +
               1j * rng.standard_normal(ga.fnmodes))
    for g in (ga, gb):
        g.set_field("displ", u)
        if chi is not None:
            g.set_field("fdispl", chi)
    return ga, gb


def _call_both(ga, gb):
    for g in (ga, gb):
        g.set_field("stiff", np.zeros((g.nmodes, 3), complex))
        if g.fnmodes > 0:
            g.set_field("fstiff", np.zeros(g.fnmodes, complex))
    ga.computeStiffness()
    rc = gb.lib.axs3d_compute_stiffness_static(gb.ctx)
    assert rc == 0, gb.lib.axs3d_last_error(gb.ctx)
    out = []
    for name in ("stiff", "fstiff"):
        if name == "fstiff" and ga.fnmodes == 0:
            continue
        a, b = ga.get_field(name), gb.get_field(name)
        assert np.abs(a).max() > 0
        out.append((name, a, b))
    return out


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_static_scratch_matches_context_scratch(prec):
    big = _solid3d_att()
    small = _fluid3d()
    assert big.n_updates_per_step() > small.n_updates_per_step()
    a1, b1 = _pair(big, prec, 11)
    a2, b2 = _pair(small, prec, 12)
    for ga, gb, what in ((a1, b1, "3D solid, CG4"),
                         (a2, b2, "3D fluid, smaller"),
                         (a1, b1, "3D solid, second call")):
        print(f"static scratch {prec}: {what}", flush=True)
        for name, a, b in _call_both(ga, gb):
            assert np.array_equal(a, b), (what, name,
                                          np.abs(a - b).max())
    for g in (a1, b1, a2, b2):
        g.close()
