"""Host side of the adjoint harness: the parameter chain rule and the model
round trip (CPU only)."""
import numpy as np

from axisem3d_amd import adjoint, preloop


def _spec():
    return preloop.cartesian_homogeneous(
        nx=3, nz=3, width=3e3, depth=3e3, nsteps=5, src_depth=1e3,
        n_receivers=2, nodal_nr=lambda s, z: 3.0 + 4.0 * (s / 3e3) + 0.0 * z)


def test_model_round_trip():
    sp = _spec()
    rho, vp, vs = adjoint.model_of(sp)
    sp2 = adjoint.with_model(sp, rho, vp, vs)
    assert np.allclose(sp2.elem_iso, sp.elem_iso, rtol=1e-13, atol=0)
    assert np.allclose(sp2.solid_mass, sp.solid_mass, rtol=1e-13, atol=0)


def test_chain_rule_matches_finite_differences():
    """J = sum a lambda + b mu + c m(rho) with m the assembled mass: its
    gradient in (rho, vp, vs) from chain_rule(kern = (a, b, c)) against
    central differences through with_model.  A dropped density term in
    lambda or mu, or a wrong factor, fails."""
    sp = _spec()
    rho, vp, vs = adjoint.model_of(sp)
    rng = np.random.default_rng(3)
    ne = len(sp.elem_points)
    kern = rng.standard_normal((ne, 25, 3))

    def J(r, p, s):
        m = adjoint.with_model(sp, r, p, s)
        # kern[..., 2] is dJ/drho at fixed lambda and mu (through the
        # mass), so the third term is linear in rho
        return (np.sum(kern[..., 0] * m.elem_iso[:, 0, :]) +
                np.sum(kern[..., 1] * m.elem_iso[:, 1, :]) +
                np.sum(kern[..., 2] * r))

    d = adjoint.chain_rule(kern, rho, vp, vs)
    model = [rho, vp, vs]
    for which in range(3):
        dirn = rng.standard_normal((ne, 25))
        h = 1e-6 * np.abs(model[which]).mean()
        mp = [x.copy() for x in model]
        mm = [x.copy() for x in model]
        mp[which] += h * dirn
        mm[which] -= h * dirn
        fd = (J(*mp) - J(*mm)) / (2 * h)
        ad = np.sum(d[which] * dirn)
        assert abs(fd - ad) < 1e-6 * abs(ad), which
