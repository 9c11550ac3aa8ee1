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


def test_chain_rule3d_matches_finite_differences():
    """Per-slice lambda, mu of 3D elements from (rho per GLL point, vp3d,
    vs3d per slice): chain_rule3d against central differences of a linear
    functional of (lambda, mu per element GLL point, lambda, mu per slice,
    rho)."""
    sp = _spec()
    E = sp.n_solid_elements
    preloop.make_3d_isotropic(sp, lambda x, y, z: 0.0 * x,
                              mask3d=(np.arange(E) % 2) == 0)
    model = list(adjoint.model_of(sp)) + list(adjoint.model3d_of(sp))
    rng = np.random.default_rng(5)
    kern = rng.standard_normal((E, 25, 3))
    kern3d = rng.standard_normal(sp.prop3d.shape)

    def J(m):
        s = adjoint.with_model(sp, *m)
        return (np.sum(kern[..., 0] * s.elem_iso[:, 0, :]) +
                np.sum(kern[..., 1] * s.elem_iso[:, 1, :]) +
                np.sum(kern[..., 2] * m[0]) + np.sum(kern3d * s.prop3d))
    d = adjoint.chain_rule3d(sp, kern, kern3d, *model)
    for which in range(5):
        dirn = rng.standard_normal(model[which].shape)
        h = 1e-6 * np.abs(model[which]).mean()
        mp = [x.copy() for x in model]
        mm = [x.copy() for x in model]
        mp[which] += h * dirn
        mm[which] -= h * dirn
        fd = (J(mp) - J(mm)) / (2 * h)
        ad = np.sum(d[which] * dirn)
        assert abs(fd - ad) < 1e-6 * abs(ad), which


def test_with_model_rewrites_general_material_arrays():
    """A spec built by make_media (all isotropic) is uploaded from mat1d
    and a flat prop3d; with_model must write the new (lambda, mu) there
    too, else the stiffness keeps the old model while the mass takes the
    new density.  Non-isotropic media are refused."""
    sp = _spec()
    E = sp.n_solid_elements
    preloop.make_media(sp, np.zeros(E, dtype=np.uint8),
                       field=lambda x, y, z: 0.01 * np.cos(x / 700.0),
                       mask3d=(np.arange(E) % 2) == 1)
    rho, vp, vs = adjoint.model_of(sp)
    vp3, vs3 = adjoint.model3d_of(sp)
    rng = np.random.default_rng(2)
    r2 = rho * (1 + 0.01 * rng.standard_normal(rho.shape))
    p2 = vp * (1 + 0.01 * rng.standard_normal(rho.shape))
    s2 = vs * (1 + 0.01 * rng.standard_normal(rho.shape))
    v3 = vs3 * 1.02
    sp2 = adjoint.with_model(sp, r2, p2, s2, vp3, v3)
    off = np.asarray(sp2.mat1d_off)
    for e in range(E):
        if sp.elem_kind[e] == 0:
            got = sp2.mat1d[off[e]:off[e] + 50].reshape(2, 25)
            lam = r2[e] * (p2[e] ** 2 - 2 * s2[e] ** 2)
            assert np.allclose(got[0], lam, rtol=1e-14, atol=0)
            assert np.allclose(got[1], r2[e] * s2[e] ** 2, rtol=1e-14, atol=0)
    assert sp2.prop3d.shape == np.asarray(sp.prop3d).shape
    p3 = sp2.prop3d.reshape(-1, 2, 25)
    r3 = adjoint._slice_rho(sp, r2)
    assert np.allclose(p3[:, 1], r3 * v3 ** 2, rtol=1e-14, atol=0)
    # the original spec is left unchanged
    assert np.array_equal(np.asarray(sp.mat1d)[off[0]:off[0] + 50],
                          sp.elem_iso[0].reshape(-1))
    ti = _spec()
    preloop.make_media(ti, np.ones(ti.n_solid_elements, dtype=np.uint8))
    try:
        adjoint.with_model(ti, *adjoint.model_of(ti))
    except ValueError:
        pass
    else:
        raise AssertionError("TI media must be refused")


def test_with_media_round_trip_and_mass():
    """with_media on a general-media spec's own (mat1d, rho, prop3d)
    reproduces its material arrays and point masses; doubling rho doubles
    every mass (GLLPoint.cpp:54-77)."""
    import numpy as np
    from axisem3d_amd import adjoint, preloop
    sp = preloop.cartesian_homogeneous(
        nx=4, nz=4, width=4e3, depth=4e3, nsteps=10, src_depth=2e3,
        n_receivers=2, nodal_nr=lambda s, z: np.full(s.shape, 5.0))
    E = sp.n_solid_elements
    preloop.make_media(sp, (np.arange(E) % 3).astype(np.uint8), field=None,
                       mask3d=(np.arange(E) % 2) == 1, att=None, seed=3)
    media = (np.array(sp.mat1d), np.array(sp._rho1d),
             np.array(sp.prop3d).reshape(-1))
    s2 = adjoint.with_media(sp, *media)
    assert np.array_equal(s2.mat1d, sp.mat1d)
    assert np.array_equal(np.asarray(s2.prop3d).reshape(-1), media[2])
    assert np.allclose(s2.solid_mass, sp.solid_mass, rtol=1e-14)
    s3 = adjoint.with_media(sp, media[0], 2.0 * media[1], media[2])
    assert np.allclose(s3.solid_mass, 2.0 * np.asarray(sp.solid_mass),
                       rtol=1e-14)
