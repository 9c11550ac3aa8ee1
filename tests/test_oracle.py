"""Pins the FP64 oracle (oracle/axs3d_oracle.c) without the reference:
golden tables, closed forms, identities and textbook special cases, each
chosen so that a plausible slip (a dropped term, a wrong sign or index, a
transposed operand) fails.  CPU only."""
import json
import math
import os

import numpy as np
import pytest

import oracle
from axisem3d_amd import preloop, spectral

HERE = os.path.dirname(os.path.abspath(__file__))


# --------------------------------------------------------------- tables
def test_tables_equal_reference_literals():
    ref = json.load(open(os.path.join(HERE, "golden", "spectral_npol4.json")))
    xg, xj, wg, wj, gg, gj = spectral.tables()
    for name, arr in (("pGLL", xg), ("pGLJ", xj), ("wGLL", wg),
                      ("wGLJ", wj), ("gGLL", gg.ravel()), ("gGLJ", gj.ravel())):
        assert np.array_equal(np.array(ref[name]), arr), name


def test_tables_closed_form_identities():
    xg, xj, wg, wj, gg, gj = spectral.tables(exact=True)
    for x, G, w, b in ((xg, gg, wg, 0), (xj, gj, wj, 1)):
        # derivative of sum_i l_i = 1 vanishes
        assert np.abs(G.sum(axis=0)).max() < 1e-12
        # G differentiates every polynomial of degree <= 4 exactly
        for deg in range(5):
            f = x ** deg
            df = deg * x ** (deg - 1) if deg else np.zeros_like(x)
            assert np.abs(f @ G - df).max() < 1e-12
        # quadrature integrates x^k (1+x)^b exactly up to its order
        for k in range(2 * 4 - 1 - b):
            exact = np.polynomial.polynomial.polyint(
                np.polynomial.polynomial.polymul([0] * k + [1],
                                                 [1, 1] if b else [1]))
            val = np.polynomial.polynomial.polyval(1, exact) - \
                np.polynomial.polynomial.polyval(-1, exact)
            assert abs(np.sum(w * x ** k) - val) < 1e-12


def test_lucky_numbers_equal_reference_table():
    ref = json.load(open(os.path.join(HERE, "golden", "spectral_npol4.json")))
    assert list(spectral._lucky_numbers()) == ref["lucky"]
    assert spectral.next_lucky_number(17) == 18
    assert spectral.next_lucky_number(143) == 144     # 11*13 is not lucky
    assert spectral.next_lucky_number(121) == 125     # 11^2 is not lucky
    assert spectral.next_lucky_number(100001) == 100001


# --------------------------------------------------------------- gradient
def _element(axial, s0=1000.0, ds=800.0, z0=2000.0, dz=600.0):
    """One rectangular element; axial elements start on the axis."""
    s0 = 0.0 if axial else s0
    mesh = preloop.StructuredMesh(np.array([s0, s0 + ds]),
                                  np.array([z0, z0 + dz]))
    gq = preloop.gradquad_operands(mesh)
    ops = np.stack([gq["dsdxii"], gq["dsdeta"], gq["dzdxii"], gq["dzdeta"],
                    gq["inv_s"], gq["ifact"]], axis=1)[0]
    return ops, gq["sz"][0]


def _poly(rng, s, z, deg=4):
    c = rng.standard_normal((deg + 1, deg + 1)) + 1j * rng.standard_normal(
        (deg + 1, deg + 1))
    sc, zc = 1e-3, 1e-3

    def f(s, z):
        return sum(c[a, b] * (s * sc) ** a * (z * zc) ** b
                   for a in range(deg + 1) for b in range(deg + 1 - a))

    def fs(s, z):
        return sum(c[a, b] * a * sc * (s * sc) ** (a - 1) * (z * zc) ** b
                   for a in range(1, deg + 1) for b in range(deg + 1 - a))

    def fz(s, z):
        return sum(c[a, b] * b * zc * (s * sc) ** a * (z * zc) ** (b - 1)
                   for a in range(deg + 1) for b in range(1, deg + 1 - a))
    return f, fs, fz


@pytest.mark.parametrize("alpha", [0, 1, 2, 5])
def test_grad6_closed_form_off_axis(alpha):
    ops, sz = _element(False)
    s, z = sz
    rng = np.random.default_rng(alpha)
    fields = [_poly(rng, s, z) for _ in range(3)]
    u = np.array([f[0](s, z) for f in fields])
    us, up, uz = u
    d = [(f[1](s, z), f[2](s, z)) for f in fields]
    ia = 1j * alpha
    expect = np.array([
        d[0][0],                               # ss
        (ia * up + us) / s,                    # pp
        d[2][1],                               # zz
        d[1][1] + ia * uz / s,                 # pz (engineering)
        d[0][1] + d[2][0],                     # sz (engineering)
        d[1][0] + (ia * us - up) / s,          # sp (engineering)
    ])
    got = oracle.grad6(ops, 0, alpha, u)
    assert np.abs(got - expect).max() < 1e-9 * np.abs(expect).max()


@pytest.mark.parametrize("alpha", [0, 1, 2, 3])
def test_grad6_closed_form_on_axis(alpha):
    """Axis-regular fields: the lower-order numerators vanish on s = 0 and
    the L'Hospital limit is the s-derivative of the numerator."""
    ops, sz = _element(True)
    s, z = sz
    rng = np.random.default_rng(10 + alpha)
    g = [_poly(rng, s, z, deg=3) for _ in range(3)]
    # u = s * g (vanishes on the axis) except mode-1 s/phi pairs
    us = s * g[0][0](s, z)
    up = s * g[1][0](s, z)
    uz = s * g[2][0](s, z) if alpha > 0 else g[2][0](s, z)
    dus_s = g[0][0](s, z) + s * g[0][1](s, z)
    dup_s = g[1][0](s, z) + s * g[1][1](s, z)
    duz_s = (g[2][0](s, z) + s * g[2][1](s, z)) if alpha > 0 else g[2][1](s, z)
    dus_z = s * g[0][2](s, z)
    dup_z = s * g[1][2](s, z)
    duz_z = s * g[2][2](s, z) if alpha > 0 else g[2][2](s, z)
    ia = 1j * alpha
    on = s == 0.0
    sd = np.where(on, 1.0, s)
    pp = np.where(on, ia * dup_s + dus_s, (ia * up + us) / sd)
    pz = dup_z + np.where(on, ia * duz_s, ia * uz / sd)
    sp = dup_s + np.where(on, ia * dus_s - dup_s, (ia * us - up) / sd)
    expect = np.array([dus_s, pp, duz_z, pz, dus_z + duz_s, sp])
    got = oracle.grad6(ops, 1, alpha, np.array([us, up, uz]))
    assert np.abs(got - expect).max() < 1e-9 * np.abs(expect).max()


@pytest.mark.parametrize("axial", [False, True])
def test_rigid_body_null_space(axial):
    ops, sz = _element(axial)
    s, z = sz
    one, zero = np.ones(25), np.zeros(25)
    cases = [(0, [zero, zero, one]),              # translation along z
             (0, [zero, s, zero]),                # rotation about z
             (1, [0.5 * one, 0.5j * one, zero])]  # translation along x
    for alpha, u in cases:
        e = oracle.grad6(ops, int(axial), alpha, np.array(u, complex))
        # the cancelling terms are O(1); the 12-decimal G tables leave
        # O(1e-12) residue
        assert np.abs(e).max() < 1e-10


@pytest.mark.parametrize("axial", [False, True])
@pytest.mark.parametrize("alpha", [0, 1, 3])
def test_quad_is_integral_weighted_transpose_of_grad(axial, alpha):
    """<Quad(sig), v> = <sig, IF * Grad_{-alpha}(v)>: the quadrature is the
    transpose of the gradient weighted by the integral factor (test
    functions carry e^{-i alpha phi})."""
    ops, _ = _element(axial)
    rng = np.random.default_rng(alpha + 7 * axial)
    sig = rng.standard_normal((6, 25)) + 1j * rng.standard_normal((6, 25))
    v = rng.standard_normal((3, 25)) + 1j * rng.standard_normal((3, 25))
    lhs = np.sum(oracle.quad6(ops, int(axial), alpha, sig) * v)
    rhs = np.sum(ops[5] * sig * oracle.grad6(ops, int(axial), -alpha, v))
    assert abs(lhs - rhs) < 1e-10 * abs(rhs)
    ops1 = ops.copy()
    ops1[5] = 1.0
    lhs = np.sum(oracle.quad6(ops, int(axial), alpha, sig, integrate=False)
                 * v)
    rhs = np.sum(sig * oracle.grad6(ops1, int(axial), -alpha, v))
    assert abs(lhs - rhs) < 1e-10 * abs(rhs)


def test_isotropic_hooke_matches_voigt_matrix():
    rng = np.random.default_rng(1)
    lam = rng.uniform(1, 2, 25)
    mu = rng.uniform(1, 2, 25)
    eps = rng.standard_normal((6, 25)) + 1j * rng.standard_normal((6, 25))
    got = oracle.iso_stress(lam, mu, eps)
    for p in range(25):
        C = np.zeros((6, 6))
        C[:3, :3] = lam[p]
        C[[0, 1, 2], [0, 1, 2]] += 2 * mu[p]
        C[[3, 4, 5], [3, 4, 5]] = mu[p]      # engineering shears
        assert np.allclose(got[:, p], C @ eps[:, p], rtol=1e-14, atol=0)


# --------------------------------------------------------------- FFT
@pytest.mark.parametrize("nr", list(range(1, 17)) + [21, 26, 39, 45, 77, 91])
def test_fft_conventions_against_dft(nr):
    """R2C = (1/N) DFT with the Nyquist row halved; C2R doubles it and
    returns the samples: checked against a brute-force DFT."""
    rng = np.random.default_rng(nr)
    x = rng.standard_normal(nr)
    n = np.arange(nr)
    k = np.arange(nr // 2 + 1)
    dft = np.exp(-2j * np.pi * np.outer(k, n) / nr) @ x / nr
    if nr % 2 == 0:
        dft[-1] *= 0.5
    c = oracle.r2c(x)
    assert np.abs(c - dft).max() < 1e-13
    assert np.abs(oracle.c2r(c, nr) - x).max() < 1e-12
    # C2R ignores Im of the mean and of the Nyquist row
    c2 = c.copy()
    c2[0] += 0.7j
    if nr % 2 == 0:
        c2[-1] += 0.3j
    assert np.abs(oracle.c2r(c2, nr) - x).max() < 1e-12


# --------------------------------------------------------------- Newmark
def test_newmark_harmonic_oscillator():
    """Newmark with gamma = 1/2, beta = 0 on u'' = -w^2 u is the central
    difference scheme: u_n = cos(n theta), cos(theta) = 1 - (w dt)^2 / 2."""
    w, dt, n = 2.0, 0.05, 400
    u = np.array([[1.0 + 0j, 0, 0]])
    a = -w * w * u
    v = -0.5 * dt * (a + a)        # so that the velocity at t = 0 is zero
    out = []
    for _ in range(n):
        out.append(u[0, 0].real)
        s = -w * w * u
        s, u, v, a = _newmark(s, u, v, a, dt)
    theta = math.acos(1 - (w * dt) ** 2 / 2)
    ref = np.cos(theta * np.arange(n))
    assert np.abs(np.array(out) - ref).max() < 1e-12


def _newmark(s, u, v, a, dt):
    spec = preloop.cartesian_homogeneous(nx=1, nz=1, width=1e3, depth=1e3,
                                         nu=0, nsteps=1, n_receivers=1)
    o = oracle.Oracle(spec)
    o.d.dt = dt
    o.nmodes = 1
    return o.newmark(s, u, v, a)


# --------------------------------------------------------------- receivers
def test_receiver_synthesis_inverts_r2c():
    """A receiver sitting on one point recovers f(phi_k) from the
    Fourier coefficients r2c(f): pins c_0 = 1, c_alpha = 2 and the Nyquist
    convention against the FFT."""
    for nu in (2, 3):
        spec = preloop.cartesian_homogeneous(nx=2, nz=2, width=2e3,
                                             depth=2e3, nu=nu, nsteps=1,
                                             n_receivers=1)
        nr = int(spec.solid_nr[0])
        e = 3
        pts = spec.elem_points[e]
        w = np.zeros(25)
        w[12] = 1.0
        spec.rec_elem = np.array([e], np.int32)
        spec.rec_weights = w[None, :]
        rng = np.random.default_rng(nu)
        phis = 2 * np.pi * np.arange(nr) / nr
        o = oracle.Oracle(spec)
        f = rng.standard_normal((nr, 3))
        coef = np.array([oracle.r2c(f[:, c]) for c in range(3)]).T
        for k, phi in enumerate(phis):
            spec.rec_phi = np.array([phi])
            o2 = oracle.Oracle(spec)
            displ = np.zeros((o2.nmodes, 3), complex)
            off = o2.offsets[pts[12]]
            displ[off:off + nr // 2 + 1] = coef
            got = o2.record(displ)[0]
            assert np.abs(got - f[k]).max() < 1e-12


# --------------------------------------------------------------- attenuation
def _att_spec(cg4, mask3d):
    spec = preloop.spherical_regional(
        ntheta=2, nrad=2, theta_max_deg=1.0, depth=20e3, nu_min=2, nu_max=3,
        nsteps=1, n_receivers=1, attenuation=True, cg4=cg4, corr_len=10e3)
    spec._build_3d(None if mask3d is None else np.asarray(mask3d))
    preloop.finish_regional(spec, 1, src_depth=10e3, f_peak=0.5,
                            n_receivers=1, theta_max_deg=1.0)
    return spec


@pytest.mark.parametrize("cg4", [True, False])
@pytest.mark.parametrize("path", ["cardinal", "fourier"])
def test_sls_memory_variables_follow_the_relaxation_ode(cg4, path):
    """Each SLS memory variable solves m' = -w (m - y~ dsig(t)), y~ =
    y / sum(y), with dsig = iso(dlambda, dmu)(strain).  For a strain ramp
    eps(t) = E t the exact solution is m(t) = y~ D (t - (1 - e^{-w t}) / w)
    with D = iso(dlambda, dmu)(E), and the (alpha, beta, gamma) recursion
    integrates piecewise-linear input exactly.  The stress returned at step
    n has the memory of step n-1 subtracted; points outside the CG4 set
    carry no memory.  Cardinal path: one slice of a 3D element, real
    strain; Fourier path: one order of a 1D element, complex strain.  A
    swapped beta/gamma, the new dsig used in place of the old one, or a
    memory added instead of subtracted fails."""
    e = 1
    spec = _att_spec(cg4, None if path == "cardinal" else [False] * 4)
    o = oracle.Oracle(spec)
    att = spec.att
    w, y = att["w"], att["y"]
    dt = spec.dt
    npa = 4 if cg4 else 25
    pts = list(preloop.CG4_POINTS) if cg4 else list(range(25))
    rng = np.random.default_rng(3)
    if path == "cardinal":
        sl = 1
        s0 = int(spec.elem_slice_off[e]) + sl
        dl, dm = att["d3d"][s0, 0], att["d3d"][s0, 1]
        E = rng.standard_normal((6, 25)) * 1e-6
        apply = lambda eps: o.att_apply_cd(e, sl, eps, np.zeros((6, 25)))
    else:
        alpha = 1
        dl, dm = att["d1d"][e, 0], att["d1d"][e, 1]
        E = (rng.standard_normal((6, 25)) +
             1j * rng.standard_normal((6, 25))) * 1e-6
        apply = lambda eps: o.att_apply_fr(e, alpha, eps,
                                           np.zeros((6, 25), complex))
    D = np.zeros((6, npa), dtype=E.dtype)
    for q, p in enumerate(pts):
        tr = (E[0, p] + E[1, p] + E[2, p]) * dl[q]
        D[:3, q] = tr + 2 * dm[q] * E[:3, p]
        D[3:, q] = dm[q] * E[3:, p]
    yt = y / y.sum()
    for n in range(0, 400):
        got = apply(E * (n * dt))
        t = (n - 1) * dt
        if n == 0:
            msum = 0.0
        else:
            msum = sum(yt[l] * (t - (1 - math.exp(-w[l] * t)) / w[l])
                       for l in range(len(w)))
        want = np.zeros((6, 25), dtype=E.dtype)
        for q, p in enumerate(pts):
            want[:, p] = -D[:, q] * msum
        scale = np.abs(D).max() * max(msum, 1e-30)
        assert np.abs(got - want).max() <= 1e-10 * scale + 1e-300, n


def test_unrelaxed_correction_vanishes_without_attenuation():
    """Q -> infinity: the memory moduli vanish and the elastic moduli are
    returned unchanged (AttBuilder::createAttenuation limits)."""
    rng = np.random.default_rng(5)
    lam = rng.uniform(3e10, 5e10, (3, 25))
    mu = rng.uniform(2e10, 4e10, (3, 25))
    w, y = preloop.att_sls_fit(5, 0.005, 0.5)
    consts = preloop.att_constants(w, y, 0.005, 0.5, 1.0)
    q = np.full((3, 25), 1e15)
    wcg4 = rng.uniform(1, 4, (3, 4))
    for cg4 in (True, False):
        lc, mc, dl, dm = preloop.att_isotropic(lam, mu, q, q, consts, cg4,
                                               wcg4)
        assert np.allclose(lc, lam, rtol=1e-12)
        assert np.allclose(mc, mu, rtol=1e-12)
        assert np.abs(dl).max() < 1e-3 and np.abs(dm).max() < 1e-3


# --------------------------------------------------------------- media
def _voigt_to_tensor(v, engineering):
    f = 0.5 if engineering else 1.0
    t = np.empty((3, 3) + v.shape[1:], dtype=v.dtype)
    t[0, 0], t[1, 1], t[2, 2] = v[0], v[1], v[2]
    t[1, 2] = t[2, 1] = f * v[3]
    t[0, 2] = t[2, 0] = f * v[4]
    t[0, 1] = t[1, 0] = f * v[5]
    return t


def _tensor_to_voigt(t, engineering):
    f = 2.0 if engineering else 1.0
    return np.array([t[0, 0], t[1, 1], t[2, 2], f * t[1, 2], f * t[0, 2],
                     f * t[0, 1]])


def test_rtz_rotation_is_the_tensor_rotation():
    """(s, phi, z) -> (R, T, Z) at colatitude theta: e_R = cos e_s - sin
    e_z, e_T = e_phi, e_Z = sin e_s + cos e_z (the vector rule of
    CoordTransformSpherical).  Strain (engineering shears) maps as
    Q eps Q^T, stress back as Q^T sig Q; the work sig : eps is invariant."""
    rng = np.random.default_rng(11)
    theta = rng.uniform(0, np.pi, 25)
    eps = rng.standard_normal((6, 25)) + 1j * rng.standard_normal((6, 25))
    sig = rng.standard_normal((6, 25)) + 1j * rng.standard_normal((6, 25))
    got_e = oracle.spz_to_rtz6(theta, eps)
    got_s = oracle.rtz_to_spz6(theta, sig)
    for p in range(25):
        c, s = np.cos(theta[p]), np.sin(theta[p])
        Q = np.array([[c, 0, -s], [0, 1, 0], [s, 0, c]])
        te = _voigt_to_tensor(eps[:, p], True)
        want_e = _tensor_to_voigt(Q @ te @ Q.T, True)
        ts = _voigt_to_tensor(sig[:, p], False)
        want_s = _tensor_to_voigt(Q.T @ ts @ Q, False)
        assert np.allclose(got_e[:, p], want_e, rtol=0, atol=1e-13)
        assert np.allclose(got_s[:, p], want_s, rtol=0, atol=1e-13)
    # work invariance: sig_spz . eps_spz == sig_rtz . eps_rtz
    back = oracle.rtz_to_spz6(theta, sig)
    w1 = np.sum(back * eps, axis=0)
    w2 = np.sum(sig * got_e, axis=0)
    assert np.allclose(w1, w2, rtol=1e-12, atol=0)


def _iso_params(rng):
    lam = rng.uniform(1, 2, 25)
    mu = rng.uniform(1, 2, 25)
    return lam, mu


def test_ti_and_anisotropic_reduce_to_isotropic():
    rng = np.random.default_rng(12)
    lam, mu = _iso_params(rng)
    eps = rng.standard_normal((6, 25)) + 1j * rng.standard_normal((6, 25))
    want = oracle.material_stress(0, np.array([lam, mu]), eps)
    ti = np.array([lam + 2 * mu, lam + 2 * mu, lam, mu, mu])
    assert np.allclose(oracle.material_stress(1, ti, eps), want,
                       rtol=1e-14, atol=0)
    c = np.zeros((21, 25))
    for k, v in ((0, lam + 2 * mu), (6, lam + 2 * mu), (11, lam + 2 * mu),
                 (1, lam), (2, lam), (7, lam), (15, mu), (18, mu), (20, mu)):
        c[k] = v
    assert np.allclose(oracle.material_stress(2, c, eps), want,
                       rtol=1e-14, atol=0)


def _bond_z(psi):
    """Bond matrices of a rotation by psi about the Z axis acting on Voigt
    strain (engineering shears) and stress."""
    c, s = np.cos(psi), np.sin(psi)
    Q = np.array([[c, s, 0], [-s, c, 0], [0, 0, 1.0]])
    Me = np.zeros((6, 6))
    Ms = np.zeros((6, 6))
    for k in range(6):
        v = np.zeros(6)
        v[k] = 1
        Me[:, k] = _tensor_to_voigt(Q @ _voigt_to_tensor(v, True) @ Q.T, True)
        Ms[:, k] = _tensor_to_voigt(Q @ _voigt_to_tensor(v, False) @ Q.T,
                                    False)
    return Me, Ms


def test_ti_is_invariant_about_its_symmetry_axis():
    """A TI medium with symmetry axis Z: rotating the strain about Z rotates
    the stress the same way (sig(Me eps) = Ms sig(eps)) for any angle; a
    misplaced modulus (N on the TZ shear, F on the wrong normal) breaks
    it."""
    rng = np.random.default_rng(13)
    A, C, F, L, Nn = rng.uniform(1, 2, (5, 25))
    par = np.array([A, C, F, L, Nn])
    eps = rng.standard_normal((6, 25)) + 1j * rng.standard_normal((6, 25))
    for psi in (0.3, 1.1, 2.5):
        Me, Ms = _bond_z(psi)
        lhs = oracle.material_stress(1, par, Me @ eps)
        rhs = Ms @ oracle.material_stress(1, par, eps)
        assert np.allclose(lhs, rhs, rtol=1e-12, atol=1e-12)
    # and it is not isotropic: rotation about another axis breaks it
    Me, Ms = _bond_z(0.7)
    P = np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0.0]])   # axis permutation
    Mp = np.zeros((6, 6))
    for k in range(6):
        v = np.zeros(6)
        v[k] = 1
        Mp[:, k] = _tensor_to_voigt(P @ _voigt_to_tensor(v, True) @ P.T, True)
    lhs = oracle.material_stress(1, par, Mp @ eps)
    assert not np.allclose(lhs, oracle.material_stress(1, par, eps))


def test_anisotropic_is_the_symmetric_voigt_matrix():
    """21 Cij in the reference's order (C11 C12 .. C16 C22 .. C26 C33 ..
    C66) act as the symmetric 6x6 Voigt matrix; the work eps' sig(eps) is
    a symmetric bilinear form (eps2 . sig(eps1) = eps1 . sig(eps2))."""
    rng = np.random.default_rng(14)
    c21 = rng.standard_normal((21, 25))
    eps1 = rng.standard_normal((6, 25)) + 1j * rng.standard_normal((6, 25))
    eps2 = rng.standard_normal((6, 25)) + 1j * rng.standard_normal((6, 25))
    s1 = oracle.material_stress(2, c21, eps1)
    s2 = oracle.material_stress(2, c21, eps2)
    assert np.allclose(np.sum(eps2 * s1, 0), np.sum(eps1 * s2, 0),
                       rtol=1e-12, atol=1e-12)
    iu = np.triu_indices(6)
    for p in range(25):
        M = np.zeros((6, 6))
        M[iu] = c21[:, p]
        M = M + M.T - np.diag(np.diag(M))
        assert np.allclose(s1[:, p], M @ eps1[:, p], rtol=1e-13, atol=1e-13)


# --------------------------------------------------------------- fluid
@pytest.mark.parametrize("alpha", [0, 1, 3])
def test_grad3_closed_form_off_axis(alpha):
    ops, sz = _element(False)
    s, z = sz
    f, fs, fz = _poly(np.random.default_rng(20 + alpha), s, z)
    p = f(s, z)
    expect = np.array([fs(s, z), 1j * alpha * p / s, fz(s, z)])
    got = oracle.grad3(ops, 0, alpha, p)
    assert np.abs(got - expect).max() < 1e-9 * np.abs(expect).max()


@pytest.mark.parametrize("alpha", [0, 1, 2])
def test_grad3_closed_form_on_axis(alpha):
    """Axis-regular potential: alpha = 0 any field, alpha >= 1 vanishing on
    the axis (p = s g); the 1/s term takes the L'Hospital limit."""
    ops, sz = _element(True)
    s, z = sz
    g, gs, gz = _poly(np.random.default_rng(30 + alpha), s, z, deg=3)
    if alpha == 0:
        p, ps, pz = g(s, z), gs(s, z), gz(s, z)
    else:
        p = s * g(s, z)
        ps = g(s, z) + s * gs(s, z)
        pz = s * gz(s, z)
    on = s == 0.0
    mid = np.where(on, 1j * alpha * ps, 1j * alpha * p / np.where(on, 1, s))
    expect = np.array([ps, mid, pz])
    got = oracle.grad3(ops, 1, alpha, p)
    assert np.abs(got - expect).max() < 1e-9 * np.abs(expect).max()


@pytest.mark.parametrize("axial", [False, True])
@pytest.mark.parametrize("alpha", [0, 1, 3])
def test_quad3_is_integral_weighted_transpose_of_grad3(axial, alpha):
    ops, _ = _element(axial)
    rng = np.random.default_rng(40 + alpha + 7 * axial)
    q = rng.standard_normal((3, 25)) + 1j * rng.standard_normal((3, 25))
    v = rng.standard_normal(25) + 1j * rng.standard_normal(25)
    lhs = np.sum(oracle.quad3(ops, int(axial), alpha, q) * v)
    rhs = np.sum(ops[5] * q * oracle.grad3(ops, int(axial), -alpha, v))
    assert abs(lhs - rhs) < 1e-10 * abs(rhs)


def _fluid_layer_spec(nsteps=2000, nx=4, nz=6):
    mesh = preloop.StructuredMesh(np.linspace(0, 4e3, nx + 1),
                                  np.linspace(0, 6e3, nz + 1))

    def mat(r, rc):
        inl = (rc > 2e3) & (rc < 4e3)
        one = np.ones_like(r)
        return (np.where(inl, 1000.0, 2600.0) * one,
                np.where(inl, 1500.0, 6000.0) * one,
                np.where(inl, 0.0, 3464.0) * one)
    sp = preloop.build_domain(mesh, mat, lambda s, z: np.full(s.shape, 5.0),
                              courant=0.4)
    ix, iz = mesh.element_ix_iz()
    zc = (iz + 0.5) * 1e3
    preloop.make_fluid(sp, (zc > 2e3) & (zc < 4e3))
    mt = np.array([1.0, 1.0, 1.0, 0.3, 0.0, 0.0]) * 1e15
    preloop.add_axial_moment_source(sp, 1e3, mt, 1.628 / math.pi, 1.628,
                                    nsteps)
    preloop.add_receivers(sp, np.array([0.5e3]), np.array([5.5e3]),
                          np.zeros(1))
    return sp


def test_solid_fluid_normals_integrate_the_interface():
    """The assembled normals of a flat interface z = z0 point from fluid to
    solid and sum to the integral of s ds over [0, width] (the axisymmetric
    surface measure without the 2 pi)."""
    sp = _fluid_layer_spec()
    z = sp.solid_sz[sp.sf_solid][:, 1]
    top = z > 3e3
    assert np.allclose(sp.sf_normal[:, 0], 0.0, atol=1e-6)
    assert np.isclose(sp.sf_normal[top, 1].sum(), 0.5 * 4e3 ** 2, rtol=1e-12)
    assert np.isclose(sp.sf_normal[~top, 1].sum(), -0.5 * 4e3 ** 2,
                      rtol=1e-12)


def test_coupling_terms_are_adjoint():
    """Solid -> fluid adds B^T u to the fluid stiffness, fluid -> solid
    subtracts B (fluid stiff / m_fluid) from the solid one, with the same
    B: <c, B^T u> = <B c, u> for any u, c.  A sign flipped on one side, a
    missing 1/m_fluid or a swapped s/z normal fails."""
    sp = _fluid_layer_spec(nsteps=1)
    o = oracle.Oracle(sp)
    rng = np.random.default_rng(21)
    u = rng.standard_normal((o.nmodes, 3)) + 1j * rng.standard_normal(
        (o.nmodes, 3))
    c = rng.standard_normal(o.fnmodes) + 1j * rng.standard_normal(o.fnmodes)
    mf = np.repeat(sp.fluid_mass, sp.fluid_nr // 2 + 1)
    zs = np.zeros_like(u)
    _, btu = o.coupling(u, zs, np.zeros_like(c))          # B^T u
    bc, _ = o.coupling(zs, zs, mf * c)                   # -B c
    lhs = np.vdot(c, btu)
    rhs = np.vdot(-bc, u)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # the fluid side of coupled points only
    touched = np.zeros(o.fnmodes, bool)
    for f in sp.sf_fluid:
        touched[o.foffsets[f]:o.foffsets[f + 1]] = True
    assert np.all(btu[~touched] == 0) and np.any(btu[touched] != 0)


def test_coupled_energy_is_conserved():
    """E = 1/2 u'^T M_s u' + 1/2 u^T K_s u + 1/2 chi'^T K_f chi' + 1/2
    chi''^T M_f chi'' is the energy of the displacement-potential coupling
    (it follows from the two equations of motion with the adjoint
    coupling pair); after the source has stopped it stays constant up to
    the bounded oscillation of the central-difference scheme, and the fluid
    carries a substantial share of it."""
    sp = _fluid_layer_spec(nsteps=6500)
    o = oracle.Oracle(sp)
    o.run(300, record=False)
    ms = np.repeat(sp.solid_mass, sp.solid_nr // 2 + 1)[:, None]
    mf = np.repeat(sp.fluid_mass, sp.fluid_nr // 2 + 1)
    es, ef = [], []
    for _ in range(12):
        e_s = 0.5 * np.sum(ms * np.abs(o.veloc) ** 2) - 0.5 * np.real(
            np.vdot(o.displ, o.stiffness(o.displ)))
        e_f = -0.5 * np.real(np.vdot(o.fveloc, o.fluid_stiffness(o.fveloc))) \
            + 0.5 * np.sum(mf * np.abs(o.faccel) ** 2)
        es.append(e_s)
        ef.append(e_f)
        o.run(500, record=False)
    tot = np.array(es) + np.array(ef)
    # the unmodified energy of the central-difference scheme oscillates
    # by a few per cent here; an inconsistent coupling grows or decays
    assert (tot.max() - tot.min()) / tot.mean() < 0.1
    assert np.mean(np.array(ef) / tot) > 0.02


def _mass3d_spec(seed=5):
    def nodal_nr(s, z):
        return 3.0 + 9.0 * (s / 6e3) + 0.0 * z
    spec = preloop.cartesian_homogeneous(
        nx=4, nz=4, width=6e3, depth=4e3, nodal_nr=nodal_nr, nsteps=10,
        src_depth=2.0e3, n_receivers=2, seed=seed)
    pts = np.arange(0, len(spec.solid_nr), 3)
    return preloop.make_mass3d(spec, pts, amp=0.25, seed=seed), pts


def test_mass3d_constant_equals_mass1d():
    # a 3D mass equal to the 1D mass in every slice divides like Mass1D,
    # except that the C2R / R2C round trip drops the imaginary part of the
    # mean (SolverFFTW ignores it), so that part is zero on 3D points
    spec, pts = _mass3d_spec()
    rng = np.random.default_rng(1)
    o1 = oracle.Oracle(spec)
    n = o1.nmodes
    s = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    ref = oracle.Oracle(_drop_mass3d(spec)).stiff_to_accel(s)
    off = np.concatenate([[0], np.cumsum(spec.solid_nr // 2 + 1)])
    ref[off[pts]] = ref[off[pts]].real
    spec.mass3d = np.concatenate([np.full(int(spec.solid_nr[p]),
                                          spec.solid_mass[p]) for p in pts])
    got = oracle.Oracle(spec).stiff_to_accel(s)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def _drop_mass3d(spec):
    import copy
    s2 = copy.copy(spec)
    s2.mass3d_points = None
    s2.mass3d = None
    return s2


def test_mass3d_matches_numpy_fft():
    # per component: synthesis r_k = Re c_0 + 2 sum_{a>=1} Re(c_a e^{i a
    # phi_k}) (SolverFFTW C2R, Nyquist doubled), r / m, then numpy's rfft /
    # Nr with the Nyquist row halved (R2C); the Nyquist imaginary part is
    # zeroed first (SolidPoint::computeStiffToAccel)
    spec, pts = _mass3d_spec(seed=7)
    o = oracle.Oracle(spec)
    rng = np.random.default_rng(2)
    s = rng.standard_normal((o.nmodes, 3)) + 1j * rng.standard_normal(
        (o.nmodes, 3))
    got = o.stiff_to_accel(s)
    off = np.concatenate([[0], np.cumsum(spec.solid_nr // 2 + 1)])
    moff = 0
    checked = 0
    for p in pts:
        nr = int(spec.solid_nr[p])
        nu1 = nr // 2 + 1
        m = spec.mass3d[moff:moff + nr]
        moff += nr
        c = s[off[p]:off[p + 1]].copy()
        if nr % 2 == 0:
            c[-1] = c[-1].real
        phi = 2 * np.pi * np.arange(nr) / nr
        for comp in range(3):
            r = np.full(nr, c[0, comp].real)
            for a in range(1, nu1):
                r += 2 * np.real(c[a, comp] * np.exp(1j * a * phi))
            r /= m
            ref = np.fft.rfft(r)[:nu1] / nr
            if nr % 2 == 0:
                ref[-1] *= 0.5
            np.testing.assert_allclose(got[off[p]:off[p + 1], comp], ref,
                                       rtol=0, atol=1e-12 * np.abs(ref).max())
            checked += 1
    assert checked == 3 * len(pts)


def _fluid3d_spec(amp=0.3, seed=3):
    sp = _fluid_layer_spec(nsteps=1)
    nfe = len(sp.fluid_elem_nr)
    return preloop.make_fluid3d(sp, np.arange(0, nfe, 2), amp=amp,
                                seed=seed)


def _drop_fluid3d(spec):
    import copy
    s2 = copy.copy(spec)
    s2.fluid_elem_kind = None
    s2.fluid_k3d = None
    return s2


def _fluid_field(o, rng, real_mean=True):
    chi = rng.standard_normal(o.fnmodes) + 1j * rng.standard_normal(
        o.fnmodes)
    if real_mean:
        chi[o.foffsets[:-1]] = chi[o.foffsets[:-1]].real
    return chi


def test_fluid3d_constant_k_equals_fluid1d():
    """A 3D bulk modulus equal to the 1D one in every slice (but varying
    over the 25 GLL points) gives the 1D fluid stiffness on a field whose
    mean is real (Nr = 5 everywhere: no Nyquist mode).  A K table read
    with GLL point and slice swapped fails."""
    sp = _fluid3d_spec()
    assert np.all(sp.fluid_elem_nr % 2 == 1)
    chi = _fluid_field(oracle.Oracle(sp), np.random.default_rng(4))
    ref = oracle.Oracle(_drop_fluid3d(sp)).fluid_stiffness(chi)
    sp.fluid_k3d = np.concatenate([
        np.tile(sp.fluid_elem_k[e], int(sp.fluid_elem_nr[e]))
        for e in range(len(sp.fluid_elem_nr)) if sp.fluid_elem_kind[e]])
    got = oracle.Oracle(sp).fluid_stiffness(chi)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_fluid3d_stiffness_is_symmetric():
    """-int K grad chi . grad psi dV is symmetric in (chi, psi) for any
    K(phi) > 0: <psi, A chi> = <chi, A psi> in the real-field inner
    product sum_p (Re c_0 d_0 + 2 Re sum_{a>0} conj(c_a) d_a), and
    -<chi, A chi> > 0.  A C2R / R2C pair with mismatched conventions
    around the K product breaks the symmetry."""
    sp = _fluid3d_spec()
    o = oracle.Oracle(sp)
    rng = np.random.default_rng(8)
    chi, psi = _fluid_field(o, rng), _fluid_field(o, rng)
    w = np.full(o.fnmodes, 2.0)
    w[o.foffsets[:-1]] = 1.0

    def dot(a, b):
        return np.real(np.sum(w * np.conj(a) * b))
    a_chi, a_psi = o.fluid_stiffness(chi), o.fluid_stiffness(psi)
    lhs, rhs = dot(psi, a_chi), dot(chi, a_psi)
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)
    assert -dot(chi, a_chi) > 0


def test_fluid3d_rotation_equivariance():
    """Rotating K by one slice (K'(phi) = K(phi - 2 pi / Nr)) and the field
    by the same angle (c_a -> c_a exp(-i a 2 pi / Nr)) rotates the
    stiffness: the slice k of the K table sits at phi = 2 pi k / Nr, the
    synthesis convention of the C2R.  A K table read in reversed azimuth
    fails."""
    sp = _fluid3d_spec()
    nr = 5
    assert np.all(sp.fluid_elem_nr == nr) and np.all(sp.fluid_nr == nr)
    o = oracle.Oracle(sp)
    chi = _fluid_field(o, np.random.default_rng(6))
    a = np.concatenate([np.arange(n // 2 + 1) for n in sp.fluid_nr])
    rot = np.exp(-1j * a * 2 * np.pi / nr)
    base = o.fluid_stiffness(chi)
    k3 = sp.fluid_k3d.reshape(-1, nr, 25)
    sp.fluid_k3d = np.roll(k3, 1, axis=1).reshape(-1)
    got = oracle.Oracle(sp).fluid_stiffness(rot * chi)
    assert np.abs(got - rot * base).max() <= 1e-11 * np.abs(base).max()
    # the reversed azimuth does not satisfy it
    sp.fluid_k3d = np.roll(k3, -1, axis=1).reshape(-1)
    bad = oracle.Oracle(sp).fluid_stiffness(rot * chi)
    assert np.abs(bad - rot * base).max() > 1e-3 * np.abs(base).max()


def _wdot_modes(a, b, nr):
    w = np.concatenate([np.r_[1.0, np.full(n // 2, 2.0)] for n in nr])
    return float(np.sum(w * np.real(np.conj(a) * b)))


def _real_mean(x, offsets):
    x = x.copy()
    x[offsets[:-1]] = x[offsets[:-1]].real
    return x


def test_fluid_mass3d_constant_equals_mass1d_and_matches_numpy():
    sp = _fluid_layer_spec(nsteps=1)
    pts = np.arange(0, len(sp.fluid_nr), 3)
    rng = np.random.default_rng(12)
    o1 = oracle.Oracle(sp)
    s = rng.standard_normal(o1.fnmodes) + 1j * rng.standard_normal(
        o1.fnmodes)
    ref1 = o1.fluid_stiff_to_accel(s)
    # constant 3D mass: Mass1D except the imaginary part of the mean
    sp.fluid_mass3d_points = pts
    sp.fluid_mass3d = np.concatenate([np.full(int(sp.fluid_nr[p]),
                                              sp.fluid_mass[p]) for p in pts])
    got = oracle.Oracle(sp).fluid_stiff_to_accel(s)
    ref1[o1.foffsets[pts]] = ref1[o1.foffsets[pts]].real
    assert np.abs(got - ref1).max() <= 1e-12 * np.abs(ref1).max()
    # varying 3D mass against numpy: synthesis, division, rfft / Nr
    preloop.make_fluid_mass3d(sp, pts, amp=0.3, seed=4)
    got = oracle.Oracle(sp).fluid_stiff_to_accel(s)
    moff = 0
    for p in pts:
        nr = int(sp.fluid_nr[p])
        nu1 = nr // 2 + 1
        m = sp.fluid_mass3d[moff:moff + nr]
        moff += nr
        c = s[o1.foffsets[p]:o1.foffsets[p + 1]].copy()
        phi = 2 * np.pi * np.arange(nr) / nr
        r = np.full(nr, c[0].real)
        for a in range(1, nu1):
            r += 2 * np.real(c[a] * np.exp(1j * a * phi))
        ref = np.fft.rfft(r / m)[:nu1] / nr
        np.testing.assert_allclose(got[o1.foffsets[p]:o1.foffsets[p + 1]],
                                   ref, rtol=0,
                                   atol=1e-12 * np.abs(ref).max())


def _coupling3d_spec(amp=0.2, tilt=0.2):
    sp = _fluid_layer_spec(nsteps=1)
    return preloop.make_coupling3d(sp, amp=amp, tilt=tilt, seed=6)


def test_coupling3d_constant_normal_equals_coupling1d():
    """3D normals equal to the 1D ones in every slice (no phi component)
    couple like the 1D pairs on fields with a real mean (Nr = 5: no
    Nyquist).  A normal table read with slice and component swapped
    fails."""
    sp = _coupling3d_spec()
    o1 = oracle.Oracle(sp)
    rng = np.random.default_rng(13)
    u = rng.standard_normal((o1.nmodes, 3)) + 1j * rng.standard_normal(
        (o1.nmodes, 3))
    u = _real_mean(u, o1.offsets)
    c = _real_mean(rng.standard_normal(o1.fnmodes) +
                   1j * rng.standard_normal(o1.fnmodes), o1.foffsets)
    s0 = rng.standard_normal((o1.nmodes, 3)) + 0j
    s0 = _real_mean(s0, o1.offsets)
    sp1 = _drop_coupling3d(sp)
    ws, wf = oracle.Oracle(sp1).coupling(u, s0, c)
    nr = sp.solid_nr[sp.sf_solid]
    sp.sf3d_normal = np.concatenate([
        np.tile(np.r_[sp.sf_normal[k][0], 0.0, sp.sf_normal[k][1]],
                (int(nr[k]), 1)) for k in sp.sf3d_pairs]).reshape(-1)
    gs, gf = oracle.Oracle(sp).coupling(u, s0, c)
    assert np.abs(gf - wf).max() <= 1e-12 * np.abs(wf).max()
    assert np.abs(gs - ws).max() <= 1e-12 * np.abs(ws).max()


def _drop_coupling3d(spec):
    import copy
    s2 = copy.copy(spec)
    s2.sf3d_pairs = None
    s2.sf3d_normal = None
    return s2


def test_coupling3d_terms_are_adjoint():
    """With per-slice normals n(phi): <c, T u> = <S(m c), u> in the
    real-field inner product, T u = R2C(n . C2R u) the solid -> fluid
    term and S f = R2C(C2R(f) n / m) the fluid -> solid one (1D fluid
    mass).  A sign, a missing 1 / m or a phi component applied on one side
    only fails."""
    sp = _coupling3d_spec()
    o = oracle.Oracle(sp)
    rng = np.random.default_rng(21)
    u = _real_mean(rng.standard_normal((o.nmodes, 3)) +
                   1j * rng.standard_normal((o.nmodes, 3)), o.offsets)
    c = _real_mean(rng.standard_normal(o.fnmodes) +
                   1j * rng.standard_normal(o.fnmodes), o.foffsets)
    mf = np.repeat(sp.fluid_mass, sp.fluid_nr // 2 + 1)
    zs = np.zeros_like(u)
    _, tu = o.coupling(u, zs, np.zeros_like(c))
    smc, _ = o.coupling(zs, zs, mf * c)
    lhs = _wdot_modes(c, tu, sp.fluid_nr)
    rhs = -sum(_wdot_modes(smc[:, k], u[:, k], sp.solid_nr)
               for k in range(3))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_coupling3d_rotation_equivariance():
    """Normals rotated by one slice and fields rotated by 2 pi / Nr rotate
    both coupling terms (slice k at phi = 2 pi k / Nr); the reversed
    azimuth does not."""
    sp = _coupling3d_spec()
    nr = 5
    assert np.all(sp.solid_nr[sp.sf_solid] == nr)
    o = oracle.Oracle(sp)
    rng = np.random.default_rng(23)
    u = _real_mean(rng.standard_normal((o.nmodes, 3)) +
                   1j * rng.standard_normal((o.nmodes, 3)), o.offsets)
    c = _real_mean(rng.standard_normal(o.fnmodes) +
                   1j * rng.standard_normal(o.fnmodes), o.foffsets)
    a_s = np.concatenate([np.arange(n // 2 + 1) for n in sp.solid_nr])
    a_f = np.concatenate([np.arange(n // 2 + 1) for n in sp.fluid_nr])
    rs = np.exp(-1j * a_s * 2 * np.pi / nr)[:, None]
    rf = np.exp(-1j * a_f * 2 * np.pi / nr)
    zs = np.zeros_like(u)
    bs, bf = o.coupling(u, zs, c)
    n3 = sp.sf3d_normal.reshape(-1, nr, 3)
    for shift, ok in ((1, True), (-1, False)):
        sp.sf3d_normal = np.roll(n3, shift, axis=1).reshape(-1)
        gs, gf = oracle.Oracle(sp).coupling(rs * u, zs, rf * c)
        close = (np.abs(gf - rf * bf).max() <= 1e-11 * np.abs(bf).max() and
                 np.abs(gs - rs * bs).max() <= 1e-11 * np.abs(bs).max())
        assert close == ok
