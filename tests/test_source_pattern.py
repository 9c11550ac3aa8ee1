"""The Fourier mapping of the axial moment-source pattern
(preloop.moment_pattern_axial, after Mechanism.cpp:160-218), pinned by
geometry: a moment tensor on the axis is one Cartesian tensor M; its
components in the cylindrical frame (s, phi, z) at azimuth phi are
e_a(phi) . M e_b(phi), and the pattern of order alpha is their Fourier
coefficient (1/N) sum_k f(phi_k) exp(-i alpha phi_k) (the R2C convention of
the solver), times the in-plane weight over 2 pi.  Orders above 2 vanish
(a rotated second-order tensor has no higher harmonics)."""
import math

import numpy as np

from axisem3d_amd import preloop

VOIGT = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]   # ss pp zz pz sz sp


def _cylindrical_series(M, nphi):
    """[k][voigt] components of the Cartesian M in the frame (s, phi, z)
    at phi_k = 2 pi k / nphi (s along x and phi along y at phi = 0)."""
    out = np.zeros((nphi, 6))
    for k in range(nphi):
        p = 2.0 * math.pi * k / nphi
        E = np.array([[math.cos(p), math.sin(p), 0.0],
                      [-math.sin(p), math.cos(p), 0.0],
                      [0.0, 0.0, 1.0]])           # rows e_s, e_phi, e_z
        C = E @ M @ E.T
        out[k] = [C[a, b] for a, b in VOIGT]
    return out


def test_axial_pattern_is_the_fourier_series_of_the_rotated_tensor():
    rng = np.random.default_rng(11)
    for trial in range(5):
        A = rng.standard_normal((3, 3))
        M = A + A.T                                   # x, y, z Cartesian
        # [Mzz, Mss, Mpp, Mzs, Mzp, Msp] at phi = 0 (s = x, p = y)
        mt6 = [M[2, 2], M[0, 0], M[1, 1], M[0, 2], M[1, 2], M[0, 1]]
        w = rng.uniform(0.1, 1.0, 25)
        pat = preloop.moment_pattern_axial(mt6, w, nu1_elem=5)
        assert pat.shape == (3, 6, 25)
        nphi = 16
        coef = np.fft.rfft(_cylindrical_series(M, nphi), axis=0) / nphi
        assert np.abs(coef[3:]).max() < 1e-12 * np.abs(M).max()
        for alpha in range(3):
            for v in range(6):
                want = coef[alpha, v] * w[:5] / (2.0 * math.pi)
                np.testing.assert_allclose(pat[alpha, v, :5], want,
                                           rtol=0, atol=1e-14)
        # only the axis row (ipol == 0) carries the source
        assert np.all(pat[:, :, 5:] == 0)


def test_axial_pattern_orders_by_tensor_part():
    w = np.ones(25)
    iso = preloop.moment_pattern_axial([1, 1, 1, 0, 0, 0], w, 3)
    assert np.abs(iso[1:]).max() == 0 and np.abs(iso[0]).max() > 0
    shear_z = preloop.moment_pattern_axial([0, 0, 0, 1, 0.5, 0], w, 3)
    assert np.abs(shear_z[[0, 2]]).max() == 0 and np.abs(shear_z[1]).max() > 0
    horiz = preloop.moment_pattern_axial([0, 1, -1, 0, 0, 0.3], w, 3)
    assert np.abs(horiz[:2]).max() == 0 and np.abs(horiz[2]).max() > 0
    # an element whose Nr keeps only orders 0 and 1 gets the first two
    assert preloop.moment_pattern_axial([0, 1, -1, 0.2, 0, 0.3], w, 2).shape \
        == (2, 6, 25)
