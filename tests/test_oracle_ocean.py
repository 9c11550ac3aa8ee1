"""The oracle's ocean-load mass (MassOceanLoad1D/3D) pinned by what the
physics fixes: the ocean moves with the normal motion of the sea floor
only, so the normal component of a force is divided by m + m0 and the
tangential ones by m; the ocean masses sum to the integrated density
times the top surface area (per radian of azimuth: W^2 / 2 for a flat top
of width W); the 3D form with a constant mass and ocean equals the 1D
form; with an ocean that varies in azimuth the 3D form is the per-slice
operator between numpy's transforms."""
import math

import numpy as np
import pytest

import oracle
from axisem3d_amd import preloop

SRD = 1.03e6          # 1 km of sea water, kg/m^2


def _box():
    return preloop.cartesian_homogeneous(
        nx=4, nz=3, width=4e3, depth=3e3, nsteps=1, src_depth=1e3,
        n_receivers=2,
        nodal_nr=lambda s, z: 3.0 + 6.0 * (s / 4e3) + 0.0 * z)


def _sector():
    sp = preloop.spherical_regional(
        ntheta=4, nrad=3, theta_max_deg=8.0, depth=60e3, nu_min=3,
        nu_max=8, nsteps=1, n_receivers=2, attenuation=False)
    preloop.finish_regional(sp, 1, src_depth=20e3, f_peak=0.5,
                            n_receivers=2, theta_max_deg=8.0)
    return sp


def _accel(sp, seed=3):
    o = oracle.Oracle(sp)
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((o.nmodes, 3)) + 1j * rng.standard_normal(
        (o.nmodes, 3))
    off = np.concatenate([[0], np.cumsum(sp.solid_nr // 2 + 1)])
    return f, o.stiff_to_accel(f.copy()), off


def _nyquist_real(c, nr):
    c = c.copy()
    if nr % 2 == 0:
        c[-1] = c[-1].real
    return c


@pytest.mark.parametrize("make", [_box, _sector])
def test_ocean1d_divides_normal_by_m_plus_m0(make):
    sp = preloop.add_ocean_load(make(), lambda s, z, phi: SRD)
    oc = sp.ocean_load
    assert len(oc["point"]) > 0 and np.all(oc["nval"] == 1)
    f, a, off = _accel(sp)
    checked = 0
    for p in oc["point"]:
        nr = int(sp.solid_nr[p])
        m = float(sp.solid_mass[p])
        s, z = sp.solid_sz[p]
        r = math.hypot(s, z)
        # the top of the box is flat (normal +z); the sector's is radial
        t = 0.0 if make is _box else math.acos(z / r)
        # m0 = SRD |nTop|: recover it from the entry's 1/(m + m0)
        o = oc["data"][oc["off"][p == oc["point"]][0] * 4:][:4]
        c, si = math.cos(t), math.sin(t)
        fs, fp, fz = (_nyquist_real(f[off[p]:off[p + 1], k], nr)
                      for k in range(3))
        fn = fs * si + fz * c
        ft = fs * c - fz * si
        im0 = o[0] * si * si + o[1] * c * c + 2 * o[2] * c * si
        # 1/(m + m0) from the factors must be below 1/m
        assert im0 < 1.0 / m
        an = fn * im0
        at = ft / m
        np.testing.assert_allclose(a[off[p]:off[p + 1], 0],
                                   an * si + at * c, rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(a[off[p]:off[p + 1], 2],
                                   an * c - at * si, rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(a[off[p]:off[p + 1], 1], fp / m,
                                   rtol=1e-12, atol=1e-300)
        checked += 1
    assert checked == len(oc["point"])


def test_ocean_masses_sum_to_density_times_area():
    # per radian of azimuth the flat top of width W has area W^2 / 2, so
    # the ocean masses m0 = SRD |nTop| sum to SRD W^2 / 2
    sp = preloop.add_ocean_load(_box(), lambda s, z, phi: SRD)
    oc = sp.ocean_load
    total = 0.0
    for i, p in enumerate(oc["point"]):
        o = oc["data"][4 * oc["off"][i]:4 * oc["off"][i] + 4]
        m = float(sp.solid_mass[p])
        # flat top: t = 0, o[1] = 1/(m + m0)
        assert abs(o[2]) == 0.0 and o[0] == pytest.approx(1.0 / m)
        total += 1.0 / o[1] - m
    assert total == pytest.approx(SRD * 4e3 ** 2 / 2, rel=1e-12)


def test_ocean3d_constant_equals_ocean1d():
    sp1 = preloop.add_ocean_load(_sector(), lambda s, z, phi: SRD)
    sp3 = _sector()
    top = sp1.ocean_load["point"]
    preloop.make_mass3d(sp3, top, amp=0.0, seed=1)
    preloop.add_ocean_load(sp3, lambda s, z, phi: SRD)
    assert np.all(sp3.ocean_load["nval"] == sp3.solid_nr[top])
    assert len(sp3.mass3d_points) == 0
    _, a1, off = _accel(sp1)
    _, a3, _ = _accel(sp3)
    for p in top:
        # the slices carry no imaginary part of order 0 (nor of an even
        # Nr's Nyquist order): the 3D form drops it, the 1D form keeps it
        ref = a1[off[p]:off[p + 1]].copy()
        ref[0] = ref[0].real
        if sp1.solid_nr[p] % 2 == 0:
            ref[-1] = ref[-1].real
        np.testing.assert_allclose(a3[off[p]:off[p + 1]], ref, rtol=0,
                                   atol=1e-13 * np.abs(ref).max())


def test_ocean3d_varying_matches_numpy():
    sp = _sector()
    srd = lambda s, z, phi: SRD * (1.0 + 0.3 * np.cos(phi) +
                                   0.1 * np.sin(2 * phi))
    preloop.add_ocean_load(sp, srd)
    oc = sp.ocean_load
    f, a, off = _accel(sp, seed=5)
    checked = 0
    for i, p in enumerate(oc["point"]):
        nr = int(sp.solid_nr[p])
        nu1 = nr // 2 + 1
        assert oc["nval"][i] == nr
        q = oc["data"][4 * oc["off"][i]:4 * (oc["off"][i] + nr)].reshape(nr, 4)
        c = _nyquist_real(f[off[p]:off[p + 1]], nr)
        phi = 2 * np.pi * np.arange(nr) / nr
        F = np.zeros((nr, 3))
        for k in range(3):
            F[:, k] = c[0, k].real
            for al in range(1, nu1):
                F[:, k] += 2 * np.real(c[al, k] * np.exp(1j * al * phi))
        dot = np.sum(F * q[:, 1:], axis=1)
        A = q[:, :1] * F - dot[:, None] * q[:, 1:]
        for k in range(3):
            ref = np.fft.rfft(A[:, k])[:nu1] / nr
            if nr % 2 == 0:
                ref[-1] *= 0.5
            np.testing.assert_allclose(a[off[p]:off[p + 1], k], ref, rtol=0,
                                       atol=1e-12 * np.abs(ref).max())
        # the normal of the sector's top is radial: k lies along it
        s, z = sp.solid_sz[p]
        n = np.array([s, 0.0, z]) / math.hypot(s, z)
        kn = q[:, 1:] / np.linalg.norm(q[:, 1:], axis=1)[:, None]
        np.testing.assert_allclose(kn, np.broadcast_to(n, kn.shape),
                                   atol=1e-12)
        checked += 1
    assert checked == len(oc["point"]) > 0
