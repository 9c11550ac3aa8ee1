"""Pins of the oracle's absorbing boundaries (oracle_clayton after
ClaytonSolid1D/3D.cpp and ClaytonFluid1D/3D.cpp, oracle_sponge after
Sponge.hpp:74-96) and of the preloop that builds their operands
(preloop.add_absorbing after Quad.cpp:142-231 and GLLPoint.cpp:152-215):

* physics: a plane P pulse at normal incidence on a Clayton-Enquist
  boundary leaves almost nothing behind (the condition is exact for that
  wave), where a free boundary reflects all of it; a Kosloff sponge layer
  damps what crosses it;
* the per-slice (3D) forms with constant slices reduce to the 1D forms:
  ClaytonSolid3D's rsa V + (V.k) k with k = sqrt((rho vp - rho vs) area) n
  against ClaytonSolid1D's theta form for a normal with n_s >= 0, the fluid
  forms likewise, and the 3D sponge against the 1D one up to the C2R/R2C
  round trip of the acceleration (imaginary parts of the mean and Nyquist
  orders dropped);
* dt: elements with a Clayton edge take half the time step
  (factorDtForBC, Quad.cpp:330-349)."""
import copy
import math

import numpy as np

import oracle
from axisem3d_amd import preloop


def _column(clayton, sponge, nx=30, nz=20):
    # a box of uniform solid, order 0 only: a plane P wave in z, watched
    # near the axis before the free side at s = 30 km can disturb it
    mesh = preloop.StructuredMesh(np.linspace(0, nx * 1e3, nx + 1),
                                  np.linspace(0, nz * 1e3, nz + 1))
    sp = preloop.build_domain(
        mesh, lambda r: (2600.0 + 0 * r, 6000.0 + 0 * r, 3464.0 + 0 * r),
        lambda s, z: np.ones_like(s), courant=0.5)
    if clayton or sponge:
        preloop.add_absorbing(sp, keys=("BOTTOM",), clayton=clayton,
                              sponge=sponge, rel_spans=[0.15], t0=1.0)
    return sp


def _pulse_left_behind(sp):
    """A downgoing Gaussian pulse u_z = f(z + vp t) from z = 15 km; the
    largest |u_z| left at s < 6 km, z > 5 km once the pulse (and any
    reflection from the bottom) had time to come back there (the side's
    disturbance, 22.8 km away at vp by then, has not arrived)."""
    o = oracle.Oracle(sp)
    z = sp.solid_sz[:, 1]
    z0, w = 15e3, 2.0e3
    f = np.exp(-((z - z0) / w) ** 2)
    off = o.offsets
    o.displ[off[:-1], 2] = f
    o.veloc[off[:-1], 2] = 6000.0 * (-2 * (z - z0) / w ** 2 * f)
    t_end = (15e3 + 8e3) / 6000.0
    o.run(int(round(t_end / sp.dt)), record=False)
    m = (z > 5e3) & (sp.solid_sz[:, 0] < 6e3)
    return float(np.abs(o.displ[off[:-1], 2][m]).max())


def test_clayton_absorbs_a_normally_incident_p_pulse():
    free = _pulse_left_behind(_column(False, False))
    abc = _pulse_left_behind(_column(True, False))
    print(f"left behind: free bottom {free:.3f}, Clayton {abc:.4f}")
    # free surface: the pulse comes back whole
    assert free > 0.9
    # Clayton-Enquist at normal incidence: a few percent (discretisation)
    assert abc < 0.05


def test_sponge_damps_what_crosses_it():
    free = _pulse_left_behind(_column(False, False))
    sponge = _pulse_left_behind(_column(False, True))
    print(f"left behind: free bottom {free:.3f}, sponge {sponge:.3f}")
    assert sponge < 0.85 * free


def test_clayton_halves_dt():
    a = _column(False, False)
    b = _column(True, False)
    assert abs(b.dt / a.dt - 0.5) < 2e-3


def _with_lists(spec, **lists):
    sp = copy.copy(spec)
    for k, v in lists.items():
        setattr(sp, k, v)
    return sp


def _entries(points, nval, vals, width):
    nval = np.asarray(nval, dtype=np.int32)
    off = np.zeros(len(nval), dtype=np.int64)
    off[1:] = np.cumsum(nval)[:-1]
    data = np.concatenate([np.asarray(v, dtype=np.float64).reshape(-1)
                           for v in vals])
    assert len(data) == width * nval.sum()
    return {"point": np.asarray(points, dtype=np.int32), "nval": nval,
            "off": off, "data": data}


def _small():
    return preloop.cartesian_homogeneous(
        nx=4, nz=4, width=4e3, depth=4e3, nsteps=4, n_receivers=1,
        nodal_nr=lambda s, z: 3.0 + 6.0 * (s / 4e3) + 0.0 * z)


def test_clayton_solid3d_with_constant_slices_is_clayton1d():
    sp = _small()
    rng = np.random.default_rng(3)
    pts = rng.choice(len(sp.solid_nr), 12, replace=False)
    rp = rng.uniform(1e7, 2e7, 12)
    rs = rp * rng.uniform(0.4, 0.7, 12)
    area = rng.uniform(10.0, 1e3, 12)
    theta = rng.uniform(0.0, math.pi, 12)     # n_s = sin(theta) >= 0
    c, s = np.cos(theta), np.sin(theta)
    one = [np.array([rs[i] * area[i] * c[i] ** 2 + rp[i] * area[i] * s[i] ** 2,
                     rs[i] * area[i] * s[i] ** 2 + rp[i] * area[i] * c[i] ** 2,
                     (rp[i] - rs[i]) * area[i] * c[i] * s[i],
                     rs[i] * area[i]]) for i in range(12)]
    nr = sp.solid_nr[pts]
    three = []
    for i in range(12):
        k = math.sqrt((rp[i] - rs[i]) * area[i]) * np.array([s[i], 0.0, c[i]])
        row = np.concatenate([[rs[i] * area[i]], k])
        three.append(np.tile(row, (nr[i], 1)))
    s1 = _with_lists(sp, clayton_solid=_entries(pts, np.ones(12), one, 4))
    s3 = _with_lists(sp, clayton_solid=_entries(pts, nr, three, 4))
    o1, o3 = oracle.Oracle(s1), oracle.Oracle(s3)
    v = rng.standard_normal((o1.nmodes, 3)) + 1j * rng.standard_normal(
        (o1.nmodes, 3))
    z = np.zeros_like(v)
    a1, _ = o1.clayton(v, z)
    a3, _ = o3.clayton(v, z)
    # the per-slice form acts on the real field: the mean's and the even
    # Nyquist order's imaginary parts do not enter
    proj = _project(sp.solid_nr, o1.offsets, v)
    a1p, _ = o1.clayton(proj, z)
    assert np.abs(a1).max() > 0
    assert np.allclose(a3, a1p, rtol=0, atol=1e-12 * np.abs(a1).max())


def _project(nr_of, offsets, x):
    """Zero the imaginary parts of order 0 and of an even Nr's Nyquist
    order (what C2R ignores)."""
    x = np.array(x, dtype=np.complex128, copy=True)
    for p, nr in enumerate(nr_of):
        o = offsets[p]
        x[o] = x[o].real
        if nr % 2 == 0:
            x[o + nr // 2] = x[o + nr // 2].real
    return x


def test_sponge3d_with_constant_gamma_is_the_projected_sponge1d():
    sp = _small()
    rng = np.random.default_rng(4)
    pts = rng.choice(len(sp.solid_nr), 10, replace=False)
    g = rng.uniform(0.1, 2.0, 10)
    nr = sp.solid_nr[pts]
    s1 = _with_lists(sp, sponge_solid=_entries(pts, np.ones(10), g, 1))
    s3 = _with_lists(sp, sponge_solid=_entries(
        pts, nr, [np.full(n, x) for n, x in zip(nr, g)], 1))
    o1, o3 = oracle.Oracle(s1), oracle.Oracle(s3)
    u, v, a = (rng.standard_normal((o1.nmodes, 3)) + 1j *
               rng.standard_normal((o1.nmodes, 3)) for _ in range(3))
    r1, _ = o1.sponge(u, v, a)
    r3, _ = o3.sponge(u, v, a)
    # 1D: a - 2 gamma v - gamma^2 u exactly, on every order
    m = np.zeros(len(sp.solid_nr), bool)
    m[pts] = True
    sel = np.repeat(m, sp.solid_nr // 2 + 1)
    gm = np.zeros(len(sp.solid_nr))
    gm[pts] = g
    gg = np.repeat(gm, sp.solid_nr // 2 + 1)[:, None]
    assert np.allclose(r1[sel], (a - 2 * gg * v - gg ** 2 * u)[sel],
                       rtol=1e-14, atol=1e-14)
    assert np.array_equal(r1[~sel], a[~sel])
    # 3D with constant gamma: the round trip of the acceleration projects
    # it, the rest is the 1D sponge of the projected fields
    ref = _project(sp.solid_nr, o1.offsets, r1)
    ref[~sel] = a[~sel]
    assert np.allclose(r3, ref, rtol=0, atol=1e-12 * np.abs(a).max())


def _fluid_small():
    mesh = preloop.StructuredMesh(np.linspace(0, 4e3, 5),
                                  np.linspace(0, 4e3, 5))
    sp = preloop.build_domain(
        mesh, lambda r: (1000.0 + 0 * r, 1500.0 + 0 * r, 0.0 * r),
        lambda s, z: 3.0 + 6.0 * (s / 4e3) + 0.0 * z, courant=0.4)
    ix, iz = mesh.element_ix_iz()
    preloop.make_fluid(sp, iz >= 2)
    return sp


def test_clayton_fluid3d_with_constant_slices_is_clayton_fluid1d():
    sp = _fluid_small()
    rng = np.random.default_rng(5)
    nf = len(sp.fluid_nr)
    pts = rng.choice(nf, 8, replace=False)
    c = rng.uniform(1e-8, 1e-6, 8)
    nr = sp.fluid_nr[pts]
    s1 = _with_lists(sp, clayton_fluid=_entries(pts, np.ones(8), c, 1))
    s3 = _with_lists(sp, clayton_fluid=_entries(
        pts, nr, [np.full(n, x) for n, x in zip(nr, c)], 1))
    o1, o3 = oracle.Oracle(s1), oracle.Oracle(s3)
    fv = rng.standard_normal(o1.fnmodes) + 1j * rng.standard_normal(
        o1.fnmodes)
    z3 = np.zeros((o1.nmodes, 3), complex)
    _, f1 = o1.clayton(z3, z3, fv, np.zeros_like(fv))
    _, f3 = o3.clayton(z3, z3, fv, np.zeros_like(fv))
    fvp = _project(sp.fluid_nr, o1.foffsets, fv[:, None])[:, 0]
    _, f1p = o1.clayton(z3, z3, fvp, np.zeros_like(fv))
    assert np.abs(f1).max() > 0
    assert np.allclose(f3, f1p, rtol=0, atol=1e-12 * np.abs(f1).max())


def test_add_absorbing_builds_entries_on_the_boundary_edges():
    """Clayton entries sit on the RIGHT and BOTTOM edges, one per
    boundary element edge point; their areas (integrated normals times s,
    Quad::computeNormal) add up to the axisymmetric side areas / 2 pi:
    int s dz = W H on the right side, int s ds = W^2 / 2 on the bottom;
    the normals give theta = pi / 2 on the right side (c0 = rho vp area)
    and pi on the bottom (c1 = rho vp area).  Sponge points lie within
    the layer, and dt shrinks."""
    W, H = 5e3, 6e3
    sp = preloop.cartesian_homogeneous(
        nx=5, nz=6, width=W, depth=H, nsteps=4, n_receivers=1)
    sp2 = preloop.cartesian_homogeneous(
        nx=5, nz=6, width=W, depth=H, nsteps=4, n_receivers=1)
    sp2.stf = None
    sp2.src_elem = -1
    preloop.add_absorbing(sp2, keys=("RIGHT", "BOTTOM"), sponge=True,
                          rel_spans=[0.2, 0.2], t0=1.0)
    cs = sp2.clayton_solid
    assert len(cs["point"]) == (6 + 5) * 5
    s, z = sp2.solid_sz[cs["point"], 0], sp2.solid_sz[cs["point"], 1]
    assert np.all((np.abs(s - W) < 1e-6) | (np.abs(z - 0.0) < 1e-6))
    d = cs["data"].reshape(-1, 4)
    rpa_, rsa_ = 2600.0 * 6000.0, 2600.0 * 3464.0
    area = d[:, 3] / rsa_
    right = d[:, 0] > d[:, 1]
    assert np.allclose(d[right, 0], rpa_ * area[right], rtol=1e-12)
    assert np.allclose(d[~right, 1], rpa_ * area[~right], rtol=1e-12)
    assert np.allclose(d[:, 2], 0.0, atol=1e-9 * np.abs(d).max())
    assert abs(area[right].sum() - W * H) / (W * H) < 1e-12
    assert abs(area[~right].sum() - W * W / 2) / (W * W / 2) < 1e-12
    ss = sp2.sponge_solid
    assert len(ss["point"]) > 0
    ps, pz = sp2.solid_sz[ss["point"], 0], sp2.solid_sz[ss["point"], 1]
    assert np.all((pz <= 0.2 * H + 1e-6) | (ps >= W * 0.8 - 1e-6))
    assert np.all(ss["data"] > 0)
    assert sp2.dt < sp.dt
