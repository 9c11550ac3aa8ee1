"""Pins of the oracle's particle relabelling transformation (PRT) branch of
the element stiffness (SolidElement::displToStiff, SolidElement.cpp:
115-159; PRT.hpp; Undulation::createPRT, Undulation.cpp:124-168) and of
preloop.add_undulation.

A relabelling that moves every point by a displacement linear in (s, z)
maps a mesh of straight-sided (bilinear) elements onto another such mesh
exactly.  The stiffness of the original mesh with the PRT must then equal
the stiffness of the moved mesh without it, point by point, and the PRT's
mass Jacobian must give the moved mesh's masses:
* Cartesian: dz = b z (stretch: X3, J0 = 1) on a box with axial
  elements, and dz = a s + b z (shear and stretch: X1 and X3) on a box
  off the axis, 1D and 3D (cardinal, nine series through the FFT)
  elements, with the moved mesh built as an unstructured quad mesh.  (On
  the axis the PRT keeps the s-derivative of the L'Hospital terms at
  constant z, SolidElement.cpp:118 with GradientQuadrature.hpp:426-437,
  which is the moved mesh's only where dz/ds = 0, as an axisymmetric
  undulation has there);
* spherical: dz = c r (a radial scaling of the sector: X0 = X3, J0 = 1 +
  dz / r, the gradient and its RTZ rotation), against the sector of radii
  scaled by 1 + c, isotropic, TI and anisotropic.
The identity relabelling leaves the stiffness unchanged bit for bit (the
9-component gradient and quadrature against the 6-component ones)."""
import numpy as np
import pytest

import oracle
from axisem3d_amd import preloop


def _mat(r):
    one = np.ones_like(r)
    return 2600.0 * one, 6000.0 * one, 3464.0 * one


def _nr(s, z):
    return 3.0 + 8.0 * (s / 4e3) + 0.0 * z


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _box(s0=0.0):
    mesh = preloop.StructuredMesh(np.linspace(s0, s0 + 4e3, 5),
                                  np.linspace(0, 4.8e3, 5))
    return preloop.build_domain(mesh, _mat, _nr, courant=0.4,
                                bound_nr=False)


def _random_u(o, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((o.nmodes, 3)) +
            1j * rng.standard_normal((o.nmodes, 3)))


def _match(sp_a, sp_b, moved):
    """Index of each point of sp_a in sp_b, whose coordinates are those of
    sp_a moved by `moved` (a function of (s, z))."""
    sa = np.asarray(sp_a.solid_sz)
    s2, z2 = moved(sa[:, 0], sa[:, 1])
    key_b = {(round(x, 4), round(y, 4)): i
             for i, (x, y) in enumerate(np.asarray(sp_b.solid_sz))}
    return np.array([key_b[(round(x, 4), round(y, 4))]
                     for x, y in zip(s2, z2)])


def _permute(sp_from, sp_to, idx, u):
    """u on sp_from's modes -> sp_to's modes (point i of sp_from is point
    idx[i] of sp_to, with the same Nr)."""
    nf = np.asarray(sp_from.solid_nr) // 2 + 1
    nt = np.asarray(sp_to.solid_nr) // 2 + 1
    of = np.concatenate([[0], np.cumsum(nf)])
    ot = np.concatenate([[0], np.cumsum(nt)])
    out = np.zeros((ot[-1], 3), dtype=complex)
    for i, j in enumerate(idx):
        assert nf[i] == nt[j]
        out[ot[j]:ot[j] + nt[j]] = u[of[i]:of[i] + nf[i]]
    return out


def test_identity_prt_is_bitwise_the_plain_stiffness():
    sp = _box()
    o0 = oracle.Oracle(sp)
    u = _random_u(o0, 1)
    k0 = o0.stiffness(u)
    preloop.add_undulation(sp, np.zeros((sp.n_solid_elements, 25)),
                           elems=np.arange(sp.n_solid_elements))
    assert np.all(sp.prt["x"][:, 0] == 1.0)
    assert np.array_equal(oracle.Oracle(sp).stiffness(u), k0)


def _sheared(a, b, mask3d=None, s0=0.0):
    """The box with the PRT of dz = a s + b z, and the moved box as an
    unstructured quad mesh without PRT (same Nr per point)."""
    sp = _box(s0)
    mesh = sp._mesh
    if mask3d is not None:
        preloop.make_3d_isotropic(sp, lambda x, y, z: 0.0 * x, None,
                                  mask3d(sp.n_solid_elements))
    preloop.add_undulation(sp, lambda s, z: a * s + b * z)
    nodes = np.array([[s, z] for z in mesh.node_b for s in mesh.node_a])
    nodes[:, 1] += a * nodes[:, 0] + b * nodes[:, 1]
    ix, _ = mesh.element_ix_iz()
    side = ({"x0": {int(e): 3 for e in np.nonzero(ix == 0)[0]}}
            if s0 == 0.0 else {})
    qm = preloop.QuadMesh(nodes, mesh.corner_nodes(), side)
    sq = preloop.build_domain(qm, _mat, _nr, courant=0.4, bound_nr=False)
    if mask3d is not None:
        preloop.make_3d_isotropic(sq, lambda x, y, z: 0.0 * x, None,
                                  mask3d(sq.n_solid_elements))
    idx = _match(sp, sq, lambda s, z: (s, z + a * s + b * z))
    return sp, sq, idx


def _check_moved(sp, sq, idx, tol):
    assert _rel(np.asarray(sp.solid_mass),
                np.asarray(sq.solid_mass)[idx]) < tol
    o_p, o_q = oracle.Oracle(sp), oracle.Oracle(sq)
    u = _random_u(o_p, 7)
    kp = o_p.stiffness(u)
    kq = o_q.stiffness(_permute(sp, sq, idx, u))
    err = _rel(_permute(sp, sq, idx, kp), kq)
    print(f"PRT vs moved mesh: relative stiffness difference {err:.2e}")
    assert err < tol


def test_prt_stretch_with_axis_equals_the_moved_mesh():
    sp, sq, idx = _sheared(0.0, 0.11)
    assert np.any(np.asarray(sp.elem_axial) == 1)
    _check_moved(sp, sq, idx, 1e-12)


def test_prt_shear_and_stretch_equal_the_moved_mesh_1d():
    sp, sq, idx = _sheared(0.07, 0.11, s0=2e3)
    assert np.any(sp.prt["x"][:, 1] != 0.0)
    _check_moved(sp, sq, idx, 1e-12)


def test_prt_shear_and_stretch_equal_the_moved_mesh_3d():
    # every other element 3D (cardinal branch: nine series through the FFT)
    sp, sq, idx = _sheared(-0.05, 0.09, lambda E: (np.arange(E) % 2) == 0,
                           s0=2e3)
    assert np.any(np.asarray(sp.elem_kind) == 1)
    _check_moved(sp, sq, idx, 1e-12)


def _sector(scale, mtype=0):
    th = np.linspace(0.0, np.radians(2.0), 5)
    rr = np.linspace(6331e3, 6371e3, 5) * scale
    mesh = preloop.StructuredMesh(th, rr, spherical=True)
    sp = preloop.build_domain(mesh, _mat,
                              lambda s, z: np.full(s.shape, 7.0),
                              courant=0.4, bound_nr=False)
    if mtype:
        E = sp.n_solid_elements
        preloop.make_media(sp, np.full(E, mtype, np.uint8), seed=4,
                           perturb=0.1)
    return sp


@pytest.mark.parametrize("mtype", [0, 1, 2])
def test_prt_radial_scaling_equals_the_scaled_sector(mtype):
    # TI and anisotropic media (seeded, the same on both sectors) take the
    # 9-component rotation to RTZ of the PRT branch on one side and the
    # Voigt rotation of the plain branch on the other
    c = 0.013
    sp = _sector(1.0, mtype)
    preloop.add_undulation(sp, lambda s, z: c * np.hypot(s, z))
    # the gradient of dz (about 8e4 m) is c to about 1e-11: the GLL
    # contraction cancels terms 1e4 times larger than the result
    x = sp.prt["x"]
    assert np.allclose(x[:, 1], 0.0, atol=1e-10)
    assert np.allclose(x[:, 0], 1.0 / (1.0 + c), rtol=1e-13)
    assert np.allclose(x[:, 3], 1.0 / (1.0 + c), rtol=1e-10)
    sq = _sector(1.0 + c, mtype)
    idx = _match(sp, sq, lambda s, z: (s * (1 + c), z * (1 + c)))
    assert np.array_equal(idx, np.arange(len(idx)))
    _check_moved(sp, sq, idx, 1e-10)


def test_undulation_refuses_a_distorted_element():
    sp = _box()
    try:
        preloop.add_undulation(sp, lambda s, z: -1.5 * z)
    except ValueError as exc:
        assert "Jacobian" in str(exc)
    else:
        raise AssertionError("a folded element was accepted")


def test_undulation_refuses_fluids_and_boundaries():
    mesh = preloop.StructuredMesh(np.linspace(0, 4e3, 5),
                                  np.linspace(0, 4.8e3, 5))
    sp = preloop.build_domain(mesh, _mat, _nr, courant=0.4)
    ix, iz = mesh.element_ix_iz()
    preloop.make_fluid(sp, iz == 3)
    with pytest.raises(ValueError, match="solid domains"):
        preloop.add_undulation(sp, lambda s, z: 0.01 * z)
    sp = _box()
    preloop.add_absorbing(sp, keys=("RIGHT",))
    with pytest.raises(ValueError, match="solid domains"):
        preloop.add_undulation(sp, lambda s, z: 0.01 * z)


def test_identity_prt_leaves_the_moment_source_bitwise():
    # the source's PRT branch (SolidElement.cpp:293-331) with X = 1, no
    # rotation: the 9-component quadrature of the mirrored stress
    sp = _box()
    mt = np.array([1.0, 0.7, -0.4, 0.3, 0.2, -0.1]) * 1e15
    preloop.add_axial_moment_source(sp, 2.0e3, mt, 0.26, 1.628, 50)
    o0 = oracle.Oracle(sp)
    f0 = o0.apply_source(np.zeros((o0.nmodes, 3), complex), 20)
    preloop.add_undulation(sp, np.zeros((sp.n_solid_elements, 25)),
                           elems=np.arange(sp.n_solid_elements))
    o1 = oracle.Oracle(sp)
    f1 = o1.apply_source(np.zeros((o1.nmodes, 3), complex), 20)
    assert np.abs(f0).max() > 0
    assert np.array_equal(f0, f1)
