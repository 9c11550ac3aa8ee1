"""Exodus mesh ingestion (axisem3d_amd.exodus, preloop.QuadMesh,
from_exodus, load_inparam, from_inparam) on the reference's SEG salt
local example (fixture tests/golden/salt_local_mesh.npz and
salt_local_inparam.json, written by make_exodus_fixture.py).

Pins: the HDF5 reader returns what the committed fixture holds (where
the reference tree is present); GLL points numbered through the
connectivity count nodes + 3 edges + 9 elements and coincide in space; the
integral factors sum to the closed-form volume per radian W^2 H / 2; the
solid-fluid normals integrate the flat sea floor (W^2 / 2 per radian);
a structured Cartesian mesh read as unstructured quads gives the same
domain, bit for bit up to the point numbering, and the same oracle
seismograms."""
import json
import os

import numpy as np
import pytest

import oracle
from axisem3d_amd import exodus, preloop

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
SRC = "/root/reference/examples/03_salt_body_SEG_local/input1D"


def fixture_mesh():
    with np.load(os.path.join(GOLD, "salt_local_mesh.npz")) as a:
        return exodus.ExodusMesh(arrays=dict(a))


def fixture_inparam():
    with open(os.path.join(GOLD, "salt_local_inparam.json")) as f:
        return json.load(f)


@pytest.mark.skipif(not os.path.isdir(SRC),
                    reason="the reference tree is not on this machine")
def test_reader_matches_fixture():
    inp = preloop.load_inparam(SRC)
    assert inp == fixture_inparam()
    ex = exodus.ExodusMesh(os.path.join(SRC, inp["exodus_mesh"]))
    fx = fixture_mesh()
    assert ex.nelem == 2000 and len(ex.coords) == 2091
    assert np.array_equal(ex.connectivity, fx.connectivity)
    assert np.array_equal(ex.coords, fx.coords)
    assert ex.side_sets == fx.side_sets
    for k, v in fx.elem_vars.items():
        assert np.array_equal(ex.elem_vars[k], v)
    assert ex.global_variables == fx.global_variables
    assert ex.records["crdsys"] == "cartesian"


def test_quadmesh_numbering_and_geometry():
    ex = fixture_mesh()
    qm = preloop.QuadMesh.from_exodus(ex)
    edges = set()
    for nd in qm.connectivity:
        for k in range(4):
            u, v = int(nd[k]), int(nd[(k + 1) % 4])
            edges.add((min(u, v), max(u, v)))
    assert qm.npoint_grid == qm.nnodes + 3 * len(edges) + 9 * qm.nelem
    gq = preloop.gradquad_operands(qm)
    sz = gq["sz"].transpose(0, 2, 1).reshape(-1, 2)
    ep = qm.element_points().reshape(-1)
    at = np.zeros((qm.npoint_grid, 2))
    at[ep] = sz
    assert np.array_equal(at[ep], sz)
    W = float(np.ptp(ex.coords[:, 0]))
    H = float(np.ptp(ex.coords[:, 1]))
    assert gq["ifact"].sum() == pytest.approx(W * W / 2 * H, rel=1e-12)
    assert qm.is_axial().sum() == len(ex.side_sets["x0"])


def test_salt_domain_from_inputs():
    ex, inp = fixture_mesh(), fixture_inparam()
    sp = preloop.from_inparam(inp, ex, 10)
    W = float(np.ptp(ex.coords[:, 0]))
    z_floor = float(ex.coords[:, 1].max()) - 200.0
    # the 200 m water layer: pairs on the flat sea floor, normals from the
    # fluid to the solid (down) integrating its area per radian
    sz = sp.solid_sz[sp.sf_solid]
    assert np.allclose(sz[:, 1], z_floor)
    assert np.abs(sp.sf_normal[:, 0]).max() < 1e-9 * W
    assert sp.sf_normal[:, 1].sum() == pytest.approx(-W * W / 2, rel=1e-12)
    assert np.all(sp.elem_nr == 1)
    assert len(sp.clayton_solid["point"]) > 0
    assert len(sp.sponge_solid["point"]) > 0
    assert sp.src_elem >= 0 and np.abs(sp.stf).max() > 0
    with pytest.raises(NotImplementedError):
        preloop.from_inparam(dict(inp, attenuation="CG4"), ex, 10)


def _as_quads(mesh):
    nodes = mesh.corner_nodes()
    na = mesh.nx + 1
    ids = np.arange(mesh.nnodes)
    coords = np.stack([mesh.node_a[ids % na], mesh.node_b[ids // na]],
                      axis=1)
    ix, iz = mesh.element_ix_iz()
    sides = {"x0": {int(e): 3 for e in np.nonzero(ix == 0)[0]},
             "x1": {int(e): 1 for e in np.nonzero(ix == mesh.nx - 1)[0]},
             "y0": {int(e): 0 for e in np.nonzero(iz == 0)[0]},
             "y1": {int(e): 2 for e in np.nonzero(iz == mesh.nz - 1)[0]}}
    return preloop.QuadMesh(coords, nodes, sides)


def test_quadmesh_equals_structured_mesh():
    mesh = preloop.StructuredMesh(np.linspace(0, 4e3, 5),
                                  np.linspace(0, 3e3, 4))
    mat = lambda r: (2600.0 + 0 * r, 6000.0 + 0 * r, 3464.0 + 0 * r)
    nr = lambda s, z: 3.0 + 6.0 * (s / 4e3) + 0.0 * z
    a = preloop.build_domain(mesh, mat, nr)
    b = preloop.build_domain(_as_quads(mesh), mat, nr)
    # the same points under another numbering
    ka = {tuple(np.round(x, 6)): i for i, x in enumerate(a.solid_sz)}
    perm = np.array([ka[tuple(np.round(x, 6))] for x in b.solid_sz])
    assert sorted(perm) == list(range(len(perm)))
    assert np.array_equal(a.elem_points[:, :], perm[b.elem_points])
    assert np.array_equal(a.solid_mass[perm], b.solid_mass)
    assert np.array_equal(a.solid_nr[perm], b.solid_nr)
    assert np.array_equal(a.elem_gq, b.elem_gq)
    assert a.dt == b.dt
    assert np.array_equal(np.sort(perm[b.axial_points]),
                          np.sort(a.axial_points))
    for sp in (a, b):
        mt = np.array([1.0, 0.7, -0.4, 0.3, 0.2, -0.1]) * 1e15
        preloop.add_axial_moment_source(sp, 1.5e3, mt, 0.26, 1.628, 80)
        preloop.add_receivers(sp, np.array([0.5e3, 3e3]),
                              np.array([3e3, 3e3]), np.array([0.3, 2.0]))
    sa = oracle.Oracle(a).run(80)
    sb = oracle.Oracle(b).run(80)
    assert np.abs(sa).max() > 0
    np.testing.assert_allclose(sb, sa, rtol=0, atol=1e-13 * np.abs(sa).max())
