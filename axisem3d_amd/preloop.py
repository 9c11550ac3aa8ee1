"""Preloop: synthetic meshes and the exact operands of the stiffness path.

The reference builds its operands once, in FP64, from an Exodus mesh
(``SOLVER/src/preloop``).  This module builds the same operands from
seeded synthetic structured meshes (Cartesian boxes and spherical sectors),
following the reference formulas:

* element mapping and Jacobian: ``MappingLinear.hpp:23-58``,
  ``MappingSpherical.hpp:35-86`` (curved outer edge = edge 2);
* inverse-Jacobian entries and 1/s: ``Quad.hpp:157-184``;
* integral factor with the axial L'Hospital rule: ``Quad.cpp:404-435``;
* 1D material interpolated bilinearly from the element corners:
  ``Material.cpp:26-60`` and ``spectrals.hpp:330-370``;
* isotropic moduli lambda = rho (vp^2 - 2 vs^2), mu = rho vs^2:
  ``Material.cpp:566-582``; mass = rho * IF assembled on points:
  ``Material.cpp:308-346``, ``Quad.cpp:100-121``;
* dt = courant * min_e hmin / vmax, rounded to 4 significant figures:
  ``Quad.cpp:271-349``, ``SE_Model.cpp:178-204``;
* nodal Nr -> GLL Nr (round, lucky numbers) -> element Nr = max of its
  points: ``Quad.cpp:50-63``, ``Element.cpp:28-35``.

Element point index: ``ipnt = ipol * 5 + jpol`` with xi along ipol.  On an
axial element xi = -1 (ipol = 0) lies on the axis and uses GLJ nodes.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Callable, Optional

import numpy as np

from . import spectral

NPED = spectral.NPED
NPEM = spectral.NPEM


def round_half_away(x):
    """std::round semantics (Eigen ``.round()``): halves away from zero."""
    x = np.asarray(x, dtype=np.float64)
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


# ---------------------------------------------------------------- geometry
@dataclasses.dataclass
class StructuredMesh:
    """nx elements along xi (s or theta) by nz along eta (z or r).

    ``node_a`` (nx+1) lateral node coordinates (s for Cartesian, theta in
    radians for spherical); ``node_b`` (nz+1) vertical node coordinates
    (z for Cartesian, r for spherical).  Element (ix, iz) has corner nodes
    0:(ix,iz) 1:(ix+1,iz) 2:(ix+1,iz+1) 3:(ix,iz+1), i.e. xi along ix and
    eta along iz, matching ``gNodeIPnt = {0, 20, 24, 4}``.
    """
    node_a: np.ndarray
    node_b: np.ndarray
    spherical: bool = False
    fluid_rows: Optional[np.ndarray] = None   # bool per iz, True = fluid

    @property
    def nx(self):
        return len(self.node_a) - 1

    @property
    def nz(self):
        return len(self.node_b) - 1

    @property
    def nelem(self):
        return self.nx * self.nz

    def element_ix_iz(self):
        e = np.arange(self.nelem)
        return e % self.nx, e // self.nx

    def is_axial(self):
        ix, _ = self.element_ix_iz()
        return (ix == 0) & (abs(self.node_a[0]) < 1e-12)

    def nodal_sz(self):
        """(E, 2, 4) nodal (s, z) of each element."""
        ix, iz = self.element_ix_iz()
        a0, a1 = self.node_a[ix], self.node_a[ix + 1]
        b0, b1 = self.node_b[iz], self.node_b[iz + 1]
        aa = np.stack([a0, a1, a1, a0], axis=1)
        bb = np.stack([b0, b0, b1, b1], axis=1)
        if self.spherical:
            s = bb * np.sin(aa)
            z = bb * np.cos(aa)
        else:
            s, z = aa, bb
        return np.stack([s, z], axis=1)

    def map_points(self, xieta, axial):
        """Mapping and Jacobian at reference coords.

        xieta: (2, 25).  Returns sz (E, 2, 25) and J (E, 25, 2, 2) with
        J = d(s,z)/d(xi,eta).
        """
        ix, iz = self.element_ix_iz()
        xi = xieta[0][None, :]
        eta = xieta[1][None, :]
        if not self.spherical:
            # MappingLinear: bilinear shape functions on the 4 nodes
            nsz = self.nodal_sz()
            xip, xim, etp, etm = 1 + xi, 1 - xi, 1 + eta, 1 - eta
            shp = np.stack([xim * etm, xip * etm, xip * etp, xim * etp],
                           axis=1) * 0.25  # (1,4,25)
            sz = np.einsum('eij,xjp->eip', nsz, shp)
            d_xi = np.stack([-etm, etm, etp, -etp], axis=1) * 0.25
            d_eta = np.stack([-xim, -xip, xip, xim], axis=1) * 0.25
            J = np.empty((self.nelem, NPEM, 2, 2))
            dxi = np.einsum('eij,xjp->eip', nsz, d_xi)
            deta = np.einsum('eij,xjp->eip', nsz, d_eta)
            J[:, :, 0, 0] = dxi[:, 0]
            J[:, :, 1, 0] = dxi[:, 1]
            J[:, :, 0, 1] = deta[:, 0]
            J[:, :, 1, 1] = deta[:, 1]
            return sz, J
        # MappingSpherical with curved outer edge 2 (identity rotation):
        # t0 = t3 = theta_ix, t1 = t2 = theta_ix+1, r0 = r_iz, r2 = r_iz+1
        t0 = self.node_a[ix][:, None]
        t1 = self.node_a[ix + 1][:, None]
        t2, t3 = t1, t0
        r0 = self.node_b[iz][:, None]
        r2 = self.node_b[iz + 1][:, None]
        t01 = ((1. - xi) * t0 + (1. + xi) * t1) * .5
        t32 = ((1. - xi) * t3 + (1. + xi) * t2) * .5
        r0m = (1. - eta) * r0 * .5
        r2p = (1. + eta) * r2 * .5
        sz = np.stack([r0m * np.sin(t01) + r2p * np.sin(t32),
                       r0m * np.cos(t01) + r2p * np.cos(t32)], axis=1)
        t01_xi = (t1 - t0) * .5
        t32_xi = (t2 - t3) * .5
        r0m_eta = -r0 * .5
        r2p_eta = r2 * .5
        J = np.empty((self.nelem, NPEM, 2, 2))
        J[:, :, 0, 0] = r0m * np.cos(t01) * t01_xi + r2p * np.cos(t32) * t32_xi
        J[:, :, 1, 0] = -(r0m * np.sin(t01) * t01_xi +
                          r2p * np.sin(t32) * t32_xi)
        J[:, :, 0, 1] = r0m_eta * np.sin(t01) + r2p_eta * np.sin(t32)
        J[:, :, 1, 1] = r0m_eta * np.cos(t01) + r2p_eta * np.cos(t32)
        return sz, J

    def element_points(self):
        """(E, 25) global point index on the structured GLL grid.

        Point id = gz * (4 nx + 1) + gs with gs = 4 ix + ipol and
        gz = 4 iz + jpol.
        """
        ix, iz = self.element_ix_iz()
        ipol = np.repeat(np.arange(NPED), NPED)
        jpol = np.tile(np.arange(NPED), NPED)
        gs = 4 * ix[:, None] + ipol[None, :]
        gz = 4 * iz[:, None] + jpol[None, :]
        return (gz * (4 * self.nx + 1) + gs).astype(np.int64)

    @property
    def npoint_grid(self):
        return (4 * self.nx + 1) * (4 * self.nz + 1)

    def corner_nodes(self):
        """(E, 4) node ids of the element corners (nodes 0..3)."""
        ix, iz = self.element_ix_iz()
        na = self.nx + 1
        return np.stack([iz * na + ix, iz * na + ix + 1,
                         (iz + 1) * na + ix + 1, (iz + 1) * na + ix], axis=1)

    @property
    def nnodes(self):
        return (self.nx + 1) * (self.nz + 1)

    def boundary_edges(self, key):
        """(element, edge) of the mesh side `key` (LEFT, RIGHT, BOTTOM,
        TOP)."""
        ix, iz = self.element_ix_iz()
        sel = {"LEFT": ix == 0, "RIGHT": ix == self.nx - 1,
               "BOTTOM": iz == 0, "TOP": iz == self.nz - 1}[key]
        edge = {"LEFT": 3, "RIGHT": 1, "BOTTOM": 0, "TOP": 2}[key]
        return [(int(m), edge) for m in np.nonzero(sel)[0]]

    def interface_edges(self, fluid):
        """(element, edge) of every element edge whose neighbour across it
        has the other medium (fluid: bool per element)."""
        ix, iz = self.element_ix_iz()
        cell = -np.ones((self.nx, self.nz), dtype=np.int64)
        cell[ix, iz] = np.arange(self.nelem)
        out = []
        for e in range(self.nelem):
            for edge, (dx, dz) in ((0, (0, -1)), (1, (1, 0)), (2, (0, 1)),
                                   (3, (-1, 0))):
                jx, jz = ix[e] + dx, iz[e] + dz
                if 0 <= jx < self.nx and 0 <= jz < self.nz and \
                        fluid[cell[jx, jz]] != fluid[e]:
                    out.append((e, edge))
        return out

    def edge_points(self, edge):
        """Element point indices on an edge (``vicinity.hpp:52-56``)."""
        return {0: [0, 5, 10, 15, 20], 1: [20, 21, 22, 23, 24],
                2: [24, 19, 14, 9, 4], 3: [4, 3, 2, 1, 0]}[edge]


# element-local points of the corners (nodes 0..3) and, per edge u -> v
# (edge k joins node k and node k+1 mod 4), its three interior points in
# order from u
_CORNER_IPNT = (0, 20, 24, 4)
_EDGE_INNER = {0: (5, 10, 15), 1: (21, 22, 23), 2: (19, 14, 9),
               3: (3, 2, 1)}
_INNER_IPNT = tuple(i * NPED + j for i in range(1, 4) for j in range(1, 4))


# coordinate rotations that bring an element's curved outer edge k to edge
# 2 (Mapping::sOrthogQ2)
_Q2 = (np.array([[-1.0, 0.0], [0.0, -1.0]]), np.array([[0.0, -1.0],
                                                         [1.0, 0.0]]),
       np.eye(2), np.array([[0.0, 1.0], [-1.0, 0.0]]))


def _angles_close(a, b):
    """Mapping::makeAnglesClose: shift one of two angles by 2 pi so that
    they differ by at most pi."""
    if a - b > math.pi:
        b += 2.0 * math.pi
    if b - a > math.pi:
        a += 2.0 * math.pi
    return a, b


def _curved_outer(nsz, semi):
    """The curved outer edge of a spherical (semi=False: edges 0 and 2 or
    1 and 3 on circles) or semi-spherical element (one edge on a circle),
    as the MappingSpherical / MappingSemiSpherical constructors find it."""
    r = np.hypot(nsz[0], nsz[1])
    edge = min(np.hypot(*(nsz[:, k] - nsz[:, (k + 1) % 4])) for k in range(4))
    tol = edge * 1e-3
    if not semi:
        if abs(r[0] - r[1]) < tol and abs(r[2] - r[3]) < tol:
            return 2 if r[2] > r[0] else 0
        if abs(r[1] - r[2]) < tol and abs(r[3] - r[0]) < tol:
            return 3 if r[3] > r[1] else 1
    else:
        for c in range(4):
            if abs(r[c] - r[(c + 1) % 4]) < tol and r[c] > r[(c + 2) % 4]:
                return c
    raise ValueError("QuadMesh: cannot identify the curved outer edge of a "
                     "(semi-)spherical element")


def _map_curved(nsz, xieta, semi):
    """MappingSpherical / MappingSemiSpherical of one element: rotate so
    the curved outer edge is edge 2, interpolate the angle along the outer
    edge (and along the inner edge for the spherical form, the straight
    inner edge for the semi-spherical one), rotate back.  Returns sz (2,
    25) and J (25, 2, 2) = d(s,z)/d(xi,eta)."""
    c = _curved_outer(nsz, semi)
    Q = _Q2[c]
    szq = Q @ nsz
    xq = Q @ xieta
    rr = np.hypot(szq[0], szq[1])
    th = 0.5 * math.pi - np.arctan2(szq[1], szq[0])
    k = lambda d: (c + d) % 4
    xi, eta = xq[0], xq[1]
    t2, t3 = float(th[k(0)]), float(th[k(1)])
    t3, t2 = _angles_close(t3, t2)
    r2 = float(rr[k(0)])
    t32 = ((1.0 - xi) * t3 + (1.0 + xi) * t2) * 0.5
    t32_xi = (t2 - t3) * 0.5
    r2p = (1.0 + eta) * r2 * 0.5
    Jq = np.empty((xieta.shape[1], 2, 2))
    if not semi:
        r0 = float(rr[k(-2)])
        t0, t1 = float(th[k(-2)]), float(th[k(-1)])
        t0, t1 = _angles_close(t0, t1)
        t01 = ((1.0 - xi) * t0 + (1.0 + xi) * t1) * 0.5
        t01_xi = (t1 - t0) * 0.5
        r0m = (1.0 - eta) * r0 * 0.5
        s = r0m * np.sin(t01) + r2p * np.sin(t32)
        z = r0m * np.cos(t01) + r2p * np.cos(t32)
        Jq[:, 0, 0] = r0m * np.cos(t01) * t01_xi + r2p * np.cos(t32) * t32_xi
        Jq[:, 1, 0] = -(r0m * np.sin(t01) * t01_xi +
                        r2p * np.sin(t32) * t32_xi)
        Jq[:, 0, 1] = -0.5 * r0 * np.sin(t01) + 0.5 * r2 * np.sin(t32)
        Jq[:, 1, 1] = -0.5 * r0 * np.cos(t01) + 0.5 * r2 * np.cos(t32)
    else:
        s0, z0 = szq[0, k(-2)], szq[1, k(-2)]
        s1, z1 = szq[0, k(-1)], szq[1, k(-1)]
        s01 = ((1.0 - xi) * s0 + (1.0 + xi) * s1) * 0.5
        z01 = ((1.0 - xi) * z0 + (1.0 + xi) * z1) * 0.5
        etam = (1.0 - eta) * 0.5
        s = etam * s01 + r2p * np.sin(t32)
        z = etam * z01 + r2p * np.cos(t32)
        Jq[:, 0, 0] = etam * (s1 - s0) * 0.5 + r2p * np.cos(t32) * t32_xi
        Jq[:, 1, 0] = etam * (z1 - z0) * 0.5 - r2p * np.sin(t32) * t32_xi
        Jq[:, 0, 1] = -0.5 * s01 + 0.5 * r2 * np.sin(t32)
        Jq[:, 1, 1] = -0.5 * z01 + 0.5 * r2 * np.cos(t32)
    sz = Q.T @ np.stack([s, z])
    J = np.einsum("ji,pjk,kl->pil", Q, Jq, Q)
    return sz, J


class QuadMesh:
    """An unstructured mesh of 4-node quads, as an AxiSEM3D Exodus mesh
    gives it (ExodusMesh.cpp:205-619): corner nodes and coordinates, the
    geometry type per element (0 spherical, 1 linear, 2 semi-spherical;
    MappingSpherical/Linear/SemiSpherical.hpp), axial elements (a side in
    the LEFT side set, x0 or t0, which must be element side 3, xi = -1),
    side sets for the boundaries (x0/x1/y0/y1 in Cartesian meshes,
    t0/t1/r0/r1 in spherical ones), the solid-fluid interface and the
    medium per element.  GLL points are numbered through the connectivity
    (corner nodes, edges by their two nodes, interiors), which is exact for
    a conforming mesh."""

    def __init__(self, coords, connectivity, side_sets, fluid=None,
                 geometry_type=None, spherical=False):
        self.coords = np.asarray(coords, dtype=np.float64)
        self.connectivity = np.asarray(connectivity, dtype=np.int64)
        self.side_sets = side_sets
        self.spherical = bool(spherical)
        E = len(self.connectivity)
        self.geometry_type = (np.ones(E, np.int64) if geometry_type is None
                              else np.asarray(geometry_type, np.int64))
        self.fluid = (np.zeros(E, bool) if fluid is None else
                      np.asarray(fluid, dtype=bool))
        self._keys = ({"LEFT": "t0", "RIGHT": "t1", "BOTTOM": "r0",
                       "TOP": "r1"} if self.spherical else
                      {"LEFT": "x0", "RIGHT": "x1", "BOTTOM": "y0",
                       "TOP": "y1"})
        self._axial = np.zeros(E, dtype=bool)
        for e, side in side_sets.get(self._keys["LEFT"], {}).items():
            if side != 3:
                raise ValueError("QuadMesh: an axial element must have the "
                                 "axis on its side 3 (xi = -1)")
            self._axial[e] = True
        self._number_points()

    @classmethod
    def from_exodus(cls, ex):
        return cls(ex.coords, ex.connectivity, ex.side_sets, ex.fluid,
                   ex.geometry_type, spherical=not ex.cartesian)

    @property
    def nelem(self):
        return len(self.connectivity)

    def is_axial(self):
        return self._axial

    def corner_nodes(self):
        return self.connectivity

    @property
    def nnodes(self):
        return len(self.coords)

    def nodal_sz(self):
        return self.coords[self.connectivity].transpose(0, 2, 1).copy()

    def map_points(self, xieta, axial):
        """The element mappings at reference coords (2, 25): sz (E, 2, 25)
        and J (E, 25, 2, 2) = d(s,z)/d(xi,eta); MappingLinear (bilinear
        shape functions on the 4 nodes) for geometry type 1, the curved
        forms for 0 and 2."""
        sz, J = self._map_linear(xieta)
        nsz = self.nodal_sz()
        for e in np.nonzero(self.geometry_type != 1)[0]:
            sz[e], J[e] = _map_curved(nsz[e], xieta,
                                      self.geometry_type[e] == 2)
        return sz, J

    def _map_linear(self, xieta):
        xi = xieta[0][None, :]
        eta = xieta[1][None, :]
        nsz = self.nodal_sz()
        xip, xim, etp, etm = 1 + xi, 1 - xi, 1 + eta, 1 - eta
        shp = np.stack([xim * etm, xip * etm, xip * etp, xim * etp],
                       axis=1) * 0.25
        sz = np.einsum('eij,xjp->eip', nsz, shp)
        d_xi = np.stack([-etm, etm, etp, -etp], axis=1) * 0.25
        d_eta = np.stack([-xim, -xip, xip, xim], axis=1) * 0.25
        J = np.empty((self.nelem, xieta.shape[1], 2, 2))
        dxi = np.einsum('eij,xjp->eip', nsz, d_xi)
        deta = np.einsum('eij,xjp->eip', nsz, d_eta)
        J[:, :, 0, 0] = dxi[:, 0]
        J[:, :, 1, 0] = dxi[:, 1]
        J[:, :, 0, 1] = deta[:, 0]
        J[:, :, 1, 1] = deta[:, 1]
        return sz, J

    def _number_points(self):
        E = self.nelem
        ep = np.empty((E, NPEM), dtype=np.int64)
        keys = {}

        def pid(key):
            if key not in keys:
                keys[key] = len(keys)
            return keys[key]
        for e in range(E):
            nd = self.connectivity[e]
            for k in range(4):
                ep[e, _CORNER_IPNT[k]] = pid(("n", int(nd[k])))
            for k in range(4):
                u, v = int(nd[k]), int(nd[(k + 1) % 4])
                for i, ip in enumerate(_EDGE_INNER[k]):
                    pos = i if u < v else 2 - i
                    ep[e, ip] = pid(("e", min(u, v), max(u, v), pos))
            for i, ip in enumerate(_INNER_IPNT):
                ep[e, ip] = pid(("i", e, i))
        self._ep = ep
        self._npoints = len(keys)

    def element_points(self):
        return self._ep.copy()

    @property
    def npoint_grid(self):
        return self._npoints

    def boundary_edges(self, key):
        name = self._keys[key]
        return sorted((int(e), int(s)) for e, s in
                      self.side_sets.get(name, {}).items())

    def interface_edges(self, fluid):
        fluid = np.asarray(fluid, dtype=bool)
        owner = {}
        for e in range(self.nelem):
            nd = self.connectivity[e]
            for k in range(4):
                u, v = int(nd[k]), int(nd[(k + 1) % 4])
                owner.setdefault((min(u, v), max(u, v)), []).append((e, k))
        out = []
        for lst in owner.values():
            if len(lst) == 2 and fluid[lst[0][0]] != fluid[lst[1][0]]:
                out.extend(lst)
        return sorted(out)


# ------------------------------------------------------------ GradQuad ops
def gradquad_operands(mesh: StructuredMesh):
    """Per-element GradQuad operands, each (E, 25) float64.

    Returns dict with dsdxii, dsdeta, dzdxii, dzdeta, inv_s, ifact, sz, J.
    Names follow ``Quad::createGradient`` (``Quad.hpp:166-178``): they are
    inverse-Jacobian entries named after the Jacobian entries.
    """
    axial = mesh.is_axial()
    E = mesh.nelem
    out = {k: np.empty((E, NPEM)) for k in
           ("dsdxii", "dsdeta", "dzdxii", "dzdeta", "inv_s", "ifact")}
    sz_all = np.empty((E, 2, NPEM))
    J_all = np.empty((E, NPEM, 2, 2))
    for ax in (False, True):
        sel = np.nonzero(axial == ax)[0]
        if len(sel) == 0:
            continue
        xieta = spectral.xieta_element(ax)
        w = spectral.weights_element(ax)
        sz, J = mesh.map_points(xieta, ax)
        sz, J = sz[sel], J[sel]
        det = J[..., 0, 0] * J[..., 1, 1] - J[..., 0, 1] * J[..., 1, 0]
        out["dsdxii"][sel] = J[..., 0, 0] / det
        out["dsdeta"][sel] = -J[..., 0, 1] / det
        out["dzdxii"][sel] = -J[..., 1, 0] / det
        out["dzdeta"][sel] = J[..., 1, 1] / det
        s = sz[:, 0, :]
        ifact = w[None, :] * det
        if ax:
            on_axis = (np.arange(NPEM) // NPED) == 0
            inv_s = np.where(on_axis[None, :], 0.0,
                             1.0 / np.where(on_axis[None, :], 1.0, s))
            one_xi = np.where(on_axis, 1.0, 1.0 + xieta[0])[None, :]
            ifact = np.where(on_axis[None, :], ifact * J[..., 0, 0],
                             ifact * s / one_xi)
        else:
            inv_s = 1.0 / s
            ifact = ifact * s
        out["inv_s"][sel] = inv_s
        out["ifact"][sel] = ifact
        sz_all[sel] = sz
        J_all[sel] = J
    out["sz"] = sz_all
    out["J"] = J_all
    return out


def interpolate_gll(nodal, axial):
    """Bilinear interpolation of nodal values (E, 4) to the 25 GLL points
    (``spectrals::interpolateGLL``, ``spectrals.hpp:330-370``)."""
    nodal = np.asarray(nodal, dtype=np.float64)
    out = np.empty((nodal.shape[0], NPEM))
    for ax in (False, True):
        sel = np.nonzero(axial == ax)[0]
        if len(sel) == 0:
            continue
        xi, eta = spectral.xieta_element(ax)
        lsf = np.stack([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta),
                        (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)]) / 4.0
        out[sel] = nodal[sel] @ lsf
    return out


def compute_dt(sz, vmax, courant, factor=None):
    """dt = courant * min_e (hmin_e / vmax_e) * factor_e, then rounded to
    4 significant figures (``Quad.cpp:271-349``, ``SE_Model.cpp:198-203``).

    sz: (E, 2, 25) point coords; vmax: (E,) max velocity per element.
    """
    ipol = np.repeat(np.arange(NPED), NPED)
    jpol = np.tile(np.arange(NPED), NPED)
    hmin = np.full(sz.shape[0], np.inf)
    for p0 in range(NPEM - 1):
        for p1 in range(p0 + 1, NPEM):
            if abs(ipol[p1] - ipol[p0]) <= 1 and abs(jpol[p1] - jpol[p0]) <= 1:
                d = np.hypot(sz[:, 0, p1] - sz[:, 0, p0],
                             sz[:, 1, p1] - sz[:, 1, p0])
                hmin = np.minimum(hmin, d)
    dt_e = hmin / vmax
    if factor is not None:
        dt_e = dt_e * factor
    dt = float(np.min(dt_e) * courant)
    return float("%.3e" % dt)


# ------------------------------------------------------------------ source
def ricker_stf(t, half_duration, decay_factor, time_shift=0.0):
    """GaussianSTF, derivative order 2 (``GaussianSTF.hpp:35-58``), FP64."""
    return gaussian_stf(t, half_duration, decay_factor, time_shift, 2)


# use_derivative_integral of inparam.source.yaml -> derivative order
# (STF.cpp:35-39)
STF_ORDER = {"ERF": -1, "GAUSSIAN": 0, "FIRST_DERIVATIVE": 1, "RICKER": 2}


def gaussian_stf(t, half_duration, decay_factor, time_shift=0.0, order=2):
    """GaussianSTF::getValue (``GaussianSTF.hpp:35-58``), FP64, with f =
    decay / half_duration and t' = t - time_shift: order -1 erf(f t') / 2 +
    1/2; 0 f exp(-(f t')^2) / sqrt(pi); 1 -2 t' f^3 exp(-(f t')^2) /
    sqrt(pi) (t', not f t', as the reference writes it); 2 (4 (f t')^2 -
    2) f^3 exp(-(f t')^2) / sqrt(pi)."""
    f = decay_factor / half_duration
    tt = np.asarray(t, dtype=np.float64) - time_shift
    ft = f * tt
    g = np.exp(-ft * ft) / math.sqrt(math.pi)
    if order == -1:
        from scipy.special import erf
        return erf(ft) * 0.5 + 0.5
    if order == 0:
        return f * g
    if order == 1:
        return f ** 3 * g * (-2.0 * tt)
    if order == 2:
        return f ** 3 * g * (4.0 * ft * ft - 2.0)
    raise ValueError("gaussian_stf: derivative order must be -1, 0, 1 or 2")


def moment_pattern_axial(mt6, inplane, nu1_elem):
    """Fourier pattern of an axial moment-tensor source in an axial element.

    mt6 = [Mzz, Mss, Mpp, Mzs, Mzp, Msp] (``Mechanism.cpp:131-145``);
    inplane: (25,) interpolation weights.  Returns complex (nu1, 6, 25) with
    Voigt order [ss, pp, zz, pz, sz, sp] (``Mechanism.cpp:160-218``).
    """
    nu1 = min(nu1_elem, 3)
    mzz, mss, mpp, mzs, mzp, msp = [float(v) for v in mt6]
    pat = np.zeros((nu1, 6, NPEM), dtype=np.complex128)
    row0 = slice(0, NPED)   # ipol == 0 points, on the axis
    w = inplane[row0]
    pat[0, 0, row0] = (mss * w + mpp * w) / 2.0
    pat[0, 1, row0] = (mss * w + mpp * w) / 2.0
    pat[0, 2, row0] = mzz * w
    if nu1 >= 2:
        pat[1, 4, row0] = (mzs * w - 1j * mzp * w) / 2.0
        pat[1, 3, row0] = 1j * pat[1, 4, row0]
    if nu1 >= 3:
        pat[2, 0, row0] = (mss * w - mpp * w - 2.0j * msp * w) / 4.0
        pat[2, 1, row0] = -pat[2, 0, row0]
        pat[2, 5, row0] = 1j * pat[2, 0, row0]
    pat /= 2.0 * math.pi
    return pat


# --------------------------------------------------------------- the spec
@dataclasses.dataclass
class DomainSpec:
    """All operands of the hot path, FP64, in the layout the C ABI takes."""
    # points
    solid_nr: np.ndarray                 # int32 (Us,)
    solid_mass: np.ndarray               # float64 (Us,)
    solid_sz: np.ndarray                 # float64 (Us, 2)
    # solid elements
    elem_points: np.ndarray              # int32 (Es, 25) -> solid point
    elem_axial: np.ndarray               # uint8 (Es,)
    elem_gq: np.ndarray                  # float64 (Es, 6, 25)
    elem_nr: np.ndarray                  # int32 (Es,)
    # material, isotropic 1D: (Es, 2, 25) [lambda, mu]
    elem_iso: np.ndarray
    # boundaries
    axial_points: np.ndarray             # int32 solid points on the axis
    # source
    src_elem: int = -1
    src_pattern: Optional[np.ndarray] = None   # complex (nu1, 6, 25)
    stf: Optional[np.ndarray] = None           # float64 (nsteps,)
    # receivers
    rec_elem: Optional[np.ndarray] = None      # int32 (R,)
    rec_weights: Optional[np.ndarray] = None   # float64 (R, 25)
    rec_phi: Optional[np.ndarray] = None       # float64 (R,)
    # time
    t0: float = 0.0
    dt: float = 0.0
    nsteps: int = 0
    name: str = ""

    @property
    def n_solid_points(self):
        return len(self.solid_nr)

    @property
    def n_solid_elements(self):
        return self.elem_points.shape[0]

    def n_updates_per_step(self):
        """N_upd = sum over points of (Nr/2 + 1), solid and fluid points
        counted separately."""
        n = int(np.sum(self.solid_nr // 2 + 1))
        fnr = getattr(self, "fluid_nr", None)
        if fnr is not None:
            n += int(np.sum(np.asarray(fnr) // 2 + 1))
        return n


def _compact_points(elem_points_grid, npoint_grid):
    used = np.zeros(npoint_grid, dtype=bool)
    used[elem_points_grid.reshape(-1)] = True
    remap = -np.ones(npoint_grid, dtype=np.int64)
    remap[used] = np.arange(int(used.sum()))
    return remap, used


def inplane_nr_bound(mesh: StructuredMesh):
    """Upper bound of the nodal Nr by the in-plane resolution, per element
    corner (E, 4): round(2 pi s / h) with h the mean edge length of the
    quads sharing the node divided by npol, and 5 on the axis
    (ExodusMesh::formNrAtNodes with bound_Nr_by_inplane,
    src/preloop/exodus/ExodusMesh.cpp:663-713)."""
    nsz = mesh.nodal_sz()                                   # (E, 2, 4)
    edges = np.zeros(mesh.nelem)
    for a, b in ((0, 1), (1, 2), (2, 3), (3, 0)):
        edges += np.hypot(nsz[:, 0, a] - nsz[:, 0, b],
                          nsz[:, 1, a] - nsz[:, 1, b])
    nodes = mesh.corner_nodes()
    tot = np.zeros(mesh.nnodes)
    cnt = np.zeros(mesh.nnodes)
    for c in range(4):
        np.add.at(tot, nodes[:, c], edges)
        np.add.at(cnt, nodes[:, c], 1.0)
    spacing = tot / np.maximum(cnt * 4 * spectral.NPOL, 1e-300)
    if mesh.spherical:
        tol = 1e-6 * float(np.max(np.hypot(nsz[:, 0], nsz[:, 1])))
    else:
        tol = 1e-6 * max(float(np.ptp(nsz[:, 0])), float(np.ptp(nsz[:, 1])))
    out = np.empty((mesh.nelem, 4))
    for c in range(4):
        h = spacing[nodes[:, c]]
        s = nsz[:, 0, c]
        up = np.maximum(round_half_away(2.0 * math.pi * s / h), 1.0)
        out[:, c] = np.where(s < tol, 5.0, up)
    return out


def build_domain(mesh: StructuredMesh,
                 rho_vp_vs: Callable,
                 nodal_nr: Callable,
                 courant: float = 0.6,
                 lucky: bool = True,
                 bound_nr: bool = True):
    """Points, solid elements, 1D isotropic material, mass, dt.

    rho_vp_vs(r) -> (rho, vp, vs) arrays evaluated at the element corners'
    radial coordinate (z for Cartesian, |(s,z)| for spherical), or the
    tuple of those (E, 4) arrays themselves; nodal_nr(s, z) -> Nr at
    element corners.
    """
    axial = mesh.is_axial()
    E = mesh.nelem
    gq = gradquad_operands(mesh)
    nsz = mesh.nodal_sz()
    rcorner = (np.hypot(nsz[:, 0], nsz[:, 1]) if mesh.spherical
               else nsz[:, 1])
    if not callable(rho_vp_vs):
        # values at the element corners, (E, 4) each (an Exodus mesh's
        # element-node storage)
        rho_n, vp_n, vs_n = (np.asarray(x, dtype=np.float64)
                             for x in rho_vp_vs)
    else:
        try:
            rho_n, vp_n, vs_n = rho_vp_vs(rcorner, rcorner.mean(
                axis=1, keepdims=True))
        except TypeError:
            rho_n, vp_n, vs_n = rho_vp_vs(rcorner)
    rho = interpolate_gll(rho_n, axial)
    vp = interpolate_gll(vp_n, axial)
    vs = interpolate_gll(vs_n, axial)
    lam = rho * vp ** 2 - 2.0 * (rho * vs ** 2)
    mu = rho * vs ** 2

    # points
    ep_grid = mesh.element_points()
    remap, used = _compact_points(ep_grid, mesh.npoint_grid)
    elem_points = remap[ep_grid].astype(np.int32)
    U = int(used.sum())
    mass = np.zeros(U)
    np.add.at(mass, elem_points.reshape(-1), (rho * gq["ifact"]).reshape(-1))
    sz_pts = np.zeros((U, 2))
    sz_pts[elem_points.reshape(-1)] = gq["sz"].transpose(0, 2, 1).reshape(-1, 2)

    # Nr: nodal (bounded by the in-plane resolution) -> GLL (round half
    # away) -> lucky -> element max
    nr_nodal = nodal_nr(nsz[:, 0], nsz[:, 1])
    if bound_nr:
        nr_nodal = np.minimum(nr_nodal, inplane_nr_bound(mesh))
    nr_gll = round_half_away(interpolate_gll(nr_nodal, axial)).astype(np.int64)
    if lucky:
        uniq = np.unique(nr_gll)
        lut = {int(n): spectral.next_lucky_number(int(n)) for n in uniq}
        nr_gll = np.vectorize(lambda n: lut[int(n)])(nr_gll)
    point_nr = np.zeros(U, dtype=np.int64)
    point_nr[elem_points.reshape(-1)] = nr_gll.reshape(-1)
    elem_nr = point_nr[elem_points].max(axis=1)

    # axial points: ipol == 0 points of axial elements
    ax_e = np.nonzero(axial)[0]
    axial_points = np.unique(elem_points[ax_e][:, :NPED].reshape(-1))

    # dt
    vmax = vp.max(axis=1)
    dt = compute_dt(gq["sz"], vmax, courant)

    elem_gq = np.stack([gq["dsdxii"], gq["dsdeta"], gq["dzdxii"],
                        gq["dzdeta"], gq["inv_s"], gq["ifact"]], axis=1)
    spec = DomainSpec(
        solid_nr=point_nr.astype(np.int32), solid_mass=mass, solid_sz=sz_pts,
        elem_points=elem_points, elem_axial=axial.astype(np.uint8),
        elem_gq=np.ascontiguousarray(elem_gq),
        elem_nr=elem_nr.astype(np.int32),
        elem_iso=np.ascontiguousarray(np.stack([lam, mu], axis=1)),
        axial_points=axial_points.astype(np.int32), dt=dt)
    spec._mesh = mesh
    spec._gq = gq
    spec._gq_mesh = gq
    spec._vp1d = vp
    spec._rho1d = rho
    spec._courant = courant
    spec._elem_of_mesh = np.arange(E)
    return spec


# element-local points of each edge, in the order of _edge_normals
_EDGE_IPNT = {0: [0, 5, 10, 15, 20], 1: [20, 21, 22, 23, 24],
              2: [24, 19, 14, 9, 4], 3: [4, 3, 2, 1, 0]}


def _edge_normals(J, axial, edge, w_gll, w_glj, sz):
    """Integrated outward normals of one element edge at its 5 points
    (Quad::computeNormal, Quad.cpp:438-484): the edge tangent rotated
    outward (length = edge Jacobian), times the quadrature weight along
    the edge and the radius s (axial element: GLJ weights and the
    L'Hospital factor ds/dxi on the axis, s / (1 + xi) off it).  J (25,
    2, 2) = d(s,z)/d(xi,eta); returns (points (5,), n (5, 2))."""
    pts = {0: [0, 5, 10, 15, 20], 1: [20, 21, 22, 23, 24],
           2: [24, 19, 14, 9, 4], 3: [4, 3, 2, 1, 0]}[edge]
    xi = spectral.xieta_element(axial)[0]
    out = np.zeros((5, 2))
    for k, p in enumerate(pts):
        ipol, jpol = p // NPED, p % NPED
        if edge in (0, 2):     # along xi
            ts, tz = J[p, 0, 0], J[p, 1, 0]
            n = np.array([tz, -ts]) if edge == 0 else np.array([-tz, ts])
            n = n * (w_glj[ipol] if axial else w_gll[ipol])
        else:                  # along eta
            ts, tz = J[p, 0, 1], J[p, 1, 1]
            n = np.array([tz, -ts]) if edge == 1 else np.array([-tz, ts])
            n = n * w_gll[jpol]
        if axial:
            n = n * (J[p, 0, 0] if ipol == 0 else sz[0, p] / (1.0 + xi[p]))
        else:
            n = n * sz[0, p]
        out[k] = n
    return pts, out


def make_fluid(spec, fluid_elem):
    """Turn the mesh elements in fluid_elem (bool, per mesh element) into
    fluid elements (acoustic, displacement potential; Acoustic K = 1/rho,
    Material.cpp:350-360; fluid mass IF / (rho vp^2), Material.cpp:
    318-333).  GLL points shared by solid and fluid elements become a
    solid point and a fluid point joined by a solid-fluid coupling with
    the assembled normal pointing from fluid to solid (Quad.cpp:123-140).
    Fluid points on the top of the mesh form the fluid free surface.
    Call on a fresh build_domain spec, before sources, receivers and 3D
    models; solid arrays are renumbered."""
    mesh = spec._mesh
    gq = spec._gq
    fl = np.asarray(fluid_elem, dtype=bool)
    so = ~fl
    E = mesh.nelem
    ep_grid = mesh.element_points()
    remap_all, _ = _compact_points(ep_grid, mesh.npoint_grid)
    nr_grid = np.zeros(mesh.npoint_grid, dtype=np.int64)
    nr_grid[ep_grid.reshape(-1)] = spec.solid_nr[
        remap_all[ep_grid.reshape(-1)]]
    sol_ids = np.unique(ep_grid[so])
    flu_ids = np.unique(ep_grid[fl])
    rs = -np.ones(mesh.npoint_grid, dtype=np.int64)
    rs[sol_ids] = np.arange(len(sol_ids))
    rf = -np.ones(mesh.npoint_grid, dtype=np.int64)
    rf[flu_ids] = np.arange(len(flu_ids))
    ifact = gq["ifact"]
    rho, vp = spec._rho1d, spec._vp1d
    axial = mesh.is_axial()
    # solid part
    se = np.nonzero(so)[0]
    s_points = rs[ep_grid[se]].astype(np.int32)
    s_mass = np.zeros(len(sol_ids))
    np.add.at(s_mass, s_points.reshape(-1), (rho[se] * ifact[se]).reshape(-1))
    sz_grid = np.zeros((mesh.npoint_grid, 2))
    sz_grid[ep_grid.reshape(-1)] = gq["sz"].transpose(0, 2, 1).reshape(-1, 2)
    # fluid part
    fe = np.nonzero(fl)[0]
    f_points = rf[ep_grid[fe]].astype(np.int32)
    f_mass = np.zeros(len(flu_ids))
    np.add.at(f_mass, f_points.reshape(-1),
              (ifact[fe] / (rho[fe] * vp[fe] ** 2)).reshape(-1))
    # axis
    s_axial = np.unique(s_points[axial[se]][:, :NPED].reshape(-1)) \
        if np.any(axial[se]) else np.zeros(0, np.int64)
    f_axial = np.unique(f_points[axial[fe]][:, :NPED].reshape(-1)) \
        if np.any(axial[fe]) else np.zeros(0, np.int64)
    # interface normals, accumulated per grid point (fluid side + n/2,
    # solid side - n/2, both pointing from fluid to solid)
    w_gll = spectral.weights_1d(False)
    w_glj = spectral.weights_1d(True)
    nsum = {}
    for e, edge in mesh.interface_edges(fl):
        pts, n = _edge_normals(gq["J"][e], bool(axial[e]), edge, w_gll,
                               w_glj, gq["sz"][e])
        sign = 0.5 if fl[e] else -0.5
        for k, p in enumerate(pts):
            g = int(ep_grid[e, p])
            nsum[g] = nsum.get(g, np.zeros(2)) + sign * n[k]
    sf_grid = np.array(sorted(nsum.keys()), dtype=np.int64)
    spec.sf_solid = rs[sf_grid].astype(np.int32)
    spec.sf_fluid = rf[sf_grid].astype(np.int32)
    spec.sf_normal = np.array([nsum[g] for g in sf_grid]).reshape(-1, 2)
    assert (spec.sf_solid >= 0).all() and (spec.sf_fluid >= 0).all()
    # fluid free surface: fluid points on the top edges of the mesh
    top = [ep_grid[e, p] for e, edge in mesh.boundary_edges("TOP") if fl[e]
           for p in _EDGE_IPNT[edge]]
    spec.fluid_surface_points = np.unique(
        rf[np.asarray(top, dtype=np.int64)]).astype(np.int32)
    # fluid arrays
    spec.fluid_nr = nr_grid[flu_ids].astype(np.int32)
    spec.fluid_mass = f_mass
    spec.fluid_sz = sz_grid[flu_ids]
    spec.fluid_elem_points = f_points
    spec.fluid_elem_axial = axial[fe].astype(np.uint8)
    spec.fluid_elem_gq = np.ascontiguousarray(spec.elem_gq[
        spec._elem_of_mesh[fe]])
    spec.fluid_elem_nr = spec.fluid_nr[f_points].max(axis=1).astype(np.int32)
    spec.fluid_elem_k = np.ascontiguousarray(1.0 / rho[fe])
    spec.fluid_axial_points = f_axial.astype(np.int32)
    # solid arrays renumbered
    old = spec._elem_of_mesh[se]
    spec.solid_nr = nr_grid[sol_ids].astype(np.int32)
    spec.solid_mass = s_mass
    spec.solid_sz = sz_grid[sol_ids]
    spec.elem_points = s_points
    spec.elem_axial = axial[se].astype(np.uint8)
    spec.elem_gq = np.ascontiguousarray(spec.elem_gq[old])
    spec.elem_nr = spec.solid_nr[s_points].max(axis=1).astype(np.int32)
    spec.elem_iso = np.ascontiguousarray(spec.elem_iso[old])
    spec.axial_points = s_axial.astype(np.int32)
    spec._gq = {k: (v[se] if isinstance(v, np.ndarray) and v.shape[0] == E
                    else v) for k, v in gq.items()}
    spec._vp1d = vp[se]
    spec._rho1d = rho[se]
    spec._fvp1d = vp[fe]
    spec._frho1d = rho[fe]
    m = -np.ones(E, dtype=np.int64)
    m[se] = np.arange(len(se))
    spec._elem_of_mesh = m
    mf = -np.ones(E, dtype=np.int64)
    mf[fe] = np.arange(len(fe))
    spec._felem_of_mesh = mf
    return spec


# ------------------------------------------------ absorbing boundaries
def gamma_solid_default(vp, vs, rho, span, t0):
    """Kosloff sponge strength in solids, the template's gamma_expr_solid
    (inparam.model.yaml: 1.1 / T0 * (VS / VP)^2 * exp(-0.04 * SPAN / (VP *
    T0)), Haindl et al. 2020)."""
    return 1.1 / t0 * (vs / vp) ** 2 * np.exp(-0.04 * span / (vp * t0))


def gamma_fluid_default(vp, rho, span, t0):
    """The template's gamma_expr_fluid: 0.88 / T0 * exp(-0.04 * SPAN / (VP
    * T0))."""
    return 0.88 / t0 * np.exp(-0.04 * span / (vp * t0))




def _solid_rho_vp_vs(spec, e):
    """rho, vp, vs at the GLL points of solid element e
    (Material::getPointwiseRhoVpVs, Material.cpp:250-300): (25,) arrays,
    or (Nr, 25) for an element with a 3D medium; TI and anisotropic media
    through the Voigt averages of kappa and mu."""
    rho = np.asarray(spec._rho1d[e], dtype=np.float64)
    mtype = getattr(spec, "elem_mtype", None)
    kind = getattr(spec, "elem_kind", None)
    mt = 0 if mtype is None else int(mtype[e])
    if kind is not None and int(kind[e]) == 1:
        nr = int(spec.elem_nr[e])
        npar = MAT_NPARAMS[mt]
        if mtype is None:
            s0 = int(spec.elem_slice_off[e])
            par = np.asarray(spec.prop3d)[s0:s0 + nr].reshape(nr, 2, NPEM)
        else:
            o = int(spec.prop_off[e])
            par = np.asarray(spec.prop3d)[o:o + nr * npar * NPEM].reshape(
                nr, npar, NPEM)
        kp, mu = _voigt_kappa_mu(mt, np.transpose(par, (1, 0, 2)))
    else:
        if mtype is None:
            par = spec.elem_iso[e]
        else:
            o = int(spec.mat1d_off[e])
            par = np.asarray(spec.mat1d)[o:o + MAT_NPARAMS[mt] * NPEM]
            par = par.reshape(MAT_NPARAMS[mt], NPEM)
        kp, mu = _voigt_kappa_mu(mt, par)
    vp = np.sqrt((kp + mu * (4.0 / 3.0)) / rho)
    vs = np.sqrt(np.maximum(mu, 0.0) / rho)
    return rho, vp, vs


def _fluid_rho_vp(spec, f):
    """rho, vp at the GLL points of fluid element f: (25,), or (Nr, 25)
    with a 3D medium (rho = 1 / K per slice, vp of the 1D model)."""
    vp = np.asarray(spec._fvp1d[f], dtype=np.float64)
    kind = getattr(spec, "fluid_elem_kind", None)
    if kind is not None and int(kind[f]) == 1:
        nr = int(spec.fluid_elem_nr[f])
        off = 0
        for g in range(f):
            if kind[g]:
                off += int(spec.fluid_elem_nr[g]) * NPEM
        k3 = np.asarray(spec.fluid_k3d)[off:off + nr * NPEM].reshape(nr, NPEM)
        return 1.0 / k3, np.broadcast_to(vp[None, :], (nr, NPEM))
    return np.asarray(spec._frho1d[f], dtype=np.float64), vp


def _at_point(x, ipnt, nr_point):
    """Values of GLL point ipnt at the point's own Nr slices: a 1D value
    stays one value; an element's Nr slices are sampled at the nearest
    slice of each of the point's (slice k of Nr_p at phi = 2 pi k / Nr_p)."""
    x = np.asarray(x)
    if x.ndim == 1:
        return np.array([x[ipnt]])
    nre = x.shape[0]
    if nre == nr_point:
        return x[:, ipnt].copy()
    k = np.arange(nr_point)
    src = np.rint(k * nre / nr_point).astype(np.int64) % nre
    return x[src, ipnt]


def add_undulation(spec, dz, elems=None):
    """Undulated geometry by particle relabelling (Undulation,
    src/preloop/se_model/quad/physical/Undulation.cpp): every GLL point of
    the solid elements `elems` (default: those where dz is not zero) moves
    by dz along the vertical (Cartesian z, or the radius on spherical
    meshes), axisymmetric (the same for all slices).  dz is an array
    (n_elements, 25) at the element GLL points or a function of (s, z).

    Per element: the gradient of dz (computeGrad3 of order 0, Undulation::
    finished3D, :37-81), rotated to (R, Z) on spherical meshes; J1 = dZ/dR,
    J2 = 0, J3 = 1 + dZ/dZ, J0 = 1 (Cartesian) or 1 + dz / r (r > 0, else
    J3); X0 = 1 / J0, X1 = -J1 / (J0 J3), X2 = 0, X3 = 1 / J3, XJ = J0^2 J3
    (createPRT, :124-168).  The point masses are multiplied by XJ
    (getMassJacobian, :84-121).  Fills spec.prt = dict(elems, x) with x
    (n, 5, 25) = X0, X1, X2, X3, XJ, and on spherical meshes the GLL
    colatitudes spec.elem_theta for the RTZ rotation.  Solid-only domains
    without absorbing boundaries, ocean load or 3D mass (their normals and
    masses would need the undulation too)."""
    if (len(getattr(spec, "fluid_nr", ())) > 0 or
            getattr(spec, "_abc_abi", None) is not None or
            getattr(spec, "ocean_load", None) is not None or
            len(getattr(spec, "mass3d_points", ())) > 0):
        raise ValueError("add_undulation: solid domains without absorbing "
                         "boundaries, ocean load or 3D mass")
    gq = spec._gq
    sz = gq["sz"]
    E = spec.n_solid_elements
    if callable(dz):
        dzv = np.asarray(dz(sz[:, 0], sz[:, 1]), dtype=np.float64)
    else:
        dzv = np.asarray(dz, dtype=np.float64)
    dzv = np.broadcast_to(dzv, (E, NPEM))
    if elems is None:
        elems = np.nonzero(np.any(dzv != 0.0, axis=1))[0]
    elems = np.asarray(elems, dtype=np.int64)
    t = spectral.tables()
    g_gll, g_glj = t[4], t[5]
    spherical = bool(getattr(spec._mesh, "spherical", False))
    theta = gll_colatitude(sz) if spherical else None
    eg = np.asarray(spec.elem_gq)
    x = np.zeros((len(elems), 5, NPEM))
    jac = np.ones((E, NPEM))
    for n, e in enumerate(elems):
        F = dzv[e].reshape(NPED, NPED)
        gxi = g_glj if spec.elem_axial[e] else g_gll
        xx = (gxi.T @ F).reshape(-1)
        yy = (F @ g_gll).reshape(-1)
        dsdxii, dsdeta, dzdxii, dzdeta = eg[e, 0], eg[e, 1], eg[e, 2], eg[e, 3]
        d_s = xx * dzdeta + yy * dzdxii
        d_z = xx * dsdeta + yy * dsdxii
        if spherical:
            c, s = np.cos(theta[e]), np.sin(theta[e])
            j1 = d_s * c - d_z * s
            j3 = 1.0 + (d_z * c + d_s * s)
            r = np.hypot(sz[e, 0], sz[e, 1])
            j0 = np.where(r > np.finfo(float).eps,
                          1.0 + dzv[e] / np.where(r > 0, r, 1.0), j3)
        else:
            j1 = d_s
            j3 = 1.0 + d_z
            j0 = np.ones(NPEM)
        xj = j0 ** 2 * j3
        if xj.min() < np.finfo(float).eps:
            raise ValueError("add_undulation: negative Jacobian or distorted "
                             "element")
        x[n, 0] = 1.0 / j0
        x[n, 1] = -j1 / (j0 * j3)
        x[n, 3] = 1.0 / j3
        x[n, 4] = xj
        jac[e] = xj
    rho = np.asarray(spec._rho1d, dtype=np.float64)
    mass = np.zeros(len(spec.solid_nr))
    np.add.at(mass, np.asarray(spec.elem_points).reshape(-1),
              (rho * eg[:, 5, :] * jac).reshape(-1))
    spec.solid_mass = mass
    spec.prt = {"elems": elems.astype(np.int32), "x": x}
    if spherical and getattr(spec, "elem_theta", None) is None:
        spec.elem_theta = theta
    return spec


def add_ocean_load(spec, sum_rho_depth):
    """Ocean load on the solid points of the mesh top (Quad::setupGLL,
    Quad.cpp:235-246; GLLPoint::addOceanLoad and release, GLLPoint.hpp:
    88-93, GLLPoint.cpp:54-87).  Per top point: nTop, the integrated
    outward normals of the TOP edges of the solid elements that share it,
    summed (Quad::computeNormal), and sum_rho_depth(s, z, phi) (kg/m^2,
    the ocean's integrated density above the point, a scalar or one value
    per slice phi_k = 2 pi k / Nr).  With a 1D mass and one sum_rho_depth
    for all slices the point gets MassOceanLoad1D (m0 = sum_rho_depth
    |nTop|, t = acos(nz / |nTop|) as geodesy::sz2rtheta gives it), else
    MassOceanLoad3D (per slice 1/m and k = sqrt(m0 / (m (m + m0))) nTop /
    |nTop|, m its 3D mass or its 1D mass on every slice); a point's 3D
    mass moves into its ocean-load entry.  Call after the medium and the
    3D masses are set.  Fills spec.ocean_load = dict(point, nval, off,
    data): 4 numbers per value, as the oracle and axs3d_set_ocean_load
    take them."""
    mesh = spec._mesh
    gq = spec._gq_mesh
    axial = mesh.is_axial()
    w_gll = spectral.weights_1d(False)
    w_glj = spectral.weights_1d(True)
    s_of_mesh = np.asarray(spec._elem_of_mesh)
    ntop = {}
    for m, edge in mesh.boundary_edges("TOP"):
        e = int(s_of_mesh[m])
        if e < 0:
            continue
        pts, n = _edge_normals(gq["J"][m], bool(axial[m]), edge, w_gll, w_glj,
                               gq["sz"][m])
        for k, ipnt in enumerate(pts):
            p = int(spec.elem_points[e][ipnt])
            ntop[p] = ntop.get(p, np.zeros(2)) + n[k]
    m3p = getattr(spec, "mass3d_points", None)
    m3 = {}
    if m3p is not None and len(m3p) > 0:
        nr_all = np.asarray(spec.solid_nr, dtype=np.int64)
        cum = 0
        for p in np.asarray(m3p, dtype=np.int64):
            m3[int(p)] = np.asarray(spec.mass3d[cum:cum + nr_all[p]])
            cum += nr_all[p]
    points, nval, off, data = [], [], [], []
    cum = 0
    for p in sorted(ntop):
        nr = int(spec.solid_nr[p])
        phi = 2 * np.pi * np.arange(nr) / nr
        s, z = spec.solid_sz[p]
        srd = np.broadcast_to(np.asarray(
            sum_rho_depth(s, z, phi), dtype=np.float64), (nr,)).copy()
        n = ntop[p]
        area = float(np.hypot(n[0], n[1]))
        if p not in m3 and np.all(srd == srd[0]):
            m = float(spec.solid_mass[p])
            m0 = float(srd[0]) * area
            t = 0.0 if area < np.finfo(float).eps else math.acos(n[1] / area)
            c, si = math.cos(t), math.sin(t)
            ih, iv = 1.0 / m, 1.0 / (m + m0)
            vals = np.array([[ih * c * c + iv * si * si,
                              ih * si * si + iv * c * c,
                              (iv - ih) * c * si, ih]])
        else:
            m = m3.get(p, np.full(nr, float(spec.solid_mass[p])))
            m0 = srd * area
            unit = np.array([n[0], 0.0, n[1]]) / area
            kk = np.sqrt(m0 / (m * (m + m0)))
            vals = np.concatenate([(1.0 / m)[:, None],
                                   kk[:, None] * unit[None, :]], axis=1)
        points.append(p)
        nval.append(len(vals))
        off.append(cum)
        data.append(vals.reshape(-1))
        cum += len(vals)
    if m3:
        # the 3D masses of ocean-load points move into their entries
        keep = [p for p in np.asarray(m3p, dtype=np.int64)
                if int(p) not in ntop]
        spec.mass3d_points = np.array(keep, dtype=np.int32)
        spec.mass3d = (np.concatenate([m3[int(p)] for p in keep])
                       if keep else np.zeros(0))
    spec.ocean_load = {
        "point": np.array(points, dtype=np.int32),
        "nval": np.array(nval, dtype=np.int32),
        "off": np.array(off, dtype=np.int64),
        "data": np.concatenate(data) if data else np.zeros(0)}
    return spec


def add_absorbing(spec, keys=("RIGHT", "BOTTOM"), clayton=True,
                  sponge=False, rel_spans=None, t0=None,
                  gamma_solid=gamma_solid_default,
                  gamma_fluid=gamma_fluid_default):
    """Absorbing boundaries on the mesh sides `keys` (RIGHT: s max or
    theta max; BOTTOM: z min or r min; TOP), after the medium (1D or 3D)
    is set and before sources and receivers (dt changes):

    * Clayton-Enquist (Quad::setupGLL, Quad.cpp:142-163; GLLPoint::
      release, GLLPoint.cpp:152-192): one entry per boundary element edge
      point, with the element's integrated edge normal nABC
      (Quad::computeNormal) and its rho vp, rho vs; 1D elements give
      ClaytonSolid1D (rho vp, rho vs, area = |nABC|, theta = acos(nz /
      |nABC|) as geodesy::sz2rtheta gives it) or ClaytonFluid1D, 3D media
      the per-slice forms (rho vp, rho vs, area, unit normal per slice of
      the point);
    * Kosloff sponge (Quad.cpp:165-231, GLLPoint.cpp:193-215): at every
      element GLL point within rel_span * mesh span of a boundary, gamma =
      gamma_expr(vp, vs, rho, |span|, T0) cos^2(pi/2 d), d the relative
      distance to the boundary (the nearest boundary wins); a point's
      gamma is the mean over the elements that give it one;
    * dt: elements with a Clayton edge take half the time step
      (factorDtForBC, Quad.cpp:330-349); the attenuation coefficients are
      recomputed for the new dt.

    Fills spec.clayton_solid / clayton_fluid / sponge_solid /
    sponge_fluid: dicts of point (entries), nval (1 or Nr), off (value
    offset) and data (4 or 1 numbers per value for the solid Clayton:
    1D coefficients or (rsa, k_s, k_phi, k_z) per slice; 1 number per
    value otherwise: area / (rho vp), or gamma), as the oracle takes them,
    plus the raw per-entry inputs of the C ABI under 'abi'."""
    if getattr(spec, "stf", None) is not None:
        raise ValueError("add_absorbing: set the boundaries before the "
                         "source (dt changes)")
    mesh = spec._mesh
    gq = spec._gq_mesh
    keys = list(keys)
    if rel_spans is None:
        rel_spans = [0.05] * len(keys)
    E = mesh.nelem
    axial = mesh.is_axial()
    w_gll = spectral.weights_1d(False)
    w_glj = spectral.weights_1d(True)
    s_of_mesh = np.asarray(spec._elem_of_mesh)
    f_of_mesh = getattr(spec, "_felem_of_mesh", -np.ones(E, np.int64))
    mat_s = {}
    mat_f = {}

    def material(m):
        if s_of_mesh[m] >= 0:
            e = int(s_of_mesh[m])
            if e not in mat_s:
                mat_s[e] = _solid_rho_vp_vs(spec, e)
            return False, e, mat_s[e]
        f = int(f_of_mesh[m])
        if f not in mat_f:
            r, v = _fluid_rho_vp(spec, f)
            mat_f[f] = (r, v, np.zeros_like(v))
        return True, f, mat_f[f]

    def point_of(fluid, e, ipnt):
        if fluid:
            p = int(spec.fluid_elem_points[e][ipnt])
            return p, int(spec.fluid_nr[p])
        p = int(spec.elem_points[e][ipnt])
        return p, int(spec.solid_nr[p])

    lists = {k: {"point": [], "nval": [], "vals": []} for k in
             ("clayton_solid", "clayton_fluid", "sponge_solid",
              "sponge_fluid")}
    abi = {"clayton_solid": [], "clayton_solid3d": [],
           "clayton_fluid": [], "clayton_fluid3d": [],
           "sponge_solid": {}, "sponge_fluid": {}}
    factor = np.ones(E)
    if clayton:
        for key in keys:
            for m, edge in mesh.boundary_edges(key):
                fluid, e, (rho, vp, vs) = material(m)
                factor[m] = 0.5
                pts, nabc = _edge_normals(gq["J"][m], bool(axial[m]), edge,
                                          w_gll, w_glj, gq["sz"][m])
                for k, ipnt in enumerate(pts):
                    p, nrp = point_of(fluid, e, ipnt)
                    n2 = nabc[k]
                    area = float(np.hypot(n2[0], n2[1]))
                    rp = _at_point(rho * vp, ipnt, nrp)
                    rs = _at_point(rho * vs, ipnt, nrp)
                    three = len(rp) > 1
                    if fluid:
                        vals = area / rp
                        lists["clayton_fluid"]["point"].append(p)
                        lists["clayton_fluid"]["nval"].append(len(rp))
                        lists["clayton_fluid"]["vals"].append(vals)
                        abi["clayton_fluid3d" if three else
                            "clayton_fluid"].append((p, rp, np.full(len(rp),
                                                                    area)))
                        continue
                    if not three:
                        th = math.acos(n2[1] / area) if area > 0 else 0.0
                        c, s = math.cos(th), math.sin(th)
                        rsa, rpa = rs[0] * area, rp[0] * area
                        vals = np.array([rsa * c * c + rpa * s * s,
                                         rsa * s * s + rpa * c * c,
                                         (rpa - rsa) * c * s, rsa])
                        abi["clayton_solid"].append((p, rp[0], rs[0], area,
                                                     th))
                    else:
                        un = np.array([n2[0], 0.0, n2[1]]) / area
                        kk = np.sqrt((rp - rs) * area)[:, None] * un[None, :]
                        vals = np.concatenate([(rs * area)[:, None], kk],
                                              axis=1).reshape(-1)
                        abi["clayton_solid3d"].append(
                            (p, rp, rs, np.full(nrp, area),
                             np.broadcast_to(un, (nrp, 3)).copy()))
                    lists["clayton_solid"]["point"].append(p)
                    lists["clayton_solid"]["nval"].append(len(rp))
                    lists["clayton_solid"]["vals"].append(vals)
    if sponge:
        if t0 is None:
            raise ValueError("add_absorbing: the sponge needs the mesh "
                             "period t0")
        info = {}
        sz_m = gq["sz"]
        for key, rel in zip(keys, rel_spans):
            if mesh.spherical:
                r = np.hypot(sz_m[:, 0], sz_m[:, 1])
                th = np.arccos(np.clip(sz_m[:, 1] / np.maximum(r, 1e-300),
                                       -1, 1))
                if key == "RIGHT":
                    outer = float(th.max())
                    span = rel * outer
                elif key == "BOTTOM":
                    outer = float(r.min())
                    span = rel * (float(r.min()) - float(r.max()))
                else:
                    outer = float(r.max())
                    span = rel * (float(r.max()) - float(r.min()))
            else:
                s, z = sz_m[:, 0], sz_m[:, 1]
                if key == "RIGHT":
                    outer = float(s.max())
                    span = rel * outer
                elif key == "BOTTOM":
                    outer = float(z.min())
                    span = rel * (float(z.min()) - float(z.max()))
                else:
                    outer = float(z.max())
                    span = rel * (float(z.max()) - float(z.min()))
            info[key] = (outer, span)
        acc_s, acc_f = {}, {}
        for m in range(E):
            fluid, e, (rho, vp, vs) = material(m)
            for ipnt in range(NPEM):
                s_, z_ = gq["sz"][m, 0, ipnt], gq["sz"][m, 1, ipnt]
                if mesh.spherical:
                    rr = math.hypot(s_, z_)
                    tt = math.acos(max(-1.0, min(1.0, z_ / rr))) \
                        if rr > 0 else 0.0
                best = 1.0
                gam = None
                for key in keys:
                    outer, span = info[key]
                    if mesh.spherical:
                        coord = tt if key == "RIGHT" else rr
                        radius = rr
                    else:
                        coord = s_ if key == "RIGHT" else z_
                        radius = z_
                    d = (outer - coord) / span
                    if d > best:
                        continue
                    best = d
                    pat = math.cos(0.5 * math.pi * d) ** 2
                    sp = span * radius if (mesh.spherical and
                                           key == "RIGHT") else span
                    rv = rho[..., ipnt] if np.ndim(rho) == 2 else rho[ipnt]
                    pv = vp[..., ipnt] if np.ndim(vp) == 2 else vp[ipnt]
                    sv = vs[..., ipnt] if np.ndim(vs) == 2 else vs[ipnt]
                    if fluid:
                        g = gamma_fluid(pv, rv, abs(sp), t0)
                    else:
                        g = gamma_solid(pv, sv, rv, abs(sp), t0)
                    gam = np.atleast_1d(np.asarray(g, dtype=np.float64)) * pat
                if gam is None or np.linalg.norm(gam) <= 2.220446049250313e-16:
                    continue
                p, nrp = point_of(fluid, e, ipnt)
                if len(gam) > 1 and len(gam) != nrp:
                    gam = _at_point(gam[:, None], 0, nrp)
                acc = acc_f if fluid else acc_s
                tot, cnt = acc.get(p, (np.zeros(1), 0))
                if len(gam) > 1 and len(tot) == 1:
                    tot = np.full(nrp, tot[0])
                if len(gam) == 1 and len(tot) > 1:
                    gam = np.full(len(tot), gam[0])
                acc[p] = (tot + gam, cnt + 1)
        for fluid, acc in ((False, acc_s), (True, acc_f)):
            name = "sponge_fluid" if fluid else "sponge_solid"
            for p in sorted(acc):
                tot, cnt = acc[p]
                g = tot / cnt
                lists[name]["point"].append(p)
                lists[name]["nval"].append(len(g))
                lists[name]["vals"].append(g)
                abi[name][p] = g
    for name, d in lists.items():
        nval = np.asarray(d["nval"], dtype=np.int32)
        width = 4 if name == "clayton_solid" else 1
        off = np.zeros(len(nval), dtype=np.int64)
        if len(nval):
            off[1:] = np.cumsum(nval)[:-1]
        data = (np.concatenate([np.asarray(v, np.float64).reshape(-1)
                                for v in d["vals"]])
                if d["vals"] else np.zeros(0))
        assert len(data) == width * int(nval.sum())
        setattr(spec, name, {"point": np.asarray(d["point"], np.int32),
                             "nval": nval, "off": off, "data": data})
    spec._abc_abi = abi
    # dt: half the step for the elements with a Clayton edge
    if clayton:
        vmax = np.zeros(E)
        for m in range(E):
            fluid, e, (rho, vp, vs) = material(m)
            vmax[m] = float(np.max(vp))
        old = spec.dt
        spec.dt = compute_dt(gq["sz"], vmax, spec._courant, factor)
        att = getattr(spec, "att", None)
        if att is not None and spec.dt != old:
            a, b, g = att_alpha_beta_gamma(att["w"], att["y"], spec.dt)
            att.update(alpha=a, beta=b, gamma=g)
    return spec


def _locate_quads(mesh, s, z):
    """locate() on an unstructured linear mesh: the first element in index
    order whose closed quad contains (s, z), with (xi, eta) from Newton's
    method on the bilinear map."""
    nsz = mesh.nodal_sz()
    lo, hi = nsz.min(axis=2), nsz.max(axis=2)
    tol = 1e-9 * max(float(np.ptp(nsz[:, 0])), float(np.ptp(nsz[:, 1])))
    cand = np.nonzero((lo[:, 0] - tol <= s) & (s <= hi[:, 0] + tol) &
                      (lo[:, 1] - tol <= z) & (z <= hi[:, 1] + tol))[0]
    for e in cand:
        x = nsz[e]
        xi = eta = 0.0
        for _ in range(50):
            xip, xim, etp, etm = 1 + xi, 1 - xi, 1 + eta, 1 - eta
            shp = 0.25 * np.array([xim * etm, xip * etm, xip * etp,
                                   xim * etp])
            r = x @ shp - np.array([s, z])
            dxi = 0.25 * np.array([-etm, etm, etp, -etp])
            det_ = 0.25 * np.array([-xim, -xip, xip, xim])
            Jm = np.stack([x @ dxi, x @ det_], axis=1)
            step = np.linalg.solve(Jm, r)
            xi, eta = xi - step[0], eta - step[1]
            if abs(step).max() < 1e-14:
                break
        if abs(xi) <= 1 + 1e-9 and abs(eta) <= 1 + 1e-9:
            return int(e), float(np.clip(xi, -1, 1)), float(np.clip(eta, -1,
                                                                    1))
    raise ValueError(f"locate: ({s}, {z}) is outside the mesh")


def locate(mesh: StructuredMesh, s, z):
    """Element and (xi, eta) of a point (s, z); first element in index order
    whose closed cell contains it."""
    if isinstance(mesh, QuadMesh):
        return _locate_quads(mesh, s, z)
    if mesh.spherical:
        a = math.atan2(s, z)
        b = math.hypot(s, z)
    else:
        a, b = s, z
    na, nb = mesh.node_a, mesh.node_b
    tol = 1e-9 * max(abs(na[-1] - na[0]), abs(nb[-1] - nb[0]))
    ix = int(np.clip(np.searchsorted(na, a - tol, side="right") - 1,
                     0, mesh.nx - 1))
    iz = int(np.clip(np.searchsorted(nb, b - tol, side="right") - 1,
                     0, mesh.nz - 1))
    xi = 2.0 * (a - na[ix]) / (na[ix + 1] - na[ix]) - 1.0
    eta = 2.0 * (b - nb[iz]) / (nb[iz + 1] - nb[iz]) - 1.0
    xi = float(np.clip(xi, -1.0, 1.0))
    eta = float(np.clip(eta, -1.0, 1.0))
    return iz * mesh.nx + ix, xi, eta


def inplane_factor(mesh, elem, xi, eta):
    """Lagrange weights of (xi, eta) at the 25 element points
    (``SE_Model::computeInplaneFactor``, ``SE_Model.cpp:604-618``)."""
    axial = bool(mesh.is_axial()[elem])
    wx = spectral.interp_lagrange(xi, spectral.positions(axial))
    we = spectral.interp_lagrange(eta, spectral.positions(False))
    return np.outer(wx, we).reshape(-1)


def time_axis(t0, dt, nsteps):
    """Times of the steps as the solver accumulates them: t = t0, then
    t += dt after every step (``NewmarkTimeScheme.cpp:41,107``)."""
    out = np.empty(nsteps)
    t = float(t0)
    for n in range(nsteps):
        out[n] = t
        t += dt
    return out


def add_axial_moment_source(spec: DomainSpec, depth_z: float, mt6,
                            half_duration, decay_factor, nsteps,
                            time_shift=0.0, stf_order=2):
    """Axial moment source at (s=0, z) with a GaussianSTF of derivative
    order stf_order (2: Ricker); sets the time axis t0 = floor(t_start /
    dt) dt (``main.cpp:574-602``); hdur >= 5 dt (``STF.cpp:32``)."""
    mesh = spec._mesh
    elem_m, xi, eta = locate(mesh, 0.0, depth_z)
    w = inplane_factor(mesh, elem_m, xi, eta)
    elem = int(spec._elem_of_mesh[elem_m])
    if elem < 0:
        raise ValueError("moment source inside a fluid element")
    nu1 = int(spec.elem_nr[elem]) // 2 + 1
    spec.src_elem = int(elem)
    spec.src_pattern = moment_pattern_axial(mt6, w, nu1)
    hdur = max(half_duration, spec.dt * 5.0)
    t_start = -hdur * 2.5 + time_shift
    step0 = int(math.floor(t_start / spec.dt))
    spec.t0 = step0 * spec.dt
    spec.nsteps = int(nsteps)
    spec.stf = gaussian_stf(time_axis(spec.t0, spec.dt, nsteps), hdur,
                            decay_factor, time_shift, stf_order)
    spec._stf_params = (hdur, decay_factor, time_shift)
    return spec


def add_receivers(spec: DomainSpec, s_list, z_list, phi_list):
    mesh = spec._mesh
    elems, ws = [], []
    for s, z in zip(s_list, z_list):
        e, xi, eta = locate(mesh, s, z)
        if spec._elem_of_mesh[e] < 0:
            raise ValueError("receiver inside a fluid element")
        elems.append(int(spec._elem_of_mesh[e]))
        ws.append(inplane_factor(mesh, e, xi, eta))
    spec.rec_elem = np.array(elems, dtype=np.int32)
    spec.rec_weights = np.array(ws, dtype=np.float64)
    spec.rec_phi = np.asarray(phi_list, dtype=np.float64)
    return spec


def make_fluid3d(spec: DomainSpec, elems, amp=0.1, orders=3, seed=0):
    """3D acoustic medium for the given fluid elements: K(phi) = K_1d (1 +
    d(phi)) per GLL point with d a seeded low-order Fourier series in phi
    (|d| <= amp), evaluated at the element's Nr slices (Material::
    createAcoustic with a 3D density, Material.cpp:350-360).  Sets
    spec.fluid_elem_kind (1 for these elements) and spec.fluid_k3d
    ([Nr][25] per 3D element, in element order)."""
    rng = np.random.default_rng(seed)
    nfe = len(spec.fluid_elem_nr)
    kind = np.zeros(nfe, dtype=np.uint8)
    kind[np.asarray(elems, dtype=np.int64)] = 1
    out = []
    for e in range(nfe):
        if not kind[e]:
            continue
        nr = int(spec.fluid_elem_nr[e])
        phi = 2.0 * np.pi * np.arange(nr) / nr
        d = np.zeros((nr, 25))
        for n in range(1, orders + 1):
            a, b = rng.uniform(-1.0, 1.0, (2, 25))
            d += a[None, :] * np.cos(n * phi)[:, None] + \
                b[None, :] * np.sin(n * phi)[:, None]
        d *= amp / max(np.abs(d).max(), 1e-300)
        out.append(spec.fluid_elem_k[e][None, :] * (1.0 + d))
    spec.fluid_elem_kind = kind
    spec.fluid_k3d = (np.concatenate([o.reshape(-1) for o in out])
                      if out else np.zeros(0))
    return spec


def make_mass3d(spec: DomainSpec, points, amp=0.2, orders=3, seed=0):
    """3D mass of the given solid points for a density rho_1d (1 + d(phi))
    with d a seeded low-order Fourier series in phi (|d| <= amp): the
    assembled mass of slice k is m_1d (1 + d(2 pi k / Nr)), since the
    integral factors are the same for every slice (GLLPoint::release with
    a 3D density, src/preloop/se_model/gll/GLLPoint.cpp:54-77).  Sets
    spec.mass3d_points and spec.mass3d (the points' Nr values, in order)."""
    rng = np.random.default_rng(seed)
    points = np.asarray(points, dtype=np.int32)
    out = []
    for p in points:
        nr = int(spec.solid_nr[p])
        phi = 2.0 * np.pi * np.arange(nr) / nr
        d = np.zeros(nr)
        for n in range(1, orders + 1):
            a, b = rng.uniform(-1.0, 1.0, 2)
            d += a * np.cos(n * phi) + b * np.sin(n * phi)
        d *= amp / max(np.abs(d).max(), 1e-300)
        out.append(spec.solid_mass[p] * (1.0 + d))
    spec.mass3d_points = points
    spec.mass3d = np.concatenate(out) if out else np.zeros(0)
    return spec


# ----------------------------------------------------------- workloads
def cartesian_homogeneous(nx=80, nz=100, width=80e3, depth=80e3, nu=2,
                          vp=6000.0, vs=3464.0, rho=2600.0, nsteps=3000,
                          src_depth=10e3, n_receivers=10, seed=0,
                          nodal_nr=None, m0=1e16):
    """Cartesian box, homogeneous isotropic solid, axial moment source with
    a 2 Hz Ricker, surface receivers (the 8000-element nu=2 workload)."""
    mesh = StructuredMesh(np.linspace(0.0, width, nx + 1),
                          np.linspace(0.0, depth, nz + 1))
    nr = 2 * nu + 1

    def material(r):
        one = np.ones_like(r)
        return rho * one, vp * one, vs * one

    if nodal_nr is None:
        nodal_nr = lambda s, z: np.full(s.shape, nr)
    spec = build_domain(mesh, material, nodal_nr)
    rng = np.random.default_rng(seed)
    mt = rng.standard_normal(6)
    mt *= m0 / np.linalg.norm(mt)
    f = 2.0 * math.pi * 1.0     # exp(-f^2 t^2) with f = pi * 2 Hz
    decay = 1.628
    add_axial_moment_source(spec, depth - src_depth, mt, decay / f, decay,
                            nsteps)
    rng1 = np.random.default_rng(seed + 1)
    s_rec = np.linspace(width / n_receivers, width, n_receivers)
    phi = rng1.uniform(0.0, 2.0 * math.pi, n_receivers)
    add_receivers(spec, s_rec, np.full(n_receivers, depth), phi)
    spec.name = f"cartesian_iso1d_E{mesh.nelem}_nu{nu}"
    spec._mt = mt
    return spec


# --------------------------------------------------- Exodus + inparam
def from_exodus(ex, nodal_nr, courant=0.6, bound_nr=True, lucky=True,
                attenuation="NONE"):
    """A domain from an AxiSEM3D Exodus mesh (exodus.ExodusMesh): its quads
    with their linear, spherical or semi-spherical mappings, the model
    stored per element node (RHO with VP, VS or the transversely isotropic
    VPV, VPH, VSV, VSH, ETA; QKAPPA, QMU for attenuation), the points,
    elements, mass and dt of build_domain, the fluid elements through
    make_fluid (solid-fluid pairs from the edges whose neighbour has the
    other medium, the free surface from the TOP side set).  A transversely
    isotropic element stays one unless its velocities reduce to isotropic
    (Material.cpp:80-101): A = rho vph^2, C = rho vpv^2, L = rho vsv^2, N =
    rho vsh^2, F = eta (A - 2 L) at the GLL points (Material::createTISO,
    Material.cpp:498-520).  attenuation CG4 or FULL takes the mesh's SLS
    (w_l, y_l, nr_lin_solids) and band (f_min, f_max, f_ref) and corrects
    every medium through its Voigt averages (make_media)."""
    ti = any(k in ex.elem_vars for k in ("VPV_0", "VPV"))
    nv = ex.nodal_variable
    rho = nv("RHO")
    if ti:
        vpv, vph, vsv, vsh, eta = (nv(k) for k in
                                   ("VPV", "VPH", "VSV", "VSH", "ETA"))
        vp_b, vs_b = np.maximum(vpv, vph), np.maximum(vsv, vsh)
    else:
        vp_b, vs_b = nv("VP"), nv("VS")
    qm = QuadMesh.from_exodus(ex)
    sp = build_domain(qm, (rho, vp_b, vs_b), nodal_nr, courant=courant,
                      lucky=lucky, bound_nr=bound_nr)
    if ex.fluid.any():
        make_fluid(sp, ex.fluid)
    att_kind = str(attenuation).upper()
    if ti or att_kind != "NONE":
        ax = qm.is_axial()
        g = lambda a: interpolate_gll(a, ax)
        e_of_m = np.asarray(sp._elem_of_mesh)
        solid = np.nonzero(e_of_m >= 0)[0]
        mesh_of = solid[np.argsort(e_of_m[solid])]    # solid elem -> mesh
        rg = g(rho)[mesh_of]
        mtype = np.zeros(len(mesh_of), dtype=np.uint8)
        params = []
        if ti:
            gv = {k: g(a)[mesh_of] for k, a in
                  (("vpv", vpv), ("vph", vph), ("vsv", vsv), ("vsh", vsh),
                   ("eta", eta))}
            eps = np.finfo(float).eps
            for i, m in enumerate(mesh_of):
                iso = (np.linalg.norm(vpv[m] - vph[m]) <
                       np.linalg.norm(vpv[m]) * eps and
                       np.linalg.norm(vsv[m] - vsh[m]) <
                       np.linalg.norm(vsv[m]) * eps and
                       np.linalg.norm(eta[m] - 1.0) < 2.0 * eps)
                if iso:
                    vp2, vs2 = gv["vpv"][i] ** 2, gv["vsv"][i] ** 2
                    params.append(np.array([rg[i] * (vp2 - 2.0 * vs2),
                                            rg[i] * vs2]))
                    continue
                mtype[i] = 1
                A = rg[i] * gv["vph"][i] ** 2
                C = rg[i] * gv["vpv"][i] ** 2
                L = rg[i] * gv["vsv"][i] ** 2
                N = rg[i] * gv["vsh"][i] ** 2
                params.append(np.array([A, C, gv["eta"][i] * (A - 2.0 * L),
                                        L, N]))
        else:
            params = [sp.elem_iso[i] for i in range(len(mesh_of))]
        att = None
        if att_kind != "NONE":
            gl = ex.global_variables
            nsls = int(round(gl["nr_lin_solids"]))
            att = {"nsls": nsls, "cg4": att_kind == "CG4",
                   "f_min": gl["f_min"], "f_max": gl["f_max"],
                   "f_ref": gl["f_ref"],
                   "w": [gl[f"w_{k}"] for k in range(nsls)],
                   "y": [gl[f"y_{k}"] for k in range(nsls)],
                   "qkp": g(nv("QKAPPA"))[mesh_of],
                   "qmu": g(nv("QMU"))[mesh_of]}
        make_media(sp, mtype, att=att, params1d=params)
    sp._exodus = ex
    return sp


def load_inparam(input_dir):
    """The keys of a run's inparam.model/nr/source.yaml that the GPU path
    takes (a subset; the reference reads them in main.cpp, NrField.cpp,
    ABC.cpp and Source/STF builders), as a plain dict."""
    import yaml

    def rd(name):
        with open(os.path.join(input_dir, name)) as f:
            return yaml.safe_load(f)
    model, nr, src = (rd("inparam.model.yaml"), rd("inparam.nr.yaml"),
                      rd("inparam.source.yaml"))
    abc = model.get("absorbing_boundary") or {}
    kk = abc.get("Kosloff_Kosloff") or {}
    first = list(src["list_of_sources"][0].values())[0]
    loc, mech, stf = (first["location"], first["mechanism"],
                      first["source_time_function"])
    geo = model.get("geodesy") or {}
    return {
        "exodus_mesh": model["model1D"]["exodus_mesh"],
        "attenuation": model.get("attenuation", "NONE"),
        "pole": str(geo.get("lat_lon_north_pole_mesh", "SOURCE")),
        "abc": {"boundaries": list(abc.get("boundaries") or []),
                "clayton": bool(abc.get("enable_Clayton_Enquist", False)),
                "sponge": bool(kk.get("enable", False)),
                "relative_spans": list(kk.get("relative_spans") or []),
                "gamma_expr_solid": kk.get("gamma_expr_solid"),
                "gamma_expr_fluid": kk.get("gamma_expr_fluid")},
        "nr": {"type": nr["type_Nr"], "constant": nr.get("constant"),
               "bound_by_inplane": bool(nr.get("bound_Nr_by_inplane",
                                                True))},
        "time_axis": {k: src["time_axis"].get(k) for k in
                      ("record_length", "enforced_dt", "Courant_number",
                       "integrator")},
        "source": {"latitude_longitude": list(loc["latitude_longitude"]),
                   "depth": float(loc["depth"]),
                   "ellipticity": bool(loc.get("ellipticity", False)),
                   "undulated_geometry": bool(loc.get("undulated_geometry",
                                                      False)),
                   "mechanism": mech["type"],
                   "data": [float(x) for x in mech["data"]],
                   "unit": float(mech["unit"]),
                   "stf_class": stf["class_name"],
                   "half_duration": float(stf["half_duration"]),
                   "decay_factor": float(stf["decay_factor"]),
                   "time_shift": float(stf["time_shift"]),
                   "derivative": stf["use_derivative_integral"]}}


# the template's Kosloff gamma expressions (inparam.model.yaml), which
# gamma_solid_default / gamma_fluid_default restate
_GAMMA_SOLID_EXPR = "1.1 / T0 * (VS / VP)^2 * exp(-0.04 * SPAN / (VP * T0))"
_GAMMA_FLUID_EXPR = "0.88 / T0 * exp(-0.04 * SPAN / (VP * T0))"


def from_inparam(inp, ex, nsteps, receivers_s=None, receivers_z=None,
                 receivers_phi=None):
    """A run set up as the reference sets it up from its inputs (inp:
    load_inparam's dict; ex: the exodus.ExodusMesh it names): constant Nr
    (bounded by the in-plane resolution if asked), the Courant number,
    Clayton-Enquist and Kosloff boundaries with the template's gamma
    expressions and the mesh period T0, Newmark or the symplectic scheme,
    one moment-tensor source on the axis with a GaussianSTF (depth below
    the mesh top, tensor data times unit), and receivers at (s, z, phi)
    (the inputs' station groups are not read); attenuation NONE, CG4 or
    FULL as from_exodus takes it.  Raises NotImplementedError for what the
    GPU path does not build from inputs (other Nr types, sources off the
    axis, ellipticity, undulated geometry)."""
    nr = inp["nr"]
    if nr["type"] != "CONSTANT":
        raise NotImplementedError("from_inparam: type_Nr other than "
                                  "CONSTANT")
    const = float(nr["constant"])
    ta = inp["time_axis"]
    sp = from_exodus(ex, lambda s, z: np.full(np.shape(s), const),
                     courant=float(ta["Courant_number"]),
                     bound_nr=nr["bound_by_inplane"],
                     attenuation=inp["attenuation"])
    abc = inp["abc"]
    if abc["boundaries"]:
        if abc["sponge"] and (abc["gamma_expr_solid"] != _GAMMA_SOLID_EXPR
                              or abc["gamma_expr_fluid"] !=
                              _GAMMA_FLUID_EXPR):
            raise NotImplementedError("from_inparam: Kosloff gamma "
                                      "expressions other than the "
                                      "template's")
        add_absorbing(sp, keys=tuple(abc["boundaries"]),
                      clayton=abc["clayton"], sponge=abc["sponge"],
                      rel_spans=abc["relative_spans"] or None,
                      t0=ex.global_variables["minimum_period"])
    if ta.get("enforced_dt") not in (None, "NONE"):
        # main.cpp:431-480: an enforced dt replaces the computed one
        sp.dt = float(ta["enforced_dt"])
    if str(ta.get("integrator", "NEWMARK")).upper() == "SYMPLECTIC":
        sp.time_scheme = "symplectic"
    src = inp["source"]
    if src["mechanism"] != "MOMENT_TENSOR" or src["stf_class"] != \
            "GaussianSTF":
        raise NotImplementedError("from_inparam: moment tensors with a "
                                  "GaussianSTF only")
    # the mesh's north pole at the source puts it on the axis
    if (abs(float(src["latitude_longitude"][0]) - 90.0) > 1e-12 and
            inp.get("pole", "SOURCE").upper() != "SOURCE"):
        raise NotImplementedError("from_inparam: sources on the axis "
                                  "(latitude 90, or the mesh's pole at "
                                  "the source) only")
    if src.get("ellipticity") or src.get("undulated_geometry"):
        raise NotImplementedError("from_inparam: ellipticity and "
                                  "undulated geometry")
    ztop = float(ex.coords[:, 1].max())
    mt = np.asarray(src["data"]) * src["unit"]
    add_axial_moment_source(sp, ztop - src["depth"], mt,
                            src["half_duration"], src["decay_factor"],
                            nsteps, time_shift=src["time_shift"],
                            stf_order=STF_ORDER[src["derivative"]])
    if receivers_s is not None:
        add_receivers(sp, np.asarray(receivers_s, dtype=np.float64),
                      np.asarray(receivers_z, dtype=np.float64),
                      np.asarray(receivers_phi, dtype=np.float64))
    return sp


# ------------------------------------------------------------- 3D media
class RandomField:
    """Seeded Gaussian random field with Gaussian covariance
    exp(-|r|^2 / (2 l^2)) by random Fourier features:
    d(x) = amp sqrt(2/K) sum_k cos(w_k . x + b_k), w_k ~ N(0, I / l^2),
    clipped to [-clip, clip]."""

    def __init__(self, seed, corr_len=50e3, nfeat=16, amp=0.015, clip=0.03):
        rng = np.random.default_rng(seed)
        self.w = rng.standard_normal((nfeat, 3)) / corr_len
        self.b = rng.uniform(0.0, 2.0 * math.pi, nfeat)
        self.amp = float(amp)
        self.clip = float(clip)

    @property
    def features(self):
        """[K][4] = wx, wy, wz, b (the layout the C++ generator takes)."""
        return np.ascontiguousarray(np.concatenate([self.w, self.b[:, None]],
                                                   axis=1))

    def __call__(self, x, y, z):
        ph = (x[..., None] * self.w[:, 0] + y[..., None] * self.w[:, 1] +
              z[..., None] * self.w[:, 2] + self.b)
        d = self.amp * math.sqrt(2.0 / len(self.b)) * np.cos(ph).sum(-1)
        return np.clip(d, -self.clip, self.clip)


def att_sls_fit(nsls, f_min, f_max):
    """Relaxation frequencies w_l (log-spaced over the band) and weights
    y_l fitting Q^-1(omega) = sum_l y_l omega w_l / (omega^2 + w_l^2) = 1
    over the band in least squares.  (The reference reads w_l, y_l from
    its external mesher; any such pair defines the SLS model the solver
    then integrates, AttBuilder.cpp:25-36.)"""
    w = 2 * math.pi * np.logspace(math.log10(f_min), math.log10(f_max), nsls)
    om = 2 * math.pi * np.logspace(math.log10(f_min), math.log10(f_max), 100)
    A = om[:, None] * w[None, :] / (om[:, None] ** 2 + w[None, :] ** 2)
    y = np.linalg.lstsq(A, np.ones_like(om), rcond=None)[0]
    return w, y


def att_alpha_beta_gamma(w, y, dt):
    """AttBuilder::setAlphaBetaGamma (AttBuilder.cpp:183-198)."""
    yun = y / y.sum()
    wdt = w * dt
    alpha = np.exp(-wdt)
    beta = yun * ((1.0 - alpha) / wdt - alpha)
    gamma = yun * ((alpha - 1.0) / wdt + 1.0)
    return alpha, beta, gamma


def att_constants(w, y, f_min, f_max, f_ref):
    """Constants of AttBuilder::createAttenuation (AttBuilder.cpp:107-112):
    ysum, w10f, fact."""
    ysum = float(y.sum())
    w0 = f_ref * 2 * math.pi
    w1 = math.sqrt(f_min * f_max) * 2 * math.pi
    w10f = 2.0 * math.log(w1 / w0) / math.pi
    fact = float(np.sum(y * w ** 2 / (w ** 2 + w1 * w1)) / ysum)
    return ysum, w10f, fact


CG4_POINTS = (6, 8, 16, 18)   # (1,1), (1,3), (3,1), (3,3)


def cg4_weights(ifact):
    """Quad::computeWeightsCG4 (Quad.cpp:487-511); ifact (E, 25)."""
    f = ifact.reshape(-1, 5, 5)
    w = np.empty((f.shape[0], 4))
    w[:, 0] = (f[:, 0, 0] + f[:, 0, 1] + f[:, 1, 0] + f[:, 1, 1] +
               0.50 * (f[:, 0, 2] + f[:, 1, 2] + f[:, 2, 0] + f[:, 2, 1]) +
               0.25 * f[:, 2, 2]) / f[:, 1, 1]
    w[:, 1] = (f[:, 0, 3] + f[:, 0, 4] + f[:, 1, 3] + f[:, 1, 4] +
               0.50 * (f[:, 0, 2] + f[:, 1, 2] + f[:, 2, 3] + f[:, 2, 4]) +
               0.25 * f[:, 2, 2]) / f[:, 1, 3]
    w[:, 2] = (f[:, 3, 0] + f[:, 3, 1] + f[:, 4, 0] + f[:, 4, 1] +
               0.50 * (f[:, 2, 0] + f[:, 2, 1] + f[:, 3, 2] + f[:, 4, 2]) +
               0.25 * f[:, 2, 2]) / f[:, 3, 1]
    w[:, 3] = (f[:, 3, 3] + f[:, 3, 4] + f[:, 4, 3] + f[:, 4, 4] +
               0.50 * (f[:, 2, 3] + f[:, 2, 4] + f[:, 3, 2] + f[:, 4, 2]) +
               0.25 * f[:, 2, 2]) / f[:, 3, 3]
    return w


def att_isotropic(lam, mu, qkp, qmu, consts, cg4, wcg4=None):
    """Unrelaxed-modulus correction and memory-variable moduli of an
    isotropic medium (Material::createIsotropic with an AttBuilder,
    Material.cpp:585-598, and AttBuilder::createAttenuation,
    AttBuilder.cpp:103-178).  Arrays (..., 25) broadcast together; wcg4
    (..., 4).  Returns (lam, mu, dlam, dmu) with dlam/dmu (..., 4) for CG4
    (weighted) or (..., 25) for FULL."""
    ysum, w10f, fact = consts
    kp = lam + (2.0 / 3.0) * mu
    kp_na = 1.0 + w10f / qkp
    dkp = kp_na / (qkp / ysum + (1.0 - fact))
    kp_at = kp_na + dkp * fact
    dkp = dkp * kp
    mu_na = 1.0 + w10f / qmu
    dmu = mu_na / (qmu / ysum + (1.0 - fact))
    mu_at = mu_na + dmu * fact
    dmu = dmu * mu
    dlm = dkp - (2.0 / 3.0) * dmu
    if cg4:
        idx = list(CG4_POINTS)
        dlm4 = wcg4 * dlm[..., idx]
        dmu4 = wcg4 * dmu[..., idx]
        kp = kp * kp_na
        mu = mu * mu_na
        kp[..., idx] *= 1.0 + wcg4 * (kp_at[..., idx] / kp_na[..., idx] - 1.0)
        mu[..., idx] *= 1.0 + wcg4 * (mu_at[..., idx] / mu_na[..., idx] - 1.0)
        out_d = (dlm4, dmu4)
    else:
        kp = kp * kp_at
        mu = mu * mu_at
        out_d = (dlm, dmu)
    lam_c = kp - (2.0 / 3.0) * mu
    return lam_c, mu, out_d[0], out_d[1]


def make_3d_isotropic(spec, field, att=None, mask3d=None):
    """Turn the solid elements selected by mask3d (all by default) into 3D
    isotropic elements: vp and vs of the 1D model scaled by (1 + d(x, y,
    z)) with d the random field at the GLL points of slice phi_k = 2 pi k /
    Nr (Model3D.cpp:32,87), rho kept 1D.  The other elements keep their 1D
    material (Fourier path).  att: None or dict(nsls, cg4, f_min, f_max,
    f_ref, qkp, qmu) with qkp, qmu (E, 25); 1D elements then get the 1D
    attenuation (unrelaxed moduli and dlambda, dmu per element).  Numpy
    version for small domains; the C++ generator (synth.py) computes the
    same 3D arrays for large ones."""
    E = spec.n_solid_elements
    if mask3d is None:
        mask3d = np.ones(E, dtype=bool)
    mask3d = np.asarray(mask3d, dtype=bool)
    nr = np.where(mask3d, spec.elem_nr.astype(np.int64), 0)
    off = np.zeros(E + 1, dtype=np.int64)
    off[1:] = np.cumsum(nr)
    iso = spec.elem_iso.copy()
    lam1, mu1 = iso[:, 0], iso[:, 1]
    sz = spec._gq["sz"]
    prop = np.empty((off[-1], 2, NPEM))
    if att is not None:
        w, y = att_sls_fit(att["nsls"], att["f_min"], att["f_max"])
        consts = att_constants(w, y, att["f_min"], att["f_max"],
                               att["f_ref"])
        npa = 4 if att["cg4"] else NPEM
        d3d = np.empty((off[-1], 2, npa))
        d1d = np.zeros((E, 2, npa))
        wcg4 = cg4_weights(spec._gq["ifact"])
    vmax = np.zeros(E)
    for e in range(E):
        if not mask3d[e]:
            vmax[e] = spec._vp1d[e].max()
            if att is not None:
                lc, mc, dl, dm = att_isotropic(
                    lam1[e].copy(), mu1[e].copy(), att["qkp"][e],
                    att["qmu"][e], consts, att["cg4"], wcg4[e])
                spec.elem_iso[e, 0] = lc
                spec.elem_iso[e, 1] = mc
                d1d[e, 0] = dl
                d1d[e, 1] = dm
            continue
        phi = 2 * math.pi * np.arange(nr[e]) / nr[e]
        s, z = sz[e, 0], sz[e, 1]
        x = s[None, :] * np.cos(phi)[:, None]
        yy = s[None, :] * np.sin(phi)[:, None]
        zz = np.broadcast_to(z[None, :], x.shape)
        sc = (1.0 + field(x, yy, zz)) ** 2
        lam = lam1[e][None, :] * sc
        mu = mu1[e][None, :] * sc
        vmax[e] = (spec._vp1d[e][None, :] * np.sqrt(sc)).max()
        if att is not None:
            qk = np.broadcast_to(att["qkp"][e][None, :], lam.shape)
            qm = np.broadcast_to(att["qmu"][e][None, :], lam.shape)
            lam, mu, dl, dm = att_isotropic(lam.copy(), mu.copy(), qk, qm,
                                            consts, att["cg4"],
                                            wcg4[e][None, :])
            d3d[off[e]:off[e + 1], 0] = dl
            d3d[off[e]:off[e + 1], 1] = dm
        prop[off[e]:off[e + 1], 0] = lam
        prop[off[e]:off[e + 1], 1] = mu
    spec.elem_kind = mask3d.astype(np.uint8)
    spec.elem_slice_off = off[:-1].copy()
    spec.prop3d = prop
    # dt from the 3D velocities (max over slices, Quad.cpp:306-339)
    spec.dt = compute_dt(sz, vmax, spec._courant)
    if att is not None:
        spec.att = {"nsls": att["nsls"], "cg4": int(att["cg4"]),
                    "w": w, "y": y, "d3d": d3d,
                    "d1d": d1d if not mask3d.all() else None,
                    "f_min": att["f_min"], "f_max": att["f_max"],
                    "f_ref": att["f_ref"]}
        a, b, g = att_alpha_beta_gamma(w, y, spec.dt)
        spec.att.update(alpha=a, beta=b, gamma=g)
    return spec


MAT_NPARAMS = (2, 5, 21)   # isotropic, TI, anisotropic


def _iso_to_type(mt, lam, mu, u):
    """Parameters (np, ...) of material type mt built from an isotropic
    (lam, mu) and perturbed by the factors (1 + u) (TI, u (5, ...)) or by
    u (lam + 2 mu) on every Cij (anisotropic, u (21, ...))."""
    if mt == 0:
        return np.array([lam, mu])
    if mt == 1:
        base = np.array([lam + 2 * mu, lam + 2 * mu, lam, mu, mu])
        return base * (1.0 + u)
    c = np.zeros((21,) + np.shape(lam))
    for k in (0, 6, 11):
        c[k] = lam + 2 * mu
    for k in (1, 2, 7):
        c[k] = lam
    for k in (15, 18, 20):
        c[k] = mu
    return c + u * (lam + 2 * mu)


def _voigt_kappa_mu(mt, par):
    """Voigt averages of Material::createTISO / createAnisotropic
    (Material.cpp:447-458, 533-536)."""
    if mt == 0:
        lam, mu = par
        return lam + (2.0 / 3.0) * mu, mu
    if mt == 1:
        A, C, F, L, N = par
        return (A * 4. + C + F * 4. - N * 4.) / 9., \
            (A + C - F * 2. + L * 6. + N * 5.) / 15.
    c = par
    kp = (c[0] + c[6] + c[11] + (c[1] + c[7] + c[2]) * 2.) / 9.
    mu = (c[0] + c[6] + c[11] - (c[1] + c[7] + c[2]) +
          (c[15] + c[18] + c[20]) * 3.) / 15.
    return kp, mu


def _att_correct(mt, par, qkp, qmu, consts, cg4, wcg4):
    """Unrelaxed-modulus correction of any medium through its Voigt
    averages (Material.cpp:438-460 and 525-545): the isotropic part
    (kappa + 4/3 mu, kappa - 2/3 mu, mu) is taken out, corrected by
    AttBuilder::createAttenuation and put back.  Returns (par, dlam, dmu)."""
    par = np.array(par, dtype=np.float64)
    kp, mu = _voigt_kappa_mu(mt, par)
    lam_c, mu_c, dl, dm = att_isotropic(kp - (2.0 / 3.0) * mu, mu.copy(),
                                        qkp, qmu, consts, cg4, wcg4)
    kp_c = lam_c + (2.0 / 3.0) * mu_c
    d_n = (kp_c + 4. / 3. * mu_c) - (kp + 4. / 3. * mu)
    d_o = (kp_c - 2. / 3. * mu_c) - (kp - 2. / 3. * mu)
    d_s = mu_c - mu
    if mt == 0:
        par = np.array([kp_c - (2.0 / 3.0) * mu_c, mu_c])
    elif mt == 1:
        par[0] += d_n
        par[1] += d_n
        par[2] += d_o
        par[3] += d_s
        par[4] += d_s
    else:
        for k in (0, 6, 11):
            par[k] += d_n
        for k in (1, 7, 2):
            par[k] += d_o
        for k in (15, 18, 20):
            par[k] += d_s
    return par, dl, dm


def gll_colatitude(sz):
    """Colatitude theta of GLL points (s, z) (geodesy::sz2rtheta, 0 at the
    centre); sz (E, 2, 25) -> (E, 25)."""
    s, z = sz[:, 0], sz[:, 1]
    r = np.hypot(s, z)
    return np.where(r < np.finfo(float).eps, 0.0,
                    np.arccos(np.clip(z / np.maximum(r, 1e-300), -1, 1)))


def make_media(spec, mtype, field=None, mask3d=None, att=None, seed=0,
               perturb=0.05, params1d=None):
    """General solid media: per element material type mtype (0 isotropic, 1
    TI with vertical (radial) axis, 2 anisotropic), 1D or, where mask3d
    holds, 3D (the isotropic 1D model scaled per slice by (1 + d)^2 with
    the random field, as make_3d_isotropic).  TI and anisotropic moduli
    are the isotropic ones perturbed by +-perturb (seeded, per GLL point
    and slice).  Attenuation corrects every medium through its Voigt
    averages.  Spherical meshes get the GLL colatitudes for the RTZ
    rotation of TI and anisotropic elements.  params1d, if given, holds
    each 1D element's own parameters ((np, 25) per element: lambda, mu or
    A, C, F, L, N or the 21 Cij) in place of the perturbed isotropic ones;
    att may bring its SLS (w, y) (keys "w", "y") in place of the fit over
    its band.  Fills spec.elem_mtype, mat1d/mat1d_off, prop3d/prop_off,
    elem_kind/elem_slice_off, att."""
    E = spec.n_solid_elements
    mtype = np.asarray(mtype, dtype=np.uint8)
    if mask3d is None:
        mask3d = np.zeros(E, dtype=bool)
    mask3d = np.asarray(mask3d, dtype=bool)
    rng = np.random.default_rng(seed)
    npar = np.array(MAT_NPARAMS)[mtype]
    nr = np.where(mask3d, spec.elem_nr.astype(np.int64), 0)
    slice_off = np.zeros(E + 1, dtype=np.int64)
    slice_off[1:] = np.cumsum(nr)
    mat1d_off = np.zeros(E + 1, dtype=np.int64)
    mat1d_off[1:] = np.cumsum(npar * NPEM)
    prop_off = np.zeros(E + 1, dtype=np.int64)
    prop_off[1:] = np.cumsum(nr * npar * NPEM)
    mat1d = np.zeros(mat1d_off[-1])
    prop3d = np.zeros(prop_off[-1])
    iso = spec.elem_iso.copy()
    sz = spec._gq["sz"]
    if att is not None:
        if att.get("w") is not None:
            w = np.asarray(att["w"], dtype=np.float64)
            y = np.asarray(att["y"], dtype=np.float64)
        else:
            w, y = att_sls_fit(att["nsls"], att["f_min"], att["f_max"])
        consts = att_constants(w, y, att["f_min"], att["f_max"],
                               att["f_ref"])
        npa = 4 if att["cg4"] else NPEM
        d3d = np.zeros((slice_off[-1], 2, npa))
        d1d = np.zeros((E, 2, npa))
        wcg4 = cg4_weights(spec._gq["ifact"])
    vmax = np.zeros(E)
    grow = math.sqrt(1.0 + 21 * perturb)
    for e in range(E):
        mt = int(mtype[e])
        lam, mu = iso[e, 0], iso[e, 1]
        if not mask3d[e]:
            if params1d is not None:
                par = np.array(params1d[e], dtype=np.float64)
            else:
                u = rng.uniform(-perturb, perturb, (npar[e], NPEM))
                par = _iso_to_type(mt, lam, mu, u)
            if att is not None:
                par, dl, dm = _att_correct(mt, par, att["qkp"][e],
                                           att["qmu"][e], consts,
                                           att["cg4"], wcg4[e])
                d1d[e, 0], d1d[e, 1] = dl, dm
            mat1d[mat1d_off[e]:mat1d_off[e + 1]] = par.reshape(-1)
            vmax[e] = spec._vp1d[e].max() * (grow if mt and params1d is None
                                             else 1.0)
            continue
        k = np.arange(nr[e])
        phi = 2 * math.pi * k / nr[e]
        x = sz[e, 0][None, :] * np.cos(phi)[:, None]
        yy = sz[e, 0][None, :] * np.sin(phi)[:, None]
        zz = np.broadcast_to(sz[e, 1][None, :], x.shape)
        sc = (1.0 + field(x, yy, zz)) ** 2 if field is not None else \
            np.ones_like(x)
        lam3 = lam[None, :] * sc
        mu3 = mu[None, :] * sc
        u = rng.uniform(-perturb, perturb, (npar[e], nr[e], NPEM))
        par = _iso_to_type(mt, lam3, mu3, u)           # (np, Nr, 25)
        if att is not None:
            qk = np.broadcast_to(att["qkp"][e][None, :], lam3.shape)
            qm = np.broadcast_to(att["qmu"][e][None, :], lam3.shape)
            par, dl, dm = _att_correct(mt, par, qk, qm, consts, att["cg4"],
                                       wcg4[e][None, :])
            d3d[slice_off[e]:slice_off[e + 1], 0] = dl
            d3d[slice_off[e]:slice_off[e + 1], 1] = dm
        prop3d[prop_off[e]:prop_off[e + 1]] = \
            np.transpose(par, (1, 0, 2)).reshape(-1)
        vmax[e] = (spec._vp1d[e][None, :] * np.sqrt(sc)).max() * \
            (grow if mt else 1.0)
    spec.elem_mtype = mtype
    spec.mat1d = mat1d
    spec.mat1d_off = mat1d_off[:-1].copy()
    spec.prop3d = prop3d
    spec.prop_off = prop_off[:-1].copy()
    spec.elem_kind = mask3d.astype(np.uint8)
    spec.elem_slice_off = slice_off[:-1].copy()
    # isotropic 1D operands for callers that only know them
    for e in range(E):
        if mtype[e] == 0 and not mask3d[e]:
            spec.elem_iso[e] = mat1d[mat1d_off[e]:mat1d_off[e + 1]].reshape(
                2, NPEM)
    if getattr(spec._mesh, "spherical", False):
        spec.elem_theta = gll_colatitude(sz)
    spec.dt = compute_dt(sz, vmax, spec._courant)
    if att is not None:
        spec.att = {"nsls": att["nsls"], "cg4": int(att["cg4"]),
                    "w": w, "y": y, "d3d": d3d, "d1d": d1d,
                    "f_min": att["f_min"], "f_max": att["f_max"],
                    "f_ref": att["f_ref"]}
        a, b, g = att_alpha_beta_gamma(w, y, spec.dt)
        spec.att.update(alpha=a, beta=b, gamma=g)
    else:
        spec.att = None
    return spec


def layered_profile(seed, r_top, r_bot, nlayer=6, vs=(3200.0, 4800.0),
                    vpvs=(1.73, 1.80), rho=(2600.0, 3600.0)):
    """Seeded 1D model of `nlayer` equal-thickness layers, linear within
    each layer: vs and rho increase with depth (sorted draws), vp/vs drawn
    per layer.  Returns (interfaces, f) with f(r, r_center) ->
    (rho, vp, vs) evaluated in the layer holding r_center."""
    rng = np.random.default_rng(seed)
    vs_v = np.sort(rng.uniform(vs[0], vs[1], 2 * nlayer))
    rho_v = np.sort(rng.uniform(rho[0], rho[1], 2 * nlayer))
    k = rng.uniform(vpvs[0], vpvs[1], nlayer)
    ifc = np.linspace(r_top, r_bot, nlayer + 1)

    def f(r, rc):
        li = np.clip(((r_top - rc) / (r_top - r_bot) * nlayer).astype(int),
                     0, nlayer - 1)
        t = (ifc[li] - r) / (ifc[li] - ifc[li + 1])
        vsr = vs_v[2 * li] + t * (vs_v[2 * li + 1] - vs_v[2 * li])
        rhr = rho_v[2 * li] + t * (rho_v[2 * li + 1] - rho_v[2 * li])
        return rhr, k[li] * vsr, vsr
    return ifc, f


def spherical_regional(ntheta=556, nrad=216, theta_max_deg=30.0,
                       r_outer=6371e3, depth=1200e3, nu_min=10, nu_max=400,
                       seed=11, attenuation=True, cg4=True, nsls=5,
                       nsteps=10000, n_receivers=100, src_depth=50e3,
                       f_peak=0.2, field=None, corr_len=50e3, courant=0.5,
                       bound_nr=True):
    """Spherical sector (theta in [0, theta_max] from the axis, r in
    [r_outer - depth, r_outer]) with a seeded 6-layer 1D model, per-node
    Fourier order nu from nu_min (next to the axis, bounded there by the
    in-plane resolution) to nu_max (at the sector edge), a +-3% 3D random field on vp and vs with correlation length
    corr_len, and 5-SLS attenuation with Q_mu ~ U[80, 600] per layer
    (Q_kappa = 57823).  Axial moment-tensor source with a Ricker STF and
    surface receivers.  Courant number 0.5: the GLJ points next to the axis
    carry orders alpha whose azimuthal stiffness (~ vp alpha / s) adds to
    the in-plane one, and at the template's 0.6 the axial elements exceed
    the Newmark stability limit for alpha ~ 4 (measured by power iteration
    of M^-1 K_alpha), as the reference solver would with the same input."""
    th = np.linspace(0.0, math.radians(theta_max_deg), ntheta + 1)
    rr = np.linspace(r_outer - depth, r_outer, nrad + 1)
    mesh = StructuredMesh(th, rr, spherical=True)
    r_bot = r_outer - depth
    ifc, prof = layered_profile(seed, r_outer, r_bot)
    s_max = r_outer * math.sin(math.radians(theta_max_deg))

    def nodal_nr(s, z):
        # nu grows with the distance from the axis; the in-plane bound of
        # build_domain caps it near the axis (Nr = 5 on it)
        nu = nu_min + (nu_max - nu_min) * np.clip(s / s_max, 0.0, 1.0)
        return 2.0 * nu + 1.0

    spec = build_domain(mesh, prof, nodal_nr, courant=courant,
                        bound_nr=bound_nr)
    if field is None:
        field = RandomField(seed + 1, corr_len=corr_len)
    att = None
    if attenuation:
        rng = np.random.default_rng(seed + 2)
        qmu_layer = rng.uniform(80.0, 600.0, 6)
        rc = np.hypot(spec._gq["sz"][:, 0], spec._gq["sz"][:, 1]).mean(1)
        li = np.clip(((r_outer - rc) / depth * 6).astype(int), 0, 5)
        qmu = np.repeat(qmu_layer[li][:, None], NPEM, axis=1)
        att = dict(nsls=nsls, cg4=cg4, f_min=0.005, f_max=0.5, f_ref=1.0,
                   qkp=np.full_like(qmu, 57823.0), qmu=qmu)
    spec._field = field
    spec._att_in = att
    spec._build_3d = lambda mask3d=None: make_3d_isotropic(spec, field, att,
                                                           mask3d)
    return spec


def finish_regional(spec, nsteps, src_depth=50e3, f_peak=0.2,
                    n_receivers=100, seed=11, theta_max_deg=30.0,
                    r_outer=6371e3):
    """Source (axial moment tensor, Ricker) and surface receivers of the
    spherical workload, after its 3D model fixed dt."""
    rng = np.random.default_rng(seed + 3)
    mt = rng.standard_normal(6)
    mt *= 1e16 / np.linalg.norm(mt)
    f = math.pi * f_peak
    decay = 1.628
    add_axial_moment_source(spec, r_outer - src_depth, mt, decay / f, decay,
                            nsteps)
    th = np.linspace(math.radians(theta_max_deg) / n_receivers,
                     math.radians(theta_max_deg), n_receivers)
    phi = np.random.default_rng(seed + 4).uniform(0, 2 * math.pi,
                                                  n_receivers)
    add_receivers(spec, r_outer * np.sin(th), r_outer * np.cos(th), phi)
    spec._mt = mt
    return spec


def subdomain(spec, elems, prop3d=None, d3d=None):
    """A DomainSpec holding only the given solid elements (and their
    points, renumbered), for timing a bounded sample of a large workload.
    3D operands of the sampled elements may be passed in element order
    ([sum Nr][2][25] and [sum Nr][2][npa])."""
    elems = np.asarray(elems, dtype=np.int64)
    ep = spec.elem_points[elems]
    uniq, inv = np.unique(ep.reshape(-1), return_inverse=True)
    sub = DomainSpec(
        solid_nr=spec.solid_nr[uniq].copy(),
        solid_mass=spec.solid_mass[uniq].copy(),
        solid_sz=spec.solid_sz[uniq].copy(),
        elem_points=inv.reshape(ep.shape).astype(np.int32),
        elem_axial=spec.elem_axial[elems].copy(),
        elem_gq=spec.elem_gq[elems].copy(),
        elem_nr=spec.elem_nr[elems].copy(),
        elem_iso=spec.elem_iso[elems].copy(),
        axial_points=np.nonzero(np.isin(uniq, spec.axial_points))[0].astype(
            np.int32),
        dt=spec.dt, name=spec.name + "_sample")
    if getattr(spec, "elem_kind", None) is not None:
        sub.elem_kind = spec.elem_kind[elems].copy()
        nr = sub.elem_nr.astype(np.int64)
        off = np.zeros(len(elems) + 1, dtype=np.int64)
        off[1:] = np.cumsum(nr)
        sub.elem_slice_off = off[:-1]
        sub.prop3d = prop3d
        att = getattr(spec, "att", None)
        if att is not None:
            sub.att = dict(att)
            sub.att["d3d"] = d3d
    return sub


# ------------------------------------------------------ batched instances
def local_instance(seed, n_elem, nu, nsteps, n_receivers=4, h=1.0e3,
                   courant=0.5, m0=1e16, monitors=0):
    """One independent local simulation: a Cartesian box of about n_elem
    elements of size h, Fourier order nu (bounded by the in-plane
    resolution as the reference bounds it), a seeded isotropic two-layer
    model, a seeded axial moment tensor with a Ricker STF, surface
    receivers and `monitors` receivers 1.5 km apart at the source depth
    (signal within a few hundred steps on large boxes)."""
    rng = np.random.default_rng(seed)
    nx = max(2, int(round(math.sqrt(n_elem))))
    nz = max(2, int(round(n_elem / nx)))
    width, depth = nx * h, nz * h
    mesh = StructuredMesh(np.linspace(0.0, width, nx + 1),
                          np.linspace(0.0, depth, nz + 1))
    vs = rng.uniform(2800.0, 3600.0, 2)
    vp = vs * rng.uniform(1.70, 1.80, 2)
    rho = rng.uniform(2400.0, 3000.0, 2)
    z_if = depth * rng.uniform(0.4, 0.7)

    def material(r, rc):
        top = rc > z_if
        one = np.ones_like(r)
        return (np.where(top, rho[0], rho[1]) * one,
                np.where(top, vp[0], vp[1]) * one,
                np.where(top, vs[0], vs[1]) * one)
    nr = 2 * int(nu) + 1
    spec = build_domain(mesh, material, lambda s, z: np.full(s.shape, nr),
                        courant=courant)
    mt = rng.standard_normal(6)
    mt *= m0 / np.linalg.norm(mt)
    f = 2.0 * math.pi * 1.0
    z_src = depth - 0.3 * depth
    add_axial_moment_source(spec, z_src, mt, 1.628 / f, 1.628, nsteps)
    s_rec = np.linspace(width / n_receivers, width, n_receivers)
    z_rec = np.full(n_receivers, depth)
    if monitors:
        s_rec = np.concatenate([s_rec, 1.5e3 * np.arange(1, monitors + 1)])
        z_rec = np.concatenate([z_rec, np.full(monitors, z_src)])
    phi = rng.uniform(0.0, 2.0 * math.pi, len(s_rec))
    add_receivers(spec, s_rec, z_rec, phi)
    spec.name = f"local_E{mesh.nelem}_nu{nu}_seed{seed}"
    return spec


def batch_instances(specs):
    """One domain holding independent solid instances (1D isotropic, no
    fluid): points, elements, sources and receivers concatenated, each
    instance keeping its own time step (spec.point_instance,
    spec.instance_dt).  spec._instances lists, per instance, its point,
    element and receiver ranges."""
    if not specs:
        raise ValueError("no instances")
    nsteps = specs[0].nsteps
    for s in specs:
        if getattr(s, "fluid_nr", None) is not None or \
                getattr(s, "elem_kind", None) is not None or \
                getattr(s, "elem_mtype", None) is not None:
            raise ValueError("batched instances must be 1D isotropic solids")
        if s.nsteps != nsteps or len(s.stf) != len(specs[0].stf):
            raise ValueError("instances need the same number of steps")
    p0 = np.cumsum([0] + [len(s.solid_nr) for s in specs])
    e0 = np.cumsum([0] + [s.n_solid_elements for s in specs])
    r0 = np.cumsum([0] + [0 if s.rec_elem is None else len(s.rec_elem)
                          for s in specs])
    cat = np.concatenate
    out = DomainSpec(
        solid_nr=cat([s.solid_nr for s in specs]).astype(np.int32),
        solid_mass=cat([s.solid_mass for s in specs]),
        solid_sz=cat([s.solid_sz for s in specs]),
        elem_points=cat([s.elem_points + p for s, p in
                         zip(specs, p0)]).astype(np.int32),
        elem_axial=cat([s.elem_axial for s in specs]),
        elem_gq=np.ascontiguousarray(cat([s.elem_gq for s in specs])),
        elem_nr=cat([s.elem_nr for s in specs]).astype(np.int32),
        elem_iso=np.ascontiguousarray(cat([s.elem_iso for s in specs])),
        axial_points=cat([s.axial_points + p for s, p in
                          zip(specs, p0)]).astype(np.int32),
        dt=float(specs[0].dt))
    out.src_elem = int(specs[0].src_elem)
    out.src_pattern = specs[0].src_pattern
    out.stf = specs[0].stf
    out.extra_sources = [(int(s.src_elem + e), s.src_pattern, s.stf)
                         for s, e in zip(specs[1:], e0[1:])]
    out.rec_elem = cat([s.rec_elem + e for s, e in
                        zip(specs, e0)]).astype(np.int32)
    out.rec_weights = cat([s.rec_weights for s in specs])
    out.rec_phi = cat([s.rec_phi for s in specs])
    out.nsteps = nsteps
    out.point_instance = cat([np.full(len(s.solid_nr), i, np.int32)
                              for i, s in enumerate(specs)])
    out.instance_dt = np.array([s.dt for s in specs])
    out._instances = [((int(p0[i]), int(p0[i + 1])),
                       (int(e0[i]), int(e0[i + 1])),
                       (int(r0[i]), int(r0[i + 1])))
                      for i in range(len(specs))]
    out.name = f"batch_{len(specs)}"
    return out


# ------------------------------------------- 3D fluid mass and coupling
def _phi_series(rng, nr, orders, width):
    """Seeded low-order Fourier series in phi at the nr slices, (nr, width),
    scaled to max |d| = 1."""
    phi = 2.0 * np.pi * np.arange(nr) / nr
    d = np.zeros((nr, width))
    for n in range(1, orders + 1):
        a, b = rng.uniform(-1.0, 1.0, (2, width))
        d += a[None, :] * np.cos(n * phi)[:, None] + \
            b[None, :] * np.sin(n * phi)[:, None]
    return d / max(np.abs(d).max(), 1e-300)


def make_fluid_mass3d(spec, points, amp=0.2, orders=3, seed=0):
    """3D mass of the given fluid points, m_1d (1 + d(phi)) per slice with
    d a seeded low-order Fourier series (|d| <= amp): the fluid mass IF /
    kappa of a 3D bulk modulus (GLLPoint.cpp:78-112 -> Mass3D).  Sets
    spec.fluid_mass3d_points and spec.fluid_mass3d."""
    rng = np.random.default_rng(seed)
    points = np.asarray(points, dtype=np.int32)
    out = []
    for p in points:
        nr = int(spec.fluid_nr[p])
        d = _phi_series(rng, nr, orders, 1)[:, 0]
        out.append(spec.fluid_mass[p] * (1.0 + amp * d))
    spec.fluid_mass3d_points = points
    spec.fluid_mass3d = np.concatenate(out) if out else np.zeros(0)
    return spec


def make_coupling3d(spec, pairs=None, amp=0.1, tilt=0.1, orders=2,
                    seed=0):
    """3D solid-fluid coupling: per slice normals (s, phi, z) of an
    undulated interface, the 1D normal scaled by (1 + d(phi)) with a phi
    component tilt |n| d'(phi) (seeded series, |d| <= amp).  Every pair
    whose fluid point has a 3D mass becomes 3D as well (GLLPoint.cpp:
    129-143).  Sets spec.sf3d_pairs (ascending) and spec.sf3d_normal
    ([Nr][3] per 3D pair, in pair order)."""
    rng = np.random.default_rng(seed)
    sel = set(range(len(spec.sf_solid)) if pairs is None else
              (int(k) for k in pairs))
    fm3 = getattr(spec, "fluid_mass3d_points", None)
    if fm3 is not None:
        f3 = set(int(p) for p in fm3)
        sel |= {k for k in range(len(spec.sf_fluid))
                if int(spec.sf_fluid[k]) in f3}
    sel = sorted(sel)
    out = []
    for k in sel:
        nr = int(spec.solid_nr[spec.sf_solid[k]])
        ns, nz = spec.sf_normal[k]
        mag = math.hypot(ns, nz)
        d = _phi_series(rng, nr, orders, 3)
        n3 = np.stack([ns * (1.0 + amp * d[:, 0]),
                       tilt * mag * d[:, 1],
                       nz * (1.0 + amp * d[:, 2])], axis=1)
        out.append(n3)
    spec.sf3d_pairs = np.asarray(sel, dtype=np.int32)
    spec.sf3d_normal = (np.concatenate(out).reshape(-1) if out
                        else np.zeros(0))
    return spec
