"""Spectral-element constants for npol = 4 (5 x 5 GLL points per element).

Nodes, weights and the differentiation matrix G are computed here from
their closed forms:

* GLL nodes: -1, +1 and the roots of P'_4 (Legendre); weights
  2 / (N (N+1) P_N(x)^2).
* GLJ(0,1) nodes used along xi on axial elements: -1, +1 and the roots of
  the Jacobi polynomial P_{N-1}^{(1,2)}; weights w_i = int l_i(x) (1+x) dx.
* G(i, j) = l_i'(x_j), the derivative of the i-th Lagrange polynomial at
  the j-th node.

The reference stores these as 12-decimal literals
(``SOLVER/src/preloop/se_model/quad/spectrals.hpp:73-102``, salvus tables)
and the solver uses those literals, not the exact values.  To keep the
operands identical to the reference's, every table below is rounded to 12
decimals after it is computed; the exact values remain available through
``exact=True``.  Layout follows the reference: ``G[i, j]`` row-major,
element point index ``ipnt = ipol * 5 + jpol`` with xi along ipol
(``spectrals.hpp:296-313``).
"""
from __future__ import annotations

import functools

import numpy as np
from numpy.polynomial import legendre as npleg

NPOL = 4
NPED = NPOL + 1          # points per direction
NPEM = NPED * NPED       # points per element


def _jacobi_poly(n: int, a: float, b: float) -> np.poly1d:
    """Jacobi polynomial P_n^{(a,b)} built by the three-term recurrence."""
    p0 = np.poly1d([1.0])
    if n == 0:
        return p0
    p1 = np.poly1d([0.5 * (a + b + 2.0), 0.5 * (a - b)])
    for k in range(2, n + 1):
        c = 2.0 * k + a + b
        a1 = 2.0 * k * (k + a + b) * (c - 2.0)
        a2 = (c - 1.0) * (a * a - b * b)
        a3 = (c - 2.0) * (c - 1.0) * c
        a4 = 2.0 * (k + a - 1.0) * (k + b - 1.0) * c
        p2 = (np.poly1d([a3, a2]) * p1 - a4 * p0) / a1
        p0, p1 = p1, p2
    return p1


def _polish_roots(poly: np.poly1d, roots: np.ndarray) -> np.ndarray:
    d = poly.deriv()
    r = np.array(sorted(np.real(roots)), dtype=np.float64)
    for _ in range(8):
        r = r - poly(r) / d(r)
    return r


def _lagrange_deriv_matrix(x: np.ndarray) -> np.ndarray:
    """G[i, j] = l_i'(x_j) for Lagrange polynomials through nodes x."""
    n = len(x)
    G = np.zeros((n, n))
    for i in range(n):
        others = np.delete(x, i)
        li = np.poly1d(others, r=True) / np.prod(x[i] - others)
        G[i, :] = li.deriv()(x)
    return G


def _weights_by_integration(x: np.ndarray, jacobi_b: int) -> np.ndarray:
    """w_i = int_{-1}^{1} l_i(t) (1+t)^b dt, exact via Gauss-Legendre."""
    gx, gw = npleg.leggauss(12)
    n = len(x)
    w = np.zeros(n)
    for i in range(n):
        others = np.delete(x, i)
        li = np.poly1d(others, r=True) / np.prod(x[i] - others)
        w[i] = np.sum(gw * li(gx) * (1.0 + gx) ** jacobi_b)
    return w


@functools.lru_cache(maxsize=None)
def _exact_tables():
    N = NPOL
    # GLL: interior nodes are roots of P_N'
    pn = np.poly1d(npleg.leg2poly([0] * N + [1])[::-1])
    inner = _polish_roots(pn.deriv(), pn.deriv().roots)
    x_gll = np.concatenate([[-1.0], inner, [1.0]])
    w_gll = 2.0 / (N * (N + 1) * pn(x_gll) ** 2)
    # GLJ(0,1): interior nodes are roots of P_{N-1}^{(1,2)}
    pj = _jacobi_poly(N - 1, 1.0, 2.0)
    inner_j = _polish_roots(pj, pj.roots)
    x_glj = np.concatenate([[-1.0], inner_j, [1.0]])
    w_glj = _weights_by_integration(x_glj, 1)
    g_gll = _lagrange_deriv_matrix(x_gll)
    g_glj = _lagrange_deriv_matrix(x_glj)
    return x_gll, x_glj, w_gll, w_glj, g_gll, g_glj


def _round12(a: np.ndarray) -> np.ndarray:
    out = np.round(np.asarray(a, dtype=np.float64), 12)
    out[out == 0.0] = 0.0   # drop negative zeros
    return out


def tables(exact: bool = False):
    """Return (x_gll, x_glj, w_gll, w_glj, G_gll, G_glj) as float64 arrays."""
    t = _exact_tables()
    if exact:
        return tuple(a.copy() for a in t)
    return tuple(_round12(a) for a in t)


def positions(axial: bool) -> np.ndarray:
    x_gll, x_glj = tables()[0:2]
    return x_glj if axial else x_gll


def xieta_element(axial: bool) -> np.ndarray:
    """(2, 25) reference coordinates of the element points, ipnt = ipol*5+jpol
    (``spectrals.hpp:296-316``)."""
    x_gll, x_glj = tables()[0:2]
    xi_nodes = x_glj if axial else x_gll
    xi = np.repeat(xi_nodes, NPED)
    eta = np.tile(x_gll, NPED)
    return np.stack([xi, eta])


def weights_1d(axial: bool) -> np.ndarray:
    """(5,) GLJ (axial) or GLL quadrature weights."""
    _, _, w_gll, w_glj, _, _ = tables()
    return (w_glj if axial else w_gll).copy()


def weights_element(axial: bool) -> np.ndarray:
    """(25,) tensor-product weights (``spectrals.hpp:319-327``)."""
    _, _, w_gll, w_glj, _, _ = tables()
    wxi = w_glj if axial else w_gll
    return np.outer(wxi, w_gll).reshape(-1)


def interp_lagrange(x: float, nodes: np.ndarray) -> np.ndarray:
    """Lagrange interpolation weights of point x on the given nodes."""
    n = len(nodes)
    w = np.ones(n)
    for i in range(n):
        for j in range(n):
            if i != j:
                w[i] *= (x - nodes[j]) / (nodes[i] - nodes[j])
    return w


# ---------------------------------------------------------------- lucky Nr
@functools.lru_cache(maxsize=None)
def _lucky_numbers(limit: int = 100000) -> np.ndarray:
    """FFT-friendly sizes: largest prime factor <= 13 and at most one factor
    of 11 or 13 in total (rule of ``NrField::isLuckyNumber``,
    ``SOLVER/src/preloop/nr_field/NrField.cpp:63-92``), up to 100000."""
    out = []
    for n in range(1, limit + 1):
        m = n
        for p in (2, 3, 5, 7):
            while m % p == 0:
                m //= p
        e = 0
        while m % 11 == 0:
            m //= 11
            e += 1
        while m % 13 == 0:
            m //= 13
            e += 1
        if m == 1 and e <= 1:
            out.append(n)
    return np.array(out, dtype=np.int64)


def next_lucky_number(n: int) -> int:
    """Smallest lucky number >= n; n itself beyond the table
    (``NrField::nextLuckyNumber``, ``NrField.cpp:95-103``)."""
    tab = _lucky_numbers()
    k = int(np.searchsorted(tab, n, side="left"))
    if k >= len(tab):
        return int(n)
    return int(tab[k])
