"""ctypes wrapper of the FP64 CPU oracle (``oracle/axs3d_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product
path (``axisem3d_amd``) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build(force=False):
    """Compile the oracle with its Makefile (gcc)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _setup(_lib)
    return _lib


_p = C.c_void_p


class OracleDomain(C.Structure):
    _fields_ = [
        ("n_points", C.c_int), ("point_nr", _p), ("point_mass", _p),
        ("offsets", _p),
        ("n_elem", C.c_int), ("elem_points", _p), ("elem_nr", _p),
        ("elem_axial", _p), ("elem_gq", _p), ("elem_iso", _p),
        ("elem_kind", _p), ("elem_slice_off", _p), ("prop3d", _p),
        ("nsls", C.c_int), ("att_cg4", C.c_int), ("att_abg", _p),
        ("att1d", _p), ("att3d", _p), ("att_state", _p),
        ("n_axial", C.c_int), ("axial_points", _p),
        ("src_elem", C.c_int), ("src_nu1", C.c_int), ("src_pattern", _p),
        ("stf", _p),
        ("n_rec", C.c_int), ("rec_elem", _p), ("rec_weights", _p),
        ("rec_phi", _p),
        ("dt", C.c_double),
        ("elem_mtype", _p), ("mat1d", _p), ("mat1d_off", _p),
        ("prop_off", _p), ("elem_theta", _p),
        ("n_fpoints", C.c_int), ("fpoint_nr", _p), ("fpoint_mass", _p),
        ("foffsets", _p),
        ("n_felem", C.c_int), ("felem_points", _p), ("felem_nr", _p),
        ("felem_axial", _p), ("felem_gq", _p), ("felem_k", _p),
        ("n_faxial", C.c_int), ("faxial_points", _p),
        ("n_fsurf", C.c_int), ("fsurf_points", _p),
        ("n_sf", C.c_int), ("sf_solid", _p), ("sf_fluid", _p),
        ("sf_normal", _p),
        ("mass3d_off", _p), ("mass3d", _p),
        ("felem_kind", _p), ("fk3d", _p), ("fk3d_off", _p),
        ("fmass3d_off", _p), ("fmass3d", _p),
        ("sf3d_off", _p), ("sf3d_normal", _p),
        ("n_clayton_s", C.c_int), ("clayton_s_point", _p),
        ("clayton_s_nval", _p), ("clayton_s_off", _p),
        ("clayton_s_data", _p),
        ("n_clayton_f", C.c_int), ("clayton_f_point", _p),
        ("clayton_f_nval", _p), ("clayton_f_off", _p),
        ("clayton_f_data", _p),
        ("n_sponge_s", C.c_int), ("sponge_s_point", _p),
        ("sponge_s_nval", _p), ("sponge_s_off", _p), ("sponge_s_data", _p),
        ("n_sponge_f", C.c_int), ("sponge_f_point", _p),
        ("sponge_f_nval", _p), ("sponge_f_off", _p), ("sponge_f_data", _p),
        ("ocean_off", _p), ("ocean_nval", _p), ("ocean_data", _p),
        ("prt_off", _p), ("prt_data", _p),
    ]


def _setup(L):
    L.oracle_set_gmat.argtypes = [_p, _p]
    for name in ("oracle_grad6", "oracle_quad6", "oracle_quad6_noint"):
        getattr(L, name).argtypes = [_p, C.c_int, C.c_int, _p, _p]
    L.oracle_iso_stress.argtypes = [_p, _p, _p, _p]
    L.oracle_solid_element_stiff.argtypes = [C.POINTER(OracleDomain), C.c_int,
                                             _p, _p]
    L.oracle_run.argtypes = [C.POINTER(OracleDomain), C.c_int, C.c_int,
                             _p, _p, _p, _p, _p]
    L.oracle_run.restype = C.c_int
    L.oracle_record.argtypes = [C.POINTER(OracleDomain), _p, _p]
    L.oracle_apply_moment_source.argtypes = [C.POINTER(OracleDomain),
                                             C.c_double, _p]
    L.oracle_axial_solid.argtypes = [C.POINTER(OracleDomain), _p]
    L.oracle_stiff_to_accel.argtypes = [C.POINTER(OracleDomain), _p]
    L.oracle_newmark.argtypes = [C.c_int64, C.c_double, _p, _p, _p, _p]
    L.oracle_r2c.argtypes = [_p, C.c_int, _p]
    L.oracle_c2r.argtypes = [_p, C.c_int, _p]
    L.oracle_att_alloc.argtypes = [C.POINTER(OracleDomain)]
    L.oracle_att_alloc.restype = _p
    L.oracle_att_free.argtypes = [_p, C.c_int]
    L.oracle_att_apply_cd.argtypes = [C.POINTER(OracleDomain), C.c_int,
                                      C.c_int, _p, _p]
    L.oracle_att_apply_fr.argtypes = [C.POINTER(OracleDomain), C.c_int,
                                      C.c_int, _p, _p]
    L.oracle_spz_to_rtz6.argtypes = [_p, _p]
    for name in ("oracle_grad9", "oracle_quad9"):
        getattr(L, name).argtypes = [_p, C.c_int, C.c_int, _p, _p]
    for name in ("oracle_spz_to_rtz9", "oracle_rtz_to_spz9",
                 "oracle_prt_sph_to_und6", "oracle_prt_und_to_sph9"):
        getattr(L, name).argtypes = [_p, _p, _p][:2 if "rtz" in name else 3]
    for name in ("oracle_grad3", "oracle_quad3"):
        getattr(L, name).argtypes = [_p, C.c_int, C.c_int, _p, _p]
    L.oracle_fluid_element_stiff.argtypes = [C.POINTER(OracleDomain),
                                             C.c_int, _p, _p]
    L.oracle_coupling.argtypes = [C.POINTER(OracleDomain), _p, _p, _p]
    L.oracle_fluid_bc.argtypes = [C.POINTER(OracleDomain), _p]
    L.oracle_fluid_stiff_to_accel.argtypes = [C.POINTER(OracleDomain), _p]
    L.oracle_run_coupled.argtypes = [C.POINTER(OracleDomain), C.c_int,
                                     C.c_int] + [_p] * 9
    L.oracle_run_coupled.restype = C.c_int
    L.oracle_material_stress.argtypes = [C.c_int, _p, _p, _p]
    L.oracle_rtz_to_spz6.argtypes = [_p, _p]
    L.oracle_set_threads.argtypes = [C.c_int]
    L.oracle_clayton.argtypes = [C.POINTER(OracleDomain), _p, _p, _p, _p]
    L.oracle_run_symplectic.argtypes = [C.POINTER(OracleDomain), C.c_int,
                                        C.c_int] + [_p] * 7
    L.oracle_run_symplectic.restype = C.c_int
    L.oracle_sponge.argtypes = [C.POINTER(OracleDomain)] + [_p] * 6
    # serial unless a caller asks for threads (set_threads)
    L.oracle_set_threads(1)
    from axisem3d_amd import spectral
    t = spectral.tables()
    g_gll = np.ascontiguousarray(t[4])
    g_glj = np.ascontiguousarray(t[5])
    L.oracle_set_gmat(g_gll.ctypes.data, g_glj.ctypes.data)


def set_threads(n):
    """OpenMP threads of the oracle's element loop (results are the same
    for any count)."""
    lib().oracle_set_threads(int(n))


def _ptr(a):
    return None if a is None else a.ctypes.data


# ------------------------------------------------------------ stage level
def grad6(gq_e, axial, alpha, u):
    """u complex (3, 25) -> strain complex (6, 25)."""
    u = np.ascontiguousarray(u, dtype=np.complex128)
    gq_e = np.ascontiguousarray(gq_e, dtype=np.float64)
    out = np.zeros((6, 25), dtype=np.complex128)
    lib().oracle_grad6(gq_e.ctypes.data, int(axial), int(alpha),
                       u.ctypes.data, out.ctypes.data)
    return out


def quad6(gq_e, axial, alpha, sig, integrate=True):
    sig = np.ascontiguousarray(sig, dtype=np.complex128)
    gq_e = np.ascontiguousarray(gq_e, dtype=np.float64)
    out = np.zeros((3, 25), dtype=np.complex128)
    fn = lib().oracle_quad6 if integrate else lib().oracle_quad6_noint
    fn(gq_e.ctypes.data, int(axial), int(alpha), sig.ctypes.data,
       out.ctypes.data)
    return out


def iso_stress(lam, mu, eps):
    eps = np.ascontiguousarray(eps, dtype=np.complex128)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    out = np.zeros((6, 25), dtype=np.complex128)
    lib().oracle_iso_stress(lam.ctypes.data, mu.ctypes.data, eps.ctypes.data,
                            out.ctypes.data)
    return out


def material_stress(mt, par, eps):
    """Stress (6, 25) of material type mt (0 isotropic [lambda, mu], 1 TI
    [A, C, F, L, N], 2 anisotropic [21 Cij]); par (np, 25)."""
    eps = np.ascontiguousarray(eps, dtype=np.complex128)
    par = np.ascontiguousarray(par, dtype=np.float64)
    out = np.zeros((6, 25), dtype=np.complex128)
    lib().oracle_material_stress(int(mt), par.ctypes.data, eps.ctypes.data,
                                 out.ctypes.data)
    return out


def spz_to_rtz6(theta, eps):
    """Voigt strain (engineering shears) (6, 25) from (s, phi, z) to
    (R, T, Z) at colatitudes theta (25,)."""
    e = np.ascontiguousarray(eps, dtype=np.complex128).copy()
    th = np.ascontiguousarray(theta, dtype=np.float64)
    lib().oracle_spz_to_rtz6(th.ctypes.data, e.ctypes.data)
    return e


def rtz_to_spz6(theta, sig):
    """Voigt stress (6, 25) from (R, T, Z) back to (s, phi, z)."""
    s = np.ascontiguousarray(sig, dtype=np.complex128).copy()
    th = np.ascontiguousarray(theta, dtype=np.float64)
    lib().oracle_rtz_to_spz6(th.ctypes.data, s.ctypes.data)
    return s


def grad9(gq_e, axial, alpha, u):
    """computeGrad9: u (3, 25) complex -> (9, 25), g[3 i + j] = du_i/dx_j."""
    u = np.ascontiguousarray(u, dtype=np.complex128)
    g = np.zeros((9, 25), dtype=np.complex128)
    lib().oracle_grad9(_ptr(gq_e), int(axial), int(alpha), _ptr(u), _ptr(g))
    return g


def quad9(gq_e, axial, alpha, f):
    """computeQuad9: f (9, 25) complex -> (3, 25)."""
    f = np.ascontiguousarray(f, dtype=np.complex128)
    fi = np.zeros((3, 25), dtype=np.complex128)
    lib().oracle_quad9(_ptr(gq_e), int(axial), int(alpha), _ptr(f), _ptr(fi))
    return fi


def spz_to_rtz9(theta, g):
    g = np.array(g, dtype=np.complex128)
    lib().oracle_spz_to_rtz9(_ptr(np.ascontiguousarray(theta, np.float64)),
                             _ptr(g))
    return g


def rtz_to_spz9(theta, g):
    g = np.array(g, dtype=np.complex128)
    lib().oracle_rtz_to_spz9(_ptr(np.ascontiguousarray(theta, np.float64)),
                             _ptr(g))
    return g


def prt_sph_to_und6(x, g9):
    """PRT::sphericalToUndulated6 with x = (X0, X1, X2, X3) (4, 25)."""
    g9 = np.ascontiguousarray(g9, dtype=np.complex128)
    e6 = np.zeros((6, 25), dtype=np.complex128)
    lib().oracle_prt_sph_to_und6(_ptr(np.ascontiguousarray(x, np.float64)),
                                 _ptr(g9), _ptr(e6))
    return e6


def prt_und_to_sph9(xj, s6):
    """PRT::undulatedToSpherical6 with xj = (X0 XJ, .., X3 XJ) (4, 25)."""
    s6 = np.ascontiguousarray(s6, dtype=np.complex128)
    f9 = np.zeros((9, 25), dtype=np.complex128)
    lib().oracle_prt_und_to_sph9(_ptr(np.ascontiguousarray(xj, np.float64)),
                                 _ptr(s6), _ptr(f9))
    return f9


def grad3(gq_e, axial, alpha, p):
    """p complex (25,) -> (dp/ds, i alpha p / s, dp/dz) complex (3, 25)."""
    p = np.ascontiguousarray(p, dtype=np.complex128)
    gq_e = np.ascontiguousarray(gq_e, dtype=np.float64)
    out = np.zeros((3, 25), dtype=np.complex128)
    lib().oracle_grad3(gq_e.ctypes.data, int(axial), int(alpha),
                       p.ctypes.data, out.ctypes.data)
    return out


def quad3(gq_e, axial, alpha, qi):
    qi = np.ascontiguousarray(qi, dtype=np.complex128)
    gq_e = np.ascontiguousarray(gq_e, dtype=np.float64)
    out = np.zeros(25, dtype=np.complex128)
    lib().oracle_quad3(gq_e.ctypes.data, int(axial), int(alpha),
                       qi.ctypes.data, out.ctypes.data)
    return out


def r2c(x):
    """computeR2C of one real series (reference FFT conventions)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(len(x) // 2 + 1, dtype=np.complex128)
    lib().oracle_r2c(x.ctypes.data, len(x), out.ctypes.data)
    return out


def c2r(c, nr):
    """computeC2R of one Hermitian series of nr // 2 + 1 coefficients."""
    c = np.ascontiguousarray(c, dtype=np.complex128)
    assert len(c) == nr // 2 + 1
    out = np.zeros(nr)
    lib().oracle_c2r(c.ctypes.data, nr, out.ctypes.data)
    return out


# ------------------------------------------------------------ domain level
def rotate_station(seis, wcs, theta, baz, phi, cartesian):
    """Station::rotateField for the displacement (Station.hpp:79-118),
    which arrives in (s, phi, z) (SolidElement::displInRTZ is false), with
    Station::rotate (Station.cpp:41-55): about axis k by angle a, v_i' =
    v_i cos a + v_j sin a, v_j' = -v_i sin a + v_j cos a, i = k + 1, j = k +
    2 (mod 3).  seis [sample][station][3]; per station wcs (0 spz, 1 RTZ,
    2 ENZ, 3 xyz), theta, back azimuth baz and azimuth phi."""
    out = np.array(seis, dtype=np.float64, copy=True)

    def rot(r, k, a):
        i, j = (k + 1) % 3, (k + 2) % 3
        vi, vj = out[:, r, i].copy(), out[:, r, j].copy()
        out[:, r, i] = vi * np.cos(a) + vj * np.sin(a)
        out[:, r, j] = -vi * np.sin(a) + vj * np.cos(a)

    for r in range(out.shape[1]):
        w, bz90 = int(wcs[r]), float(baz[r]) + np.pi * 0.5
        if cartesian:
            if w == 2:
                rot(r, 2, bz90)
            elif w == 3:
                rot(r, 2, -float(phi[r]))
        else:
            if w == 1:
                rot(r, 1, float(theta[r]))
            elif w == 2:
                rot(r, 1, float(theta[r]))
                rot(r, 2, bz90)
            elif w == 3:
                rot(r, 2, -float(phi[r]))
    return out


class Oracle:
    """FP64 restatement of the time loop on a ``DomainSpec``."""

    def __init__(self, spec):
        self.spec = spec
        self._keep = []
        nr = np.ascontiguousarray(spec.solid_nr, dtype=np.int32)
        self.offsets = np.zeros(len(nr) + 1, dtype=np.int64)
        self.offsets[1:] = np.cumsum(nr // 2 + 1)
        self.nmodes = int(self.offsets[-1])
        d = OracleDomain()
        d.n_points = len(nr)
        d.point_nr = self._k(nr)
        d.point_mass = self._k(np.ascontiguousarray(spec.solid_mass,
                                                    dtype=np.float64))
        d.offsets = self._k(self.offsets)
        d.n_elem = spec.n_solid_elements
        d.elem_points = self._k(np.ascontiguousarray(spec.elem_points,
                                                     dtype=np.int32))
        d.elem_nr = self._k(np.ascontiguousarray(spec.elem_nr,
                                                 dtype=np.int32))
        d.elem_axial = self._k(np.ascontiguousarray(spec.elem_axial,
                                                    dtype=np.uint8))
        d.elem_gq = self._k(np.ascontiguousarray(spec.elem_gq,
                                                 dtype=np.float64))
        d.elem_iso = self._k(np.ascontiguousarray(spec.elem_iso,
                                                  dtype=np.float64))
        ax = np.ascontiguousarray(spec.axial_points, dtype=np.int32)
        d.n_axial = len(ax)
        d.axial_points = self._k(ax)
        if spec.src_elem >= 0 and spec.src_pattern is not None:
            d.src_elem = int(spec.src_elem)
            d.src_nu1 = spec.src_pattern.shape[0]
            d.src_pattern = self._k(np.ascontiguousarray(
                spec.src_pattern, dtype=np.complex128))
            d.stf = self._k(np.ascontiguousarray(spec.stf, dtype=np.float64))
        else:
            d.src_elem = -1
        if spec.rec_elem is not None:
            d.n_rec = len(spec.rec_elem)
            d.rec_elem = self._k(np.ascontiguousarray(spec.rec_elem,
                                                      dtype=np.int32))
            d.rec_weights = self._k(np.ascontiguousarray(spec.rec_weights,
                                                         dtype=np.float64))
            d.rec_phi = self._k(np.ascontiguousarray(spec.rec_phi,
                                                     dtype=np.float64))
        kind = getattr(spec, "elem_kind", None)
        if kind is not None:
            d.elem_kind = self._k(np.ascontiguousarray(kind, dtype=np.uint8))
            d.elem_slice_off = self._k(np.ascontiguousarray(
                spec.elem_slice_off, dtype=np.int64))
            d.prop3d = self._k(np.ascontiguousarray(spec.prop3d,
                                                    dtype=np.float64))
        mtype = getattr(spec, "elem_mtype", None)
        if mtype is not None:
            d.elem_mtype = self._k(np.ascontiguousarray(mtype, dtype=np.uint8))
            d.mat1d = self._k(np.ascontiguousarray(spec.mat1d,
                                                   dtype=np.float64))
            d.mat1d_off = self._k(np.ascontiguousarray(spec.mat1d_off,
                                                       dtype=np.int64))
            if getattr(spec, "prop_off", None) is not None:
                d.prop_off = self._k(np.ascontiguousarray(spec.prop_off,
                                                          dtype=np.int64))
        theta = getattr(spec, "elem_theta", None)
        if theta is not None:
            d.elem_theta = self._k(np.ascontiguousarray(theta,
                                                        dtype=np.float64))
        att = getattr(spec, "att", None)
        if att is not None:
            d.nsls = int(att["nsls"])
            d.att_cg4 = int(att["cg4"])
            d.att_abg = self._k(np.ascontiguousarray(
                np.concatenate([att["alpha"], att["beta"], att["gamma"]]),
                dtype=np.float64))
            if att.get("d1d") is not None:
                d.att1d = self._k(np.ascontiguousarray(att["d1d"],
                                                       dtype=np.float64))
            if att.get("d3d") is not None:
                d.att3d = self._k(np.ascontiguousarray(att["d3d"],
                                                       dtype=np.float64))
        fnr = getattr(spec, "fluid_nr", None)
        self.fnmodes = 0
        if fnr is not None and len(fnr) > 0:
            fnr = np.ascontiguousarray(fnr, dtype=np.int32)
            self.foffsets = np.zeros(len(fnr) + 1, dtype=np.int64)
            self.foffsets[1:] = np.cumsum(fnr // 2 + 1)
            self.fnmodes = int(self.foffsets[-1])
            d.n_fpoints = len(fnr)
            d.fpoint_nr = self._k(fnr)
            d.fpoint_mass = self._k(np.ascontiguousarray(spec.fluid_mass,
                                                         dtype=np.float64))
            d.foffsets = self._k(self.foffsets)
            fe = np.ascontiguousarray(spec.fluid_elem_points, dtype=np.int32)
            d.n_felem = fe.shape[0]
            d.felem_points = self._k(fe)
            d.felem_nr = self._k(np.ascontiguousarray(spec.fluid_elem_nr,
                                                      dtype=np.int32))
            d.felem_axial = self._k(np.ascontiguousarray(
                spec.fluid_elem_axial, dtype=np.uint8))
            d.felem_gq = self._k(np.ascontiguousarray(spec.fluid_elem_gq,
                                                      dtype=np.float64))
            d.felem_k = self._k(np.ascontiguousarray(spec.fluid_elem_k,
                                                     dtype=np.float64))
            fa = np.ascontiguousarray(spec.fluid_axial_points, dtype=np.int32)
            d.n_faxial = len(fa)
            d.faxial_points = self._k(fa)
            fs = np.ascontiguousarray(spec.fluid_surface_points,
                                      dtype=np.int32)
            d.n_fsurf = len(fs)
            d.fsurf_points = self._k(fs)
            d.n_sf = len(spec.sf_solid)
            d.sf_solid = self._k(np.ascontiguousarray(spec.sf_solid,
                                                      dtype=np.int32))
            d.sf_fluid = self._k(np.ascontiguousarray(spec.sf_fluid,
                                                      dtype=np.int32))
            d.sf_normal = self._k(np.ascontiguousarray(spec.sf_normal,
                                                       dtype=np.float64))
        fkind = getattr(spec, "fluid_elem_kind", None)
        if fkind is not None and getattr(spec, "fluid_nr", None) is not None:
            fkind = np.ascontiguousarray(fkind, dtype=np.uint8)
            fnr = np.asarray(spec.fluid_elem_nr, dtype=np.int64)
            off = np.full(len(fkind), -1, dtype=np.int64)
            cum = 0
            for e in range(len(fkind)):
                if fkind[e] == 1:
                    off[e] = cum
                    cum += fnr[e] * 25
            d.felem_kind = self._k(fkind)
            d.fk3d = self._k(np.ascontiguousarray(spec.fluid_k3d,
                                                  dtype=np.float64))
            d.fk3d_off = self._k(off)
        m3p = getattr(spec, "mass3d_points", None)
        if m3p is not None and len(m3p) > 0:
            off = np.full(len(spec.solid_nr), -1, dtype=np.int64)
            nr = np.asarray(spec.solid_nr, dtype=np.int64)
            cum = 0
            for p in np.asarray(m3p, dtype=np.int64):
                off[p] = cum
                cum += nr[p]
            d.mass3d_off = self._k(off)
            d.mass3d = self._k(np.ascontiguousarray(spec.mass3d,
                                                    dtype=np.float64))
        fm3 = getattr(spec, "fluid_mass3d_points", None)
        if fm3 is not None and len(fm3) > 0:
            off = np.full(len(spec.fluid_nr), -1, dtype=np.int64)
            nr = np.asarray(spec.fluid_nr, dtype=np.int64)
            cum = 0
            for p in np.asarray(fm3, dtype=np.int64):
                off[p] = cum
                cum += nr[p]
            d.fmass3d_off = self._k(off)
            d.fmass3d = self._k(np.ascontiguousarray(spec.fluid_mass3d,
                                                     dtype=np.float64))
        c3 = getattr(spec, "sf3d_pairs", None)
        if c3 is not None and len(c3) > 0:
            off = np.full(len(spec.sf_solid), -1, dtype=np.int64)
            nr = np.asarray(spec.solid_nr, dtype=np.int64)
            cum = 0
            for k in np.asarray(c3, dtype=np.int64):
                off[k] = cum
                cum += nr[spec.sf_solid[k]]
            d.sf3d_off = self._k(off)
            d.sf3d_normal = self._k(np.ascontiguousarray(
                spec.sf3d_normal, dtype=np.float64))
        # absorbing boundaries (preloop.add_absorbing)
        for key, cname in (("clayton_solid", "clayton_s"),
                           ("clayton_fluid", "clayton_f"),
                           ("sponge_solid", "sponge_s"),
                           ("sponge_fluid", "sponge_f")):
            b = getattr(spec, key, None)
            if b is None or len(b["point"]) == 0:
                continue
            setattr(d, "n_" + cname, len(b["point"]))
            setattr(d, cname + "_point", self._k(np.ascontiguousarray(
                b["point"], dtype=np.int32)))
            setattr(d, cname + "_nval", self._k(np.ascontiguousarray(
                b["nval"], dtype=np.int32)))
            setattr(d, cname + "_off", self._k(np.ascontiguousarray(
                b["off"], dtype=np.int64)))
            setattr(d, cname + "_data", self._k(np.ascontiguousarray(
                b["data"], dtype=np.float64)))
        # ocean load (preloop.add_ocean_load)
        oc = getattr(spec, "ocean_load", None)
        if oc is not None and len(oc["point"]) > 0:
            off = np.full(len(spec.solid_nr), -1, dtype=np.int64)
            nv = np.zeros(len(spec.solid_nr), dtype=np.int32)
            off[oc["point"]] = oc["off"]
            nv[oc["point"]] = oc["nval"]
            d.ocean_off = self._k(off)
            d.ocean_nval = self._k(nv)
            d.ocean_data = self._k(np.ascontiguousarray(oc["data"],
                                                        dtype=np.float64))
        # undulated geometry (preloop.add_undulation)
        prt = getattr(spec, "prt", None)
        if prt is not None and len(prt["elems"]) > 0:
            off = np.full(len(spec.elem_points), -1, dtype=np.int64)
            off[np.asarray(prt["elems"])] = np.arange(len(prt["elems"])) * 125
            d.prt_off = self._k(off)
            d.prt_data = self._k(np.ascontiguousarray(prt["x"],
                                                      dtype=np.float64))
        d.dt = float(spec.dt)
        self.d = d
        self._att_state = None
        if d.nsls > 0:
            self._att_state = lib().oracle_att_alloc(C.byref(d))
            d.att_state = self._att_state
        z = lambda: np.zeros((self.nmodes, 3), dtype=np.complex128)
        self.displ, self.veloc, self.accel, self.stiff = z(), z(), z(), z()
        zf = lambda: np.zeros(self.fnmodes, dtype=np.complex128)
        self.fdispl, self.fveloc, self.faccel, self.fstiff = (zf(), zf(),
                                                              zf(), zf())
        self.step = 0

    def _k(self, a):
        self._keep.append(a)
        return a.ctypes.data

    def __del__(self):
        try:
            if self._att_state:
                lib().oracle_att_free(self._att_state, self.d.n_elem)
                self._att_state = None
        except Exception:
            pass

    def run(self, nsteps, record=True):
        nrec = self.d.n_rec
        seis = np.zeros((nsteps, max(nrec, 1), 3)) if record else None
        if getattr(self.spec, "time_scheme", "newmark") == "symplectic":
            f = self.fnmodes > 0
            rc = lib().oracle_run_symplectic(
                C.byref(self.d), self.step, nsteps, self.displ.ctypes.data,
                self.veloc.ctypes.data, self.stiff.ctypes.data,
                self.fdispl.ctypes.data if f else None,
                self.fveloc.ctypes.data if f else None,
                self.fstiff.ctypes.data if f else None, _ptr(seis))
        elif self.fnmodes > 0:
            rc = lib().oracle_run_coupled(
                C.byref(self.d), self.step, nsteps, self.displ.ctypes.data,
                self.veloc.ctypes.data, self.accel.ctypes.data,
                self.stiff.ctypes.data, self.fdispl.ctypes.data,
                self.fveloc.ctypes.data, self.faccel.ctypes.data,
                self.fstiff.ctypes.data, _ptr(seis))
        else:
            rc = lib().oracle_run(C.byref(self.d), self.step, nsteps,
                                  self.displ.ctypes.data,
                                  self.veloc.ctypes.data,
                                  self.accel.ctypes.data,
                                  self.stiff.ctypes.data, _ptr(seis))
        if rc != 0:
            raise RuntimeError("Domain::checkStability || Simulation has "
                               "blown up (oracle).")
        self.step += nsteps
        return seis[:, :nrec] if record else None

    def element_stiff(self, e, displ):
        """Stiffness of one element (subtracted into a zero array)."""
        displ = np.ascontiguousarray(displ, dtype=np.complex128)
        out = np.zeros_like(displ)
        lib().oracle_solid_element_stiff(C.byref(self.d), int(e),
                                         displ.ctypes.data, out.ctypes.data)
        return out

    def stiffness(self, displ):
        """Assembled stiffness of all solid elements."""
        displ = np.ascontiguousarray(displ, dtype=np.complex128)
        out = np.zeros_like(displ)
        for e in range(self.d.n_elem):
            lib().oracle_solid_element_stiff(C.byref(self.d), e,
                                             displ.ctypes.data,
                                             out.ctypes.data)
        return out

    def fluid_stiffness(self, chi):
        """Assembled stiffness of all fluid elements."""
        chi = np.ascontiguousarray(chi, dtype=np.complex128)
        out = np.zeros_like(chi)
        for e in range(self.d.n_felem):
            lib().oracle_fluid_element_stiff(C.byref(self.d), e,
                                             chi.ctypes.data, out.ctypes.data)
        return out

    def coupling(self, displ, stiff, fstiff):
        """Solid-fluid coupling on (solid stiff, fluid stiff)."""
        displ = np.ascontiguousarray(displ, dtype=np.complex128)
        stiff = np.ascontiguousarray(stiff, dtype=np.complex128).copy()
        fstiff = np.ascontiguousarray(fstiff, dtype=np.complex128).copy()
        lib().oracle_coupling(C.byref(self.d), displ.ctypes.data,
                              stiff.ctypes.data, fstiff.ctypes.data)
        return stiff, fstiff

    # ------------------------------------------------- phase level
    def apply_source(self, stiff, tstep):
        stiff = np.ascontiguousarray(stiff, dtype=np.complex128).copy()
        lib().oracle_apply_moment_source(C.byref(self.d),
                                         float(self.spec.stf[tstep]),
                                         stiff.ctypes.data)
        return stiff

    def axial(self, stiff):
        stiff = np.ascontiguousarray(stiff, dtype=np.complex128).copy()
        lib().oracle_axial_solid(C.byref(self.d), stiff.ctypes.data)
        return stiff

    def stiff_to_accel(self, stiff):
        stiff = np.ascontiguousarray(stiff, dtype=np.complex128).copy()
        lib().oracle_stiff_to_accel(C.byref(self.d), stiff.ctypes.data)
        return stiff

    def fluid_stiff_to_accel(self, fstiff):
        fstiff = np.ascontiguousarray(fstiff, dtype=np.complex128).copy()
        lib().oracle_fluid_stiff_to_accel(C.byref(self.d), fstiff.ctypes.data)
        return fstiff

    def clayton(self, veloc, stiff, fveloc=None, fstiff=None):
        """Clayton-Enquist boundaries on copies of (solid stiff, fluid
        stiff) from the velocities."""
        veloc = np.ascontiguousarray(veloc, dtype=np.complex128)
        stiff = np.ascontiguousarray(stiff, dtype=np.complex128).copy()
        fv = None if fveloc is None else np.ascontiguousarray(
            fveloc, dtype=np.complex128)
        fs = None if fstiff is None else np.ascontiguousarray(
            fstiff, dtype=np.complex128).copy()
        lib().oracle_clayton(C.byref(self.d), veloc.ctypes.data,
                             stiff.ctypes.data, _ptr(fv), _ptr(fs))
        return stiff, fs

    def sponge(self, displ, veloc, accel, fdispl=None, fveloc=None,
               faccel=None):
        """Kosloff sponge on copies of (solid accel, fluid accel)."""
        c = lambda a: None if a is None else np.ascontiguousarray(
            a, dtype=np.complex128)
        displ, veloc = c(displ), c(veloc)
        accel = c(accel).copy()
        fd, fv = c(fdispl), c(fveloc)
        fa = None if faccel is None else c(faccel).copy()
        lib().oracle_sponge(C.byref(self.d), displ.ctypes.data,
                            veloc.ctypes.data, accel.ctypes.data, _ptr(fd),
                            _ptr(fv), _ptr(fa))
        return accel, fa

    def newmark(self, stiff, displ, veloc, accel):
        """Returns updated (stiff, displ, veloc, accel) copies."""
        arrs = [np.ascontiguousarray(a, dtype=np.complex128).copy()
                for a in (stiff, displ, veloc, accel)]
        lib().oracle_newmark(self.nmodes, float(self.d.dt),
                             *[a.ctypes.data for a in arrs])
        return arrs

    def record(self, displ):
        displ = np.ascontiguousarray(displ, dtype=np.complex128)
        out = np.zeros((max(self.d.n_rec, 1), 3))
        lib().oracle_record(C.byref(self.d), displ.ctypes.data,
                            out.ctypes.data)
        return out[:self.d.n_rec]

    def att_apply_fr(self, e, alpha, strain, stress):
        """Attenuation of one Fourier order of a 1D element (strain,
        stress: (6, 25) complex)."""
        strain = np.ascontiguousarray(strain, dtype=np.complex128)
        stress = np.ascontiguousarray(stress, dtype=np.complex128).copy()
        lib().oracle_att_apply_fr(C.byref(self.d), int(e), int(alpha),
                                  strain.ctypes.data, stress.ctypes.data)
        return stress

    def att_apply_cd(self, e, slice_, strain, stress):
        """Attenuation of one slice (strain, stress: (6, 25) real)."""
        strain = np.ascontiguousarray(strain, dtype=np.float64)
        stress = np.ascontiguousarray(stress, dtype=np.float64).copy()
        lib().oracle_att_apply_cd(C.byref(self.d), int(e), int(slice_),
                                  strain.ctypes.data, stress.ctypes.data)
        return stress
