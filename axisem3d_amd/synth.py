"""Streaming preloop for large 3D workloads.

The per-slice 3D isotropic model of a large domain (tens of millions of
slices) is generated in chunks by the C++/OpenMP generator
``libaxs3d_synth.so`` (csrc/synth_model.cpp, the same formulas as
``preloop.make_3d_isotropic``) and uploaded to the device chunk by chunk
through ``axs3d_set_solid_iso3d`` / ``axs3d_set_attenuation3d``, so the
host never holds the whole 3D model.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from . import preloop

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libaxs3d_synth.so")
SRC = os.path.join(HERE, "csrc", "synth_model.cpp")

_lib = None


def build():
    tmp = LIB + ".tmp"
    subprocess.check_call(["g++", "-O3", "-fopenmp", "-fPIC", "-shared",
                           "-std=c++17", SRC, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("libaxs3d_synth.so is not built")
        L = C.CDLL(LIB)
        p = C.c_void_p
        L.axs3d_synth_iso3d.argtypes = [
            C.c_int, p, p, p, p, C.c_int, p, C.c_double, C.c_double,
            C.c_int, p, p, p, C.c_double, C.c_double, C.c_double, p, p, p]
        L.axs3d_synth_iso3d.restype = C.c_int
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def generate(spec, e0, e1, field, att_in=None, consts=None, wcg4=None,
             vmax_only=False):
    """3D operands of elements [e0, e1): (prop [S][2][25], att [S][2][npa]
    or None, vmax [e1-e0])."""
    nr = _c(spec.elem_nr[e0:e1], np.int32)
    ns = int(nr.astype(np.int64).sum())
    sz = _c(spec._gq["sz"][e0:e1], np.float64)
    lm = _c(spec.elem_iso[e0:e1], np.float64)
    vp = _c(spec._vp1d[e0:e1], np.float64)
    feat = field.features
    prop = np.empty((ns, 2, 25))
    vmax = np.empty(e1 - e0)
    att_mode, att = -1, None
    qk = qm = w4 = None
    ysum = w10f = fact = 0.0
    if att_in is not None and not vmax_only:
        att_mode = 1 if att_in["cg4"] else 0
        npa = 4 if att_in["cg4"] else 25
        att = np.empty((ns, 2, npa))
        qk = _c(att_in["qkp"][e0:e1], np.float64)
        qm = _c(att_in["qmu"][e0:e1], np.float64)
        w4 = _c(wcg4[e0:e1], np.float64)
        ysum, w10f, fact = consts
    ptr = lambda a: None if a is None else a.ctypes.data
    lib().axs3d_synth_iso3d(
        e1 - e0, nr.ctypes.data, sz.ctypes.data, lm.ctypes.data,
        vp.ctypes.data, feat.shape[0], feat.ctypes.data, field.amp,
        field.clip, att_mode, ptr(qk), ptr(qm), ptr(w4), ysum, w10f, fact,
        prop.ctypes.data, ptr(att), vmax.ctypes.data)
    return prop, att, vmax


def element_hmin(sz):
    """Per-element minimum distance between neighbouring GLL points
    (Quad::computeDt, Quad.cpp:306-325)."""
    ipol = np.repeat(np.arange(5), 5)
    jpol = np.tile(np.arange(5), 5)
    hmin = np.full(sz.shape[0], np.inf)
    for p0 in range(24):
        for p1 in range(p0 + 1, 25):
            if abs(ipol[p1] - ipol[p0]) <= 1 and abs(jpol[p1] - jpol[p0]) <= 1:
                hmin = np.minimum(hmin, np.hypot(sz[:, 0, p1] - sz[:, 0, p0],
                                                 sz[:, 1, p1] - sz[:, 1, p0]))
    return hmin


def prepare_3d_streaming(spec, chunk=4096):
    """Streaming counterpart of ``spec._build_3d()``: fixes dt from the 3D
    velocities (evaluating the 3D model only on the elements that can hold
    the minimum), sets the attenuation coefficients, and installs
    ``spec._upload_3d`` which generates and uploads the 3D operands in
    chunks."""
    field, att_in = spec._field, spec._att_in
    hmin = element_hmin(spec._gq["sz"])
    vp1max = spec._vp1d.max(axis=1)
    clip = field.clip
    # dt_e = hmin / vmax_e with vp1max <= vmax_e <= vp1max (1 + clip)
    upper = hmin / vp1max
    lower = hmin / (vp1max * (1.0 + clip))
    cand = np.nonzero(lower <= upper.min())[0]
    dt_min = np.inf
    for e in cand:
        _, _, vm = generate(spec, int(e), int(e) + 1, field, vmax_only=True)
        dt_min = min(dt_min, hmin[e] / vm[0])
    spec.dt = float("%.3e" % (dt_min * spec._courant))
    spec.elem_kind = np.ones(spec.n_solid_elements, dtype=np.uint8)
    spec.elem_slice_off = None
    spec.prop3d = None
    consts = wcg4 = None
    if att_in is not None:
        w, y = preloop.att_sls_fit(att_in["nsls"], att_in["f_min"],
                                   att_in["f_max"])
        consts = preloop.att_constants(w, y, att_in["f_min"],
                                       att_in["f_max"], att_in["f_ref"])
        wcg4 = preloop.cg4_weights(spec._gq["ifact"])
        a, b, g = preloop.att_alpha_beta_gamma(w, y, spec.dt)
        spec.att = {"nsls": att_in["nsls"], "cg4": int(att_in["cg4"]),
                    "w": w, "y": y, "alpha": a, "beta": b, "gamma": g,
                    "d3d": None, "d1d": None}
    else:
        spec.att = None

    def upload(solver):
        L, ctx = solver.lib, solver.ctx
        E = spec.n_solid_elements
        for e0 in range(0, E, chunk):
            e1 = min(E, e0 + chunk)
            prop, att, _ = generate(spec, e0, e1, field, att_in, consts,
                                    wcg4)
            solver._chk(L.axs3d_set_solid_iso3d(ctx, e0, e1,
                                                prop.ctypes.data),
                        "set_iso3d")
            if att is not None:
                solver._chk(L.axs3d_set_attenuation3d(ctx, e0, e1,
                                                      att.ctypes.data),
                            "set_attenuation3d")

    spec._upload_3d = upload
    return spec
