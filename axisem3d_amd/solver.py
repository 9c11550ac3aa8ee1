"""Host driver of the HIP path: uploads a ``DomainSpec`` through the C ABI
and runs the Newmark loop on the device.

The phase methods keep the names of the reference's ``Domain`` and
``NewmarkTimeScheme`` (``src/core/domain/Domain.cpp:285-345``,
``src/core/time/NewmarkTimeScheme.cpp:18-131``) so stage-level parity tests
read like the reference's loop.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi, spectral

PRECISIONS = {"fp64": capi.AXS3D_FP64, "fp32": capi.AXS3D_FP32_FP64}


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class GpuSolver:
    def __init__(self, spec, precision="fp64", device=0, max_samples=None):
        self.lib = capi.load()
        self.spec = spec
        self.precision = precision
        ctx = C.c_void_p()
        rc = self.lib.axs3d_ctx_create(device, PRECISIONS[precision],
                                       C.byref(ctx))
        if rc != 0 or not ctx.value:
            raise capi.Axs3dError("axs3d_ctx_create failed (no HIP device "
                                  "or invalid precision)")
        self.ctx = ctx
        self._upload(max_samples)
        n = C.c_int64()
        self._chk(self.lib.axs3d_num_modes(self.ctx, C.byref(n)))
        self.nmodes = n.value
        self._chk(self.lib.axs3d_num_fluid_modes(self.ctx, C.byref(n)))
        self.fnmodes = n.value
        self.step = 0

    # ------------------------------------------------------------------
    def _chk(self, rc, what=""):
        capi.check(self.ctx, rc, what)

    def _upload(self, max_samples):
        L, s, ctx = self.lib, self.spec, self.ctx
        t = spectral.tables()
        g_gll, g_glj = _c(t[4], np.float64), _c(t[5], np.float64)
        self._chk(L.axs3d_set_gmat(ctx, g_gll.ctypes.data,
                                   g_glj.ctypes.data), "set_gmat")
        nr = _c(s.solid_nr, np.int32)
        mass = _c(s.solid_mass, np.float64)
        self._chk(L.axs3d_set_solid_points(ctx, len(nr), nr.ctypes.data,
                                           mass.ctypes.data), "set_points")
        ep = _c(s.elem_points, np.int32)
        ax = _c(s.elem_axial, np.uint8)
        gq = _c(s.elem_gq, np.float64)
        self._chk(L.axs3d_set_solid_elements(ctx, ep.shape[0], ep.ctypes.data,
                                             ax.ctypes.data, gq.ctypes.data),
                  "set_elements")
        kind = getattr(s, "elem_kind", None)
        mtype = getattr(s, "elem_mtype", None)
        if mtype is not None:
            mt = _c(mtype, np.uint8)
            self._chk(L.axs3d_set_material_types(ctx, mt.ctypes.data),
                      "set_material_types")
            m1 = _c(s.mat1d, np.float64)
            self._chk(L.axs3d_set_solid_material1d(ctx, m1.ctypes.data),
                      "set_material1d")
        elif kind is None or np.any(np.asarray(kind) == 0):
            iso = _c(s.elem_iso, np.float64)
            self._chk(L.axs3d_set_solid_iso1d(ctx, iso.ctypes.data),
                      "set_iso1d")
        theta = getattr(s, "elem_theta", None)
        if theta is not None:
            th = _c(theta, np.float64)
            self._chk(L.axs3d_set_rtz_theta(ctx, th.ctypes.data),
                      "set_rtz_theta")
        att = getattr(s, "att", None)
        if kind is not None:
            kd = _c(kind, np.uint8)
            self._chk(L.axs3d_set_element_kinds(ctx, kd.ctypes.data),
                      "set_kinds")
        if att is not None:
            a, b, g = (_c(att[k], np.float64)
                       for k in ("alpha", "beta", "gamma"))
            self._chk(L.axs3d_set_attenuation(
                ctx, int(att["nsls"]), int(att["cg4"]), a.ctypes.data,
                b.ctypes.data, g.ctypes.data), "set_attenuation")
            if att.get("d1d") is not None:
                d1 = _c(att["d1d"], np.float64)
                self._chk(L.axs3d_set_attenuation1d(ctx, d1.ctypes.data),
                          "set_attenuation1d")
        if kind is not None:
            if getattr(s, "prop3d", None) is not None:
                p3 = _c(s.prop3d, np.float64)
                fn = (L.axs3d_set_solid_material3d if mtype is not None
                      else L.axs3d_set_solid_iso3d)
                self._chk(fn(ctx, 0, ep.shape[0], p3.ctypes.data),
                          "set_material3d")
                if att is not None and att.get("d3d") is not None:
                    d3 = _c(att["d3d"], np.float64)
                    self._chk(L.axs3d_set_attenuation3d(
                        ctx, 0, ep.shape[0], d3.ctypes.data),
                        "set_attenuation3d")
            elif getattr(s, "_upload_3d", None) is not None:
                # large domains stream their 3D operands in chunks
                s._upload_3d(self)
        if getattr(s, "fluid_nr", None) is not None and len(s.fluid_nr):
            fnr = _c(s.fluid_nr, np.int32)
            fm = _c(s.fluid_mass, np.float64)
            self._chk(L.axs3d_set_fluid_points(ctx, len(fnr), fnr.ctypes.data,
                                               fm.ctypes.data),
                      "set_fluid_points")
            fp = _c(s.fluid_elem_points, np.int32)
            fa = _c(s.fluid_elem_axial, np.uint8)
            fg = _c(s.fluid_elem_gq, np.float64)
            self._chk(L.axs3d_set_fluid_elements(
                ctx, fp.shape[0], fp.ctypes.data, fa.ctypes.data,
                fg.ctypes.data), "set_fluid_elements")
            fk = _c(s.fluid_elem_k, np.float64)
            self._chk(L.axs3d_set_fluid_k1d(ctx, fk.ctypes.data),
                      "set_fluid_k1d")
            fkind = getattr(s, "fluid_elem_kind", None)
            if fkind is not None:
                fkind = _c(fkind, np.uint8)
                self._chk(L.axs3d_set_fluid_kinds(ctx, fkind.ctypes.data),
                          "set_fluid_kinds")
                k3 = _c(s.fluid_k3d, np.float64)
                self._chk(L.axs3d_set_fluid_k3d(ctx, 0, len(fkind),
                                                k3.ctypes.data),
                          "set_fluid_k3d")
            ax = _c(s.fluid_axial_points, np.int32)
            self._chk(L.axs3d_set_fluid_axial_points(ctx, len(ax),
                                                     ax.ctypes.data),
                      "set_fluid_axial")
            su = _c(s.fluid_surface_points, np.int32)
            self._chk(L.axs3d_set_fluid_surface_points(ctx, len(su),
                                                       su.ctypes.data),
                      "set_fluid_surface")
            cs = _c(s.sf_solid, np.int32)
            cf = _c(s.sf_fluid, np.int32)
            nn = _c(np.concatenate([s.sf_normal, s.sf_normal], axis=1),
                    np.float64)
            self._chk(L.axs3d_set_solid_fluid_coupling(
                ctx, len(cs), cs.ctypes.data, cf.ctypes.data, nn.ctypes.data),
                "set_coupling")
            fm3 = getattr(s, "fluid_mass3d_points", None)
            if fm3 is not None and len(fm3) > 0:
                fm3 = _c(fm3, np.int32)
                fmv = _c(s.fluid_mass3d, np.float64)
                self._chk(L.axs3d_set_fluid_mass3d(
                    ctx, len(fm3), fm3.ctypes.data, fmv.ctypes.data),
                    "set_fluid_mass3d")
            c3 = getattr(s, "sf3d_pairs", None)
            if c3 is not None and len(c3) > 0:
                # one rank: the unassembled and assembled normals coincide
                c3 = _c(c3, np.int32)
                n3 = _c(s.sf3d_normal, np.float64)
                self._chk(L.axs3d_set_coupling3d(
                    ctx, len(c3), c3.ctypes.data, n3.ctypes.data,
                    n3.ctypes.data), "set_coupling3d")
        axp = _c(s.axial_points, np.int32)
        self._chk(L.axs3d_set_axial_points(ctx, len(axp), axp.ctypes.data),
                  "set_axial")
        m3p = getattr(s, "mass3d_points", None)
        if m3p is not None and len(m3p) > 0:
            m3p = _c(m3p, np.int32)
            m3 = _c(s.mass3d, np.float64)
            self._chk(L.axs3d_set_solid_mass3d(ctx, len(m3p), m3p.ctypes.data,
                                               m3.ctypes.data), "set_mass3d")
        oc = getattr(s, "ocean_load", None)
        if oc is not None and len(oc["point"]) > 0:
            op = _c(oc["point"], np.int32)
            onv = _c(oc["nval"], np.int32)
            ov = _c(oc["data"], np.float64)
            self._chk(L.axs3d_set_ocean_load(ctx, len(op), op.ctypes.data,
                                             onv.ctypes.data, ov.ctypes.data),
                      "set_ocean_load")
        prt = getattr(s, "prt", None)
        if prt is not None and len(prt["elems"]) > 0:
            pe = _c(prt["elems"], np.int32)
            px = _c(prt["x"], np.float64)
            self._chk(L.axs3d_set_solid_prt(ctx, len(pe), pe.ctypes.data,
                                            px.ctypes.data), "set_solid_prt")
        if s.src_elem >= 0 and s.src_pattern is not None:
            pat = _c(s.src_pattern, np.complex128)
            stf = _c(s.stf, np.float64)
            self._chk(L.axs3d_set_moment_source(
                ctx, int(s.src_elem), pat.shape[0], pat.ctypes.data,
                len(stf), stf.ctypes.data), "set_source")
        # further sources of batched instances
        for (e, pat, stf) in getattr(s, "extra_sources", None) or []:
            pat = _c(pat, np.complex128)
            stf = _c(stf, np.float64)
            self._chk(L.axs3d_add_moment_source(
                ctx, int(e), pat.shape[0], pat.ctypes.data, len(stf),
                stf.ctypes.data), "add_source")
        pinst = getattr(s, "point_instance", None)
        if pinst is not None:
            pinst = _c(pinst, np.int32)
            idt = _c(s.instance_dt, np.float64)
            self._chk(L.axs3d_set_instances(ctx, len(idt), pinst.ctypes.data,
                                            idt.ctypes.data), "set_instances")
        self.nrec = 0 if s.rec_elem is None else len(s.rec_elem)
        if self.nrec:
            ms = int(max_samples if max_samples is not None else
                     max(s.nsteps, 1))
            self.max_samples = ms
            re = _c(s.rec_elem, np.int32)
            rw = _c(s.rec_weights, np.float64)
            rp = _c(s.rec_phi, np.float64)
            self._chk(L.axs3d_set_receivers(ctx, self.nrec, re.ctypes.data,
                                            rw.ctypes.data, rp.ctypes.data,
                                            ms), "set_receivers")
            fr = getattr(s, "rec_frames", None)
            if fr is not None:
                w = _c(fr["wcs"], np.int32)
                th = _c(fr["theta"], np.float64)
                bz = _c(fr["baz"], np.float64)
                self._chk(L.axs3d_set_receiver_frames(
                    ctx, self.nrec, w.ctypes.data, th.ctypes.data,
                    bz.ctypes.data, int(fr["cartesian"])),
                    "set_receiver_frames")
        self._upload_abc(getattr(s, "_abc_abi", None))
        if getattr(s, "time_scheme", "newmark") == "symplectic":
            self._chk(L.axs3d_set_time_scheme(ctx, 1), "set_time_scheme")
        self._chk(L.axs3d_set_dt(ctx, float(s.dt)), "set_dt")
        self._chk(L.axs3d_finalize(ctx), "finalize")

    def _upload_abc(self, abi):
        """Absorbing-boundary entries as the reference constructs them
        (preloop.add_absorbing): Clayton 1D/3D on solid and fluid points,
        sponges."""
        if not abi:
            return
        L, ctx = self.lib, self.ctx
        cs = abi.get("clayton_solid") or []
        if cs:
            pts = _c([e[0] for e in cs], np.int32)
            a = [_c([e[k] for e in cs], np.float64) for k in (1, 2, 3, 4)]
            self._chk(L.axs3d_add_clayton_solid(
                ctx, len(cs), pts.ctypes.data, *[x.ctypes.data for x in a]),
                "add_clayton_solid")
        c3 = abi.get("clayton_solid3d") or []
        if c3:
            pts = _c([e[0] for e in c3], np.int32)
            a = [_c(np.concatenate([np.asarray(e[k]).reshape(-1)
                                    for e in c3]), np.float64)
                 for k in (1, 2, 3, 4)]
            self._chk(L.axs3d_add_clayton_solid3d(
                ctx, len(c3), pts.ctypes.data, *[x.ctypes.data for x in a]),
                "add_clayton_solid3d")
        for key, per in (("clayton_fluid", 0), ("clayton_fluid3d", 1)):
            cf = abi.get(key) or []
            if not cf:
                continue
            pts = _c([e[0] for e in cf], np.int32)
            rp = _c(np.concatenate([np.atleast_1d(e[1]) for e in cf]),
                    np.float64)
            ar = _c(np.concatenate([np.atleast_1d(e[2]) for e in cf]),
                    np.float64)
            self._chk(L.axs3d_add_clayton_fluid(
                ctx, len(cf), pts.ctypes.data, per, rp.ctypes.data,
                ar.ctypes.data), "add_clayton_fluid")
        for fluid, key in ((0, "sponge_solid"), (1, "sponge_fluid")):
            sp = abi.get(key) or {}
            for per in (0, 1):
                ent = [(p, g) for p, g in sp.items()
                       if (len(g) > 1) == bool(per)]
                if not ent:
                    continue
                pts = _c([p for p, _ in ent], np.int32)
                g = _c(np.concatenate([np.asarray(x).reshape(-1)
                                       for _, x in ent]), np.float64)
                self._chk(L.axs3d_add_sponge(ctx, fluid, len(ent),
                                             pts.ctypes.data, per,
                                             g.ctypes.data), "add_sponge")

    def close(self):
        if getattr(self, "ctx", None) is not None and self.ctx.value:
            self.lib.axs3d_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------- bulk loop
    def run(self, nsteps):
        self._chk(self.lib.axs3d_run(self.ctx, self.step, int(nsteps)), "run")
        self.step += nsteps

    def use_graph(self, enable=True):
        """Replay one captured step graph per step in run() (the default),
        or launch every kernel from the host."""
        self._chk(self.lib.axs3d_set_graph(self.ctx, int(bool(enable))),
                  "set_graph")

    def seismograms(self, nsamples=None):
        n = self.step if nsamples is None else nsamples
        out = np.zeros((n, max(self.nrec, 1), 3))
        if self.nrec and n:
            self._chk(self.lib.axs3d_read_seismograms(self.ctx,
                                                      out.ctypes.data, n))
        return out[:, :self.nrec]

    # ------------------------------------------------------- phases
    def applySources(self, tstep):
        self._chk(self.lib.axs3d_apply_sources(self.ctx, tstep))

    def computeStiffness(self):
        self._chk(self.lib.axs3d_compute_stiffness(self.ctx))

    def applyBC_BeforeAssemblingStiff(self):
        self._chk(self.lib.axs3d_apply_bc_before(self.ctx))

    def recordOutput(self, isample):
        self._chk(self.lib.axs3d_record(self.ctx, isample))

    def applyBC_AfterAssemblingStiff(self):
        self._chk(self.lib.axs3d_apply_bc_after(self.ctx))

    def applyBC_AfterComputingAccel(self):
        self._chk(self.lib.axs3d_apply_bc_after_accel(self.ctx))

    def computeStiffToAccel(self):
        self._chk(self.lib.axs3d_stiff_to_accel(self.ctx))

    def newmark(self):
        self._chk(self.lib.axs3d_newmark(self.ctx))

    def symplecticLaunch(self):
        self._chk(self.lib.axs3d_symplectic_launch(self.ctx))

    def symplecticUpdate(self, isub):
        self._chk(self.lib.axs3d_symplectic_update(self.ctx, int(isub)))

    def checkStability(self, tstep=0):
        self._chk(self.lib.axs3d_check_stability(self.ctx, tstep))

    def synchronize(self):
        self._chk(self.lib.axs3d_synchronize(self.ctx))

    # ------------------------------------------------------- fields
    _FIELDS = {"displ": capi.FIELD_DISPL, "veloc": capi.FIELD_VELOC,
               "accel": capi.FIELD_ACCEL, "stiff": capi.FIELD_STIFF,
               "fdispl": capi.FIELD_FLUID_DISPL,
               "fveloc": capi.FIELD_FLUID_VELOC,
               "faccel": capi.FIELD_FLUID_ACCEL,
               "fstiff": capi.FIELD_FLUID_STIFF}

    def _shape(self, name):
        return (self.fnmodes,) if name.startswith("f") else (self.nmodes, 3)

    def get_field(self, name):
        out = np.zeros(self._shape(name), dtype=np.complex128)
        self._chk(self.lib.axs3d_get_field(self.ctx, self._FIELDS[name],
                                           out.ctypes.data))
        return out

    def set_field(self, name, values):
        v = _c(values, np.complex128)
        if v.shape != self._shape(name):
            raise ValueError("field shape mismatch")
        self._chk(self.lib.axs3d_set_field(self.ctx, self._FIELDS[name],
                                           v.ctypes.data))

    # --------------------------------------------------------- adjoint
    def adjoint_run(self, djdr, checkpoint_interval=None):
        """Discrete adjoint of run() from rest over len(djdr) steps
        (axs3d_adjoint_run): djdr[n, r, c] = dJ / d seismogram[n, r, c].
        Returns dJ/d(lambda, mu, rho) per element GLL point, shape
        (n_elements, 25, 3), and, with 3D elements, dJ/d(lambda, mu) per
        slice in the layout of spec.prop3d (else None).  The wavefields
        are left undefined."""
        g = _c(djdr, np.float64)
        nsteps = g.shape[0]
        if checkpoint_interval is None:
            checkpoint_interval = max(1, int(round(np.sqrt(nsteps))))
        kern = np.zeros((len(self.spec.elem_points), 25, 3))
        kind = getattr(self.spec, "elem_kind", None)
        k3 = None
        if kind is not None and np.any(np.asarray(kind) == 1):
            k3 = np.zeros(np.asarray(self.spec.prop3d).shape)
        self._chk(self.lib.axs3d_adjoint_run(
            self.ctx, int(nsteps), int(checkpoint_interval), g.ctypes.data,
            kern.ctypes.data, None if k3 is None else k3.ctypes.data),
            "adjoint_run")
        return kern, k3

    def adjoint_run_media(self, djdr, checkpoint_interval=None):
        """The adjoint run for any solid media (axs3d_adjoint_run_media):
        dJ/d(material parameters) packed as spec.mat1d, dJ/drho per element
        GLL point (n_elements, 25), and with 3D elements the per-slice
        parameters packed as spec.prop3d (else None)."""
        g = _c(djdr, np.float64)
        nsteps = g.shape[0]
        if checkpoint_interval is None:
            checkpoint_interval = max(1, int(round(np.sqrt(nsteps))))
        s = self.spec
        mtype = getattr(s, "elem_mtype", None)
        ne = len(s.elem_points)
        npar = (np.full(ne, 2) if mtype is None else
                np.array([2, 5, 21])[np.asarray(mtype)])
        kp = np.zeros(int(npar.sum()) * 25)
        kr = np.zeros((ne, 25))
        kind = getattr(s, "elem_kind", None)
        k3 = None
        if kind is not None and np.any(np.asarray(kind) == 1):
            k3 = np.zeros(np.asarray(s.prop3d).size)
        self._chk(self.lib.axs3d_adjoint_run_media(
            self.ctx, int(nsteps), int(checkpoint_interval), g.ctypes.data,
            kp.ctypes.data, kr.ctypes.data,
            None if k3 is None else k3.ctypes.data), "adjoint_run_media")
        return kp, kr, k3

    def adjoint_density(self, psi, u, a):
        """Unscaled per-step sensitivity densities (axs3d_adjoint_density)
        of complex (n_modes, 3) fields; shape (n_elements, 25, 3)."""
        psi, u, a = (_c(x, np.complex128) for x in (psi, u, a))
        kern = np.zeros((len(self.spec.elem_points), 25, 3))
        self._chk(self.lib.axs3d_adjoint_density(
            self.ctx, psi.ctypes.data, u.ctypes.data, a.ctypes.data,
            kern.ctypes.data), "adjoint_density")
        return kern

    # ------------------------------------------------------- profiling
    def profile(self, enable=True):
        self._chk(self.lib.axs3d_profile_enable(self.ctx, int(enable)))

    def profile_reset(self):
        self._chk(self.lib.axs3d_profile_reset(self.ctx))

    def profile_read(self, kernel):
        ms = C.c_double()
        cnt = C.c_int64()
        self._chk(self.lib.axs3d_profile_read(self.ctx, kernel, C.byref(ms),
                                              C.byref(cnt)))
        return ms.value, cnt.value

    def autotune_info(self):
        occ, blocks = C.c_int(), C.c_int()
        ms = (C.c_double * 3)()
        self._chk(self.lib.axs3d_autotune_info(self.ctx, C.byref(occ),
                                               C.byref(blocks), ms))
        rep = (C.c_double * 32)()
        nw = C.c_int()
        self._chk(self.lib.axs3d_autotune_report(self.ctx, rep, 32,
                                                 C.byref(nw)))
        r = list(rep[:nw.value])
        return {"cardinal_occupancy": occ.value,
                "cardinal_blocks": blocks.value,
                "cardinal_pass_ms": ms[0],
                "cardinal_scratch_mib": ms[1],
                "scratch_sizes_timed": int(ms[2]),
                "fft_max_radix": int(r[3]),
                "fft_radices_timed": {int(r[5 + 2 * i]): r[6 + 2 * i]
                                      for i in range(int(r[4]))}}
