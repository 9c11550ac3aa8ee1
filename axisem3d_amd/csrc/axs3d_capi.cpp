// axs3d_capi.cpp -- extern "C" entry points declared in include/axs3d.h.
#include <exception>
#include <memory>
#include <string>

#include "axs3d.h"
#include "axs3d_domain.hpp"
#include "axs3d_fft.hpp"

struct axs3d_ctx {
    std::unique_ptr<axs3d::DomainBase> dom;
    std::string err;
    int device = 0;
};

namespace {

template <typename F>
int guarded(axs3d_ctx *ctx, F &&f) {
    if (!ctx) return 2;
    try {
        (void)hipSetDevice(ctx->device);
        // a kernel launch that the runtime rejects (grid, LDS) leaves its
        // error here only: clear it first, check it after the call
        (void)hipGetLastError();
        f(*ctx->dom);
        const hipError_t le = hipGetLastError();
        if (le != hipSuccess)
            throw axs3d::Error(std::string("axs3d || HIP launch error: ") +
                               hipGetErrorString(le));
        ctx->err.clear();
        return 0;
    } catch (const std::exception &e) {
        ctx->err = e.what();
        return 1;
    } catch (...) {
        ctx->err = "axs3d || Unknown exception.";
        return 1;
    }
}


// Set-up calls: the device buffers, the captured step graph and the adjoint
// caches are built once by finalize, so every set-up call after it (and a
// second finalize) is refused instead of leaving them stale.
template <typename F>
int setup(axs3d_ctx *ctx, const char *where, F &&f) {
    return guarded(ctx, [&](axs3d::DomainBase &d) {
        if (d.finalized())
            throw axs3d::Error(std::string(where) +
                               " || Domain is already finalized.");
        f(d);
    });
}

}  // namespace

extern "C" {

int axs3d_ctx_create(int device, int precision, axs3d_ctx **out) {
    if (!out) return 2;
    *out = nullptr;
    try {
        auto ctx = std::make_unique<axs3d_ctx>();
        ctx->device = device;
        ctx->dom = axs3d::make_domain(device, precision);
        *out = ctx.release();
        return 0;
    } catch (...) {
        return 1;
    }
}

int axs3d_ctx_destroy(axs3d_ctx *ctx) {
    if (!ctx) return 0;
    (void)hipSetDevice(ctx->device);
    delete ctx;
    return 0;
}

const char *axs3d_last_error(const axs3d_ctx *ctx) {
    return ctx ? ctx->err.c_str() : "axs3d || Null context.";
}

int axs3d_precision_of(const axs3d_ctx *ctx) {
    return ctx ? ctx->dom->precision() : -1;
}

int axs3d_set_gmat(axs3d_ctx *ctx, const double *g_gll, const double *g_glj) {
    return setup(ctx, "Domain::set_gmat", [&](axs3d::DomainBase &d) { d.setGMat(g_gll, g_glj); });
}

int axs3d_set_solid_points(axs3d_ctx *ctx, int n, const int32_t *nr,
                           const double *mass) {
    return setup(ctx, "Domain::set_solid_points",
                   [&](axs3d::DomainBase &d) { d.setSolidPoints(n, nr, mass); });
}

int axs3d_set_solid_elements(axs3d_ctx *ctx, int n, const int32_t *points,
                             const uint8_t *axial, const double *gq) {
    return setup(ctx, "Domain::set_solid_elements", [&](axs3d::DomainBase &d) {
        d.setSolidElements(n, points, axial, gq);
    });
}

int axs3d_set_solid_iso1d(axs3d_ctx *ctx, const double *lam_mu) {
    return setup(ctx, "Domain::set_solid_iso1d", [&](axs3d::DomainBase &d) { d.setSolidIso1D(lam_mu); });
}

int axs3d_set_element_kinds(axs3d_ctx *ctx, const uint8_t *kind) {
    return setup(ctx, "Domain::set_element_kinds", [&](axs3d::DomainBase &d) { d.setElementKinds(kind); });
}

int axs3d_set_solid_iso3d(axs3d_ctx *ctx, int e_begin, int e_end,
                          const double *lam_mu) {
    return setup(ctx, "Domain::set_solid_iso3d", [&](axs3d::DomainBase &d) {
        d.setSolidIso3D(e_begin, e_end, lam_mu);
    });
}

int axs3d_set_attenuation(axs3d_ctx *ctx, int nsls, int cg4,
                          const double *alpha, const double *beta,
                          const double *gamma) {
    return setup(ctx, "Domain::set_attenuation", [&](axs3d::DomainBase &d) {
        d.setAttenuation(nsls, cg4, alpha, beta, gamma);
    });
}

int axs3d_set_attenuation3d(axs3d_ctx *ctx, int e_begin, int e_end,
                            const double *dlam_dmu) {
    return setup(ctx, "Domain::set_attenuation3d", [&](axs3d::DomainBase &d) {
        d.setAttenuation3D(e_begin, e_end, dlam_dmu);
    });
}

int axs3d_set_fluid_points(axs3d_ctx *ctx, int n, const int32_t *nr,
                           const double *mass) {
    return setup(ctx, "Domain::set_fluid_points", [&](axs3d::DomainBase &d) {
        d.setFluidPoints(n, nr, mass);
    });
}

int axs3d_set_fluid_elements(axs3d_ctx *ctx, int n, const int32_t *points,
                             const uint8_t *axial, const double *gq) {
    return setup(ctx, "Domain::set_fluid_elements", [&](axs3d::DomainBase &d) {
        d.setFluidElements(n, points, axial, gq);
    });
}

int axs3d_set_fluid_k1d(axs3d_ctx *ctx, const double *k) {
    return setup(ctx, "Domain::set_fluid_k1d", [&](axs3d::DomainBase &d) { d.setFluidK1D(k); });
}

int axs3d_set_fluid_axial_points(axs3d_ctx *ctx, int n,
                                 const int32_t *points) {
    return setup(ctx, "Domain::set_fluid_axial_points", [&](axs3d::DomainBase &d) {
        d.setFluidAxialPoints(n, points);
    });
}

int axs3d_set_fluid_surface_points(axs3d_ctx *ctx, int n,
                                   const int32_t *points) {
    return setup(ctx, "Domain::set_fluid_surface_points", [&](axs3d::DomainBase &d) {
        d.setFluidSurfacePoints(n, points);
    });
}

int axs3d_set_solid_fluid_coupling(axs3d_ctx *ctx, int n,
                                   const int32_t *solid_points,
                                   const int32_t *fluid_points,
                                   const double *normals) {
    return setup(ctx, "Domain::set_solid_fluid_coupling", [&](axs3d::DomainBase &d) {
        d.setSolidFluidCoupling(n, solid_points, fluid_points, normals);
    });
}

int axs3d_num_fluid_modes(const axs3d_ctx *ctx, int64_t *nmodes) {
    if (!ctx || !nmodes) return 2;
    *nmodes = ctx->dom->numFluidModes();
    return 0;
}

int axs3d_set_material_types(axs3d_ctx *ctx, const uint8_t *mtype) {
    return setup(ctx, "Domain::set_material_types", [&](axs3d::DomainBase &d) {
        d.setMaterialTypes(mtype);
    });
}

int axs3d_set_solid_material1d(axs3d_ctx *ctx, const double *params) {
    return setup(ctx, "Domain::set_solid_material1d", [&](axs3d::DomainBase &d) {
        d.setSolidMaterial1D(params);
    });
}

int axs3d_set_solid_material3d(axs3d_ctx *ctx, int e_begin, int e_end,
                               const double *params) {
    return setup(ctx, "Domain::set_solid_material3d", [&](axs3d::DomainBase &d) {
        d.setSolidMaterial3D(e_begin, e_end, params);
    });
}

int axs3d_set_rtz_theta(axs3d_ctx *ctx, const double *theta) {
    return setup(ctx, "Domain::set_rtz_theta", [&](axs3d::DomainBase &d) { d.setRtzTheta(theta); });
}

int axs3d_set_attenuation1d(axs3d_ctx *ctx, const double *dlam_dmu) {
    return setup(ctx, "Domain::set_attenuation1d", [&](axs3d::DomainBase &d) {
        d.setAttenuation1D(dlam_dmu);
    });
}

int axs3d_set_axial_points(axs3d_ctx *ctx, int n, const int32_t *points) {
    return setup(ctx, "Domain::set_axial_points",
                   [&](axs3d::DomainBase &d) { d.setAxialPoints(n, points); });
}

int axs3d_set_fluid_kinds(axs3d_ctx *ctx, const uint8_t *kind) {
    return setup(ctx, "Domain::set_fluid_kinds",
                   [&](axs3d::DomainBase &d) { d.setFluidKinds(kind); });
}

int axs3d_set_fluid_k3d(axs3d_ctx *ctx, int e_begin, int e_end,
                        const double *k) {
    return setup(ctx, "Domain::set_fluid_k3d", [&](axs3d::DomainBase &d) {
        d.setFluidK3D(e_begin, e_end, k);
    });
}

int axs3d_set_solid_mass3d(axs3d_ctx *ctx, int n, const int32_t *points,
                           const double *mass) {
    return setup(ctx, "Domain::set_solid_mass3d", [&](axs3d::DomainBase &d) {
        d.setSolidMass3D(n, points, mass);
    });
}

int axs3d_set_ocean_load(axs3d_ctx *ctx, int n, const int32_t *points,
                         const int32_t *nval, const double *values) {
    return setup(ctx, "Domain::set_ocean_load", [&](axs3d::DomainBase &d) {
        d.setOceanLoad(n, points, nval, values);
    });
}

int axs3d_set_solid_prt(axs3d_ctx *ctx, int n, const int32_t *elems,
                        const double *x) {
    return setup(ctx, "Domain::set_solid_prt", [&](axs3d::DomainBase &d) {
        d.setSolidPRT(n, elems, x);
    });
}

int axs3d_set_moment_source(axs3d_ctx *ctx, int elem, int nu1,
                            const double *pattern, int nsteps,
                            const double *stf) {
    return setup(ctx, "Domain::set_moment_source", [&](axs3d::DomainBase &d) {
        d.setMomentSource(elem, nu1, pattern, nsteps, stf);
    });
}

int axs3d_set_fluid_mass3d(axs3d_ctx *ctx, int n, const int32_t *points,
                           const double *mass) {
    return setup(ctx, "Domain::set_fluid_mass3d", [&](axs3d::DomainBase &d) {
        d.setFluidMass3D(n, points, mass);
    });
}

int axs3d_set_coupling3d(axs3d_ctx *ctx, int n, const int32_t *pairs,
                         const double *n_unassembled,
                         const double *n_assembled) {
    return setup(ctx, "Domain::set_coupling3d", [&](axs3d::DomainBase &d) {
        d.setCoupling3D(n, pairs, n_unassembled, n_assembled);
    });
}

int axs3d_add_moment_source(axs3d_ctx *ctx, int elem, int nu1,
                            const double *pattern, int nsteps,
                            const double *stf) {
    return setup(ctx, "Domain::add_moment_source", [&](axs3d::DomainBase &d) {
        d.addMomentSource(elem, nu1, pattern, nsteps, stf);
    });
}

int axs3d_set_instances(axs3d_ctx *ctx, int n, const int32_t *point_inst,
                        const double *dt) {
    return setup(ctx, "Domain::set_instances", [&](axs3d::DomainBase &d) {
        d.setInstances(n, point_inst, dt);
    });
}

int axs3d_set_receivers(axs3d_ctx *ctx, int n, const int32_t *elem,
                        const double *weights, const double *phi,
                        int max_samples) {
    return setup(ctx, "Domain::set_receivers", [&](axs3d::DomainBase &d) {
        d.setReceivers(n, elem, weights, phi, max_samples);
    });
}

int axs3d_set_receiver_frames(axs3d_ctx *ctx, int n, const int32_t *wcs,
                              const double *theta,
                              const double *back_azimuth, int cartesian) {
    if (n > 0 && (!wcs || !theta || !back_azimuth)) return 2;
    return setup(ctx, "Domain::set_receiver_frames",
                 [&](axs3d::DomainBase &d) {
                     d.setReceiverFrames(n, wcs, theta, back_azimuth,
                                         cartesian);
                 });
}

int axs3d_add_clayton_solid(axs3d_ctx *ctx, int n, const int32_t *points,
                            const double *rho_vp, const double *rho_vs,
                            const double *area, const double *theta) {
    if (n < 0 || (n > 0 && (!points || !rho_vp || !rho_vs || !area ||
                            !theta)))
        return 2;
    return setup(ctx, "Domain::add_clayton_solid", [&](axs3d::DomainBase &d) {
        d.addClaytonSolid(n, points, rho_vp, rho_vs, area, theta);
    });
}

int axs3d_add_clayton_solid3d(axs3d_ctx *ctx, int n, const int32_t *points,
                              const double *rho_vp, const double *rho_vs,
                              const double *area,
                              const double *unit_normal) {
    if (n < 0 || (n > 0 && (!points || !rho_vp || !rho_vs || !area ||
                            !unit_normal)))
        return 2;
    return setup(ctx, "Domain::add_clayton_solid3d",
                 [&](axs3d::DomainBase &d) {
                     d.addClaytonSolid3D(n, points, rho_vp, rho_vs, area,
                                         unit_normal);
                 });
}

int axs3d_add_clayton_fluid(axs3d_ctx *ctx, int n, const int32_t *points,
                            int per_slice, const double *rho_vp,
                            const double *area) {
    if (n < 0 || (n > 0 && (!points || !rho_vp || !area))) return 2;
    return setup(ctx, "Domain::add_clayton_fluid", [&](axs3d::DomainBase &d) {
        d.addClaytonFluid(n, points, per_slice, rho_vp, area);
    });
}

int axs3d_add_sponge(axs3d_ctx *ctx, int fluid, int n, const int32_t *points,
                     int per_slice, const double *gamma) {
    if (n < 0 || (n > 0 && (!points || !gamma))) return 2;
    return setup(ctx, "Domain::add_sponge", [&](axs3d::DomainBase &d) {
        d.addSponge(fluid, n, points, per_slice, gamma);
    });
}

int axs3d_set_dt(axs3d_ctx *ctx, double dt) {
    return setup(ctx, "Domain::set_dt", [&](axs3d::DomainBase &d) { d.setDt(dt); });
}

int axs3d_finalize(axs3d_ctx *ctx) {
    return setup(ctx, "Domain::finalize", [&](axs3d::DomainBase &d) { d.finalize(); });
}

int axs3d_apply_sources(axs3d_ctx *ctx, int tstep) {
    return guarded(ctx, [&](axs3d::DomainBase &d) { d.applySources(tstep); });
}

int axs3d_compute_stiffness(axs3d_ctx *ctx) {
    return guarded(ctx, [&](axs3d::DomainBase &d) { d.computeStiffness(); });
}

int axs3d_compute_stiffness_static(axs3d_ctx *ctx) {
    return guarded(ctx,
                   [&](axs3d::DomainBase &d) { d.computeStiffnessStatic(); });
}

int axs3d_apply_bc_before(axs3d_ctx *ctx) {
    return guarded(ctx, [&](axs3d::DomainBase &d) {
        d.applyBC_BeforeAssemblingStiff();
    });
}

int axs3d_record(axs3d_ctx *ctx, int isample) {
    return guarded(ctx, [&](axs3d::DomainBase &d) { d.recordOutput(isample); });
}

int axs3d_apply_bc_after(axs3d_ctx *ctx) {
    return guarded(ctx, [&](axs3d::DomainBase &d) {
        d.applyBC_AfterAssemblingStiff();
    });
}

int axs3d_set_time_scheme(axs3d_ctx *ctx, int scheme) {
    return setup(ctx, "Domain::set_time_scheme", [&](axs3d::DomainBase &d) {
        d.setTimeScheme(scheme);
    });
}

int axs3d_symplectic_launch(axs3d_ctx *ctx) {
    return guarded(ctx,
                   [&](axs3d::DomainBase &d) { d.symplecticLaunch(); });
}

int axs3d_symplectic_update(axs3d_ctx *ctx, int isub) {
    return guarded(ctx,
                   [&](axs3d::DomainBase &d) { d.symplecticUpdate(isub); });
}

int axs3d_apply_bc_after_accel(axs3d_ctx *ctx) {
    return guarded(ctx, [&](axs3d::DomainBase &d) {
        d.applyBC_AfterComputingAccel();
    });
}

int axs3d_stiff_to_accel(axs3d_ctx *ctx) {
    return guarded(ctx, [&](axs3d::DomainBase &d) { d.computeStiffToAccel(); });
}

int axs3d_newmark(axs3d_ctx *ctx) {
    return guarded(ctx, [&](axs3d::DomainBase &d) { d.newmark(); });
}

int axs3d_check_stability(axs3d_ctx *ctx, int tstep) {
    return guarded(ctx, [&](axs3d::DomainBase &d) { d.checkStability(tstep); });
}

int axs3d_run(axs3d_ctx *ctx, int step_begin, int nsteps) {
    return guarded(ctx,
                   [&](axs3d::DomainBase &d) { d.solve(step_begin, nsteps); });
}

int axs3d_set_graph(axs3d_ctx *ctx, int enable) {
    return guarded(ctx,
                   [&](axs3d::DomainBase &d) { d.useGraph(enable != 0); });
}

int axs3d_adjoint_run(axs3d_ctx *ctx, int nsteps, int checkpoint_interval,
                      const double *djdr, double *kernels,
                      double *kernels3d) {
    return guarded(ctx, [&](axs3d::DomainBase &d) {
        if (!kernels) throw axs3d::Error("Domain::adjoint || Null output.");
        d.adjointRun(nsteps, checkpoint_interval, djdr, kernels, kernels3d);
    });
}

int axs3d_adjoint_run_media(axs3d_ctx *ctx, int nsteps,
                            int checkpoint_interval, const double *djdr,
                            double *kernels_par, double *kernels_rho,
                            double *kernels_par3d) {
    return guarded(ctx, [&](axs3d::DomainBase &d) {
        if (!kernels_par || !kernels_rho)
            throw axs3d::Error("Domain::adjoint || Null output.");
        d.adjointRunMedia(nsteps, checkpoint_interval, djdr, kernels_par,
                          kernels_rho, kernels_par3d);
    });
}

int axs3d_adjoint_density(axs3d_ctx *ctx, const double *psi, const double *u,
                          const double *a, double *kernels) {
    return guarded(ctx, [&](axs3d::DomainBase &d) {
        if (!psi || !u || !a || !kernels)
            throw axs3d::Error("Domain::adjointDensity || Null pointer.");
        d.adjointDensity(psi, u, a, kernels);
    });
}

int axs3d_set_record_origin(axs3d_ctx *ctx, int record_step0) {
    return guarded(ctx, [&](axs3d::DomainBase &d) {
        d.setRecordOrigin(record_step0);
    });
}

int axs3d_read_seismograms(axs3d_ctx *ctx, double *out, int nsamples) {
    return guarded(ctx, [&](axs3d::DomainBase &d) {
        d.readSeismograms(out, nsamples);
    });
}

int axs3d_num_modes(const axs3d_ctx *ctx, int64_t *nmodes) {
    if (!ctx || !nmodes) return 2;
    *nmodes = ctx->dom->numModes();
    return 0;
}

int axs3d_get_field(axs3d_ctx *ctx, int which, double *out) {
    return guarded(ctx, [&](axs3d::DomainBase &d) { d.getField(which, out); });
}

int axs3d_set_field(axs3d_ctx *ctx, int which, const double *in) {
    return guarded(ctx, [&](axs3d::DomainBase &d) { d.setField(which, in); });
}

int axs3d_synchronize(axs3d_ctx *ctx) {
    return guarded(ctx, [&](axs3d::DomainBase &d) { d.synchronize(); });
}

int axs3d_profile_enable(axs3d_ctx *ctx, int enable) {
    return guarded(ctx,
                   [&](axs3d::DomainBase &d) { d.profileEnable(enable != 0); });
}

int axs3d_profile_read(axs3d_ctx *ctx, int kernel, double *total_ms,
                       int64_t *count) {
    return guarded(ctx, [&](axs3d::DomainBase &d) {
        d.profileRead(kernel, total_ms, count);
    });
}

int axs3d_autotune_info(const axs3d_ctx *ctx, int *cardinal_occupancy,
                        int *cardinal_blocks, double *trial_ms3) {
    if (!ctx || !cardinal_occupancy || !cardinal_blocks || !trial_ms3)
        return 2;
    ctx->dom->autotuneInfo(cardinal_occupancy, cardinal_blocks, trial_ms3);
    return 0;
}

int axs3d_fft_check(int nr, int max_radix, int inverse, double *data,
                    int *radices, int *npasses) {
    if (!data || !radices || !npasses || nr < 1 || nr > axs3d::FFT_NR_CAP ||
        max_radix < 2 || max_radix > 16)
        return 2;
    using axs3d::cpx;
    const std::vector<int> rad = axs3d::fft_choose_radices(nr, max_radix);
    if (rad.empty() && nr > 1) return 1;
    std::vector<uint32_t> plan(axs3d::FFT_PLAN_W);
    std::vector<double> tw;
    std::vector<uint16_t> pos;
    axs3d::fft_build_plan(nr, rad, plan.data(), tw, pos);
    std::vector<cpx<double>> twc(tw.size() / 2);
    for (size_t i = 0; i < twc.size(); i++) twc[i] = {tw[2 * i], tw[2 * i + 1]};
    std::vector<cpx<double>> buf(axs3d::padded(nr));
    cpx<double> *x = reinterpret_cast<cpx<double> *>(data);
    if (inverse) {
        for (int k = 0; k < nr; k++) buf[axs3d::pad(k)] = x[k];
        axs3d::fft_host_run<double>(buf.data(), plan.data(), true, twc.data());
        for (int k = 0; k < nr; k++) x[k] = buf[axs3d::pad(pos[k])];
    } else {
        for (int k = 0; k < nr; k++) buf[axs3d::pad(pos[k])] = x[k];
        axs3d::fft_host_run<double>(buf.data(), plan.data(), false,
                                    twc.data());
        for (int k = 0; k < nr; k++) x[k] = buf[axs3d::pad(k)];
    }
    *npasses = (int)rad.size();
    for (size_t i = 0; i < rad.size(); i++) radices[i] = rad[i];
    return 0;
}

int axs3d_autotune_report(const axs3d_ctx *ctx, double *out, int n,
                          int *written) {
    if (!ctx || !out || !written || n < 0) return 2;
    *written = ctx->dom->autotuneReport(out, n);
    return 0;
}

int axs3d_profile_reset(axs3d_ctx *ctx) {
    return guarded(ctx, [&](axs3d::DomainBase &d) { d.profileReset(); });
}

}  // extern "C"
