// gpu_domain.hpp -- the seam a C++ host (the reference's Domain) keeps
// between its time loop and libaxs3d.so: one axs3d_ctx per Domain, one
// method per Domain phase under the reference's names (src/core/domain/
// Domain.hpp:160-207), each a single C-ABI call that throws the library's
// "Class::method || detail" message on failure.  The time loop below is
// the order of NewmarkTimeScheme::solve (src/core/time/
// NewmarkTimeScheme.cpp:41-108).
#ifndef AXS3D_GPU_DOMAIN_HPP
#define AXS3D_GPU_DOMAIN_HPP

#include <stdexcept>
#include <string>

#include "axs3d.h"

class GpuDomain {
public:
    // device, AXS3D_FP64 or AXS3D_FP32_FP64
    GpuDomain(int device, int precision) {
        if (axs3d_ctx_create(device, precision, &mCtx) != 0 || !mCtx)
            throw std::runtime_error("GpuDomain || axs3d_ctx_create failed");
    }
    ~GpuDomain() { axs3d_ctx_destroy(mCtx); }
    GpuDomain(const GpuDomain &) = delete;
    GpuDomain &operator=(const GpuDomain &) = delete;

    // the context, for the set-up calls made where the reference builds
    // its points, elements, sources and stations (GLLPoint::release,
    // Quad::release, Source/Station builders)
    axs3d_ctx *context() const { return mCtx; }

    // throws with the library's message unless rc == 0
    void check(int rc, const char *what) const {
        if (rc != 0)
            throw std::runtime_error(std::string(what) + ": " +
                                     axs3d_last_error(mCtx));
    }

    void applySources(int tstep) const {
        check(axs3d_apply_sources(mCtx, tstep), "applySources");
    }
    void computeStiffness() const {
        check(axs3d_compute_stiffness(mCtx), "computeStiffness");
    }
    void applyBC_BeforeAssemblingStiff() const {
        check(axs3d_apply_bc_before(mCtx), "applyBC_BeforeAssemblingStiff");
    }
    void recordOutput(int isample) const {
        check(axs3d_record(mCtx, isample), "recordOutput");
    }
    void applyBC_AfterAssemblingStiff() const {
        check(axs3d_apply_bc_after(mCtx), "applyBC_AfterAssemblingStiff");
    }
    void computeStiffToAccel() const {
        check(axs3d_stiff_to_accel(mCtx), "computeStiffToAccel");
    }
    void applyBC_AfterComputingAccel() const {
        check(axs3d_apply_bc_after_accel(mCtx),
              "applyBC_AfterComputingAccel");
    }
    void checkStability(int tstep) const {
        check(axs3d_check_stability(mCtx, tstep), "checkStability");
    }
    // NewmarkTimeScheme::update (NewmarkTimeScheme.hpp:48-71)
    void newmarkUpdate() const { check(axs3d_newmark(mCtx), "newmark"); }

private:
    axs3d_ctx *mCtx = nullptr;
};

// NewmarkTimeScheme::solve with every step recorded and the stability
// check every `stability_interval` steps (0: never), naming step tstep + 1
// as the reference does (NewmarkTimeScheme.cpp:72-74).
inline void solveNewmark(const GpuDomain &domain, int nsteps,
                         int stability_interval) {
    for (int tstep = 0; tstep < nsteps; tstep++) {
        domain.applySources(tstep);
        domain.computeStiffness();
        domain.applyBC_BeforeAssemblingStiff();
        domain.recordOutput(tstep);
        if (stability_interval > 0 && tstep % stability_interval == 0)
            domain.checkStability(tstep + 1);
        domain.applyBC_AfterAssemblingStiff();
        domain.computeStiffToAccel();
        domain.applyBC_AfterComputingAccel();
        domain.newmarkUpdate();
    }
}

#endif
