// axs3d_domain.cpp -- Domain set-up, phases and the Newmark loop.
#include "axs3d_domain.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <sstream>

#include "axs3d_launch.hpp"

namespace axs3d {

void hip_check(hipError_t e, const char *where) {
    if (e != hipSuccess) {
        std::ostringstream ss;
        ss << where << " || HIP error: " << hipGetErrorString(e);
        throw Error(ss.str());
    }
}

// ------------------------------------------------------------ buffers
void DeviceBuffer::alloc(size_t bytes) {
    reset();
    if (bytes == 0) return;
    hip_check(hipMalloc(&ptr_, bytes), "DeviceBuffer::alloc");
    bytes_ = bytes;
}

void DeviceBuffer::upload(const void *src, size_t bytes, hipStream_t st) {
    if (bytes > bytes_)
        throw Error("DeviceBuffer::upload || Size exceeds allocation.");
    if (bytes == 0) return;
    hip_check(hipMemcpyAsync(ptr_, src, bytes, hipMemcpyHostToDevice, st),
              "DeviceBuffer::upload");
    hip_check(hipStreamSynchronize(st), "DeviceBuffer::upload");
}

void DeviceBuffer::zero(hipStream_t st) {
    if (bytes_ == 0) return;
    hip_check(hipMemsetAsync(ptr_, 0, bytes_, st), "DeviceBuffer::zero");
}

void DeviceBuffer::reset() {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    bytes_ = 0;
}

// ------------------------------------------------------------ base
DomainBase::DomainBase(int device, int precision)
    : mDevice(device), mPrecision(precision) {
    hip_check(hipSetDevice(device), "Domain::Domain");
    hip_check(hipStreamCreateWithFlags(&mStream, hipStreamNonBlocking),
              "Domain::Domain");
}

DomainBase::~DomainBase() {
    (void)hipSetDevice(mDevice);
    if (mStream) (void)hipStreamSynchronize(mStream);
    for (auto &p : mPending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : mEventPool) (void)hipEventDestroy(e);
    for (auto e : mOpenA)
        if (e) (void)hipEventDestroy(e);
    if (mStream) (void)hipStreamDestroy(mStream);
}

void DomainBase::requireFinalized(const char *where) const {
    if (!mFinalized)
        throw Error(std::string(where) + " || Domain is not finalized.");
}

void DomainBase::setGMat(const double *g_gll, const double *g_glj) {
    hGll.assign(g_gll, g_gll + NPEM);
    hGlj.assign(g_glj, g_glj + NPEM);
}

void DomainBase::setSolidPoints(int n, const int32_t *nr, const double *mass) {
    if (n <= 0) throw Error("Domain::setSolidPoints || No points.");
    hPointNr.assign(nr, nr + n);
    hMass.assign(mass, mass + n);
    for (int i = 0; i < n; i++) {
        if (nr[i] < 1)
            throw Error("Domain::setSolidPoints || Nr must be positive.");
        if (!(mass[i] > 0.))
            throw Error("Domain::setSolidPoints || Mass must be positive.");
    }
    mNumPoints = n;
}

void DomainBase::setSolidElements(int n, const int32_t *points,
                                  const uint8_t *axial, const double *gq) {
    if (n <= 0) throw Error("Domain::setSolidElements || No elements.");
    hElemPoints.assign(points, points + (size_t)n * NPEM);
    hElemAxial.assign(axial, axial + n);
    hElemGQ.assign(gq, gq + (size_t)n * 6 * NPEM);
    mNumElems = n;
}

void DomainBase::setSolidIso1D(const double *lam_mu) {
    if (mNumElems <= 0)
        throw Error("Domain::setSolidIso1D || Set elements first.");
    hElemIso.assign(lam_mu, lam_mu + (size_t)mNumElems * 2 * NPEM);
}

void DomainBase::setAxialPoints(int n, const int32_t *points) {
    hAxial.assign(points, points + std::max(n, 0));
}

void DomainBase::setMomentSource(int elem, int nu1, const double *pattern,
                                 int nsteps, const double *stf) {
    if (nu1 <= 0 || nsteps <= 0)
        throw Error("Domain::setMomentSource || Empty source.");
    hSrcElem = elem;
    hSrcNu1 = nu1;
    hSrcPattern.assign(pattern, pattern + (size_t)nu1 * 6 * NPEM * 2);
    hStf.assign(stf, stf + nsteps);
}

void DomainBase::setReceivers(int n, const int32_t *elem, const double *w,
                              const double *phi, int max_samples) {
    hRecElem.assign(elem, elem + std::max(n, 0));
    hRecW.assign(w, w + (size_t)std::max(n, 0) * NPEM);
    hRecPhi.assign(phi, phi + std::max(n, 0));
    hMaxSamples = std::max(max_samples, 0);
}

void DomainBase::finalize() {
    hip_check(hipSetDevice(mDevice), "Domain::finalize");
    if (hGll.size() != NPEM || hGlj.size() != NPEM)
        throw Error("Domain::finalize || G matrices are not set.");
    if (mNumPoints <= 0 || mNumElems <= 0)
        throw Error("Domain::finalize || Points or elements are not set.");
    if (hElemIso.size() != (size_t)mNumElems * 2 * NPEM)
        throw Error("Domain::finalize || Material is not set.");
    if (!(mDt > 0.)) throw Error("Domain::finalize || dt is not set.");
    // points: mode offsets and the mode -> point map
    std::vector<uint32_t> off(mNumPoints + 1, 0);
    std::vector<int32_t> nu1(mNumPoints);
    int64_t total = 0;
    for (int p = 0; p < mNumPoints; p++) {
        off[p] = (uint32_t)total;
        nu1[p] = hPointNr[p] / 2 + 1;
        total += nu1[p];
    }
    if (total >= (int64_t)UINT32_MAX / 2)
        throw Error("Domain::finalize || Too many point modes.");
    off[mNumPoints] = (uint32_t)total;
    mNumModes = total;
    std::vector<int32_t> modePoint(total);
    for (int p = 0; p < mNumPoints; p++)
        for (int a = 0; a < nu1[p]; a++) modePoint[off[p] + a] = p;
    std::vector<uint8_t> pAxial(mNumPoints, 0);
    for (int32_t p : hAxial) {
        if (p < 0 || p >= mNumPoints)
            throw Error("Domain::finalize || Axial point out of range.");
        pAxial[p] = 1;
    }
    // elements: Nr = max over points (Element::pointSet, Element.cpp:28-35)
    std::vector<int32_t> eNu1(mNumElems);
    std::vector<EMItem> em;
    for (int e = 0; e < mNumElems; e++) {
        int nrmax = 0;
        for (int k = 0; k < NPEM; k++) {
            int p = hElemPoints[(size_t)e * NPEM + k];
            if (p < 0 || p >= mNumPoints)
                throw Error("Domain::finalize || Element point out of range.");
            nrmax = std::max(nrmax, hPointNr[p]);
        }
        eNu1[e] = nrmax / 2 + 1;
        for (int a = 0; a < eNu1[e]; a++) em.push_back({e, a});
    }
    mNumEM = (int)em.size();
    // source
    if (hSrcElem >= 0) {
        if (hSrcElem >= mNumElems)
            throw Error("Domain::finalize || Source element out of range.");
        mSrcNu1c = std::min(eNu1[hSrcElem], hSrcNu1);
        mStfSteps = (int)hStf.size();
    }
    // receivers
    mNumRec = (int)hRecElem.size();
    for (int32_t e : hRecElem)
        if (e < 0 || e >= mNumElems)
            throw Error("Domain::finalize || Receiver element out of range.");

    hipStream_t st = mStream;
    auto up = [&](DeviceBuffer &b, const void *src, size_t bytes) {
        b.alloc(bytes);
        b.upload(src, bytes, st);
    };
    up(dPointOff, off.data(), off.size() * 4);
    up(dPointNu1, nu1.data(), nu1.size() * 4);
    up(dPointNr, hPointNr.data(), hPointNr.size() * 4);
    up(dPointAxial, pAxial.data(), pAxial.size());
    std::vector<double> inv(mNumPoints);
    for (int p = 0; p < mNumPoints; p++) inv[p] = 1. / hMass[p];
    up(dInvMass, inv.data(), inv.size() * 8);
    up(dModePoint, modePoint.data(), modePoint.size() * 4);
    up(dElemPoints, hElemPoints.data(), hElemPoints.size() * 4);
    up(dElemAxial, hElemAxial.data(), hElemAxial.size());
    up(dElemNu1, eNu1.data(), eNu1.size() * 4);
    up(dEMList, em.data(), em.size() * sizeof(EMItem));
    const size_t fbytes = (size_t)mNumModes * 6 * sizeof(double);
    dDispl.alloc(fbytes);
    dVeloc.alloc(fbytes);
    dAccel.alloc(fbytes);
    dDispl.zero(st);
    dVeloc.zero(st);
    dAccel.zero(st);
    if (hSrcElem >= 0) up(dStf, hStf.data(), hStf.size() * 8);
    if (mNumRec > 0) {
        up(dRecElem, hRecElem.data(), hRecElem.size() * 4);
        up(dRecW, hRecW.data(), hRecW.size() * 8);
        up(dRecPhi, hRecPhi.data(), hRecPhi.size() * 8);
        dSeismo.alloc((size_t)std::max(hMaxSamples, 1) * mNumRec * 3 * 8);
        dSeismo.zero(st);
    }
    dFlag.alloc(sizeof(int));
    dFlag.zero(st);
    uploadReal();
    hip_check(hipStreamSynchronize(st), "Domain::finalize");
    mFinalized = true;
}

void DomainBase::recordOutput(int isample) {
    requireFinalized("Domain::recordOutput");
    if (mNumRec == 0) return;
    if (isample < 0 || isample >= hMaxSamples)
        throw Error("Domain::recordOutput || Sample index out of range.");
    tick(AXS3D_KERNEL_RECORD);
    launch_record(mStream, mNumRec, dRecElem.as<int32_t>(),
                  dRecW.as<double>(), dRecPhi.as<double>(),
                  dElemPoints.as<int32_t>(), dElemNu1.as<int32_t>(),
                  dPointOff.as<uint32_t>(), dPointNu1.as<int32_t>(),
                  dDispl.as<double>(),
                  dSeismo.as<double>() + (size_t)isample * mNumRec * 3);
    tock(AXS3D_KERNEL_RECORD);
}

void DomainBase::checkStability(int tstep) {
    requireFinalized("Domain::checkStability");
    int flag = 0;
    hip_check(hipMemcpyAsync(&flag, dFlag.as<int>(), sizeof(int),
                             hipMemcpyDeviceToHost, mStream),
              "Domain::checkStability");
    hip_check(hipStreamSynchronize(mStream), "Domain::checkStability");
    if (flag) {
        std::ostringstream ss;
        ss << "Domain::checkStability || Simulation has blown up, with dt = "
           << mDt << " s. || When the instability occurred: || "
           << "* current step  =  " << tstep;
        throw Error(ss.str());
    }
}

// NewmarkTimeScheme::solve (NewmarkTimeScheme.cpp:41-108).  Stability is
// checked from the device flag at the end of the batch (the point kernel
// raises it on any non-finite displacement).
void DomainBase::solve(int step_begin, int nsteps) {
    requireFinalized("NewmarkTimeScheme::solve");
    for (int tstep = step_begin; tstep < step_begin + nsteps; tstep++) {
        applySources(tstep);
        computeStiffness();
        applyBC_BeforeAssemblingStiff();
        if (mNumRec > 0) {
            int isample = tstep - mRecordStep0;
            if (isample >= 0 && isample < hMaxSamples) recordOutput(isample);
        }
        fusedPointUpdate();
    }
    checkStability(step_begin + nsteps);
}

void DomainBase::readSeismograms(double *out, int nsamples) {
    requireFinalized("Domain::readSeismograms");
    if (mNumRec == 0 || nsamples <= 0) return;
    if (nsamples > hMaxSamples)
        throw Error("Domain::readSeismograms || Too many samples.");
    hip_check(hipMemcpyAsync(out, dSeismo.as<double>(),
                             (size_t)nsamples * mNumRec * 3 * 8,
                             hipMemcpyDeviceToHost, mStream),
              "Domain::readSeismograms");
    hip_check(hipStreamSynchronize(mStream), "Domain::readSeismograms");
}

void DomainBase::synchronize() {
    hip_check(hipStreamSynchronize(mStream), "Domain::synchronize");
}

// ------------------------------------------------------------ profiling
hipEvent_t DomainBase::takeEvent() {
    if (!mEventPool.empty()) {
        hipEvent_t e = mEventPool.back();
        mEventPool.pop_back();
        return e;
    }
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "Domain::profile");
    return e;
}

void DomainBase::tick(int kernel) {
    if (!mProfile) return;
    hipEvent_t a = takeEvent();
    hip_check(hipEventRecord(a, mStream), "Domain::profile");
    mOpenA[kernel] = a;
}

void DomainBase::tock(int kernel) {
    if (!mProfile || !mOpenA[kernel]) return;
    hipEvent_t b = takeEvent();
    hip_check(hipEventRecord(b, mStream), "Domain::profile");
    mPending.push_back({kernel, mOpenA[kernel], b});
    mOpenA[kernel] = nullptr;
    if (mPending.size() > 4096) drainPending();
}

void DomainBase::drainPending() {
    if (mPending.empty()) return;
    hip_check(hipStreamSynchronize(mStream), "Domain::profile");
    for (auto &p : mPending) {
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, p.a, p.b), "Domain::profile");
        mProfMs[p.kernel] += ms;
        mProfCount[p.kernel] += 1;
        mEventPool.push_back(p.a);
        mEventPool.push_back(p.b);
    }
    mPending.clear();
}

void DomainBase::profileRead(int kernel, double *ms, int64_t *count) {
    if (kernel < 0 || kernel >= AXS3D_NUM_KERNELS)
        throw Error("Domain::profileRead || Unknown kernel id.");
    drainPending();
    *ms = mProfMs[kernel];
    *count = mProfCount[kernel];
}

void DomainBase::profileReset() {
    drainPending();
    for (int k = 0; k < AXS3D_NUM_KERNELS; k++) {
        mProfMs[k] = 0.;
        mProfCount[k] = 0;
    }
}

// ------------------------------------------------------------ precision
template <typename Real>
class Domain final : public DomainBase {
public:
    Domain(int device, int precision) : DomainBase(device, precision) {}

    void applySources(int tstep) override {
        requireFinalized("Domain::applySources");
        if (hSrcElem < 0) return;
        if (tstep < 0 || tstep >= mStfSteps)
            throw Error("Domain::applySources || Step outside the STF.");
        tick(AXS3D_KERNEL_SOURCE);
        launch_moment_source<Real>(
            mStream, hSrcElem, mSrcNu1c, dSrcPattern.as<Real>(),
            dStf.as<double>(), tstep, dElemPoints.as<int32_t>(),
            dElemAxial.as<uint8_t>(), dElemOps.as<Real>(),
            dPointOff.as<uint32_t>(), dPointNu1.as<int32_t>(),
            dStiff.as<Real>(), dGTab.as<Real>());
        tock(AXS3D_KERNEL_SOURCE);
    }

    void computeStiffness() override {
        requireFinalized("Domain::computeStiffness");
        tick(AXS3D_KERNEL_STIFFNESS);
        launch_solid_stiff_fr<Real>(
            mStream, mNumEM, dEMList.as<EMItem>(), dElemPoints.as<int32_t>(),
            dElemAxial.as<uint8_t>(), dElemOps.as<Real>(),
            dElemMat.as<Real>(), dPointOff.as<uint32_t>(),
            dPointNu1.as<int32_t>(), dDispl.as<double>(), dStiff.as<Real>(),
            dGTab.as<Real>());
        tock(AXS3D_KERNEL_STIFFNESS);
    }

    void point(int phases) {
        tick(AXS3D_KERNEL_POINT);
        launch_point<Real>(mStream, phases, mNumModes,
                           dModePoint.as<int32_t>(), dPointOff.as<uint32_t>(),
                           dPointNr.as<int32_t>(), dPointAxial.as<uint8_t>(),
                           dInvMass.as<double>(), dStiff.as<Real>(),
                           dDispl.as<double>(), dVeloc.as<double>(),
                           dAccel.as<double>(), mDt, dFlag.as<int>());
        tock(AXS3D_KERNEL_POINT);
    }
    void applyBC_AfterAssemblingStiff() override {
        requireFinalized("Domain::applyBC_AfterAssemblingStiff");
        point(PHASE_AXIAL);
    }
    void computeStiffToAccel() override {
        requireFinalized("Domain::computeStiffToAccel");
        point(PHASE_MASS);
    }
    void newmark() override {
        requireFinalized("NewmarkTimeScheme::update");
        point(PHASE_MARCH);
    }
    void fusedPointUpdate() override {
        point(PHASE_AXIAL | PHASE_MASS | PHASE_MARCH);
    }

    void getField(int which, double *out) override {
        requireFinalized("Domain::getField");
        const size_t n = (size_t)mNumModes * 6;
        if (which == AXS3D_FIELD_STIFF) {
            std::vector<Real> tmp(n);
            hip_check(hipMemcpyAsync(tmp.data(), dStiff.as<Real>(),
                                     n * sizeof(Real), hipMemcpyDeviceToHost,
                                     mStream), "Domain::getField");
            hip_check(hipStreamSynchronize(mStream), "Domain::getField");
            for (size_t k = 0; k < n; k++) out[k] = (double)tmp[k];
            return;
        }
        hip_check(hipMemcpyAsync(out, fieldPtr(which), n * 8,
                                 hipMemcpyDeviceToHost, mStream),
                  "Domain::getField");
        hip_check(hipStreamSynchronize(mStream), "Domain::getField");
    }

    void setField(int which, const double *in) override {
        requireFinalized("Domain::setField");
        const size_t n = (size_t)mNumModes * 6;
        if (which == AXS3D_FIELD_STIFF) {
            std::vector<Real> tmp(in, in + n);
            hip_check(hipMemcpyAsync(dStiff.as<Real>(), tmp.data(),
                                     n * sizeof(Real), hipMemcpyHostToDevice,
                                     mStream), "Domain::setField");
            hip_check(hipStreamSynchronize(mStream), "Domain::setField");
            return;
        }
        hip_check(hipMemcpyAsync(fieldPtr(which), in, n * 8,
                                 hipMemcpyHostToDevice, mStream),
                  "Domain::setField");
        hip_check(hipStreamSynchronize(mStream), "Domain::setField");
    }

private:
    double *fieldPtr(int which) {
        switch (which) {
            case AXS3D_FIELD_DISPL: return dDispl.as<double>();
            case AXS3D_FIELD_VELOC: return dVeloc.as<double>();
            case AXS3D_FIELD_ACCEL: return dAccel.as<double>();
            default: throw Error("Domain::field || Unknown field id.");
        }
    }

    // Operands cast to Real after their FP64 products, as the reference's
    // GradientQuadrature constructor does (GradientQuadrature.hpp:127-137).
    void uploadReal() override {
        hipStream_t st = mStream;
        std::vector<Real> ops((size_t)mNumElems * ELEM_OPS * NPEM);
        for (int e = 0; e < mNumElems; e++) {
            const double *g = hElemGQ.data() + (size_t)e * 6 * NPEM;
            Real *o = ops.data() + (size_t)e * ELEM_OPS * NPEM;
            for (int k = 0; k < NPEM; k++) {
                const double ifk = g[5 * NPEM + k];
                for (int r = 0; r < 5; r++) {
                    o[r * NPEM + k] = (Real)g[r * NPEM + k];
                    o[(5 + r) * NPEM + k] = (Real)(g[r * NPEM + k] * ifk);
                }
            }
        }
        dElemOps.alloc(ops.size() * sizeof(Real));
        dElemOps.upload(ops.data(), ops.size() * sizeof(Real), st);
        std::vector<Real> mat(hElemIso.begin(), hElemIso.end());
        dElemMat.alloc(mat.size() * sizeof(Real));
        dElemMat.upload(mat.data(), mat.size() * sizeof(Real), st);
        std::vector<Real> gt(2 * NPEM);
        for (int k = 0; k < NPEM; k++) {
            gt[k] = (Real)hGll[k];
            gt[NPEM + k] = (Real)hGlj[k];
        }
        dGTab.alloc(gt.size() * sizeof(Real));
        dGTab.upload(gt.data(), gt.size() * sizeof(Real), st);
        dStiff.alloc((size_t)mNumModes * 6 * sizeof(Real));
        dStiff.zero(st);
        if (hSrcElem >= 0) {
            std::vector<Real> pat(hSrcPattern.begin(), hSrcPattern.end());
            dSrcPattern.alloc(pat.size() * sizeof(Real));
            dSrcPattern.upload(pat.data(), pat.size() * sizeof(Real), st);
        }
    }
};

std::unique_ptr<DomainBase> make_domain(int device, int precision) {
    if (precision == AXS3D_FP64)
        return std::make_unique<Domain<double>>(device, precision);
    if (precision == AXS3D_FP32_FP64)
        return std::make_unique<Domain<float>>(device, precision);
    throw Error("Domain::create || Unknown precision mode.");
}

}  // namespace axs3d
