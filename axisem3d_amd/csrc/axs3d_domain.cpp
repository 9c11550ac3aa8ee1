// axs3d_domain.cpp -- Domain set-up, phases and the Newmark loop.
#include "axs3d_domain.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
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
    // this device's address of the static scratch, looked up here at set-up
    // (with the device current) so that the static entry point never does
    mStaticScratch = static_scratch(device);
}

DomainBase::~DomainBase() {
    (void)hipSetDevice(mDevice);
    if (mStream) (void)hipStreamSynchronize(mStream);
    dropGraph();
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

static int mat_nparams(int mt) { return mt == 0 ? 2 : (mt == 1 ? 5 : 21); }

void DomainBase::setSolidIso1D(const double *lam_mu) {
    if (mNumElems <= 0)
        throw Error("Domain::setSolidIso1D || Set elements first.");
    for (uint8_t t : hMType)
        if (t != 0)
            throw Error("Domain::setSolidIso1D || Not all elements are "
                        "isotropic; use the general 1D material.");
    hMat1D.assign(lam_mu, lam_mu + (size_t)mNumElems * 2 * NPEM);
    hMatOff.resize(mNumElems);
    for (int e = 0; e < mNumElems; e++) hMatOff[e] = (int64_t)e * 2 * NPEM;
}

// Material::createElastic picks the medium per element (Material.cpp:
// 366-379): 0 isotropic, 1 transversely isotropic, 2 anisotropic.
void DomainBase::setMaterialTypes(const uint8_t *mtype) {
    if (mNumElems <= 0)
        throw Error("Domain::setMaterialTypes || Set elements first.");
    if (!hSliceOff.empty() || !hMat1D.empty())
        throw Error("Domain::setMaterialTypes || Types must be set before "
                    "the material.");
    hMType.assign(mtype, mtype + mNumElems);
    for (uint8_t t : hMType)
        if (t > 2)
            throw Error("Domain::setMaterialTypes || Unknown material type.");
}

void DomainBase::setSolidMaterial1D(const double *params) {
    if (mNumElems <= 0)
        throw Error("Domain::setSolidMaterial1D || Set elements first.");
    if (hMType.empty()) hMType.assign(mNumElems, 0);
    hMatOff.resize(mNumElems);
    int64_t off = 0;
    for (int e = 0; e < mNumElems; e++) {
        hMatOff[e] = off;
        off += (int64_t)mat_nparams(hMType[e]) * NPEM;
    }
    hMat1D.assign(params, params + off);
}

void DomainBase::setFluidPoints(int n, const int32_t *nr,
                                const double *mass) {
    if (n < 0) throw Error("Domain::setFluidPoints || Negative count.");
    mNumFPoints = n;
    hFPointNr.assign(nr, nr + n);
    hFMass.assign(mass, mass + n);
    for (int p = 0; p < n; p++) {
        if (hFPointNr[p] < 1)
            throw Error("Domain::setFluidPoints || Nr must be positive.");
        if (!(hFMass[p] > 0.))
            throw Error("Domain::setFluidPoints || Mass must be positive.");
    }
}

void DomainBase::setFluidElements(int n, const int32_t *points,
                                  const uint8_t *axial, const double *gq) {
    if (n < 0) throw Error("Domain::setFluidElements || Negative count.");
    mNumFElems = n;
    hFElemPoints.assign(points, points + (size_t)n * NPEM);
    hFElemAxial.assign(axial, axial + n);
    hFElemGQ.assign(gq, gq + (size_t)n * 6 * NPEM);
}

// Acoustic 1D: K = 1 / rho per GLL point (Material::createAcoustic,
// Material.cpp:350-360; Acoustic.hpp:20-25).
void DomainBase::setFluidK1D(const double *k) {
    if (mNumFElems <= 0)
        throw Error("Domain::setFluidK1D || Set fluid elements first.");
    hFElemK.assign(k, k + (size_t)mNumFElems * NPEM);
}

void DomainBase::setFluidKinds(const uint8_t *kind) {
    if (mNumFElems <= 0 || hFPointNr.empty())
        throw Error("Domain::setFluidKinds || Set fluid points and elements "
                    "first.");
    hFElemKind.assign(kind, kind + mNumFElems);
    hFElemNr.assign(mNumFElems, 0);
    hFK3DOff.assign(mNumFElems, -1);
    int64_t total = 0;
    for (int e = 0; e < mNumFElems; e++) {
        if (hFElemKind[e] > 1)
            throw Error("Domain::setFluidKinds || Unknown fluid element "
                        "kind.");
        int nrmax = 0;
        for (int k = 0; k < NPEM; k++) {
            const int p = hFElemPoints[(size_t)e * NPEM + k];
            if (p < 0 || p >= (int)hFPointNr.size())
                throw Error("Domain::setFluidKinds || Fluid element point "
                            "out of range.");
            nrmax = std::max(nrmax, hFPointNr[p]);
        }
        hFElemNr[e] = nrmax;
        if (hFElemKind[e] == 1) {
            hFK3DOff[e] = total;
            total += (int64_t)nrmax * NPEM;
        }
    }
    hFK3D.assign((size_t)total, -1.0);
}

void DomainBase::setFluidK3D(int e_begin, int e_end, const double *k) {
    if (hFElemKind.empty())
        throw Error("Domain::setFluidK3D || Set the fluid element kinds "
                    "first.");
    if (e_begin < 0 || e_end > mNumFElems || e_begin > e_end)
        throw Error("Domain::setFluidK3D || Element range out of bounds.");
    for (int e = e_begin; e < e_end; e++) {
        if (hFElemKind[e] != 1) continue;
        const size_t n = (size_t)hFElemNr[e] * NPEM;
        std::copy(k, k + n, hFK3D.begin() + hFK3DOff[e]);
        k += n;
    }
}

void DomainBase::setFluidAxialPoints(int n, const int32_t *points) {
    hFAxial.assign(points, points + std::max(n, 0));
}

void DomainBase::setFluidSurfacePoints(int n, const int32_t *points) {
    hFSurf.assign(points, points + std::max(n, 0));
}

void DomainBase::setSolidFluidCoupling(int n, const int32_t *solid,
                                       const int32_t *fluid,
                                       const double *normals) {
    if (n < 0)
        throw Error("Domain::setSolidFluidCoupling || Negative count.");
    mNumPairs = n;
    hCplSolid.assign(solid, solid + n);
    hCplFluid.assign(fluid, fluid + n);
    hCplNormal.assign(normals, normals + (size_t)n * 4);
}

void DomainBase::setFluidMass3D(int n, const int32_t *points,
                                const double *mass) {
    if (hFPointNr.empty())
        throw Error("Domain::setFluidMass3D || Set the fluid points first.");
    hFMass3DPts.clear();
    hFMass3DOff.clear();
    hFMass3D.clear();
    std::vector<uint8_t> seen(hFPointNr.size(), 0);
    for (int i = 0; i < std::max(n, 0); i++) {
        const int p = points[i];
        if (p < 0 || p >= (int)hFPointNr.size() || seen[p])
            throw Error("Domain::setFluidMass3D || Point out of range or "
                        "listed twice.");
        seen[p] = 1;
        hFMass3DPts.push_back(p);
        hFMass3DOff.push_back((int64_t)hFMass3D.size());
        for (int k = 0; k < hFPointNr[p]; k++) {
            const double m = *mass++;
            if (!(m > 0.))
                throw Error("Domain::setFluidMass3D || Mass must be "
                            "positive.");
            hFMass3D.push_back(m);
        }
    }
}

void DomainBase::setCoupling3D(int n, const int32_t *pairs, const double *n_u,
                               const double *n_a) {
    if (n < 0 || (n > 0 && (!pairs || !n_u || !n_a)))
        throw Error("Domain::setCoupling3D || Bad arguments.");
    hCpl3DPairs.assign(pairs, pairs + n);
    size_t tot = 0;
    for (int k : hCpl3DPairs) {
        if (k < 0 || k >= mNumPairs || hCplSolid[k] < 0 ||
            hCplSolid[k] >= (int)hPointNr.size())
            throw Error("Domain::setCoupling3D || Pair out of range (set "
                        "the coupling pairs and solid points first).");
        tot += (size_t)hPointNr[hCplSolid[k]] * 3;
    }
    hCpl3DNu.assign(n_u, n_u + tot);
    hCpl3DNa.assign(n_a, n_a + tot);
}

void DomainBase::setRtzTheta(const double *theta) {
    if (mNumElems <= 0)
        throw Error("Domain::setRtzTheta || Set elements first.");
    hTheta.assign(theta, theta + (size_t)mNumElems * NPEM);
}

void DomainBase::setElementKinds(const uint8_t *kind) {
    if (mNumElems <= 0)
        throw Error("Domain::setElementKinds || Set elements first.");
    if (!hSliceOff.empty())
        throw Error("Domain::setElementKinds || Kinds must be set before "
                    "3D properties.");
    hElemKind.assign(kind, kind + mNumElems);
    for (uint8_t k : hElemKind)
        if (k > 1)
            throw Error("Domain::setElementKinds || Unknown element kind.");
}

// Element Nr (Element::pointSet, Element.cpp:28-35) and the slice layout of
// 3D properties: the Nr slices of each 3D element, in element order.
void DomainBase::layout3D() {
    if (!hSliceOff.empty()) return;
    if (mNumPoints <= 0 || mNumElems <= 0)
        throw Error("Domain::layout3D || Set points and elements first.");
    if (hElemKind.empty()) hElemKind.assign(mNumElems, 0);
    hElemNr.assign(mNumElems, 0);
    hSliceOff.assign(mNumElems, -1);
    int64_t sl = 0;
    mNum3D = 0;
    for (int e = 0; e < mNumElems; e++) {
        int nr = 0;
        for (int k = 0; k < NPEM; k++) {
            int p = hElemPoints[(size_t)e * NPEM + k];
            if (p < 0 || p >= mNumPoints)
                throw Error("Domain::layout3D || Element point out of range.");
            nr = std::max(nr, hPointNr[p]);
        }
        hElemNr[e] = nr;
        if (hElemKind[e] == 1) {
            hSliceOff[e] = sl;
            sl += nr;
            mNum3D++;
        }
    }
    mNumSlices = sl;
    if (hMType.empty()) hMType.assign(mNumElems, 0);
    hPropOff.assign(mNumElems, -1);
    int64_t po = 0;
    for (int e = 0; e < mNumElems; e++)
        if (hElemKind[e] == 1) {
            hPropOff[e] = po;
            po += (int64_t)mat_nparams(hMType[e]) * NPEM * hElemNr[e];
        }
    mPropTotal = po;
    hProp3DSet.assign(mNumElems, 0);
    hAtt3DSet.assign(mNumElems, 0);
}

static void slice_range(const std::vector<uint8_t> &kind,
                        const std::vector<int64_t> &off,
                        const std::vector<int32_t> &nr, int e0, int e1,
                        int64_t &s0, int64_t &ns) {
    s0 = -1;
    ns = 0;
    for (int e = e0; e < e1; e++) {
        if (kind[e] != 1) continue;
        if (s0 < 0) s0 = off[e];
        ns += nr[e];
    }
}

void DomainBase::setSolidIso3D(int e_begin, int e_end, const double *lam_mu) {
    layout3D();
    for (int e = std::max(e_begin, 0); e < std::min(e_end, mNumElems); e++)
        if (hElemKind[e] == 1 && hMType[e] != 0)
            throw Error("Domain::setSolidIso3D || Element is not isotropic; "
                        "use the general 3D material.");
    setSolidMaterial3D(e_begin, e_end, lam_mu);
}

void DomainBase::setSolidMaterial3D(int e_begin, int e_end,
                                    const double *params) {
    layout3D();
    if (e_begin < 0 || e_end > mNumElems || e_begin > e_end)
        throw Error("Domain::setSolidMaterial3D || Element range out of "
                    "bounds.");
    int64_t s0, ns;
    slice_range(hElemKind, hSliceOff, hElemNr, e_begin, e_end, s0, ns);
    if (ns == 0) return;
    uploadProp(e_begin, e_end, params);
    for (int e = e_begin; e < e_end; e++)
        if (hElemKind[e] == 1) hProp3DSet[e] = 1;
}

void DomainBase::setAttenuation(int nsls, int cg4, const double *alpha,
                                const double *beta, const double *gamma) {
    if (nsls < 0 || nsls > 16)
        throw Error("Domain::setAttenuation || Invalid number of SLS.");
    mNsls = nsls;
    mAttCG4 = cg4 ? 1 : 0;
    hAttABG.clear();
    hAttABG.insert(hAttABG.end(), alpha, alpha + nsls);
    hAttABG.insert(hAttABG.end(), beta, beta + nsls);
    hAttABG.insert(hAttABG.end(), gamma, gamma + nsls);
}

void DomainBase::setAttenuation3D(int e_begin, int e_end,
                                  const double *dlam_dmu) {
    layout3D();
    if (mNsls <= 0)
        throw Error("Domain::setAttenuation3D || Set the SLS first.");
    if (e_begin < 0 || e_end > mNumElems || e_begin > e_end)
        throw Error("Domain::setAttenuation3D || Element range out of "
                    "bounds.");
    int64_t s0, ns;
    slice_range(hElemKind, hSliceOff, hElemNr, e_begin, e_end, s0, ns);
    if (ns == 0) return;
    uploadSlices(dAtt3D, 2 * (mAttCG4 ? 4 : NPEM), e_begin, e_end, dlam_dmu);
    for (int e = e_begin; e < e_end; e++)
        if (hElemKind[e] == 1) hAtt3DSet[e] = 1;
}

void DomainBase::setAttenuation1D(const double *dlam_dmu) {
    if (mNsls <= 0)
        throw Error("Domain::setAttenuation1D || Set the SLS first.");
    if (mNumElems <= 0)
        throw Error("Domain::setAttenuation1D || Set elements first.");
    hAtt1D.assign(dlam_dmu,
                  dlam_dmu + (size_t)mNumElems * 2 * (mAttCG4 ? 4 : NPEM));
}

void DomainBase::setAxialPoints(int n, const int32_t *points) {
    hAxial.assign(points, points + std::max(n, 0));
}

void DomainBase::setSolidMass3D(int n, const int32_t *points,
                                const double *mass) {
    if (hPointNr.empty())
        throw Error("Domain::setSolidMass3D || Set the solid points first.");
    hMass3DPts.clear();
    hMass3DOff.clear();
    hMass3D.clear();
    std::vector<uint8_t> seen(hPointNr.size(), 0);
    for (int i = 0; i < std::max(n, 0); i++) {
        const int p = points[i];
        if (p < 0 || p >= (int)hPointNr.size() || seen[p])
            throw Error("Domain::setSolidMass3D || Point out of range or "
                        "listed twice.");
        seen[p] = 1;
        hMass3DPts.push_back(p);
        hMass3DOff.push_back((int64_t)hMass3D.size());
        for (int k = 0; k < hPointNr[p]; k++) {
            const double m = *mass++;
            if (!(m > 0.))
                throw Error("Domain::setSolidMass3D || Mass must be "
                            "positive.");
            hMass3D.push_back(1. / m);
        }
    }
    mMass3DUserPts = hMass3DPts.size();
    mMass3DUserVals = hMass3D.size();
}

void DomainBase::setSolidPRT(int n, const int32_t *elems, const double *x) {
    if (n < 0 || (n > 0 && (elems == nullptr || x == nullptr)))
        throw Error("Domain::setSolidPRT || Invalid arguments.");
    hPrtElems.assign(elems, elems + n);
    hPrtX.assign(x, x + (size_t)n * 5 * NPEM);
}

void DomainBase::setOceanLoad(int n, const int32_t *points,
                              const int32_t *nval, const double *values) {
    if (hPointNr.empty())
        throw Error("Domain::setOceanLoad || Set the solid points first.");
    hOcean1DPts.clear();
    hOcean3DPts.clear();
    hOcean1D.clear();
    hOcean3D.clear();
    std::vector<uint8_t> seen(hPointNr.size(), 0);
    for (int i = 0; i < std::max(n, 0); i++) {
        const int p = points[i];
        if (p < 0 || p >= (int)hPointNr.size() || seen[p])
            throw Error("Domain::setOceanLoad || Point out of range or "
                        "listed twice.");
        seen[p] = 1;
        const int nv = nval[i];
        if (nv != 1 && nv != hPointNr[p])
            throw Error("Domain::setOceanLoad || An entry has 1 value (1D) "
                        "or one per slice of its point (3D).");
        std::vector<double> &dst = nv == 1 ? hOcean1D : hOcean3D;
        (nv == 1 ? hOcean1DPts : hOcean3DPts).push_back(p);
        for (int k = 0; k < 4 * nv; k++) dst.push_back(*values++);
    }
}

// The 3D ocean-load points join the 3D-mass list (after the points with a
// plain 3D mass): their 1 / m per slice with the masses, their k per slice
// in hMass3DK.  Idempotent: a second finalize rebuilds the same list.
void DomainBase::mergeOceanLoad() {
    hMass3DPts.resize(mMass3DUserPts);
    hMass3DOff.resize(mMass3DUserPts);
    hMass3D.resize(mMass3DUserVals);
    hMass3DKOff.assign(mMass3DUserPts, -1);
    hMass3DK.clear();
    std::vector<uint8_t> m3(hPointNr.size(), 0);
    for (int32_t p : hMass3DPts) m3[p] = 1;
    for (int32_t p : hOcean1DPts)
        if (m3[p])
            throw Error("Domain::finalize || A point with an ocean load takes "
                        "its mass from the ocean-load entry, not from "
                        "set_solid_mass3d.");
    const double *o = hOcean3D.data();
    for (int32_t p : hOcean3DPts) {
        if (m3[p])
            throw Error("Domain::finalize || A point with an ocean load takes "
                        "its mass from the ocean-load entry, not from "
                        "set_solid_mass3d.");
        hMass3DPts.push_back(p);
        hMass3DOff.push_back((int64_t)hMass3D.size());
        hMass3DKOff.push_back((int64_t)hMass3DK.size());
        for (int k = 0; k < hPointNr[p]; k++, o += 4) {
            if (!(o[0] > 0.))
                throw Error("Domain::finalize || Ocean load: 1 / m must be "
                            "positive.");
            hMass3D.push_back(o[0]);
            for (int c = 1; c < 4; c++) hMass3DK.push_back(o[c]);
        }
    }
}

void DomainBase::setMomentSource(int elem, int nu1, const double *pattern,
                                 int nsteps, const double *stf) {
    hSrcElem = -1;
    hSrcElems.clear();
    hSrcNu1s.clear();
    hSrcPatOff.clear();
    hSrcPattern.clear();
    hStf.clear();
    addMomentSource(elem, nu1, pattern, nsteps, stf);
}

// Sources of batched instances (one per instance, each in its own
// element); all STFs have the same number of steps.
void DomainBase::addMomentSource(int elem, int nu1, const double *pattern,
                                 int nsteps, const double *stf) {
    if (nu1 <= 0 || nsteps <= 0)
        throw Error("Domain::setMomentSource || Empty source.");
    if (!hSrcElems.empty() && (size_t)nsteps * hSrcElems.size() != hStf.size())
        throw Error("Domain::setMomentSource || Every source needs the same "
                    "number of STF steps.");
    for (int32_t e : hSrcElems)
        if (e == elem)
            throw Error("Domain::setMomentSource || Two sources in one "
                        "element.");
    if (hSrcElems.empty()) hSrcElem = elem;
    hSrcElems.push_back(elem);
    hSrcNu1s.push_back(nu1);
    hSrcPatOff.push_back(hSrcPattern.size() / 2);
    hSrcPattern.insert(hSrcPattern.end(), pattern,
                       pattern + (size_t)nu1 * 6 * NPEM * 2);
    hStf.insert(hStf.end(), stf, stf + nsteps);
}

void DomainBase::setInstances(int n, const int32_t *point_inst,
                              const double *dt) {
    if (n <= 0 || !point_inst || !dt)
        throw Error("Domain::setInstances || No instances.");
    hInstDt.assign(dt, dt + n);
    for (double v : hInstDt)
        if (!(v > 0.))
            throw Error("Domain::setInstances || Time steps must be "
                        "positive.");
    hPointInst.assign(point_inst, point_inst + mNumPoints);
    for (int32_t i : hPointInst)
        if (i < 0 || i >= n)
            throw Error("Domain::setInstances || Instance out of range.");
}

void DomainBase::setReceivers(int n, const int32_t *elem, const double *w,
                              const double *phi, int max_samples) {
    hRecElem.assign(elem, elem + std::max(n, 0));
    hRecW.assign(w, w + (size_t)std::max(n, 0) * NPEM);
    hRecPhi.assign(phi, phi + std::max(n, 0));
    hMaxSamples = std::max(max_samples, 0);
    hRecRot.clear();
}

// Station::rotateField for the displacement (Station.hpp:79-118), which
// the element hands over in (s, phi, z) (SolidElement::displInRTZ is
// false): up to two rotations about axis k by an angle, as
// Station::rotate(V, nrow, k, angle) (Station.cpp:41-55).
void DomainBase::setReceiverFrames(int n, const int32_t *wcs,
                                   const double *theta, const double *baz,
                                   int cartesian) {
    if (n != (int)hRecElem.size())
        throw Error("Domain::setReceiverFrames || One frame per receiver "
                    "(set the receivers first).");
    constexpr double kHalfPi = 1.57079632679489661923;
    hRecRot.assign((size_t)n * 6, 0.);
    for (int i = 0; i < n; i++) {
        double *r = hRecRot.data() + (size_t)i * 6;
        r[0] = r[3] = -1.;
        auto put = [&](int slot, int k, double ang) {
            r[3 * slot] = k;
            r[3 * slot + 1] = std::cos(ang);
            r[3 * slot + 2] = std::sin(ang);
        };
        const int w = wcs[i];
        if (w < 0 || w > 3)
            throw Error("Domain::setReceiverFrames || Unknown wavefield "
                        "coordinate system.");
        const double bazp90 = baz[i] + kHalfPi;
        if (cartesian) {
            // RTZ = SPZ
            if (w == 2) put(0, 2, bazp90);
            if (w == 3) put(0, 2, -hRecPhi[i]);
        } else {
            if (w == 1) put(0, 1, theta[i]);
            if (w == 2) {
                put(0, 1, theta[i]);
                put(1, 2, bazp90);
            }
            if (w == 3) put(0, 2, -hRecPhi[i]);
        }
    }
}

void DomainBase::setTimeScheme(int scheme) {
    if (scheme != 0 && scheme != 1)
        throw Error("Domain::setTimeScheme || Unknown time scheme.");
    mScheme = scheme;
}

// ------------------------------------------------- absorbing boundaries
int DomainBase::abcRecord(AbcHost &h, int npoints, int p, int w1) {
    if (p < 0 || p >= npoints)
        throw Error("Domain::addAbsorbingBoundary || Point out of range.");
    if (h.rec.size() != (size_t)npoints) h.rec.assign(npoints, -1);
    if (h.rec[p] < 0) {
        h.rec[p] = (int32_t)h.point.size();
        h.point.push_back(p);
        h.cl1.insert(h.cl1.end(), (size_t)w1, 0.);
        h.cl3.emplace_back();
        h.sp1.insert(h.sp1.end(), 2, 0.);
        h.sp3.emplace_back();
    }
    return h.rec[p];
}

// ClaytonSolid1D (ClaytonSolid1D.hpp:19-33): the four coefficients of its
// constructor, summed over the point's entries.
void DomainBase::addClaytonSolid(int n, const int32_t *points,
                                 const double *rho_vp, const double *rho_vs,
                                 const double *area, const double *theta) {
    if (mNumPoints <= 0)
        throw Error("Domain::addClaytonSolid || Set the points first.");
    for (int i = 0; i < n; i++) {
        const int r = abcRecord(mAbcS, mNumPoints, points[i], 4);
        const double rsa = rho_vs[i] * area[i], rpa = rho_vp[i] * area[i];
        const double c = std::cos(theta[i]), s = std::sin(theta[i]);
        double *k = mAbcS.cl1.data() + (size_t)r * 4;
        k[0] += rsa * c * c + rpa * s * s;
        k[1] += rsa * s * s + rpa * c * c;
        k[2] += (rpa - rsa) * c * s;
        k[3] += rsa;
    }
}

// ClaytonSolid3D (ClaytonSolid3D.hpp:19-27): per slice rsa = rho vs area
// and k = sqrt((rho vp - rho vs) area) n; its force rsa V + (V.k) k is the
// symmetric operator rsa I + k k^T, summed over the point's entries.
void DomainBase::addClaytonSolid3D(int n, const int32_t *points,
                                   const double *rho_vp,
                                   const double *rho_vs, const double *area,
                                   const double *unit_normal) {
    if (mNumPoints <= 0)
        throw Error("Domain::addClaytonSolid3D || Set the points first.");
    size_t v = 0;
    for (int i = 0; i < n; i++) {
        const int r = abcRecord(mAbcS, mNumPoints, points[i], 4);
        const int nr = hPointNr[points[i]];
        std::vector<double> &m = mAbcS.cl3[r];
        if (m.empty()) m.assign((size_t)nr * 6, 0.);
        for (int k = 0; k < nr; k++, v++) {
            const double rsa = rho_vs[v] * area[v];
            const double a = std::sqrt((rho_vp[v] - rho_vs[v]) * area[v]);
            const double ks = a * unit_normal[3 * v],
                         kp = a * unit_normal[3 * v + 1],
                         kz = a * unit_normal[3 * v + 2];
            double *o = m.data() + (size_t)k * 6;
            o[0] += rsa + ks * ks;
            o[1] += rsa + kp * kp;
            o[2] += rsa + kz * kz;
            o[3] += ks * kp;
            o[4] += ks * kz;
            o[5] += kp * kz;
        }
    }
}

// ClaytonFluid1D/3D (ClaytonFluid1D.hpp:19-24, ClaytonFluid3D.hpp:19-25):
// area / (rho vp), one value or one per slice of the point.
void DomainBase::addClaytonFluid(int n, const int32_t *points, int per_slice,
                                 const double *rho_vp, const double *area) {
    if (mNumFPoints <= 0)
        throw Error("Domain::addClaytonFluid || Set the fluid points first.");
    size_t v = 0;
    for (int i = 0; i < n; i++) {
        const int r = abcRecord(mAbcF, mNumFPoints, points[i], 1);
        if (!per_slice) {
            mAbcF.cl1[r] += area[v] / rho_vp[v];
            v++;
            continue;
        }
        const int nr = hFPointNr[points[i]];
        std::vector<double> &m = mAbcF.cl3[r];
        if (m.empty()) m.assign(nr, 0.);
        for (int k = 0; k < nr; k++, v++) m[k] += area[v] / rho_vp[v];
    }
}

// Sponge<SolidPoint/FluidPoint> (Sponge.hpp:30-45): gamma, one value or
// one per slice of the point; the force 2 gamma v + gamma^2 u kept as
// (2 gamma, gamma^2), summed over the point's entries.
void DomainBase::addSponge(int fluid, int n, const int32_t *points,
                           int per_slice, const double *gamma) {
    AbcHost &h = fluid ? mAbcF : mAbcS;
    const int np = fluid ? mNumFPoints : mNumPoints;
    if (np <= 0) throw Error("Domain::addSponge || Set the points first.");
    size_t v = 0;
    for (int i = 0; i < n; i++) {
        const int r = abcRecord(h, np, points[i], fluid ? 1 : 4);
        if (!per_slice) {
            h.sp1[2 * r] += 2. * gamma[v];
            h.sp1[2 * r + 1] += gamma[v] * gamma[v];
            v++;
            continue;
        }
        const int nr = fluid ? hFPointNr[points[i]] : hPointNr[points[i]];
        std::vector<double> &m = h.sp3[r];
        if (m.empty()) m.assign((size_t)nr * 2, 0.);
        for (int k = 0; k < nr; k++, v++) {
            m[2 * k] += 2. * gamma[v];
            m[2 * k + 1] += gamma[v] * gamma[v];
        }
    }
}

// Plan words, twiddles (FP64, interleaved) and slice positions of every
// Nr of mFftNrs with pass radices of at most maxRadix (axs3d_fft.hpp), in
// the layout finalize's offsets describe.
void DomainBase::buildFftTables(int maxRadix, std::vector<uint32_t> &plan,
                                std::vector<double> &tw,
                                std::vector<uint16_t> &pos) const {
    plan.assign((size_t)(mFftNrMax + 1) * CARD_PLAN_W, 0u);
    tw.clear();
    pos.clear();
    std::vector<double> tw1;
    std::vector<uint16_t> pos1;
    for (const int nr : mFftNrs) {
        const std::vector<int> rad = fft_choose_radices(nr, maxRadix);
        if (rad.empty() && nr > 1)
            throw Error("Domain::finalize || Nr of a 3D element has "
                        "a prime factor above 13 (use FFT lucky "
                        "numbers).");
        if ((int)rad.size() > FFT_MAX_PASSES)
            throw Error("Domain::finalize || Too many FFT passes.");
        uint32_t *pl = plan.data() + (size_t)nr * CARD_PLAN_W;
        fft_build_plan(nr, rad, pl, tw1, pos1);
        tw.insert(tw.end(), tw1.begin(), tw1.end());
        pos.insert(pos.end(), pos1.begin(), pos1.end());
    }
}

// Point -> element-local force offsets in ascending element order: four
// slots per point (a conforming quad mesh has at most four elements at a
// point); a point shared by more (the doubling layers of unstructured
// meshes) keeps its first three in slots 0..2 and in slot 3 the marker
// 0x80000000 | j of the rest of its list, stored from index j of the same
// array (after the 4 n slots) and closed by 0xffffffff.
template <typename OffsetFn>
std::vector<uint32_t> buildAdjacency(int npoints, int nelems,
                                     const int32_t *elem_points,
                                     OffsetFn offset) {
    std::vector<std::vector<uint32_t>> lists(npoints);
    for (int e = 0; e < nelems; e++)
        for (int k = 0; k < NPEM; k++)
            lists[elem_points[(size_t)e * NPEM + k]].push_back(offset(e, k));
    std::vector<uint32_t> adj((size_t)npoints * 4, 0xffffffffu);
    for (int p = 0; p < npoints; p++) {
        const std::vector<uint32_t> &l = lists[p];
        if (l.size() <= 4) {
            for (size_t i = 0; i < l.size(); i++) adj[(size_t)p * 4 + i] = l[i];
            continue;
        }
        for (int i = 0; i < 3; i++) adj[(size_t)p * 4 + i] = l[i];
        if (adj.size() >= 0x7fffffffu)
            throw Error("Domain::finalize || Point adjacency too large.");
        adj[(size_t)p * 4 + 3] = 0x80000000u | (uint32_t)adj.size();
        adj.insert(adj.end(), l.begin() + 3, l.end());
        adj.push_back(0xffffffffu);
    }
    return adj;
}

void DomainBase::finalize() {
    hip_check(hipSetDevice(mDevice), "Domain::finalize");
    if (const char *env = std::getenv("AXS3D_FFT_MAX_RADIX")) {
        const int r = std::atoi(env);
        if (r < 2 || r > 16)
            throw Error("Domain::finalize || AXS3D_FFT_MAX_RADIX must be in "
                        "2..16.");
        mFftMaxRadix = r;
        mFftRadixFixed = true;
    }
    mergeOceanLoad();
    if (hGll.size() != NPEM || hGlj.size() != NPEM)
        throw Error("Domain::finalize || G matrices are not set.");
    if (mNumPoints <= 0 || mNumElems <= 0)
        throw Error("Domain::finalize || Points or elements are not set.");
    if (!(mDt > 0.)) throw Error("Domain::finalize || dt is not set.");
    if (mScheme == 1 && !hInstDt.empty())
        throw Error("Domain::finalize || The symplectic scheme takes one "
                    "time step for the whole domain.");
    layout3D();
    const bool has1D = mNum3D < mNumElems;
    if (has1D && hMatOff.size() != (size_t)mNumElems)
        throw Error("Domain::finalize || Material is not set.");
    if (!hTheta.empty() && hTheta.size() != (size_t)mNumElems * NPEM)
        throw Error("Domain::finalize || Colatitudes are not set.");
    for (int e = 0; e < mNumElems; e++) {
        if (hElemKind[e] != 1) continue;
        if (!hProp3DSet[e])
            throw Error("Domain::finalize || 3D properties are not set.");
        if (mNsls > 0 && !hAtt3DSet[e])
            throw Error("Domain::finalize || 3D attenuation is not set.");
    }
    if (mNsls > 0 && has1D &&
        hAtt1D.size() != (size_t)mNumElems * 2 * (mAttCG4 ? 4 : NPEM))
        throw Error("Domain::finalize || 1D attenuation is not set.");
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
    }
    // Fourier-path element-modes grouped by material type (one launch per
    // type), in element order within a group, the axial elements last
    for (int mt = 0; mt < 3; mt++) {
        mEMTypeBegin[mt] = (int)em.size();
        for (int ax = 0; ax < 2; ax++) {
            if (ax) mEMAxBegin[mt] = (int)em.size();
            for (int e = 0; e < mNumElems; e++)
                if (hElemKind[e] != 1 && hMType[e] == mt &&
                    (hElemAxial[e] != 0) == (ax != 0))
                    for (int a = 0; a < eNu1[e]; a++) em.push_back({e, a});
        }
    }
    mEMTypeBegin[3] = (int)em.size();
    mNumEM = (int)em.size();
    // 3D elements, heaviest (largest Nr) first for the work queue
    std::vector<int32_t> list3d;
    mNu1Max3D = 0;
    for (int e = 0; e < mNumElems; e++)
        if (hElemKind[e] == 1) {
            list3d.push_back(e);
            mNu1Max3D = std::max(mNu1Max3D, eNu1[e]);
        }
    // grouped by material type (one launch per type), heaviest first
    std::stable_sort(list3d.begin(), list3d.end(), [&](int a, int b) {
        if (hMType[a] != hMType[b]) return hMType[a] < hMType[b];
        return hElemNr[a] > hElemNr[b];
    });
    for (int mt = 0; mt <= 3; mt++) {
        int n = 0;
        for (int32_t e : list3d)
            if (hMType[e] < mt) n++;
        mTypeBegin3D[mt] = n;
    }
    // In-place FFT plans, stage twiddle tables and slice positions for
    // every 3D Nr: powers of 2 as radix 8 (one 4 or 2 for the remainder),
    // then 3, 5, 7, 11, 13; stages in descending radix (decimation in
    // frequency order), which keeps the twiddle tables short.  Stage t of
    // span L_t (L_0 = Nr, L_{t+1} = m_t = L_t / R_t) has the twiddles
    // exp(-2 pi i j / L_t), j < m_t.  After the inverse (DIF) transform
    // slice s = s_0 + R_0 s_1 + R_0 R_1 s_2 + ... sits at position
    // s_0 m_0 + s_1 m_1 + ...
    std::vector<int64_t> twOff, posOff;
    std::vector<double> twTab;
    std::vector<uint16_t> posTab;
    std::vector<uint32_t> plan;
    // Nr of the 3D elements and of the points with a 3D mass
    std::vector<int> fftNr;
    for (int32_t e : list3d) fftNr.push_back(hElemNr[e]);
    for (int32_t p : hMass3DPts) fftNr.push_back(hPointNr[p]);
    for (int32_t p : hFMass3DPts) fftNr.push_back(hFPointNr[p]);
    for (int32_t k : hCpl3DPairs) fftNr.push_back(hPointNr[hCplSolid[k]]);
    // absorbing-boundary points with a per-slice part
    for (size_t r = 0; r < mAbcS.point.size(); r++)
        if (!mAbcS.cl3[r].empty() || !mAbcS.sp3[r].empty())
            fftNr.push_back(hPointNr[mAbcS.point[r]]);
    for (size_t r = 0; r < mAbcF.point.size(); r++)
        if (!mAbcF.cl3[r].empty() || !mAbcF.sp3[r].empty())
            fftNr.push_back(hFPointNr[mAbcF.point[r]]);
    // fluid elements with a 3D medium, heaviest first
    hFList3D.clear();
    if (mNumFElems > 0 && !hFElemKind.empty()) {
        for (int e = 0; e < mNumFElems; e++)
            if (hFElemKind[e] == 1) {
                for (size_t k = 0; k < (size_t)hFElemNr[e] * NPEM; k++)
                    if (!(hFK3D[hFK3DOff[e] + k] > 0.))
                        throw Error("Domain::finalize || 3D fluid medium "
                                    "not set (or not positive).");
                hFList3D.push_back(e);
                fftNr.push_back(hFElemNr[e]);
            }
        std::stable_sort(hFList3D.begin(), hFList3D.end(), [&](int a, int b) {
            return hFElemNr[a] > hFElemNr[b];
        });
    }
    mNumF3D = (int)hFList3D.size();
    mNumMass3D = (int)hMass3DPts.size();
    if (!fftNr.empty()) {
        int nrmax = 0;
        for (int nr : fftNr) nrmax = std::max(nrmax, nr);
        if (nrmax > CARD_NR_CAP)
            throw Error("Domain::finalize || Nr of a 3D element or 3D-mass "
                        "point exceeds the FFT capacity (2048).");
        twOff.assign(nrmax + 1, -1);
        posOff.assign(nrmax + 1, -1);
        mFftNrs.clear();
        mFftNrMax = nrmax;
        // the distinct Nr in order of first use; the tables of each take
        // Nr - 1 twiddles and Nr positions whatever the pass radices, so
        // these offsets hold for every plan buildFftTables makes
        int64_t tcum = 0, pcum = 0;
        for (const int nr : fftNr) {
            if (twOff[nr] >= 0) continue;
            mFftNrs.push_back(nr);
            twOff[nr] = tcum;
            posOff[nr] = pcum;
            tcum += std::max(nr - 1, 0);
            pcum += nr;
        }
        buildFftTables(mFftMaxRadix, plan, twTab, posTab);
    }
    mMass3DNrMax = 0;
    for (int32_t p : hMass3DPts)
        mMass3DNrMax = std::max(mMass3DNrMax, (int)hPointNr[p]);
    // element-local force buffer [e][ipnt][alpha]: offsets in triples
    std::vector<uint32_t> ebOff(mNumElems);
    int64_t ebTotal = 0;
    for (int e = 0; e < mNumElems; e++) {
        ebOff[e] = (uint32_t)ebTotal;
        ebTotal += (int64_t)NPEM * eNu1[e];
    }
    if (ebTotal >= (int64_t)UINT32_MAX / 2)
        throw Error("Domain::finalize || Too many element modes.");
    mNumEBTriples = ebTotal;
    // point -> (element, ipnt) adjacency in ascending element order, the
    // order in which Domain::computeStiffness visits the elements
    std::vector<uint32_t> adj = buildAdjacency(
        mNumPoints, mNumElems, hElemPoints.data(), [&](int e, int k) {
            return ebOff[e] + (uint32_t)k * eNu1[e];
        });
    // sources
    mSrcNu1c.clear();
    for (size_t i = 0; i < hSrcElems.size(); i++) {
        const int e = hSrcElems[i];
        if (e < 0 || e >= mNumElems)
            throw Error("Domain::finalize || Source element out of range.");
        mSrcNu1c.push_back(std::min(eNu1[e], hSrcNu1s[i]));
    }
    if (!hSrcElems.empty())
        mStfSteps = (int)(hStf.size() / hSrcElems.size());
    // particle relabelling: elements in range, listed once, |J| > 0
    if (!hPrtElems.empty()) {
        std::vector<uint8_t> seen(mNumElems, 0);
        for (size_t i = 0; i < hPrtElems.size(); i++) {
            const int e = hPrtElems[i];
            if (e < 0 || e >= mNumElems || seen[e])
                throw Error("Domain::finalize || PRT element out of range or "
                            "listed twice.");
            seen[e] = 1;
            for (int k = 0; k < 5 * NPEM; k++)
                if (!std::isfinite(hPrtX[i * 5 * NPEM + k]))
                    throw Error("Domain::finalize || PRT operands must be "
                                "finite.");
            for (int k = 0; k < NPEM; k++)
                if (!(hPrtX[(i * 5 + 4) * NPEM + k] > 0.))
                    throw Error("Domain::finalize || Negative Jacobian or "
                                "distorted element occurred (PRT |J| <= 0).");
        }
    }
    // batched instances
    if (!hInstDt.empty()) {
        if ((int)hPointInst.size() != mNumPoints)
            throw Error("Domain::finalize || Instances set before the "
                        "points.");
        if (mNumFElems > 0 || mNumMass3D > 0)
            throw Error("Domain::finalize || Batched instances with fluid "
                        "or 3D-mass points are not supported.");
    }
    // receivers
    mNumRec = (int)hRecElem.size();
    mRecNu1Max = 0;
    for (int32_t e : hRecElem) {
        if (e < 0 || e >= mNumElems)
            throw Error("Domain::finalize || Receiver element out of range.");
        mRecNu1Max = std::max(mRecNu1Max, eNu1[e]);
    }

    // fluid points, elements, boundaries and the solid-fluid coupling
    std::vector<uint32_t> fOff(mNumFPoints + 1, 0), fEbOff, fAdj, cSOff,
        cFOff;
    std::vector<int32_t> fNu1(mNumFPoints), fModePoint, fElemNu1,
        cSolidPair(mNumPoints, -1), cFluidPair(mNumFPoints, -1);
    std::vector<uint8_t> fBC(mNumFPoints, 0);
    std::vector<EMItem> fem;
    std::vector<double> cN, cNM;
    // 3D pairs (index j among them) and the deferred fluid points
    std::vector<int32_t> pair3d(std::max(mNumPairs, 0), -1),
        cSolidPair3D(mNumPoints, -1), fPair3D(mNumFPoints, -1), dfPts,
        dfPair, cpl3Solid;
    std::vector<int64_t> fMass3DOff(mNumFPoints, -1), dfMOff, cpl3NOff;
    std::vector<uint32_t> cpl3COff;
    std::vector<uint8_t> dfFlag;
    uint32_t cpl3Modes = 0;
    if (!hFMass3DPts.empty() && mNumFElems <= 0)
        throw Error("Domain::finalize || Fluid 3D mass without fluid "
                    "elements.");
    for (size_t i = 0; i < hFMass3DPts.size(); i++)
        fMass3DOff[hFMass3DPts[i]] = hFMass3DOff[i];
    mNumCpl3D = (int)hCpl3DPairs.size();
    for (int j = 0; j < mNumCpl3D; j++) {
        const int k = hCpl3DPairs[j];
        if (k < 0 || k >= mNumPairs || pair3d[k] >= 0)
            throw Error("Domain::finalize || 3D coupling pair out of range "
                        "or listed twice.");
        pair3d[k] = j;
    }
    if (mNumFElems > 0) {
        if (mNumFPoints <= 0)
            throw Error("Domain::finalize || Fluid points are not set.");
        if (hFElemK.size() != (size_t)mNumFElems * NPEM)
            throw Error("Domain::finalize || Fluid material is not set.");
        int64_t ft = 0;
        for (int p = 0; p < mNumFPoints; p++) {
            fOff[p] = (uint32_t)ft;
            fNu1[p] = hFPointNr[p] / 2 + 1;
            ft += fNu1[p];
        }
        if (ft >= (int64_t)UINT32_MAX / 4)
            throw Error("Domain::finalize || Too many fluid modes.");
        fOff[mNumFPoints] = (uint32_t)ft;
        mNumFModes = ft;
        fModePoint.resize(ft);
        for (int p = 0; p < mNumFPoints; p++)
            for (int a = 0; a < fNu1[p]; a++) fModePoint[fOff[p] + a] = p;
        fElemNu1.resize(mNumFElems);
        fEbOff.resize(mNumFElems);
        int64_t eb = 0;
        for (int e = 0; e < mNumFElems; e++) {
            int nrmax = 0;
            for (int k = 0; k < NPEM; k++) {
                const int p = hFElemPoints[(size_t)e * NPEM + k];
                if (p < 0 || p >= mNumFPoints)
                    throw Error("Domain::finalize || Fluid element point out "
                                "of range.");
                nrmax = std::max(nrmax, hFPointNr[p]);
            }
            fElemNu1[e] = nrmax / 2 + 1;
            if (hFElemKind.empty() || hFElemKind[e] == 0)
                for (int a = 0; a < fElemNu1[e]; a++) fem.push_back({e, a});
            fEbOff[e] = (uint32_t)eb;
            eb += (int64_t)NPEM * fElemNu1[e];
        }
        if (eb >= (int64_t)UINT32_MAX / 4)
            throw Error("Domain::finalize || Too many fluid element modes.");
        mNumFEBModes = eb;
        mNumFEM = (int)fem.size();
        fAdj = buildAdjacency(
            mNumFPoints, mNumFElems, hFElemPoints.data(), [&](int e, int k) {
                return fEbOff[e] + (uint32_t)k * fElemNu1[e];
            });
        for (int32_t p : hFAxial) {
            if (p < 0 || p >= mNumFPoints)
                throw Error("Domain::finalize || Fluid axial point out of "
                            "range.");
            fBC[p] |= 1;
        }
        for (int32_t p : hFSurf) {
            if (p < 0 || p >= mNumFPoints)
                throw Error("Domain::finalize || Fluid surface point out of "
                            "range.");
            fBC[p] |= 2;
        }
        // SolidFluidCoupling constructor and checkCompatibility
        // (SolidFluidCoupling.cpp:18-44, SolidFluidCoupling1D.cpp): the
        // fluid-side normal is taken unassembled, the solid side assembled
        // and divided by the fluid mass
        cSOff.resize(mNumPairs);
        cFOff.resize(mNumPairs);
        cN.resize((size_t)mNumPairs * 2);
        cNM.resize((size_t)mNumPairs * 2);
        std::vector<uint8_t> sSeen(mNumPoints, 0), fSeen(mNumFPoints, 0);
        for (int k = 0; k < mNumPairs; k++) {
            const int sp = hCplSolid[k], fp = hCplFluid[k];
            if (sp < 0 || sp >= mNumPoints || fp < 0 || fp >= mNumFPoints)
                throw Error("SolidFluidCoupling::checkCompatibility || Point "
                            "out of range.");
            if (hPointNr[sp] != hFPointNr[fp])
                throw Error("SolidFluidCoupling::checkCompatibility || "
                            "Incompatible sizes.");
            if (sSeen[sp] || fSeen[fp])
                throw Error("SolidFluidCoupling::checkCompatibility || Point "
                            "coupled twice.");
            sSeen[sp] = fSeen[fp] = 1;
            if (pair3d[k] >= 0) {
                cSolidPair3D[sp] = pair3d[k];
                fPair3D[fp] = pair3d[k];
                continue;
            }
            if (fMass3DOff[fp] >= 0)
                throw Error("SolidFluidCoupling::checkCompatibility || A "
                            "fluid point with a 3D mass needs a 3D "
                            "coupling.");
            cSolidPair[sp] = k;
            cFluidPair[fp] = k;
            cSOff[k] = off[sp];
            cFOff[k] = fOff[fp];
            const double *nn = hCplNormal.data() + (size_t)k * 4;
            cN[2 * k] = nn[0];
            cN[2 * k + 1] = nn[1];
            cNM[2 * k] = nn[2] / hFMass[fp];
            cNM[2 * k + 1] = nn[3] / hFMass[fp];
        }
        // the deferred fluid points, ascending
        dfFlag.assign(mNumFPoints, 0);
        for (int p = 0; p < mNumFPoints; p++)
            if (fMass3DOff[p] >= 0 || fPair3D[p] >= 0) {
                dfFlag[p] = 1;
                dfPts.push_back(p);
                dfPair.push_back(fPair3D[p]);
                dfMOff.push_back(fMass3DOff[p]);
            }
        mNumFDefer = (int)dfPts.size();
        mFDeferNrMax = 0;
        for (int32_t p : dfPts)
            mFDeferNrMax = std::max(mFDeferNrMax, (int)hFPointNr[p]);
        // 3D pairs: slice and mode offsets, assembled normals over the
        // fluid mass per slice (SolidFluidCoupling3D constructor,
        // SolidFluidCoupling3D.cpp:15-24)
        hCpl3DNam.assign(hCpl3DNa.size(), 0.);
        int64_t sl = 0;
        for (int j = 0; j < mNumCpl3D; j++) {
            const int k = hCpl3DPairs[j];
            const int sp = hCplSolid[k], fp = hCplFluid[k];
            const int nr = hPointNr[sp];
            cpl3Solid.push_back(sp);
            cpl3NOff.push_back(sl);
            cpl3COff.push_back(cpl3Modes);
            for (int s = 0; s < nr; s++) {
                const double mf = fMass3DOff[fp] >= 0
                                      ? hFMass3D[fMass3DOff[fp] + s]
                                      : hFMass[fp];
                for (int c = 0; c < 3; c++)
                    hCpl3DNam[(sl + s) * 3 + c] =
                        (1. / mf) * hCpl3DNa[(sl + s) * 3 + c];
            }
            sl += nr;
            cpl3Modes += (uint32_t)(nr / 2 + 1);
        }
    } else if (mNumPairs > 0) {
        throw Error("Domain::finalize || Coupling without fluid elements.");
    }

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
    up(dElemEBOff, ebOff.data(), ebOff.size() * 4);
    up(dAdj, adj.data(), adj.size() * 4);
    up(dElemNr, hElemNr.data(), hElemNr.size() * 4);
    up(dSliceOff, hSliceOff.data(), hSliceOff.size() * 8);
    up(dPropOff, hPropOff.data(), hPropOff.size() * 8);
    if (!list3d.empty()) up(dList3D, list3d.data(), list3d.size() * 4);
    if (!fftNr.empty()) {
        up(dTwOff, twOff.data(), twOff.size() * 8);
        up(dPlan, plan.data(), plan.size() * 4);
        up(dPosOff, posOff.data(), posOff.size() * 8);
        up(dPosTab, posTab.data(), posTab.size() * 2);
    }
    if (mNumMass3D > 0) {
        std::vector<uint8_t> flag(mNumPoints, 0);
        for (int32_t p : hMass3DPts) flag[p] = 1;
        up(dMass3DFlag, flag.data(), flag.size());
        up(dMass3DPts, hMass3DPts.data(), hMass3DPts.size() * 4);
        up(dMass3DOff, hMass3DOff.data(), hMass3DOff.size() * 8);
        if (!hMass3DK.empty())
            up(dMass3DKOff, hMass3DKOff.data(), hMass3DKOff.size() * 8);
    }
    if (!hOcean1DPts.empty()) {
        std::vector<int32_t> idx(mNumPoints, -1);
        for (size_t i = 0; i < hOcean1DPts.size(); i++)
            idx[hOcean1DPts[i]] = (int32_t)i;
        up(dOceanIdx, idx.data(), idx.size() * 4);
        up(dOcean1D, hOcean1D.data(), hOcean1D.size() * 8);
    }
    hList3D.swap(list3d);
    hTwTab.swap(twTab);
    if (mNsls > 0) up(dAttABG, hAttABG.data(), hAttABG.size() * 8);
    const size_t fbytes = (size_t)mNumModes * 6 * sizeof(double);
    dDispl.alloc(fbytes);
    dVeloc.alloc(fbytes);
    dAccel.alloc(fbytes);
    dDispl.zero(st);
    dVeloc.zero(st);
    dAccel.zero(st);
    if (hSrcElem >= 0) up(dStf, hStf.data(), hStf.size() * 8);
    if (!hInstDt.empty()) {
        std::vector<double> pdt(mNumPoints);
        for (int p = 0; p < mNumPoints; p++) pdt[p] = hInstDt[hPointInst[p]];
        up(dPointDt, pdt.data(), pdt.size() * 8);
    }
    // absorbing boundaries: records, their FP64 parts and term buffers (the
    // per-slice parts follow in the working precision in uploadReal)
    auto abcUp = [&](AbcHost &h, AbcDev &dv, int npts,
                     const std::vector<int32_t> &nrOf) {
        dv.n = (int)h.point.size();
        if (dv.n == 0) return;
        const int nc = &h == &mAbcS ? 3 : 1;
        std::vector<int64_t> toff(dv.n), c3(dv.n, -1), s3(dv.n, -1);
        int64_t t = 0, c = 0, s = 0;
        dv.nr_max = 0;
        for (int r = 0; r < dv.n; r++) {
            const int nr = nrOf[h.point[r]];
            toff[r] = t;
            t += nr / 2 + 1;
            if (!h.cl3[r].empty()) {
                c3[r] = c;
                c += nr;
            }
            if (!h.sp3[r].empty()) {
                s3[r] = s;
                s += nr;
            }
            if (c3[r] >= 0 || s3[r] >= 0) dv.nr_max = std::max(dv.nr_max, nr);
        }
        std::vector<int32_t> idx(npts, -1);
        for (int r = 0; r < dv.n; r++) idx[h.point[r]] = r;
        up(dv.idx, idx.data(), idx.size() * 4);
        up(dv.point, h.point.data(), h.point.size() * 4);
        up(dv.cl1, h.cl1.data(), h.cl1.size() * 8);
        up(dv.cl3off, c3.data(), c3.size() * 8);
        up(dv.sp1, h.sp1.data(), h.sp1.size() * 8);
        up(dv.sp3off, s3.data(), s3.size() * 8);
        up(dv.termoff, toff.data(), toff.size() * 8);
        dv.cterm.alloc((size_t)std::max<int64_t>(t, 1) * 2 * nc * 8);
        dv.sterm.alloc((size_t)std::max<int64_t>(t, 1) * 2 * nc * 8);
        dv.cterm.zero(st);
        dv.sterm.zero(st);
    };
    abcUp(mAbcS, mAbcSD, mNumPoints, hPointNr);
    if (mNumFElems > 0) abcUp(mAbcF, mAbcFD, mNumFPoints, hFPointNr);
    if (mNumRec > 0) {
        up(dRecElem, hRecElem.data(), hRecElem.size() * 4);
        up(dRecW, hRecW.data(), hRecW.size() * 8);
        up(dRecPhi, hRecPhi.data(), hRecPhi.size() * 8);
        if (!hRecRot.empty()) up(dRecRot, hRecRot.data(), hRecRot.size() * 8);
        dSeismo.alloc((size_t)std::max(hMaxSamples, 1) * mNumRec * 3 * 8);
        dSeismo.zero(st);
    }
    if (mNumFElems > 0) {
        up(dFPointOff, fOff.data(), fOff.size() * 4);
        up(dFPointNu1, fNu1.data(), fNu1.size() * 4);
        up(dFPointNr, hFPointNr.data(), hFPointNr.size() * 4);
        up(dFPointBC, fBC.data(), fBC.size());
        std::vector<double> finv(mNumFPoints);
        for (int p = 0; p < mNumFPoints; p++) finv[p] = 1. / hFMass[p];
        up(dFInvMass, finv.data(), finv.size() * 8);
        up(dFModePoint, fModePoint.data(), fModePoint.size() * 4);
        up(dFAdj, fAdj.data(), fAdj.size() * 4);
        up(dFElemPoints, hFElemPoints.data(), hFElemPoints.size() * 4);
        up(dFElemAxial, hFElemAxial.data(), hFElemAxial.size());
        up(dFElemNu1, fElemNu1.data(), fElemNu1.size() * 4);
        if (!fem.empty())
            up(dFEMList, fem.data(), fem.size() * sizeof(EMItem));
        if (mNumF3D > 0) {
            up(dFList3D, hFList3D.data(), hFList3D.size() * 4);
            up(dFElemNr, hFElemNr.data(), hFElemNr.size() * 4);
        }
        up(dFEBOff, fEbOff.data(), fEbOff.size() * 4);
        const size_t fb = (size_t)mNumFModes * 2 * sizeof(double);
        dChi.alloc(fb);
        dChiV.alloc(fb);
        dChiA.alloc(fb);
        dChi.zero(st);
        dChiV.zero(st);
        dChiA.zero(st);
        dCStiff.alloc(fb);
        dCStiff.zero(st);
        if (mNumFDefer > 0) {
            up(dFDeferPts, dfPts.data(), dfPts.size() * 4);
            up(dFDeferPair, dfPair.data(), dfPair.size() * 4);
            up(dFDeferMOff, dfMOff.data(), dfMOff.size() * 8);
            up(dFDeferFlag, dfFlag.data(), dfFlag.size());
        }
        if (mNumCpl3D > 0) {
            up(dCpl3DSolid, cpl3Solid.data(), cpl3Solid.size() * 4);
            up(dCpl3DNOff, cpl3NOff.data(), cpl3NOff.size() * 8);
            up(dCpl3DCOff, cpl3COff.data(), cpl3COff.size() * 4);
            up(dCplSolidPair3D, cSolidPair3D.data(), cSolidPair3D.size() * 4);
            dCSolid3D.alloc((size_t)cpl3Modes * 6 * sizeof(double));
            dCSolid3D.zero(st);
        }
        if (mNumPairs > 0) {
            up(dCplSolidPair, cSolidPair.data(), cSolidPair.size() * 4);
            up(dCplFluidPair, cFluidPair.data(), cFluidPair.size() * 4);
            up(dCplSOff, cSOff.data(), cSOff.size() * 4);
            up(dCplFOff, cFOff.data(), cFOff.size() * 4);
            up(dCplN, cN.data(), cN.size() * 8);
            up(dCplNM, cNM.data(), cNM.size() * 8);
        }
    }
    {
        // [0] raised on a non-finite displacement, [1] the first step
        // after which it was up
        const int flag0[2] = {0, -1};
        dFlag.alloc(sizeof(flag0));
        dFlag.upload(flag0, sizeof(flag0), st);
    }
    dStepCtr.alloc(2 * sizeof(int));
    dStepCtr.zero(st);
    uploadReal();
    hip_check(hipStreamSynchronize(st), "Domain::finalize");
    mFinalized = true;
    autotune();
}

void DomainBase::recordOutput(int isample, bool dev) {
    requireFinalized("Domain::recordOutput");
    if (mNumRec == 0) return;
    if (!dev && (isample < 0 || isample >= hMaxSamples))
        throw Error("Domain::recordOutput || Sample index out of range.");
    tick(AXS3D_KERNEL_RECORD);
    launch_record(mStream, mNumRec, dRecElem.as<int32_t>(),
                  dRecW.as<double>(), dRecPhi.as<double>(),
                  dElemPoints.as<int32_t>(), dElemNu1.as<int32_t>(),
                  dPointOff.as<uint32_t>(), dPointNu1.as<int32_t>(),
                  dDispl.as<double>(),
                  dSeismo.as<double>() +
                      (dev ? 0 : (size_t)isample * mNumRec * 3),
                  dev ? dStepCtr.as<int>() : nullptr, hMaxSamples,
                  hRecRot.empty() ? nullptr : dRecRot.as<double>());
    tock(AXS3D_KERNEL_RECORD);
}

void DomainBase::checkStability(int tstep) {
    requireFinalized("Domain::checkStability");
    int flag[2] = {0, -1};
    hip_check(hipMemcpyAsync(flag, dFlag.as<int>(), sizeof(flag),
                             hipMemcpyDeviceToHost, mStream),
              "Domain::checkStability");
    hip_check(hipStreamSynchronize(mStream), "Domain::checkStability");
    if (flag[0]) {
        // the reference checks at the start of each step and names it
        // from 1 (NewmarkTimeScheme.cpp:72-74 passes tstep + 1): the first
        // step that started from a non-finite displacement, or, when only
        // the batch's last update left one, the caller's step (the steps
        // of a batch after the blow-up ran on and are void)
        std::ostringstream ss;
        ss << "Domain::checkStability || Simulation has blown up, with dt = "
           << mDt << " s. || When the instability occurred: || "
           << "* current step  =  " << (flag[1] >= 0 ? flag[1] + 1 : tstep);
        throw Error(ss.str());
    }
}

// NewmarkTimeScheme::solve (NewmarkTimeScheme.cpp:41-108).  Stability is
// checked from the device flag at the end of the batch (the point kernel
// raises it on any non-finite displacement).  By default the steps are
// replays of one captured step graph, whose source and receiver kernels
// read the step from the device counter set here; with profiling on, or
// useGraph(false), every step is launched kernel by kernel.
void DomainBase::solve(int step_begin, int nsteps) {
    requireFinalized("NewmarkTimeScheme::solve");
    if (nsteps > 0 && hSrcElem >= 0 &&
        (step_begin < 0 || step_begin + nsteps > mStfSteps))
        throw Error("Domain::applySources || Step outside the STF.");
    if (nsteps > 0 && mUseGraph && !mProfile) {
        if (!mGraphExec) captureStep();
        launch_step_set(mStream, dStepCtr.as<int>(), step_begin,
                        mRecordStep0);
        for (int i = 0; i < nsteps; i++)
            hip_check(hipGraphLaunch(mGraphExec, mStream),
                      "NewmarkTimeScheme::solve");
    } else {
        for (int tstep = step_begin; tstep < step_begin + nsteps; tstep++) {
            int isample = -1;
            if (mNumRec > 0) {
                isample = tstep - mRecordStep0;
                if (isample < 0 || isample >= hMaxSamples) isample = -1;
            }
            step(tstep, isample, false);
        }
    }
    checkStability(step_begin + nsteps + 1);
}

// Captures one step in its device-counter form.  The captured kernels take
// only device pointers fixed at finalize, the time step and the launch
// shapes, so the graph stays valid until dt changes (setDt drops it).
void DomainBase::captureStep() {
    hip_check(hipStreamBeginCapture(mStream, hipStreamCaptureModeThreadLocal),
              "NewmarkTimeScheme::solve");
    try {
        step(0, -1, true);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(mStream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    hip_check(hipStreamEndCapture(mStream, &mGraph),
              "NewmarkTimeScheme::solve");
    const hipError_t rc =
        hipGraphInstantiate(&mGraphExec, mGraph, nullptr, nullptr, 0);
    if (rc != hipSuccess) {
        mGraphExec = nullptr;
        dropGraph();
        hip_check(rc, "NewmarkTimeScheme::solve");
    }
}

// ------------------------------------------------------------ adjoint
void DomainBase::adjointCheck(bool media) const {
    requireFinalized("Domain::adjoint");
    bool iso = true;
    for (uint8_t t : hMType) iso = iso && t == 0;
    // the media form takes transversely isotropic and anisotropic solids
    if (media) iso = true;
    bool rotated = false;
    for (size_t i = 0; i < hRecRot.size(); i += 3)
        rotated = rotated || hRecRot[i] >= 0.;
    if (mNumFElems > 0 || mNumMass3D > 0 || !iso || !hInstDt.empty() ||
        !mAbcS.point.empty() || rotated || mScheme != 0 ||
        !hOcean1DPts.empty() || !hPrtElems.empty())
        throw Error("Domain::adjoint || The adjoint run supports solid "
                    "elements with an isotropic elastic medium (1D or 3D, "
                    "with or without attenuation) and no fluid, 3D mass, "
                    "ocean load, batched instances, absorbing boundaries, "
                    "rotated station frames, undulated geometry (PRT) or "
                    "the symplectic scheme.");
}

void DomainBase::adjointOperands() {
    if (dAdjGQ.bytes() > 0) return;
    dAdjGQ.alloc(hElemGQ.size() * sizeof(double));
    dAdjGQ.upload(hElemGQ.data(), hElemGQ.size() * sizeof(double), mStream);
    std::vector<double> gt(2 * NPEM);
    for (int k = 0; k < NPEM; k++) {
        gt[k] = hGll[k];
        gt[NPEM + k] = hGlj[k];
    }
    dAdjGTab.alloc(gt.size() * sizeof(double));
    dAdjGTab.upload(gt.data(), gt.size() * sizeof(double), mStream);
    if (mNum3D > 0) {
        dAdjKind.alloc(hElemKind.size());
        dAdjKind.upload(hElemKind.data(), hElemKind.size(), mStream);
    }
    // the media form: per element type and parameter offset; the FP64
    // sin, cos of theta and 2 theta for TI and anisotropic elements on
    // spherical meshes
    hAdjMatOff.assign(mNumElems, 0);
    int64_t off = 0;
    for (int e = 0; e < mNumElems; e++) {
        hAdjMatOff[e] = off;
        off += (int64_t)mat_nparams(hMType[e]) * NPEM;
    }
    mAdjMatTotal = off;
    dAdjMType.alloc(hMType.size());
    dAdjMType.upload(hMType.data(), hMType.size(), mStream);
    dAdjMatOff.alloc(hAdjMatOff.size() * 8);
    dAdjMatOff.upload(hAdjMatOff.data(), hAdjMatOff.size() * 8, mStream);
    if (!hTheta.empty()) {
        std::vector<double> rot((size_t)mNumElems * 4 * NPEM);
        for (int e = 0; e < mNumElems; e++)
            for (int k = 0; k < NPEM; k++) {
                const double t = hTheta[(size_t)e * NPEM + k];
                double *r = rot.data() + (size_t)e * 4 * NPEM + k;
                r[0] = std::sin(t);
                r[NPEM] = std::cos(t);
                r[2 * NPEM] = std::sin(2. * t);
                r[3 * NPEM] = std::cos(2. * t);
            }
        dAdjRot.alloc(rot.size() * sizeof(double));
        dAdjRot.upload(rot.data(), rot.size() * sizeof(double), mStream);
    }
}

SensArgs DomainBase::sensArgs(const double *psi, const double *u,
                              const double *a, double *kern,
                              double scale) const {
    SensArgs s;
    s.n_elem = mNumElems;
    s.elem_points = dElemPoints.as<int32_t>();
    s.elem_axial = dElemAxial.as<uint8_t>();
    s.gq = dAdjGQ.as<double>();
    s.elem_nu1 = dElemNu1.as<int32_t>();
    s.point_off = dPointOff.as<uint32_t>();
    s.point_nu1 = dPointNu1.as<int32_t>();
    s.gtab = dAdjGTab.as<double>();
    s.psi = psi;
    s.u = u;
    s.a = a;
    s.kern = kern;
    s.scale = scale;
    s.elem_kind = mNum3D > 0 ? dAdjKind.as<uint8_t>() : nullptr;
    return s;
}

void DomainBase::adjointDensity(const double *psi, const double *u,
                                const double *a, double *kern) {
    adjointCheck();
    adjointOperands();
    const size_t nf = (size_t)mNumModes * 6 * sizeof(double);
    const size_t nk = (size_t)mNumElems * NPEM * 3 * sizeof(double);
    DeviceBuffer bp, bu, ba, bk;
    bp.alloc(nf);
    bu.alloc(nf);
    ba.alloc(nf);
    bk.alloc(nk);
    bp.upload(psi, nf, mStream);
    bu.upload(u, nf, mStream);
    ba.upload(a, nf, mStream);
    bk.zero(mStream);
    launch_sensitivity(mStream, sensArgs(bp.as<double>(), bu.as<double>(),
                                         ba.as<double>(), bk.as<double>(),
                                         1.0));
    hip_check(hipMemcpyAsync(kern, bk.as<double>(), nk,
                             hipMemcpyDeviceToHost, mStream),
              "Domain::adjointDensity");
    hip_check(hipStreamSynchronize(mStream), "Domain::adjointDensity");
}

void DomainBase::swapAdjointFields() {
    dDispl.swap(dAdjU);
    dVeloc.swap(dAdjV);
    dAccel.swap(dAdjA);
    for (int i = 0; i < 4; i++) attState(i)->swap(dAdjAtt[i]);
}

void DomainBase::dropGraph() {
    if (mGraphExec) (void)hipGraphExecDestroy(mGraphExec);
    if (mGraph) (void)hipGraphDestroy(mGraph);
    mGraphExec = nullptr;
    mGraph = nullptr;
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

    StiffArgs<Real> stiffArgs() const {
        StiffArgs<Real> a;
        a.n_em = mNumEM;
        a.em_list = dEMList.as<EMItem>();
        a.elem_points = dElemPoints.as<int32_t>();
        a.elem_axial = dElemAxial.as<uint8_t>();
        a.elem_ops = dElemOps.as<Real>();
        a.elem_mat = dElemMat.as<Real>();
        a.mat_off = dMatOff.as<uint32_t>();
        a.rot = hTheta.empty() ? nullptr : dRot.as<Real>();
        a.prt_idx = hPrtElems.empty() ? nullptr : dPrtIdx.as<int32_t>();
        a.prt = hPrtElems.empty() ? nullptr : dPrt.as<Real>();
        for (int k = 0; k < 4; k++) a.em_type_begin[k] = mEMTypeBegin[k];
        for (int k = 0; k < 3; k++) a.em_ax_begin[k] = mEMAxBegin[k];
        a.point_off = dPointOff.as<uint32_t>();
        a.point_nu1 = dPointNu1.as<int32_t>();
        a.displ = dDispl.as<double>();
        a.elem_eb_off = dElemEBOff.as<uint32_t>();
        a.elem_nu1 = dElemNu1.as<int32_t>();
        a.eb = dEB.as<Real>();
        a.gtab = dGTab.as<Real>();
        a.nsls = (mNsls > 0 && mNumEM > 0) ? mNsls : 0;
        a.npa = mAttCG4 ? 4 : NPEM;
        a.abg = dAttABG.as<double>();
        a.att1d = dAtt1D.as<Real>();
        a.mv_fr = dMemVarFR.as<double>();
        a.ds_fr = dDStressFR.as<Real>();
        return a;
    }
    CardArgs<Real> cardArgs() const {
        CardArgs<Real> c;
        c.s = stiffArgs();
        c.s.em_list = dEM3D.as<EMItem>();
        c.em_ax = dEM3DAx.as<EMItem>();
        c.n3d = mNum3D;
        c.list3d = dList3D.as<int32_t>();
        c.elem_nr = dElemNr.as<int32_t>();
        c.slice_off = dSliceOff.as<int64_t>();
        c.prop3d = dProp3D.as<Real>();
        c.prop_off = dPropOff.as<int64_t>();
        for (int k = 0; k < 4; k++) c.type_begin[k] = mTypeBegin3D[k];
        c.twiddle = dTwiddle.as<cpx<Real>>();
        c.tw_off = dTwOff.as<int64_t>();
        c.plan = dPlan.as<uint32_t>();
        c.slice_pos = dPosTab.as<uint16_t>();
        c.pos_off = dPosOff.as<int64_t>();
        c.scratch = dCardScratch.as<Real>();
        c.scr_off = dScrOff.as<uint32_t>();
        c.nsls = mNsls;
        c.npa = mAttCG4 ? 4 : NPEM;
        c.abg = dAttABG.as<double>();
        c.att3d = dAtt3D.as<Real>();
        c.memvar = dMemVar.as<double>();
        c.dstress = dDStress.as<Real>();
        return c;
    }
    PointGeom pointGeom() const {
        PointGeom g;
        g.n_modes = mNumModes;
        g.mode_point = dModePoint.as<int32_t>();
        g.point_off = dPointOff.as<uint32_t>();
        g.point_nr = dPointNr.as<int32_t>();
        g.point_axial = dPointAxial.as<uint8_t>();
        g.inv_mass = dInvMass.as<double>();
        g.mass3d = mNumMass3D > 0 ? dMass3DFlag.as<uint8_t>() : nullptr;
        if (!hOcean1DPts.empty()) {
            g.ocean_idx = dOceanIdx.as<int32_t>();
            g.ocean1d = dOcean1D.as<double>();
        }
        g.adj = dAdj.as<uint32_t>();
        if (!hInstDt.empty()) g.point_dt = dPointDt.as<double>();
        g.abc = abcGeom(mAbcSD);
        g.scheme = mScheme;
        g.sym_pi = mSymPi;
        g.sym_kappa = mSymKappa;
        return g;
    }
    // absorbing boundaries: what the point kernels read, the terms' kernel
    AbcGeom abcGeom(const AbcDev &dv) const {
        AbcGeom g;
        if (dv.n == 0) return g;
        g.idx = dv.idx.as<int32_t>();
        g.term_off = dv.termoff.as<int64_t>();
        g.sp3_off = dv.sp3off.as<int64_t>();
        g.cterm = dv.cterm.as<double>();
        g.sterm = dv.sterm.as<double>();
        return g;
    }
    AbcArgs<Real> abcArgs(const AbcDev &dv) const {
        const bool fl = &dv == &mAbcFD;
        AbcArgs<Real> a;
        a.n = dv.n;
        a.nr_max = dv.nr_max;
        a.point = dv.point.as<int32_t>();
        a.point_off = fl ? dFPointOff.as<uint32_t>() : dPointOff.as<uint32_t>();
        a.point_nr = fl ? dFPointNr.as<int32_t>() : dPointNr.as<int32_t>();
        a.displ = fl ? dChi.as<double>() : dDispl.as<double>();
        a.veloc = fl ? dChiV.as<double>() : dVeloc.as<double>();
        a.cl1 = dv.cl1.as<double>();
        a.cl3_off = dv.cl3off.as<int64_t>();
        a.cl3 = dv.cl3.as<Real>();
        a.sp1 = dv.sp1.as<double>();
        a.sp3_off = dv.sp3off.as<int64_t>();
        a.sp3 = dv.sp3.as<Real>();
        a.term_off = dv.termoff.as<int64_t>();
        a.cterm = dv.cterm.as<double>();
        a.sterm = dv.sterm.as<double>();
        a.twiddle = dTwiddle.as<cpx<Real>>();
        a.tw_off = dTwOff.as<int64_t>();
        a.plan = dPlan.as<uint32_t>();
        a.slice_pos = dPosTab.as<uint16_t>();
        a.pos_off = dPosOff.as<int64_t>();
        return a;
    }
    // the Clayton and sponge terms of this step, from the displacement and
    // velocity it starts from (before any point kernel of the step)
    void abcKernel() {
        if (mAbcSD.n > 0) launch_abc<Real, 3>(mStream, abcArgs(mAbcSD));
        if (mAbcFD.n > 0) launch_abc<Real, 1>(mStream, abcArgs(mAbcFD));
    }
    // the per-slice parts in the working precision, in record order (the
    // offsets abcUp gave them)
    void abcReal(const AbcHost &h, AbcDev &dv) {
        if (dv.n == 0) return;
        std::vector<Real> c3, s3;
        for (size_t r = 0; r < h.point.size(); r++) {
            for (double x : h.cl3[r]) c3.push_back((Real)x);
            for (double x : h.sp3[r]) s3.push_back((Real)x);
        }
        dv.cl3.alloc(std::max<size_t>(c3.size(), 1) * sizeof(Real));
        if (!c3.empty())
            dv.cl3.upload(c3.data(), c3.size() * sizeof(Real), mStream);
        dv.sp3.alloc(std::max<size_t>(s3.size(), 1) * sizeof(Real));
        if (!s3.empty())
            dv.sp3.upload(s3.data(), s3.size() * sizeof(Real), mStream);
    }
    Mass3DArgs<Real> mass3dArgs() const {
        Mass3DArgs<Real> m;
        m.n = mNumMass3D;
        m.nr_max = mMass3DNrMax;
        m.point = dMass3DPts.as<int32_t>();
        m.inv_off = dMass3DOff.as<int64_t>();
        if (!hMass3DK.empty()) {
            m.k_off = dMass3DKOff.as<int64_t>();
            m.ocean_k = dMass3DK.as<Real>();
        }
        m.inv_mass = dInvMass3D.as<Real>();
        m.point_off = dPointOff.as<uint32_t>();
        m.point_nr = dPointNr.as<int32_t>();
        m.twiddle = dTwiddle.as<cpx<Real>>();
        m.tw_off = dTwOff.as<int64_t>();
        m.plan = dPlan.as<uint32_t>();
        m.slice_pos = dPosTab.as<uint16_t>();
        m.pos_off = dPosOff.as<int64_t>();
        m.displ = dDispl.as<double>();
        m.veloc = dVeloc.as<double>();
        m.accel = dAccel.as<double>();
        m.dt = mDt;
        m.flag = dFlag.as<int>();
        m.abc = abcGeom(mAbcSD);
        m.scheme = mScheme;
        m.sym_pi = mSymPi;
        m.sym_kappa = mSymKappa;
        return m;
    }
    PointGeom fluidGeom() const {
        PointGeom g;
        g.n_modes = mNumFModes;
        g.mode_point = dFModePoint.as<int32_t>();
        g.point_off = dFPointOff.as<uint32_t>();
        g.point_nr = dFPointNr.as<int32_t>();
        g.point_axial = dFPointBC.as<uint8_t>();
        g.inv_mass = dFInvMass.as<double>();
        g.mass3d = mNumFDefer > 0 ? dFDeferFlag.as<uint8_t>() : nullptr;
        g.adj = dFAdj.as<uint32_t>();
        g.abc = abcGeom(mAbcFD);
        g.scheme = mScheme;
        g.sym_pi = mSymPi;
        g.sym_kappa = mSymKappa;
        return g;
    }
    FluidDeferArgs<Real> fluidDeferArgs() const {
        FluidDeferArgs<Real> f;
        f.n = mNumFDefer;
        f.nr_max = mFDeferNrMax;
        f.point = dFDeferPts.as<int32_t>();
        f.pair3d = dFDeferPair.as<int32_t>();
        f.minv_off = dFDeferMOff.as<int64_t>();
        f.minv = dFInvMass3D.as<Real>();
        f.inv_mass = dFInvMass.as<double>();
        f.point_off = dFPointOff.as<uint32_t>();
        f.point_nr = dFPointNr.as<int32_t>();
        f.point_bc = dFPointBC.as<uint8_t>();
        f.fstiff = dFStiff.as<Real>();
        f.chi = dChi.as<double>();
        f.chi_v = dChiV.as<double>();
        f.chi_a = dChiA.as<double>();
        f.dt = mDt;
        f.flag = dFlag.as<int>();
        f.pair_solid = dCpl3DSolid.as<int32_t>();
        f.n_off = dCpl3DNOff.as<int64_t>();
        f.nu = dCpl3DNu.as<Real>();
        f.nam = dCpl3DNam.as<Real>();
        f.solid_off = dPointOff.as<uint32_t>();
        f.solid_displ = dDispl.as<double>();
        f.csolid3d = dCSolid3D.as<double>();
        f.c3off = dCpl3DCOff.as<uint32_t>();
        f.twiddle = dTwiddle.as<cpx<Real>>();
        f.tw_off = dTwOff.as<int64_t>();
        f.plan = dPlan.as<uint32_t>();
        f.slice_pos = dPosTab.as<uint16_t>();
        f.pos_off = dPosOff.as<int64_t>();
        f.abc = abcGeom(mAbcFD);
        f.scheme = mScheme;
        f.sym_pi = mSymPi;
        f.sym_kappa = mSymKappa;
        return f;
    }
    CouplingGeom couplingGeom() const {
        CouplingGeom c;
        c.n_pairs = mNumFElems > 0 ? mNumPairs : 0;
        c.solid_pair = dCplSolidPair.as<int32_t>();
        c.fluid_pair = dCplFluidPair.as<int32_t>();
        c.soff = dCplSOff.as<uint32_t>();
        c.foff = dCplFOff.as<uint32_t>();
        c.n = dCplN.as<double>();
        c.n_m = dCplNM.as<double>();
        c.cstiff = dCStiff.as<double>();
        if (mNumCpl3D > 0) {
            c.solid_pair3d = dCplSolidPair3D.as<int32_t>();
            c.c3off = dCpl3DCOff.as<uint32_t>();
            c.csolid3d = dCSolid3D.as<double>();
        }
        return c;
    }
    FluidArgs<Real> fluidArgs() const {
        FluidArgs<Real> f;
        f.n_em = mNumFEM;
        f.em_list = dFEMList.as<EMItem>();
        f.elem_points = dFElemPoints.as<int32_t>();
        f.elem_axial = dFElemAxial.as<uint8_t>();
        f.elem_ops = dFElemOps.as<Real>();
        f.elem_k = dFElemK.as<Real>();
        f.point_off = dFPointOff.as<uint32_t>();
        f.point_nu1 = dFPointNu1.as<int32_t>();
        f.chi = dChi.as<double>();
        f.elem_eb_off = dFEBOff.as<uint32_t>();
        f.elem_nu1 = dFElemNu1.as<int32_t>();
        f.eb = dFEB.as<Real>();
        f.gtab = dGTab.as<Real>();
        return f;
    }

    FluidCardArgs<Real> fluidCardArgs() const {
        FluidCardArgs<Real> c;
        c.f = fluidArgs();
        c.f.em_list = dFEM3D.as<EMItem>();
        c.list3d = dFList3D.as<int32_t>();
        c.elem_nr = dFElemNr.as<int32_t>();
        c.k3d = dFK3D.as<Real>();
        c.k3d_off = dFK3DOff.as<int64_t>();
        c.twiddle = dTwiddle.as<cpx<Real>>();
        c.tw_off = dTwOff.as<int64_t>();
        c.plan = dPlan.as<uint32_t>();
        c.slice_pos = dPosTab.as<uint16_t>();
        c.pos_off = dPosOff.as<int64_t>();
        c.scratch = dCardScratch.as<Real>();
        c.scr_off = dFScrOff.as<uint32_t>();
        return c;
    }

    // Domain::applySources: the source force is subtracted from the point
    // stiffness (stage form).
    void applySources(int tstep) override {
        requireFinalized("Domain::applySources");
        sourceKernel(tstep, false);
    }

    void sourceKernel(int tstep, bool toElem, bool dev = false) {
        if (hSrcElem < 0) return;
        if (!dev && (tstep < 0 || tstep >= mStfSteps))
            throw Error("Domain::applySources || Step outside the STF.");
        tick(AXS3D_KERNEL_SOURCE);
        const StiffArgs<Real> sa = stiffArgs();
        for (size_t i = 0; i < hSrcElems.size(); i++) {
            // the source element's PRT operands (X0..X3 unscaled), if any
            const Real *prt_e = nullptr;
            for (size_t k = 0; k < hPrtElems.size(); k++)
                if (hPrtElems[k] == hSrcElems[i])
                    prt_e = dPrt.as<Real>() + k * 8 * NPEM;
            launch_moment_source<Real>(
                mStream, toElem, sa, hSrcElems[i], mSrcNu1c[i],
                dSrcPattern.as<Real>() + 2 * hSrcPatOff[i],
                dStf.as<double>() + i * (size_t)mStfSteps, tstep, mStfSteps,
                dev ? dStepCtr.as<int>() : nullptr, dStiff.as<Real>(), prt_e);
        }
        tock(AXS3D_KERNEL_SOURCE);
    }

    // The cardinal path, batch by batch; with timing, each of the three
    // kernels is timed under its own id.
    void cardinalPath(bool timed) {
        cardinalPathOn(cardArgs(), mCardBatches, timed);
    }
    void cardinalPathOn(const CardArgs<Real> &c,
                        const std::vector<CardBatch> &batches, bool timed) {
        for (const CardBatch &bt : batches) {
            if (timed) tick(AXS3D_KERNEL_CD_GRAD);
            launch_cd_grad<Real>(mStream, c, bt);
            if (timed) {
                tock(AXS3D_KERNEL_CD_GRAD);
                tick(AXS3D_KERNEL_CD_STRESS);
            }
            launch_cd_stress<Real>(mStream, c, bt);
            if (timed) {
                tock(AXS3D_KERNEL_CD_STRESS);
                tick(AXS3D_KERNEL_CD_QUAD);
            }
            launch_cd_quad<Real>(mStream, c, bt);
            if (timed) tock(AXS3D_KERNEL_CD_QUAD);
        }
    }

    // The element kernels; with stat, the cardinal paths run through the
    // static scratch in the batches planned for it.
    void elementKernel(bool stat = false) {
        if (mNumEM > 0) {
            tick(AXS3D_KERNEL_STIFFNESS);
            launch_solid_stiff_fr<Real>(mStream, stiffArgs());
            tock(AXS3D_KERNEL_STIFFNESS);
        }
        if (mNum3D > 0) {
            tick(AXS3D_KERNEL_STIFFNESS_3D);
            if (stat) {
                CardArgs<Real> c = cardArgs();
                c.scratch = static_cast<Real *>(mStaticScratch);
                c.scr_off = dScrOffS.as<uint32_t>();
                cardinalPathOn(c, mCardBatchesS, true);
            } else {
                cardinalPath(true);
            }
            tock(AXS3D_KERNEL_STIFFNESS_3D);
        }
        if (mNumFEM > 0 || mNumF3D > 0) {
            tick(AXS3D_KERNEL_FLUID_STIFFNESS);
            if (mNumFEM > 0)
                launch_fluid_stiff_fr<Real>(mStream, fluidArgs());
            if (mNumF3D > 0) {
                FluidCardArgs<Real> fc = fluidCardArgs();
                if (stat) {
                    fc.scratch = static_cast<Real *>(mStaticScratch);
                    fc.scr_off = dFScrOffS.as<uint32_t>();
                }
                for (const CardBatch &bt : stat ? mFCardBatchesS
                                                : mFCardBatches)
                    launch_cdf_batch<Real>(mStream, fc, bt);
            }
            tock(AXS3D_KERNEL_FLUID_STIFFNESS);
        }
    }

    void computeStiffnessStatic() override {
        requireFinalized("Domain::computeStiffnessStatic");
        if (mStaticNeedMax > STATIC_SCRATCH_BYTES)
            throw Error("Domain::computeStiffnessStatic || An element needs " +
                        std::to_string(mStaticNeedMax) +
                        " bytes of cardinal scratch; the static scratch "
                        "holds " +
                        std::to_string(STATIC_SCRATCH_BYTES) + ".");
        elementKernel(true);
        launch_assemble<Real>(mStream, pointGeom(), dEB.as<Real>(),
                              dStiff.as<Real>());
        if (mNumFElems > 0)
            launch_fluid_assemble<Real>(mStream, fluidGeom(), dFEB.as<Real>(),
                                        dFStiff.as<Real>());
    }

    // Domain::computeStiffness (stage form): element kernels, then the
    // points subtract the element-local forces.
    void computeStiffness() override {
        requireFinalized("Domain::computeStiffness");
        elementKernel();
        launch_assemble<Real>(mStream, pointGeom(), dEB.as<Real>(),
                              dStiff.as<Real>());
        if (mNumFElems > 0)
            launch_fluid_assemble<Real>(mStream, fluidGeom(), dFEB.as<Real>(),
                                        dFStiff.as<Real>());
    }

    void point(int phases) {
        tick(AXS3D_KERNEL_POINT);
        launch_point<Real>(mStream, phases, pointGeom(), dEB.as<Real>(),
                           dStiff.as<Real>(), dDispl.as<double>(),
                           dVeloc.as<double>(), dAccel.as<double>(), mDt,
                           dFlag.as<int>(), couplingGeom());
        // points with a 3D mass: the point kernel left their assembled
        // stiffness; divide slice by slice (and march)
        if ((phases & PHASE_MASS) && mNumMass3D > 0)
            launch_mass3d<Real>(mStream, (phases & PHASE_MARCH) != 0,
                                mass3dArgs(), dStiff.as<Real>());
        tock(AXS3D_KERNEL_POINT);
    }
    // fluid points; runs before the solid points of the same phase (the
    // coupling reads the solid displacement before it is marched and
    // leaves the fluid stiffness the solid points subtract)
    void fluidPoint(int phases) {
        if (mNumFElems <= 0) return;
        tick(AXS3D_KERNEL_FLUID_POINT);
        launch_fluid_point<Real>(mStream, phases, fluidGeom(), dFEB.as<Real>(),
                                 dFStiff.as<Real>(), dChi.as<double>(),
                                 dChiV.as<double>(), dChiA.as<double>(), mDt,
                                 dFlag.as<int>(), couplingGeom(),
                                 dDispl.as<double>());
        if (mNumFDefer > 0) {
            int ph = phases & ~PHASE_GATHER;
            if (mNumPairs <= 0) ph &= ~PHASE_COUPLE;
            launch_fluid_defer<Real>(mStream, ph, fluidDeferArgs());
        }
        tock(AXS3D_KERNEL_FLUID_POINT);
    }
    // SolidFluidBoundary::apply (Domain.cpp:313-318)
    // AbsorbingBoundary::applyClayton, then SolidFluidBoundary::apply
    // (Domain.cpp:313-318)
    void applyBC_BeforeAssemblingStiff() override {
        requireFinalized("Domain::applyBC_BeforeAssemblingStiff");
        const bool abc = mAbcSD.n > 0 || mAbcFD.n > 0;
        if (abc) abcKernel();
        if (!abc && (mNumFElems <= 0 || mNumPairs <= 0)) return;
        fluidPoint(PHASE_COUPLE);
        point(PHASE_COUPLE);
    }
    // AbsorbingBoundary::applySponge (Domain.cpp:330-332), on the
    // accelerations computeStiffToAccel left, with the terms of this
    // step's applyBC_BeforeAssemblingStiff
    void applyBC_AfterComputingAccel() override {
        requireFinalized("Domain::applyBC_AfterComputingAccel");
        point(PHASE_SPONGE);
        fluidPoint(PHASE_SPONGE);
    }
    void applyBC_AfterAssemblingStiff() override {
        requireFinalized("Domain::applyBC_AfterAssemblingStiff");
        point(PHASE_AXIAL);
        fluidPoint(PHASE_AXIAL);
    }
    void computeStiffToAccel() override {
        requireFinalized("Domain::computeStiffToAccel");
        point(PHASE_MASS);
        fluidPoint(PHASE_MASS);
    }
    void newmark() override {
        requireFinalized("NewmarkTimeScheme::update");
        if (mScheme != 0)
            throw Error("NewmarkTimeScheme::update || The domain marches "
                        "with the symplectic scheme.");
        point(PHASE_MARCH);
        fluidPoint(PHASE_MARCH);
    }

    // SymplecticTimeScheme (SymplecticTimeScheme.hpp:81-100): the
    // position-extended Forest-Ruth-like coefficients of Omelyan, Mryglod
    // and Folk (2002), kappa_0..4 for the positions and pi_1..4 for the
    // velocities
    static constexpr double kSymA = 0.1786178958448091;
    static constexpr double kSymB = -0.2123418310626054;
    static constexpr double kSymG = -0.06626458266981849;
    static double symKappa(int i) {
        const double k[5] = {kSymA, kSymG, 1. - 2. * (kSymG + kSymA), kSymG,
                             kSymA};
        return k[i];
    }
    static double symPi(int i) {
        const double p[5] = {0., 0.5 - kSymB, kSymB, kSymB, 0.5 - kSymB};
        return p[i];
    }
    void symplecticLaunch() override {
        requireFinalized("SymplecticTimeScheme::launch");
        if (mScheme != 1)
            throw Error("SymplecticTimeScheme::launch || The domain marches "
                        "with the Newmark scheme.");
        const double kdt = symKappa(0) * mDt;
        launch_sym_launch(mStream, dDispl.as<double>(), dVeloc.as<double>(),
                          mNumModes * 6, kdt);
        if (mNumFElems > 0)
            launch_sym_launch(mStream, dChi.as<double>(), dChiV.as<double>(),
                              mNumFModes * 2, kdt);
    }
    void symplecticUpdate(int isub) override {
        requireFinalized("SymplecticTimeScheme::update");
        if (mScheme != 1)
            throw Error("SymplecticTimeScheme::update || The domain marches "
                        "with the Newmark scheme.");
        if (isub < 1 || isub > 4)
            throw Error("SymplecticTimeScheme::update || Sub-step out of "
                        "range (1..4).");
        mSymPi = symPi(isub);
        mSymKappa = symKappa(isub);
        point(PHASE_MARCH);
        fluidPoint(PHASE_MARCH);
    }

    // One Newmark step in its bulk form: the element kernels write the
    // element-local forces, the source adds into its element's, the
    // receivers sample u(t), then one fluid and one solid point kernel
    // gather, couple, mask, divide by the mass and march.
    void step(int tstep, int isample, bool dev) override {
        const int all = PHASE_GATHER | PHASE_AXIAL | PHASE_MASS | PHASE_MARCH;
        if (mScheme == 1) {
            // SymplecticTimeScheme::solve (SymplecticTimeScheme.cpp:30-
            // 103): record, source (into the first sub-step's forces),
            // launch, then four sub-steps of stiffness, boundaries, mass
            // and update
            if (dev || isample >= 0) recordOutput(isample, dev);
            symplecticLaunch();
            for (int isub = 1; isub <= 4; isub++) {
                elementKernel();
                if (isub == 1) sourceKernel(tstep, true, dev);
                abcKernel();
                mSymPi = symPi(isub);
                mSymKappa = symKappa(isub);
                fluidPoint(all | PHASE_COUPLE);
                point(all | PHASE_COUPLE);
            }
        } else {
            elementKernel();
            sourceKernel(tstep, true, dev);
            if (dev || isample >= 0) recordOutput(isample, dev);
            abcKernel();
            fluidPoint(all | PHASE_COUPLE);
            point(all | PHASE_COUPLE);
        }
        if (dev)
            launch_step_advance(mStream, dStepCtr.as<int>(), dFlag.as<int>());
        else
            launch_flag_note(mStream, dFlag.as<int>(), tstep);
    }

    // Discrete adjoint of the bulk step (see axs3d_adjoint.hip for the
    // equations).  Forward from rest with checkpoints of (u, v, a) every
    // seg steps; then, segment by segment from the last, the forward
    // fields of the segment are recomputed from its checkpoint (u_n before
    // and a_n after each step kept on the device), and the adjoint field
    // (the same fields, swapped in) steps backwards through the segment:
    // densities of (psi_n, u_n, a_n) scaled by -dt^2 into the kernels,
    // then one step with the adjoint force instead of the source.
    void adjointRun(int nsteps, int seg, const double *djdr, double *kern,
                    double *kern3d) override {
        adjointImpl(nsteps, seg, djdr, false, kern, nullptr, kern3d);
    }
    void adjointRunMedia(int nsteps, int seg, const double *djdr,
                         double *kern_par, double *kern_rho,
                         double *kern_par3d) override {
        adjointImpl(nsteps, seg, djdr, true, kern_par, kern_rho, kern_par3d);
    }
    // media = false: kern [e][25][3] (lambda, mu, rho), kern3d [slice][2]
    // [25]; media = true: kern the packed parameters, kern_rho [e][25],
    // kern3d the packed per-slice parameters
    void adjointImpl(int nsteps, int seg, const double *djdr, bool media,
                     double *kern, double *kern_rho, double *kern3d) {
        adjointCheck(media);
        if (mNum3D > 0 && !kern3d)
            throw Error("Domain::adjoint || A domain with 3D elements needs "
                        "the per-slice output.");
        if (media && (!kern || !kern_rho))
            throw Error("Domain::adjoint || The media form needs the "
                        "parameter and density outputs.");
        if (nsteps <= 0 || seg <= 0)
            throw Error("Domain::adjoint || The number of steps and the "
                        "checkpoint interval must be positive.");
        if (hSrcElem >= 0 && nsteps > mStfSteps)
            throw Error("Domain::adjoint || Step outside the STF.");
        adjointOperands();
        const size_t nd = (size_t)mNumModes * 6;
        const size_t nf = nd * sizeof(double);
        const int nck = (nsteps + seg - 1) / seg;
        const int seglen = std::min(seg, nsteps);
        size_t freeb = 0, totalb = 0;
        hip_check(hipMemGetInfo(&freeb, &totalb), "Domain::adjoint");
        // the forward state of a checkpoint: u, v, a and the SLS memory
        // variables and stress increments
        size_t attBytes = 0;
        for (int i = 0; i < 4; i++) attBytes += attState(i)->bytes();
        const size_t ckBytes = 3 * nf + attBytes;
        const double need =
            (double)ckBytes * nck + (double)nf * (2.0 * seglen + 3.0) +
            (double)attBytes + (double)mCardScrMax * sizeof(cpx<Real>) +
            (double)mPropTotal * sizeof(double);
        if (need > 0.9 * (double)freeb) {
            std::ostringstream ss;
            ss << "Domain::adjoint || Checkpoints and segment fields need "
               << need / 1e9 << " GB, " << freeb / 1e9
               << " GB are free; a checkpoint interval near sqrt(nsteps) "
                  "needs the least.";
            throw Error(ss.str());
        }
        dAdjU.alloc(nf);
        dAdjV.alloc(nf);
        dAdjA.alloc(nf);
        for (int i = 0; i < 4; i++) dAdjAtt[i].alloc(attState(i)->bytes());
        dAdjCkpt.alloc(ckBytes * nck);
        dAdjSegU.alloc(nf * seglen);
        dAdjSegA.alloc(nf * seglen);
        const size_t ng = (size_t)nsteps * mNumRec * 3;
        dAdjG.alloc(std::max<size_t>(ng, 1) * sizeof(double));
        if (ng > 0) dAdjG.upload(djdr, ng * sizeof(double), mStream);
        const size_t nk = media ? (size_t)mAdjMatTotal
                                : (size_t)mNumElems * NPEM * 3;
        dAdjKern.alloc(nk * sizeof(double));
        dAdjKern.zero(mStream);
        const size_t nkr = (size_t)mNumElems * NPEM;
        if (media) {
            dAdjKernR.alloc(nkr * sizeof(double));
            dAdjKernR.zero(mStream);
        }
        const size_t nk3 = (size_t)mPropTotal;
        if (mNum3D > 0) {
            dAdjKern3D.alloc(nk3 * sizeof(double));
            dAdjKern3D.zero(mStream);
            dAdjScrB.alloc(mCardScrMax * sizeof(cpx<Real>));
        }
        const int all = PHASE_GATHER | PHASE_AXIAL | PHASE_MASS | PHASE_MARCH;
        char *ck = dAdjCkpt.as<char>();
        auto d2d = [&](double *dst, const double *src) {
            hip_check(hipMemcpyAsync(dst, src, nf, hipMemcpyDeviceToDevice,
                                     mStream),
                      "Domain::adjoint");
        };
        // forward state <-> checkpoint s
        auto save = [&](int s) {
            char *c = ck + (size_t)s * ckBytes;
            d2d(reinterpret_cast<double *>(c), dDispl.as<double>());
            d2d(reinterpret_cast<double *>(c + nf), dVeloc.as<double>());
            d2d(reinterpret_cast<double *>(c + 2 * nf), dAccel.as<double>());
            size_t o = 3 * nf;
            for (int i = 0; i < 4; i++) {
                DeviceBuffer *b = attState(i);
                if (b->bytes() > 0)
                    hip_check(hipMemcpyAsync(c + o, b->as<char>(), b->bytes(),
                                             hipMemcpyDeviceToDevice,
                                             mStream),
                              "Domain::adjoint");
                o += b->bytes();
            }
        };
        auto load = [&](int s) {
            const char *c = ck + (size_t)s * ckBytes;
            d2d(dDispl.as<double>(), reinterpret_cast<const double *>(c));
            d2d(dVeloc.as<double>(),
                reinterpret_cast<const double *>(c + nf));
            d2d(dAccel.as<double>(),
                reinterpret_cast<const double *>(c + 2 * nf));
            size_t o = 3 * nf;
            for (int i = 0; i < 4; i++) {
                DeviceBuffer *b = attState(i);
                if (b->bytes() > 0)
                    hip_check(hipMemcpyAsync(b->as<char>(), c + o, b->bytes(),
                                             hipMemcpyDeviceToDevice,
                                             mStream),
                              "Domain::adjoint");
                o += b->bytes();
            }
        };
        // forward from rest up to the last checkpoint
        dDispl.zero(mStream);
        dVeloc.zero(mStream);
        dAccel.zero(mStream);
        for (int i = 0; i < 4; i++) attState(i)->zero(mStream);
        for (int n = 0; n <= (nck - 1) * seg; n++) {
            if (n % seg == 0) save(n / seg);
            if (n < (nck - 1) * seg) step(n, -1, false);
        }
        // backwards, segment by segment; the adjoint field starts at rest
        // with its own (zero) SLS state
        dAdjU.zero(mStream);
        dAdjV.zero(mStream);
        dAdjA.zero(mStream);
        for (int i = 0; i < 4; i++) dAdjAtt[i].zero(mStream);
        const double dt2 = mDt * mDt;
        double *su = dAdjSegU.as<double>(), *sa = dAdjSegA.as<double>();
        for (int s = nck - 1; s >= 0; s--) {
            const int n0 = s * seg, n1 = std::min(n0 + seg, nsteps);
            load(s);
            for (int n = n0; n < n1; n++) {
                d2d(su + (size_t)(n - n0) * nd, dDispl.as<double>());
                step(n, -1, false);
                d2d(sa + (size_t)(n - n0) * nd, dAccel.as<double>());
            }
            swapAdjointFields();
            for (int n = n1 - 1; n >= n0; n--) {
                SensArgs sa_n = sensArgs(dDispl.as<double>(),
                                         su + (size_t)(n - n0) * nd,
                                         sa + (size_t)(n - n0) * nd,
                                         dAdjKern.as<double>(), -dt2);
                if (media) {
                    sa_n.elem_mtype = dAdjMType.as<uint8_t>();
                    sa_n.mat_off = dAdjMatOff.as<int64_t>();
                    sa_n.kernm = dAdjKern.as<double>();
                    sa_n.kernr = dAdjKernR.as<double>();
                    sa_n.rot = hTheta.empty() ? nullptr
                                              : dAdjRot.as<double>();
                }
                launch_sensitivity(mStream, sa_n);
                if (mNum3D > 0) {
                    // per-slice densities: strain rows of psi and u_n,
                    // batch by batch
                    const CardArgs<Real> ca = cardArgs();
                    CardArgs<Real> cb = ca;
                    cb.s.displ = su + (size_t)(n - n0) * nd;
                    cb.scratch = dAdjScrB.as<Real>();
                    for (const CardBatch &bt : mCardBatches) {
                        launch_cd_grad<Real>(mStream, ca, bt);
                        launch_cd_grad<Real>(mStream, cb, bt);
                        launch_cd_sens<Real>(mStream, ca, cb.scratch, bt,
                                             dAdjGQ.as<double>(),
                                             dAdjKern3D.as<double>(), -dt2);
                    }
                }
                elementKernel();
                launch_adjoint_source<Real>(
                    mStream, mNumRec, dRecElem.as<int32_t>(),
                    dRecW.as<double>(), dRecPhi.as<double>(),
                    dAdjG.as<double>() + (size_t)n * mNumRec * 3, 1.0 / dt2,
                    mRecNu1Max, dElemPoints.as<int32_t>(),
                    dElemNu1.as<int32_t>(), dPointNu1.as<int32_t>(),
                    dElemEBOff.as<uint32_t>(), dEB.as<Real>());
                point(all);
            }
            swapAdjointFields();
        }
        hip_check(hipMemcpyAsync(kern, dAdjKern.as<double>(),
                                 nk * sizeof(double), hipMemcpyDeviceToHost,
                                 mStream),
                  "Domain::adjoint");
        if (media)
            hip_check(hipMemcpyAsync(kern_rho, dAdjKernR.as<double>(),
                                     nkr * sizeof(double),
                                     hipMemcpyDeviceToHost, mStream),
                      "Domain::adjoint");
        if (mNum3D > 0)
            hip_check(hipMemcpyAsync(kern3d, dAdjKern3D.as<double>(),
                                     nk3 * sizeof(double),
                                     hipMemcpyDeviceToHost, mStream),
                      "Domain::adjoint");
        hip_check(hipStreamSynchronize(mStream), "Domain::adjoint");
        checkStability(nsteps + 1);
        // the segment buffers are the largest; give them back
        dAdjCkpt.reset();
        dAdjSegU.reset();
        dAdjSegA.reset();
        dAdjScrB.reset();
        for (int i = 0; i < 4; i++) dAdjAtt[i].reset();
    }

    void getField(int which, double *out) override {
        requireFinalized("Domain::getField");
        const bool fl = which >= AXS3D_FIELD_FLUID_DISPL;
        const int base = fl ? which - AXS3D_FIELD_FLUID_DISPL : which;
        const size_t n = fl ? (size_t)mNumFModes * 2 : (size_t)mNumModes * 6;
        if (n == 0) return;
        if (base == AXS3D_FIELD_STIFF) {
            std::vector<Real> tmp(n);
            hip_check(hipMemcpyAsync(tmp.data(),
                                     (fl ? dFStiff : dStiff).template as<Real>(),
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
        const bool fl = which >= AXS3D_FIELD_FLUID_DISPL;
        const int base = fl ? which - AXS3D_FIELD_FLUID_DISPL : which;
        const size_t n = fl ? (size_t)mNumFModes * 2 : (size_t)mNumModes * 6;
        if (n == 0) return;
        if (base == AXS3D_FIELD_STIFF) {
            std::vector<Real> tmp(in, in + n);
            hip_check(hipMemcpyAsync((fl ? dFStiff : dStiff).template as<Real>(),
                                     tmp.data(), n * sizeof(Real),
                                     hipMemcpyHostToDevice, mStream),
                      "Domain::setField");
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
        if (which < 0 || which > AXS3D_FIELD_FLUID_STIFF)
            throw Error("Domain::field || Unknown field id.");
        switch (which) {
            case AXS3D_FIELD_DISPL: return dDispl.as<double>();
            case AXS3D_FIELD_VELOC: return dVeloc.as<double>();
            case AXS3D_FIELD_ACCEL: return dAccel.as<double>();
            case AXS3D_FIELD_FLUID_DISPL:
                if (mNumFElems > 0) return dChi.as<double>();
                break;
            case AXS3D_FIELD_FLUID_VELOC:
                if (mNumFElems > 0) return dChiV.as<double>();
                break;
            case AXS3D_FIELD_FLUID_ACCEL:
                if (mNumFElems > 0) return dChiA.as<double>();
                break;
            default:
                break;
        }
        throw Error("Domain::field || Unknown field id.");
    }

    // Cardinal batches for a scratch of `bytes`: runs of one material
    // type whose strain and stress ([150][nu1] complex per element) fit
    // the scratch together, and the 3D element-modes in list order for the
    // gradient and quadrature kernels.  The scratch also holds the 3D
    // fluid batches' rows.
    // Batches of the 3D solid elements for a scratch of `bytes`, their
    // elements' scratch offsets (complex values) and the largest batch
    // (complex values); the element-mode lists do not depend on the
    // budget.
    void planCardBatches(size_t bytes, std::vector<CardBatch> &batches,
                         std::vector<uint32_t> &scrOff,
                         std::vector<EMItem> &em3d,
                         std::vector<EMItem> &emAx, size_t &scrMax) const {
        const size_t budget =
            std::min(bytes / sizeof(cpx<Real>), (size_t)UINT32_MAX);
        scrOff.assign(mNumElems, 0u);
        em3d.clear();
        emAx.clear();
        batches.clear();
        scrMax = 0;
        for (int mt = 0; mt < 3; mt++) {
            size_t used = 0;
            for (int w = mTypeBegin3D[mt]; w < mTypeBegin3D[mt + 1]; w++) {
                const int e = hList3D[w];
                const int nu1 = hElemNr[e] / 2 + 1;
                const size_t need = (size_t)6 * NPEM * scr_row(nu1);
                if (batches.empty() || batches.back().mtype != mt ||
                    used + need > budget) {
                    batches.push_back(
                        {mt, w, w, (int)em3d.size(), (int)em3d.size(),
                         (int)emAx.size(), (int)emAx.size()});
                    used = 0;
                }
                scrOff[e] = (uint32_t)used;
                used += need;
                batches.back().nr_max =
                    std::max(batches.back().nr_max, (int)hElemNr[e]);
                for (int al = 0; al < nu1; al++) em3d.push_back({e, al});
                if (hElemAxial[e])
                    for (int al = 0; al < nu1; al++) emAx.push_back({e, al});
                batches.back().end = w + 1;
                batches.back().em_end = (int)em3d.size();
                batches.back().ax_end = (int)emAx.size();
                scrMax = std::max(scrMax, used);
            }
        }
    }

    // Batches of the 3D fluid elements ([3][25][nu1] rows each) for a
    // scratch of `bytes`, as planCardBatches.
    void planFluidCardBatches(size_t bytes, std::vector<CardBatch> &batches,
                              std::vector<uint32_t> &fscr,
                              std::vector<EMItem> &fem3d,
                              size_t &scrMax) const {
        const size_t budget =
            std::min(bytes / sizeof(cpx<Real>), (size_t)UINT32_MAX);
        fscr.assign(mNumFElems, 0u);
        fem3d.clear();
        batches.clear();
        size_t used = 0;
        scrMax = 0;
        for (int w = 0; w < mNumF3D; w++) {
            const int e = hFList3D[w];
            const int nu1 = hFElemNr[e] / 2 + 1;
            const size_t need = (size_t)3 * NPEM * scr_row(nu1);
            if (batches.empty() || used + need > budget) {
                batches.push_back(
                    {0, w, w, (int)fem3d.size(), (int)fem3d.size()});
                used = 0;
            }
            fscr[e] = (uint32_t)used;
            used += need;
            batches.back().nr_max =
                std::max(batches.back().nr_max, (int)hFElemNr[e]);
            for (int al = 0; al < nu1; al++) fem3d.push_back({e, al});
            batches.back().end = w + 1;
            batches.back().em_end = (int)fem3d.size();
            scrMax = std::max(scrMax, used);
        }
    }

    // The batches of the static scratch (solid and fluid 3D elements), and
    // the largest scratch a single element needs; set up at finalize.
    void planStaticBatches() {
        std::vector<uint32_t> off;
        std::vector<EMItem> em, ax;
        size_t mx = 0;
        mStaticNeedMax = 0;
        for (int32_t e : hList3D)
            mStaticNeedMax = std::max(
                mStaticNeedMax, (size_t)6 * NPEM *
                                    scr_row(hElemNr[e] / 2 + 1) *
                                    sizeof(cpx<Real>));
        for (int32_t e : hFList3D)
            mStaticNeedMax = std::max(
                mStaticNeedMax, (size_t)3 * NPEM *
                                    scr_row(hFElemNr[e] / 2 + 1) *
                                    sizeof(cpx<Real>));
        if (mNum3D > 0) {
            planCardBatches(STATIC_SCRATCH_BYTES, mCardBatchesS, off, em, ax,
                            mx);
            dScrOffS.alloc(off.size() * 4);
            dScrOffS.upload(off.data(), off.size() * 4, mStream);
        }
        if (mNumF3D > 0) {
            planFluidCardBatches(STATIC_SCRATCH_BYTES, mFCardBatchesS, off,
                                 em, mx);
            dFScrOffS.alloc(off.size() * 4);
            dFScrOffS.upload(off.data(), off.size() * 4, mStream);
        }
    }

    void buildCardBatches(size_t bytes) {
        std::vector<uint32_t> scrOff;
        std::vector<EMItem> em3d, emAx;
        size_t scrMax = 0;
        planCardBatches(bytes, mCardBatches, scrOff, em3d, emAx, scrMax);
        dScrOff.alloc(scrOff.size() * 4);
        dScrOff.upload(scrOff.data(), scrOff.size() * 4, mStream);
        dEM3D.alloc(em3d.size() * sizeof(EMItem));
        dEM3D.upload(em3d.data(), em3d.size() * sizeof(EMItem), mStream);
        dEM3DAx.alloc(std::max<size_t>(emAx.size(), 1) * sizeof(EMItem));
        if (!emAx.empty())
            dEM3DAx.upload(emAx.data(), emAx.size() * sizeof(EMItem),
                           mStream);
        scrMax = std::max(scrMax, mFCardScrMax);
        if (scrMax != mCardScrMax || dCardScratch.bytes() == 0) {
            mCardScrMax = scrMax;
            dCardScratch.alloc(scrMax * sizeof(cpx<Real>));
        }
        mScratchMiB = (double)bytes / (1 << 20);
    }

    // One untimed and one timed pass of the cardinal path, on zero fields
    // (the element forces come out zero and the attenuation state stays
    // zero, so a trial leaves no trace).
    float timeCardinalPass() {
        hipEvent_t a, b;
        hip_check(hipEventCreate(&a), "Domain::autotune");
        hip_check(hipEventCreate(&b), "Domain::autotune");
        cardinalPath(false);
        hip_check(hipEventRecord(a, mStream), "Domain::autotune");
        cardinalPath(false);
        hip_check(hipEventRecord(b, mStream), "Domain::autotune");
        hip_check(hipEventSynchronize(b), "Domain::autotune");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, a, b), "Domain::autotune");
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return ms;
    }

    // Picks the cardinal scratch size (batch size) by timing one pass for
    // each of 4096, 1024 and 256 MiB (fewer, larger batches against
    // batches whose scratch stays in the 256 MB Infinity Cache), unless
    // AXS3D_CARD_SCRATCH_MB fixes it; AXS3D_AUTOTUNE=0 skips the timing.
    void autotune() override {
        if (mNum3D == 0) return;
        const char *env = std::getenv("AXS3D_AUTOTUNE");
        if (env && env[0] == '0') return;
        mCardTuneMs = timeCardinalPass();
        mTuneTrials = 1;
        if (!mScratchFixed) {
            double best_mib = mScratchMiB;
            float best = (float)mCardTuneMs;
            for (const size_t mib : {(size_t)1024, (size_t)256}) {
                buildCardBatches(mib << 20);
                const float ms = timeCardinalPass();
                mTuneTrials++;
                if (ms < best) {
                    best = ms;
                    best_mib = (double)mib;
                }
            }
            if (best_mib != mScratchMiB)
                buildCardBatches((size_t)best_mib * ((size_t)1 << 20));
            mCardTuneMs = best;
        }
        tuneFftRadix();
    }

    // The FFT plans' largest pass radix: one timed cardinal pass with the
    // plans of each of 16, 8 and 4 (the radices change the transforms'
    // instruction mix and register demand, not their results beyond
    // rounding), keeping the fastest, unless AXS3D_FFT_MAX_RADIX fixes it.
    void uploadFftTables(int maxRadix) {
        std::vector<uint32_t> plan;
        std::vector<double> tw;
        std::vector<uint16_t> pos;
        buildFftTables(maxRadix, plan, tw, pos);
        dPlan.upload(plan.data(), plan.size() * 4, mStream);
        dPosTab.upload(pos.data(), pos.size() * 2, mStream);
        std::vector<Real> twr(tw.begin(), tw.end());
        if (!twr.empty())
            dTwiddle.upload(twr.data(), twr.size() * sizeof(Real), mStream);
        hip_check(hipStreamSynchronize(mStream), "Domain::autotune");
    }
    void tuneFftRadix() {
        mRadixMs.clear();
        // a pass under 5 ms is too short to rank plans reliably, and small
        // domains keep the default plans so that their results do not
        // depend on timing noise
        if (mFftRadixFixed || mFftNrs.empty() || mCardTuneMs < 5.0) return;
        int best_r = mFftMaxRadix;
        float best = (float)mCardTuneMs;
        mRadixMs.push_back(mFftMaxRadix);
        mRadixMs.push_back(best);
        for (const int r : {8, 4}) {
            if (r == mFftMaxRadix) continue;
            uploadFftTables(r);
            const float ms = timeCardinalPass();
            mRadixMs.push_back(r);
            mRadixMs.push_back(ms);
            if (ms < best) {
                best = ms;
                best_r = r;
            }
        }
        if (best_r != mFftMaxRadix || mRadixMs.size() > 2) {
            mFftMaxRadix = best_r;
            uploadFftTables(best_r);
        }
        mCardTuneMs = best;
    }

    // 3D operands, host [slice][width] per element -> device [width][Nr]
    // per element, cast to Real; uploaded in chunks of whole elements.
    void uploadSlices(DeviceBuffer &b, int width, int e_begin, int e_end,
                      const double *src) override {
        if (b.bytes() == 0)
            b.alloc((size_t)mNumSlices * width * sizeof(Real));
        const size_t chunk = (size_t)1 << 24;
        std::vector<Real> tmp;
        tmp.reserve(chunk + (size_t)width * 1024);
        int64_t first = -1;
        auto flush = [&]() {
            if (tmp.empty()) return;
            hip_check(hipMemcpyAsync(b.as<Real>() + (size_t)first * width,
                                     tmp.data(), tmp.size() * sizeof(Real),
                                     hipMemcpyHostToDevice, mStream),
                      "Domain::upload3D");
            hip_check(hipStreamSynchronize(mStream), "Domain::upload3D");
            tmp.clear();
            first = -1;
        };
        for (int e = e_begin; e < e_end; e++) {
            if (hElemKind[e] != 1) continue;
            const int nr = hElemNr[e];
            if (first < 0) first = hSliceOff[e];
            const size_t base = tmp.size();
            tmp.resize(base + (size_t)nr * width);
            for (int k = 0; k < nr; k++)
                for (int c = 0; c < width; c++)
                    tmp[base + (size_t)c * nr + k] =
                        (Real)src[(size_t)k * width + c];
            src += (size_t)nr * width;
            if (tmp.size() >= chunk) flush();
        }
        flush();
    }

    void uploadProp(int e_begin, int e_end, const double *src) override {
        if (dProp3D.bytes() == 0)
            dProp3D.alloc((size_t)std::max<int64_t>(mPropTotal, 1) *
                          sizeof(Real));
        const size_t chunk = (size_t)1 << 24;
        std::vector<Real> tmp;
        int64_t first = -1;
        auto flush = [&]() {
            if (tmp.empty()) return;
            hip_check(hipMemcpyAsync(dProp3D.as<Real>() + first, tmp.data(),
                                     tmp.size() * sizeof(Real),
                                     hipMemcpyHostToDevice, mStream),
                      "Domain::upload3D");
            hip_check(hipStreamSynchronize(mStream), "Domain::upload3D");
            tmp.clear();
            first = -1;
        };
        for (int e = e_begin; e < e_end; e++) {
            if (hElemKind[e] != 1) continue;
            const int nr = hElemNr[e];
            const int width = mat_nparams(hMType[e]) * NPEM;
            if (first < 0) first = hPropOff[e];
            const size_t base = tmp.size();
            tmp.resize(base + (size_t)nr * width);
            for (int k = 0; k < nr; k++)
                for (int c = 0; c < width; c++)
                    tmp[base + (size_t)c * nr + k] =
                        (Real)src[(size_t)k * width + c];
            src += (size_t)nr * width;
            if (tmp.size() >= chunk) flush();
        }
        flush();
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
        if (mNumFElems > 0) {
            std::vector<Real> fops((size_t)mNumFElems * ELEM_OPS * NPEM);
            for (int e = 0; e < mNumFElems; e++) {
                const double *g = hFElemGQ.data() + (size_t)e * 6 * NPEM;
                Real *o = fops.data() + (size_t)e * ELEM_OPS * NPEM;
                for (int k = 0; k < NPEM; k++) {
                    const double ifk = g[5 * NPEM + k];
                    for (int r = 0; r < 5; r++) {
                        o[r * NPEM + k] = (Real)g[r * NPEM + k];
                        o[(5 + r) * NPEM + k] = (Real)(g[r * NPEM + k] * ifk);
                    }
                }
            }
            dFElemOps.alloc(fops.size() * sizeof(Real));
            dFElemOps.upload(fops.data(), fops.size() * sizeof(Real), st);
            std::vector<Real> fk(hFElemK.begin(), hFElemK.end());
            dFElemK.alloc(fk.size() * sizeof(Real));
            dFElemK.upload(fk.data(), fk.size() * sizeof(Real), st);
            dFEB.alloc((size_t)std::max<int64_t>(mNumFEBModes, 1) * 2 *
                       sizeof(Real));
            dFEB.zero(st);
            dFStiff.alloc((size_t)mNumFModes * 2 * sizeof(Real));
            dFStiff.zero(st);
        }
        std::vector<Real> mat(hMat1D.begin(), hMat1D.end());
        if (mat.empty()) mat.assign(2 * NPEM, (Real)0);
        dElemMat.alloc(mat.size() * sizeof(Real));
        dElemMat.upload(mat.data(), mat.size() * sizeof(Real), st);
        {
            std::vector<uint32_t> mo(mNumElems, 0);
            for (size_t e = 0; e < hMatOff.size(); e++) {
                if (hMatOff[e] >= (int64_t)UINT32_MAX)
                    throw Error("Domain::finalize || 1D material too large.");
                mo[e] = (uint32_t)hMatOff[e];
            }
            dMatOff.alloc(mo.size() * 4);
            dMatOff.upload(mo.data(), mo.size() * 4, st);
        }
        if (!hTheta.empty()) {
            // sin, cos of theta and 2 theta cast after their FP64
            // evaluation (CoordTransformSpherical constructor,
            // CoordTransformSpherical.hpp:36-45)
            std::vector<Real> rot((size_t)mNumElems * 4 * NPEM);
            for (int e = 0; e < mNumElems; e++)
                for (int k = 0; k < NPEM; k++) {
                    const double t = hTheta[(size_t)e * NPEM + k];
                    Real *r = rot.data() + (size_t)e * 4 * NPEM + k;
                    r[0] = (Real)std::sin(t);
                    r[NPEM] = (Real)std::cos(t);
                    r[2 * NPEM] = (Real)std::sin(2. * t);
                    r[3 * NPEM] = (Real)std::cos(2. * t);
                }
            dRot.alloc(rot.size() * sizeof(Real));
            dRot.upload(rot.data(), rot.size() * sizeof(Real), st);
        }
        if (!hPrtElems.empty()) {
            // X0..X3 and X0..X3 times |J| in FP64, then cast (the PRT
            // constructor, PRT.hpp:57-80)
            const size_t n = hPrtElems.size();
            std::vector<int32_t> idx(mNumElems, -1);
            std::vector<Real> px(n * 8 * NPEM);
            for (size_t i = 0; i < n; i++) {
                idx[hPrtElems[i]] = (int32_t)i;
                const double *x = hPrtX.data() + i * 5 * NPEM;
                for (int c = 0; c < 4; c++)
                    for (int k = 0; k < NPEM; k++) {
                        px[(i * 8 + c) * NPEM + k] = (Real)x[c * NPEM + k];
                        px[(i * 8 + 4 + c) * NPEM + k] =
                            (Real)(x[c * NPEM + k] * x[4 * NPEM + k]);
                    }
            }
            dPrtIdx.alloc(idx.size() * 4);
            dPrtIdx.upload(idx.data(), idx.size() * 4, st);
            dPrt.alloc(px.size() * sizeof(Real));
            dPrt.upload(px.data(), px.size() * sizeof(Real), st);
        }
        std::vector<Real> gt(2 * NPEM);
        for (int k = 0; k < NPEM; k++) {
            gt[k] = (Real)hGll[k];
            gt[NPEM + k] = (Real)hGlj[k];
        }
        dGTab.alloc(gt.size() * sizeof(Real));
        dGTab.upload(gt.data(), gt.size() * sizeof(Real), st);
        dStiff.alloc((size_t)mNumModes * 6 * sizeof(Real));
        dStiff.zero(st);
        dEB.alloc((size_t)mNumEBTriples * 6 * sizeof(Real));
        dEB.zero(st);
        if (mNsls > 0 && mNumEM > 0) {
            const int npa = mAttCG4 ? 4 : NPEM;
            std::vector<Real> a1(hAtt1D.begin(), hAtt1D.end());
            dAtt1D.alloc(a1.size() * sizeof(Real));
            dAtt1D.upload(a1.data(), a1.size() * sizeof(Real), st);
            dMemVarFR.alloc((size_t)mNumEM * mNsls * 6 * npa * 2 *
                            sizeof(double));
            dMemVarFR.zero(st);
            dDStressFR.alloc((size_t)mNumEM * 6 * npa * 2 * sizeof(Real));
            dDStressFR.zero(st);
        }
        if (!hTwTab.empty()) {
            std::vector<Real> tw(hTwTab.begin(), hTwTab.end());
            dTwiddle.alloc(tw.size() * sizeof(Real));
            dTwiddle.upload(tw.data(), tw.size() * sizeof(Real), st);
        }
        if (mNumMass3D > 0) {
            // Mass3D keeps 1 / m in the working precision (Mass3D.hpp:19-22),
            // MassOceanLoad3D also k (MassOceanLoad3D.cpp:13-20)
            std::vector<Real> inv(hMass3D.begin(), hMass3D.end());
            dInvMass3D.alloc(inv.size() * sizeof(Real));
            dInvMass3D.upload(inv.data(), inv.size() * sizeof(Real), st);
            if (!hMass3DK.empty()) {
                std::vector<Real> kk(hMass3DK.begin(), hMass3DK.end());
                dMass3DK.alloc(kk.size() * sizeof(Real));
                dMass3DK.upload(kk.data(), kk.size() * sizeof(Real), st);
            }
        }
        if (!hFMass3D.empty()) {
            std::vector<Real> inv(hFMass3D.size());
            for (size_t k = 0; k < inv.size(); k++)
                inv[k] = (Real)(1.0 / hFMass3D[k]);
            dFInvMass3D.alloc(inv.size() * sizeof(Real));
            dFInvMass3D.upload(inv.data(), inv.size() * sizeof(Real), st);
        }
        if (mNumCpl3D > 0) {
            // SolidFluidCoupling3D keeps both normal sets in Real
            std::vector<Real> nu(hCpl3DNu.begin(), hCpl3DNu.end());
            std::vector<Real> nam(hCpl3DNam.begin(), hCpl3DNam.end());
            dCpl3DNu.alloc(nu.size() * sizeof(Real));
            dCpl3DNu.upload(nu.data(), nu.size() * sizeof(Real), st);
            dCpl3DNam.alloc(nam.size() * sizeof(Real));
            dCpl3DNam.upload(nam.data(), nam.size() * sizeof(Real), st);
        }
        if (mNum3D > 0) {
            int ncu = 0;
            hip_check(hipDeviceGetAttribute(
                          &ncu, hipDeviceAttributeMultiprocessorCount,
                          mDevice), "Domain::finalize");
            int nr3d = 0;
            for (int32_t e : hList3D)
                nr3d = std::max(nr3d, (int)hElemNr[e]);
            mCardSlots = std::max(
                1, card_blocks_per_cu<Real>(mNsls > 0, nr3d) * ncu);
            size_t budget = CARD_SCRATCH_BYTES;
            if (const char *env = std::getenv("AXS3D_CARD_SCRATCH_MB")) {
                // MiB, fractions allowed (at least one element per batch)
                budget = (size_t)std::max(
                    1.0, std::atof(env) * (double)((size_t)1 << 20));
                mScratchFixed = true;
            }
            buildCardBatches(budget);
            if (mNsls > 0) {
                const int npa = mAttCG4 ? 4 : NPEM;
                dMemVar.alloc((size_t)mNumSlices * mNsls * 6 * npa *
                              sizeof(double));
                dMemVar.zero(st);
                dDStress.alloc((size_t)mNumSlices * 6 * npa * sizeof(Real));
                dDStress.zero(st);
            }
        }
        if (mNumF3D > 0) {
            // K of the 3D fluid elements, host [Nr][25] -> device [25][Nr]
            // per element, cast to Real
            std::vector<Real> k3;
            std::vector<int64_t> k3off(mNumFElems, 0);
            for (int e = 0; e < mNumFElems; e++) {
                if (hFElemKind[e] != 1) continue;
                const int nr = hFElemNr[e];
                k3off[e] = (int64_t)k3.size();
                const double *src = hFK3D.data() + hFK3DOff[e];
                for (int k = 0; k < NPEM; k++)
                    for (int sl = 0; sl < nr; sl++)
                        k3.push_back((Real)src[(size_t)sl * NPEM + k]);
            }
            dFK3D.alloc(k3.size() * sizeof(Real));
            dFK3D.upload(k3.data(), k3.size() * sizeof(Real), st);
            dFK3DOff.alloc(k3off.size() * 8);
            dFK3DOff.upload(k3off.data(), k3off.size() * 8, st);
            // batches whose [3][25][nu1] rows fit the scratch together
            size_t budget = CARD_SCRATCH_BYTES;
            if (const char *env = std::getenv("AXS3D_CARD_SCRATCH_MB"))
                budget = (size_t)std::max(
                    1.0, std::atof(env) * (double)((size_t)1 << 20));
            std::vector<uint32_t> fscr;
            std::vector<EMItem> fem3d;
            size_t scrMax = 0;
            planFluidCardBatches(budget, mFCardBatches, fscr, fem3d, scrMax);
            dFScrOff.alloc(fscr.size() * 4);
            dFScrOff.upload(fscr.data(), fscr.size() * 4, st);
            dFEM3D.alloc(fem3d.size() * sizeof(EMItem));
            dFEM3D.upload(fem3d.data(), fem3d.size() * sizeof(EMItem), st);
            // one scratch for both paths (they run one after the other)
            mFCardScrMax = scrMax;
            if (scrMax > mCardScrMax) {
                mCardScrMax = scrMax;
                dCardScratch.alloc(scrMax * sizeof(cpx<Real>));
            }
        }
        planStaticBatches();
        abcReal(mAbcS, mAbcSD);
        abcReal(mAbcF, mAbcFD);
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
