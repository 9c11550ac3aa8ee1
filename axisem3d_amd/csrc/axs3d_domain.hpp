// axs3d_domain.hpp -- host C++ façade: Domain and NewmarkTimeScheme.
//
// Keeps the reference's phase structure (src = SOLVER/src):
//   Domain::applySources / computeStiffness / applyBC_* / recordOutput /
//   computeStiffToAccel / checkStability   src/core/domain/Domain.cpp:285-612
//   NewmarkTimeScheme::solve / update      src/core/time/NewmarkTimeScheme.cpp:18-131,
//                                          NewmarkTimeScheme.hpp:48-71
// but the points and elements live in device memory as structure-of-arrays
// instead of shared_ptr objects, and each phase is one kernel launch on the
// domain's stream.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "axs3d.h"
#include "axs3d_common.hpp"
#include "axs3d_launch.hpp"

namespace axs3d {

// "Class::method || detail" errors, as the reference throws them.
struct Error : public std::runtime_error {
    explicit Error(const std::string &m) : std::runtime_error(m) {}
};

void hip_check(hipError_t e, const char *where);

// Size of the file-scope device scratch of the allocation-free stiffness
// entry point (axs3d_compute_stiffness_static), per device.
constexpr size_t STATIC_SCRATCH_BYTES = (size_t)64 << 20;
// Device address of the current device's copy of that scratch (looked up
// once per device with hipGetSymbolAddress, axs3d_scratch.hip).
void *static_scratch(int device);

// Owning device allocation.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer() { reset(); }
    void alloc(size_t bytes);
    void upload(const void *src, size_t bytes, hipStream_t st);
    void zero(hipStream_t st);
    void reset();
    template <typename T>
    T *as() const { return static_cast<T *>(ptr_); }
    size_t bytes() const { return bytes_; }
    void swap(DeviceBuffer &o) {
        std::swap(ptr_, o.ptr_);
        std::swap(bytes_, o.bytes_);
    }

private:
    void *ptr_ = nullptr;
    size_t bytes_ = 0;
};

class DomainBase {
public:
    virtual ~DomainBase();
    // set-up
    void setGMat(const double *g_gll, const double *g_glj);
    void setSolidPoints(int n, const int32_t *nr, const double *mass);
    void setSolidElements(int n, const int32_t *points, const uint8_t *axial,
                          const double *gq);
    void setSolidIso1D(const double *lam_mu);
    void setElementKinds(const uint8_t *kind);
    void setSolidIso3D(int e_begin, int e_end, const double *lam_mu);
    void setMaterialTypes(const uint8_t *mtype);
    void setSolidMaterial1D(const double *params);
    void setSolidMaterial3D(int e_begin, int e_end, const double *params);
    void setRtzTheta(const double *theta);
    // fluid (displacement potential) points, elements and boundaries
    void setFluidPoints(int n, const int32_t *nr, const double *mass);
    void setFluidElements(int n, const int32_t *points, const uint8_t *axial,
                          const double *gq);
    void setFluidK1D(const double *k);
    void setFluidKinds(const uint8_t *kind);
    void setFluidK3D(int e_begin, int e_end, const double *k);
    void setFluidAxialPoints(int n, const int32_t *points);
    void setFluidSurfacePoints(int n, const int32_t *points);
    void setSolidFluidCoupling(int n, const int32_t *solid,
                               const int32_t *fluid, const double *normals);
    void setFluidMass3D(int n, const int32_t *points, const double *mass);
    void setCoupling3D(int n, const int32_t *pairs, const double *n_u,
                       const double *n_a);
    void setAttenuation(int nsls, int cg4, const double *alpha,
                        const double *beta, const double *gamma);
    void setAttenuation3D(int e_begin, int e_end, const double *dlam_dmu);
    void setAttenuation1D(const double *dlam_dmu);
    void setAxialPoints(int n, const int32_t *points);
    void setSolidMass3D(int n, const int32_t *points, const double *mass);
    // Particle relabelling transformation of undulated geometry on solid
    // elements elems[n]: x[n][5][25] = X0, X1, X2, X3, XJ per GLL point,
    // one set for all slices (Undulation::createPRT with an axisymmetric
    // undulation, Undulation.cpp:124-168).
    void setSolidPRT(int n, const int32_t *elems, const double *x);
    void setOceanLoad(int n, const int32_t *points, const int32_t *nval,
                      const double *values);
    void setMomentSource(int elem, int nu1, const double *pattern, int nsteps,
                         const double *stf);
    void addMomentSource(int elem, int nu1, const double *pattern, int nsteps,
                         const double *stf);
    void setInstances(int n, const int32_t *point_inst, const double *dt);
    void setReceivers(int n, const int32_t *elem, const double *w,
                      const double *phi, int max_samples);
    // station frames: per receiver the wavefield coordinate system (0 spz,
    // 1 RTZ, 2 ENZ, 3 xyz), theta and back azimuth (Station.hpp:24-29)
    void setReceiverFrames(int n, const int32_t *wcs, const double *theta,
                           const double *baz, int cartesian);
    // absorbing boundaries (AbsorbingBoundary): entries as the reference
    // constructs them, merged per point at finalize
    void addClaytonSolid(int n, const int32_t *points, const double *rho_vp,
                         const double *rho_vs, const double *area,
                         const double *theta);
    void addClaytonSolid3D(int n, const int32_t *points,
                           const double *rho_vp, const double *rho_vs,
                           const double *area, const double *unit_normal);
    void addClaytonFluid(int n, const int32_t *points, int per_slice,
                         const double *rho_vp, const double *area);
    void addSponge(int fluid, int n, const int32_t *points, int per_slice,
                   const double *gamma);
    // time scheme: 0 Newmark (NewmarkTimeScheme), 1 symplectic
    // (SymplecticTimeScheme, four sub-steps per step)
    void setTimeScheme(int scheme);
    int timeScheme() const { return mScheme; }
    void setDt(double dt) {
        mDt = dt;
        dropGraph();  // dt is a kernel argument of the captured step
    }
    void finalize();
    bool finalized() const { return mFinalized; }

    // time-loop phases (Domain.cpp:285-345)
    virtual void applySources(int tstep) = 0;
    virtual void computeStiffness() = 0;
    // computeStiffness with the cardinal path through the static device
    // scratch (STATIC_SCRATCH_BYTES) in batches planned for it at
    // finalize; allocates nothing.  Calls that overlap on one device race
    // on that scratch.
    virtual void computeStiffnessStatic() = 0;
    virtual void applyBC_BeforeAssemblingStiff() = 0;
    // isample < 0 with dev: the sample follows the device step counter
    void recordOutput(int isample, bool dev = false);
    virtual void applyBC_AfterAssemblingStiff() = 0;
    // Domain::applyBC_AfterComputingAccel (Domain.cpp:330-332): sponge
    virtual void applyBC_AfterComputingAccel() = 0;
    virtual void computeStiffToAccel() = 0;
    virtual void newmark() = 0;
    // SymplecticTimeScheme::launch (u += kappa_0 dt v) and ::update of
    // sub-step isub (1..4) on all points (SymplecticTimeScheme.hpp:37-70)
    virtual void symplecticLaunch() = 0;
    virtual void symplecticUpdate(int isub) = 0;
    // one bulk Newmark step (sources, stiffness, receivers, points); with
    // dev, the step and sample come from the device step counter, which
    // the step advances (the form captured into a HIP graph)
    virtual void step(int tstep, int isample, bool dev) = 0;
    void checkStability(int tstep);

    // bulk
    void solve(int step_begin, int nsteps);
    void setRecordOrigin(int step0) { mRecordStep0 = step0; }
    // replay a captured step graph in solve (default) or launch eagerly
    void useGraph(bool on) { mUseGraph = on; }
    void readSeismograms(double *out, int nsamples);

    // adjoint (sensitivity) run of 1D isotropic elastic domains: a forward
    // run from rest over nsteps with checkpoints every seg steps, then the
    // adjoint run backwards with the force R^* djdr[n] / dt^2, recomputing
    // each segment's forward fields from its checkpoint; kern[e][k][3] =
    // dJ/d(lambda, mu, rho) at GLL point k of element e
    virtual void adjointRun(int nsteps, int seg, const double *djdr,
                            double *kern, double *kern3d) = 0;
    // the adjoint run for any solid media: dJ/d(material parameters) packed
    // as axs3d_set_solid_material1d takes them (every element, its type's
    // parameters x 25), dJ/drho [e][25], and per slice of the 3D elements
    // as axs3d_set_solid_material3d takes them
    virtual void adjointRunMedia(int nsteps, int seg, const double *djdr,
                                 double *kern_par, double *kern_rho,
                                 double *kern_par3d) = 0;
    // sensitivity densities of one (psi, u, a) triple, FP64, kern[e][k][3]
    void adjointDensity(const double *psi, const double *u, const double *a,
                        double *kern);

    // debug
    int64_t numModes() const { return mNumModes; }
    int64_t numFluidModes() const { return mNumFModes; }
    virtual void getField(int which, double *out) = 0;
    virtual void setField(int which, const double *in) = 0;
    void synchronize();

    // measurement
    void profileEnable(bool on) { mProfile = on; }
    void profileRead(int kernel, double *ms, int64_t *count);
    void profileReset();

    hipStream_t stream() const { return mStream; }
    void autotuneInfo(int *occ, int *slots, double *ms) const {
        *occ = mCardOcc;
        *slots = mCardSlots;
        ms[0] = mCardTuneMs;
        ms[1] = mScratchMiB;
        ms[2] = mTuneTrials;
    }
    int precision() const { return mPrecision; }
    // every autotuner choice and timing: [0] cardinal pass ms, [1] scratch
    // MiB, [2] scratch sizes timed, [3] FFT largest radix, [4] radices
    // timed, then (radix, ms) per radix timed; returns the count written
    int autotuneReport(double *out, int n) const {
        std::vector<double> v = {mCardTuneMs, mScratchMiB,
                                 (double)mTuneTrials, (double)mFftMaxRadix,
                                 (double)(mRadixMs.size() / 2)};
        v.insert(v.end(), mRadixMs.begin(), mRadixMs.end());
        const int k = std::min(n, (int)v.size());
        for (int i = 0; i < k; i++) out[i] = v[i];
        return k;
    }

protected:
    explicit DomainBase(int device, int precision);
    void requireFinalized(const char *where) const;
    // profiling brackets around one launch of `kernel`
    void tick(int kernel);
    void tock(int kernel);

    int mDevice = 0;
    int mPrecision = 0;
    hipStream_t mStream = nullptr;
    // one step captured as a HIP graph (solve replays it)
    bool mUseGraph = true;
    hipGraph_t mGraph = nullptr;
    hipGraphExec_t mGraphExec = nullptr;
    void captureStep();
    void dropGraph();
    bool mFinalized = false;
    double mDt = 0.;
    int mScheme = 0;
    // the march of the point kernels' next launch (PointGeom::scheme)
    double mSymPi = 0., mSymKappa = 0.;
    int mRecordStep0 = 0;

    // host staging until finalize
    std::vector<double> hGll, hGlj;
    std::vector<int32_t> hPointNr;
    std::vector<double> hMass;
    std::vector<int32_t> hElemPoints;
    std::vector<uint8_t> hElemAxial;
    std::vector<double> hElemGQ;
    // 1D material, packed per element: np(type) x 25 at hMatOff[e]
    std::vector<double> hMat1D;
    std::vector<int64_t> hMatOff;
    std::vector<uint8_t> hMType;        // 0 isotropic, 1 TI, 2 anisotropic
    std::vector<double> hTheta;         // [e][25] colatitude, or empty
    std::vector<int32_t> hPrtElems;     // elements with a PRT
    std::vector<double> hPrtX;          // [n][5][25] X0..X3, XJ
    int mEMTypeBegin[4] = {0, 0, 0, 0};
    int mEMAxBegin[3] = {0, 0, 0};
    // fluid
    std::vector<int32_t> hFPointNr, hFElemPoints, hFAxial, hFSurf;
    std::vector<double> hFMass, hFElemGQ, hFElemK;
    std::vector<uint8_t> hFElemAxial;
    std::vector<int32_t> hCplSolid, hCplFluid;
    std::vector<double> hCplNormal;     // [pair][4] ns, nz unassembled,
                                        // ns, nz assembled
    int mNumFPoints = 0, mNumFElems = 0, mNumFEM = 0, mNumPairs = 0;
    int64_t mNumFModes = 0, mNumFEBModes = 0;
    int mTypeBegin3D[4] = {0, 0, 0, 0};
    std::vector<int32_t> hAxial;
    // moment sources: hSrcElem is the first source's element (-1: none);
    // patterns and STFs of all sources back to back
    int hSrcElem = -1;
    std::vector<int32_t> hSrcElems, hSrcNu1s;
    std::vector<size_t> hSrcPatOff;     // complex values before source i
    std::vector<double> hSrcPattern, hStf;
    // batched instances: instance of each solid point, dt of each instance
    std::vector<int32_t> hPointInst;
    std::vector<double> hInstDt;
    std::vector<int32_t> hRecElem;
    std::vector<double> hRecW, hRecPhi;
    int hMaxSamples = 0;

    // sizes
    int mNumPoints = 0, mNumElems = 0, mNumEM = 0, mNumRec = 0;
    int64_t mNumModes = 0;
    int64_t mNumEBTriples = 0;
    // 3D (cardinal) elements
    std::vector<uint8_t> hElemKind;
    std::vector<int32_t> hElemNr;       // element Nr (max over its points)
    std::vector<int64_t> hSliceOff;     // first slice of each 3D element
    int64_t mNumSlices = 0;
    std::vector<int64_t> hPropOff;      // offset of each 3D element's
    int64_t mPropTotal = 0;             // parameters (np x 25 x Nr)
    int mNum3D = 0, mNu1Max3D = 0, mCardSlots = 0;
    std::vector<int32_t> hList3D;       // 3D elements, launch order
    // fluid elements with a 3D medium
    std::vector<uint8_t> hFElemKind;    // 0 1D, 1 3D
    std::vector<int32_t> hFElemNr;      // fluid element Nr
    std::vector<int64_t> hFK3DOff;      // [e] offset of [Nr][25] K, or -1
    std::vector<double> hFK3D;
    std::vector<int32_t> hFList3D;      // 3D fluid elements, launch order
    std::vector<CardBatch> mFCardBatches;
    int mNumF3D = 0;
    size_t mCardScrMax = 0;             // scratch size, complex values
    // solid points with a 3D mass: points, offsets of their Nr masses
    std::vector<int32_t> hMass3DPts;
    int mMass3DNrMax = 0;               // largest Nr of the 3D-mass points
    int mFDeferNrMax = 0;               // and of the deferred fluid points
    // largest FFT pass radix (2..16); AXS3D_FFT_MAX_RADIX to change it
    int mFftMaxRadix = 16;
    bool mFftRadixFixed = false;        // set by AXS3D_FFT_MAX_RADIX
    std::vector<int> mFftNrs;           // distinct Nr with FFT tables
    int mFftNrMax = 0;
    std::vector<double> mRadixMs;       // cardinal pass per radix timed
    void buildFftTables(int maxRadix, std::vector<uint32_t> &plan,
                        std::vector<double> &tw,
                        std::vector<uint16_t> &pos) const;
    std::vector<int64_t> hMass3DOff;
    std::vector<double> hMass3D;        // 1 / m per slice
    int mNumMass3D = 0;
    // ocean load (MassOceanLoad1D/3D): the 1D entries' four factors per
    // point for the point kernel; the 3D entries (1 / m and k per slice)
    // join the 3D-mass list at finalize, with the offsets of their k
    std::vector<int32_t> hOcean1DPts, hOcean3DPts;
    std::vector<double> hOcean1D, hOcean3D;
    size_t mMass3DUserPts = 0, mMass3DUserVals = 0;
    std::vector<int64_t> hMass3DKOff;
    std::vector<double> hMass3DK;
    void mergeOceanLoad();
    // fluid points with a 3D mass; 3D solid-fluid pairs (normals [Nr][3]
    // per pair, unassembled and assembled); the deferred fluid points
    std::vector<int32_t> hFMass3DPts;
    std::vector<int64_t> hFMass3DOff;
    std::vector<double> hFMass3D;
    std::vector<int32_t> hCpl3DPairs;
    std::vector<double> hCpl3DNu, hCpl3DNa, hCpl3DNam;
    int mNumFDefer = 0, mNumCpl3D = 0;
    std::vector<CardBatch> mCardBatches;
    int mCardOcc = CARD_OCC;            // stress kernel waves per SIMD
    double mCardTuneMs = 0.;            // one cardinal pass at finalize
    double mScratchMiB = 0.;            // cardinal scratch budget in use
    bool mScratchFixed = false;         // set by AXS3D_CARD_SCRATCH_MB
    int mTuneTrials = 0;                // scratch sizes timed
    size_t mFCardScrMax = 0;            // 3D fluid rows in the scratch
    // the same batches planned for the static scratch (this device's
    // address of it in mStaticScratch), and the largest scratch one
    // element needs (bytes; above STATIC_SCRATCH_BYTES the static entry
    // point refuses to run)
    void *mStaticScratch = nullptr;
    std::vector<CardBatch> mCardBatchesS, mFCardBatchesS;
    size_t mStaticNeedMax = 0;
    std::vector<uint8_t> hProp3DSet, hAtt3DSet;
    // attenuation
    int mNsls = 0, mAttCG4 = 0;
    std::vector<double> hAttABG;
    std::vector<double> hAtt1D;         // [e][2][npa] dlambda, dmu
    std::vector<double> hTwTab;         // interleaved twiddles, FP64
    void layout3D();                    // element Nr and slice offsets
    std::vector<int> mSrcNu1c;          // orders each source applies
    int mStfSteps = 0;

    // device: points
    DeviceBuffer dPointOff, dPointNu1, dPointNr, dPointAxial, dInvMass,
        dModePoint;
    // device: fields (FP64) and stiffness (Real)
    DeviceBuffer dDispl, dVeloc, dAccel, dStiff;
    // device: elements
    DeviceBuffer dElemPoints, dElemAxial, dElemNu1, dElemOps, dElemMat,
        dMatOff, dRot, dEMList, dGTab, dElemEBOff, dEB, dAdj, dPrtIdx, dPrt;
    // device: cardinal path
    DeviceBuffer dElemNr, dSliceOff, dPropOff, dList3D, dScrOff, dEM3D,
        dPosOff, dPosTab, dMass3DPts, dMass3DOff, dMass3DFlag, dInvMass3D,
        dOceanIdx, dOcean1D, dMass3DKOff, dMass3DK,
        dEM3DAx,
        dFList3D, dFElemNr, dFK3D, dFK3DOff, dFScrOff, dFEM3D,
        dProp3D,
        dTwiddle,
        dTwOff, dPlan, dCardScratch, dAttABG, dAtt3D, dMemVar, dDStress,
        dAtt1D, dMemVarFR, dDStressFR, dScrOffS, dFScrOffS;
    // 3D operands of elements [e_begin, e_end), host layout [slice][width]
    // in element order; stored per element as [width][Nr] so that
    // consecutive slices are consecutive in memory.
    virtual void uploadSlices(DeviceBuffer &b, int width, int e_begin,
                              int e_end, const double *src) = 0;
    // 3D material of elements [e_begin, e_end): host [Nr][np][25] per
    // element -> device [np * 25][Nr] at hPropOff[e]
    virtual void uploadProp(int e_begin, int e_end, const double *src) = 0;
    // device: fluid
    DeviceBuffer dFPointOff, dFPointNu1, dFPointNr, dFPointBC, dFInvMass,
        dFModePoint, dFAdj, dFElemPoints, dFElemAxial, dFElemNu1, dFElemOps,
        dFElemK, dFEMList, dFEBOff, dFEB, dChi, dChiV, dChiA, dFStiff;
    DeviceBuffer dCplSolidPair, dCplFluidPair, dCplSOff, dCplFOff, dCplN,
        dCplNM, dCStiff;
    DeviceBuffer dFDeferPts, dFDeferPair, dFDeferMOff, dFDeferFlag,
        dFInvMass3D, dCpl3DSolid, dCpl3DNOff, dCpl3DNu, dCpl3DNam,
        dCpl3DCOff, dCplSolidPair3D, dCSolid3D;
    // device: source and receivers
    DeviceBuffer dSrcPattern, dStf, dRecElem, dRecW, dRecPhi, dSeismo;
    // station frames: [receiver][6] two rotations (axis, cos, sin)
    std::vector<double> hRecRot;
    DeviceBuffer dRecRot;
    // absorbing boundaries of the solid (3 components) and fluid (1)
    // points: host records merged per point, their device form
    struct AbcHost {
        std::vector<int32_t> rec;              // [point] record or -1
        std::vector<int32_t> point;            // [record] point
        std::vector<double> cl1;               // [record][4 or 1]
        std::vector<std::vector<double>> cl3;  // [record] per slice 6 or 1
        std::vector<double> sp1;               // [record][2]
        std::vector<std::vector<double>> sp3;  // [record] per slice 2
    };
    struct AbcDev {
        int n = 0, nr_max = 0;
        DeviceBuffer idx, point, cl1, cl3off, cl3, sp1, sp3off, sp3,
            termoff, cterm, sterm;
    };
    AbcHost mAbcS, mAbcF;
    AbcDev mAbcSD, mAbcFD;
    // record of point p (created on first use) in h, for points of Nr
    // nrOf(p); w1 = 1D Clayton width
    int abcRecord(AbcHost &h, int npoints, int p, int w1);
    DeviceBuffer dPointDt;
    DeviceBuffer dFlag;
    DeviceBuffer dStepCtr;              // [0] step, [1] first recorded step
    // adjoint: FP64 operators, densities, adjoint fields, checkpoints and
    // the forward fields of one segment
    DeviceBuffer dAdjGQ, dAdjGTab, dAdjKern, dAdjU, dAdjV, dAdjA, dAdjCkpt,
        dAdjSegU, dAdjSegA, dAdjG, dAdjKind, dAdjKern3D, dAdjScrB;
    // the adjoint field's own SLS memory variables and stress increments
    // (cardinal and Fourier path), swapped in with its wavefields
    DeviceBuffer dAdjAtt[4];
    DeviceBuffer *attState(int i) {
        DeviceBuffer *b[4] = {&dMemVar, &dDStress, &dMemVarFR, &dDStressFR};
        return b[i];
    }
    int mRecNu1Max = 0;
    void adjointCheck(bool media = false) const;
    // media form of the adjoint: offsets of each element's parameters in
    // the packed output, their total, the types and FP64 rotations
    std::vector<int64_t> hAdjMatOff;
    int64_t mAdjMatTotal = 0;
    DeviceBuffer dAdjMType, dAdjMatOff, dAdjRot, dAdjKernR;
    void adjointOperands();
    SensArgs sensArgs(const double *psi, const double *u, const double *a,
                      double *kern, double scale) const;
    void swapAdjointFields();

    virtual void uploadReal() = 0;   // precision-specific operand upload
    virtual void autotune() = 0;     // pick kernel variants by timing

    // profiling
    bool mProfile = false;
    struct Pending {
        int kernel;
        hipEvent_t a, b;
    };
    std::vector<Pending> mPending;
    std::vector<hipEvent_t> mEventPool;
    hipEvent_t mOpenA[AXS3D_NUM_KERNELS] = {};
    double mProfMs[AXS3D_NUM_KERNELS] = {};
    int64_t mProfCount[AXS3D_NUM_KERNELS] = {};
    hipEvent_t takeEvent();
    void drainPending();
};

std::unique_ptr<DomainBase> make_domain(int device, int precision);

}  // namespace axs3d
