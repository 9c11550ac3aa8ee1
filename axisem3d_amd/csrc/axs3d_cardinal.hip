// axs3d_cardinal.hip -- cardinal-path solid stiffness (3D material) for
// gfx950.
//
// Reference (src = SOLVER/src): SolidElement::displToStiff, no PRT, 3D
// material (src/core/element/SolidElement.cpp:160-198):
//   computeGrad6 -> gFFT_N6.computeC2R -> strainToStress_CD (+ attenuation)
//   -> gFFT_N6.computeR2C -> computeQuad6
// with Isotropic::strainToStress<_3D_CD> (Isotropic.hpp:155-172),
// AttenuationFull/CG4::apply (attenuation/AttenuationFull.hpp:192-207,
// AttenuationCG4.hpp:312-330) and the FFT conventions of SolverFFTW
// (src/core/fft/SolverFFTW.hpp:101-168: R2C scaled by 1/Nr, Nyquist row
// halved; C2R with the Nyquist row doubled and the imaginary parts of the
// mean and Nyquist rows ignored).
//
// Design: the 3D elements are processed in batches (elements of one
// material type sorted by Nr, heaviest first, whose strain and stress fit
// the batch scratch, series-major [6*25][nu1] complex per element, each
// row padded to 16 orders), three launches per batch:
//  A. gradient: one thread per (element-mode, GLL point) as on the
//     Fourier path, strain (rotated to RTZ where the medium works there)
//     to the scratch.
//  B. stress: one GLL point per block of three waves.  Each wave packs
//     one of the point's three strain pairs as x + i y into its own LDS
//     buffer and transforms it in place (axs3d_fft.hpp: passes of radix 2
//     to 16 with compile-time DFTs and a table of every twiddle power;
//     plans, tables and the digit reversal per Nr, in LDS sized per
//     launch).  The three waves then form the stress slice by slice from
//     all six strains (3D moduli, SLS memory variables) and transform back
//     into the scratch.  One FFT call site for both directions keeps the
//     code small.
//  C. quadrature: as A, from the stress rows to the element-local force
//     buffer the point kernel gathers.
// Each kernel gets the register budget of its own phase.
#include "axs3d_common.hpp"
#include "axs3d_element.hpp"
#include "axs3d_fft.hpp"
#include "axs3d_launch.hpp"

namespace axs3d {

// dynamic LDS of the FFT kernels (layout of FftLds, sized per launch)
extern __shared__ __align__(16) unsigned char fft_lds[];

// The block's FFT tables of its Nr (twiddle powers and slice positions)
// into its LDS.
template <typename Real>
__device__ __forceinline__ void fft_load_tables(
    const FftLds<Real> &L, const cpx<Real> *__restrict__ twiddle,
    const int64_t *__restrict__ tw_off, const uint16_t *__restrict__ spos,
    const int64_t *__restrict__ pos_off, const uint32_t *plan, int nr,
    int t, int nthreads) {
    const cpx<Real> *twg = twiddle + tw_off[nr];
    const int ntw = (int)(plan[0] >> 8);
    for (int x = t; x < ntw; x += nthreads) L.tw[x] = twg[x];
    const uint16_t *pg = spos + pos_off[nr];
    for (int x = t; x < nr; x += nthreads) L.pos[x] = pg[x];
}

constexpr int CD_EMB = 16;             // gradient/quadrature: element-modes
constexpr int CD_EMB_THREADS = CD_EMB * NPEM;  // per block
constexpr int CD_PW = 3;               // stress kernel: waves per point
constexpr int CD_PT = CD_PW * 64;      // threads per stress block

// C2R input of the strain pair (X, Y) = rows (2q, 2q+1) of a point:
// buf[k] = Xfull[k] + i Yfull[k], with Xfull the Hermitian extension under
// SolverFFTW's C2R rules (Nyquist doubled, Im of mean and Nyquist
// ignored).
template <typename Real>
__device__ __forceinline__ void pack_pair(cpx<Real> *buf,
                                          const cpx<Real> *__restrict__ X,
                                          const cpx<Real> *__restrict__ Y,
                                          int nr, int nu1, bool even,
                                          int lane) {
    // branch-free: slice k takes row entry kk = k (k < nu1) or nr - k
    // (conjugated); the mean and an even Nr's Nyquist entry are real,
    // the latter doubled.  (Reading each entry once and writing its mirror
    // too measured 2% slower on the regional workload.)
    const int nyq = even ? nu1 - 1 : -1;
#pragma unroll 4
    for (int k = lane; k < nr; k += 64) {
        const bool lo = k < nu1;
        const int kk = lo ? k : nr - k;
        const cpx<Real> x = X[kk], y = Y[kk];
        const bool real_only = k == 0 || k == nyq;
        const Real sc = k == nyq ? (Real)2 : (Real)1;
        const Real sim = real_only ? (Real)0 : (lo ? (Real)1 : (Real)-1);
        // xf = (sc x.re, sim x.im), yf likewise; buf = xf + i yf
        buf[pad(k)] = {sc * x.re - sim * y.im, sim * x.im + sc * y.re};
    }
    wave_sync();
}

// R2C output: X_k = (Z_k + conj Z_{N-k}) / 2N, Y_k = (Z_k - conj Z_{N-k})
// / 2iN, Nyquist row halved; written to rows (2q, 2q+1).
template <typename Real>
__device__ __forceinline__ void unpack_pair(const cpx<Real> *buf,
                                            cpx<Real> *__restrict__ X,
                                            cpx<Real> *__restrict__ Y, int nr,
                                            int nu1, bool even, int lane) {
    const Real h0 = (Real)0.5 / (Real)nr;
#pragma unroll 4
    for (int k = lane; k < nu1; k += 64) {
        const cpx<Real> zk = buf[pad(k)];
        const cpx<Real> zn = buf[pad(k == 0 ? 0 : nr - k)];
        const Real h = (even && k == nu1 - 1) ? h0 * (Real)0.5 : h0;
        X[k] = {(zk.re + zn.re) * h, (zk.im - zn.im) * h};
        Y[k] = {(zk.im + zn.im) * h, (zn.re - zk.re) * h};
    }
    wave_sync();
}

// CG4 attenuation points (1,1),(1,3),(3,1),(3,3) (AttBuilder.cpp:141-147).
__device__ __forceinline__ int att_slot(int p, int npa) {
    if (npa == NPEM) return p;
    return p == 6 ? 0 : p == 8 ? 1 : p == 16 ? 2 : p == 18 ? 3 : -1;
}

// Per-slice operands of one GLL point: pointers at slice 0 of element e;
// parameter c of slice k sits at par[c * ps + k], attenuation component
// d of slice k at [d * cs + k].
template <typename Real>
struct SliceMat {
    const Real *par;             // this point's parameter 0
    size_t ps;                   // parameter step, 25 Nr
    const Real *dl, *dm;         // moduli of its attenuation slot
    double *mv;                  // memory variables [nsls][6] of the slot
    Real *ds;                    // memory stress [6] of the slot
    size_t cs;                   // component step of mv, ds: npa Nr
    __device__ Real p(int c, int k) const { return par[c * ps + k]; }
};

// SLS memory of stress component d at slice k (Attenuation::updateMem*,
// AttenuationFull/CG4::apply): sigma -= sum_l m_l; m_l = alpha_l m_l +
// beta_l ds_old + gamma_l ds_new; ds_old = ds_new, where dsn = ds_new is
// the (dlambda, dmu) isotropic stress of the strain in the frame the
// stress is formed in.  mvk, dsk point at slice k of the slot's memory
// variables and memory stress; abg = alpha, beta, gamma in the working
// precision (LDS, [3][CD_MAX_SLS]).  The solids go four at a time so
// their loads are in flight together.
constexpr int CD_MAX_SLS = 16;
// The memory variables and the recursion are FP64 in both precision
// modes (the stress increments stay in the working precision).
template <typename Real, int NS>
__device__ __forceinline__ void att_update(double *mvk, Real *dsk, int cs,
                                           int nsls, const double *abg,
                                           int d, Real dsn, Real &sig) {
    const int od = d * cs;
    const int ls = 6 * cs;
    const double dso = (double)dsk[od];
    const double dn = (double)dsn;
    double sd = (double)sig;
    if constexpr (NS > 0) {
        // the common five-solid case (or any fixed count): unrolled
        double m[NS];
#pragma unroll
        for (int l = 0; l < NS; l++) m[l] = mvk[od + l * ls];
#pragma unroll
        for (int l = 0; l < NS; l++) sd -= m[l];
#pragma unroll
        for (int l = 0; l < NS; l++)
            mvk[od + l * ls] = abg[l] * m[l] + abg[CD_MAX_SLS + l] * dso +
                               abg[2 * CD_MAX_SLS + l] * dn;
    } else {
        for (int l0 = 0; l0 < nsls; l0 += 4) {
            double m[4];
#pragma unroll
            for (int q = 0; q < 4; q++)
                m[q] = l0 + q < nsls ? mvk[od + (l0 + q) * ls] : 0.;
#pragma unroll
            for (int q = 0; q < 4; q++) sd -= m[q];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int l = l0 + q;
                if (l < nsls)
                    mvk[od + l * ls] = abg[l] * m[q] +
                                       abg[CD_MAX_SLS + l] * dso +
                                       abg[2 * CD_MAX_SLS + l] * dn;
            }
        }
    }
    sig = (Real)sd;
    dsk[od] = dsn;
}

// A: gradient of the element-modes [em0, em1) of the 3D element-mode list
// into the scratch rows [d*25 + ipnt][alpha] of each element, rotated to
// RTZ for TI and anisotropic media on spherical meshes (SolidElement.cpp:
// 177-193).  One thread per (element-mode, GLL point), ten element-modes
// per block, the Fourier order fastest so that a row's consecutive orders
// are written by consecutive lanes (CD_EMB = 16 orders: 128-byte runs in
// FP32).
// AX: the launch over the axial elements' element-modes (a.em_ax) that
// forms the strain of their axis row (GLJ, L'Hospital terms) only; the
// launch over all element-modes leaves those points out, which keeps the
// axis code and its registers out of the kernel that does nearly all the
// work.
// PRT: elements with a particle relabelling transformation (an
// axisymmetric undulation, one operand set for all slices) write the
// undulated strain: the 9-component gradient, rotated to RTZ on spherical
// meshes, through the inverse Jacobian.  The operands do not depend on
// phi, so applying them to the Fourier coefficients equals applying them
// slice by slice after the C2R (SolidElement.cpp:139-141).
template <typename Real, bool ROT, bool AX, bool PRT>
__global__ __launch_bounds__(CD_EMB_THREADS) void k_cd_grad(CardArgs<Real> a,
                                                         int em0, int em1) {
    __shared__ cpx<Real> sU[CD_EMB][3][NPEM];
    __shared__ Real sG[2][NPEM];
    const StiffArgs<Real> &s = a.s;
    const int t = threadIdx.x;
    if (t < 2 * NPEM) sG[t / NPEM][t % NPEM] = s.gtab[t];
    const int eml = t % CD_EMB;
    const int ipnt = t < CD_EMB * NPEM ? t / CD_EMB : 0;
    const int em = em0 + xcd_block() * CD_EMB + eml;
    const bool active = t < CD_EMB * NPEM && em < em1;
    const int i = ipnt / NPED, j = ipnt - (ipnt / NPED) * NPED;
    int e = 0, alpha = 0;
    cpx<Real> u[3] = {{0, 0}, {0, 0}, {0, 0}};
    if (active) {
        const EMItem it = (AX ? a.em_ax : s.em_list)[em];
        e = it.elem;
        alpha = it.alpha;
        const int p = s.elem_points[e * NPEM + ipnt];
        if (alpha < s.point_nu1[p]) {
            const double *src = s.displ + (size_t)(s.point_off[p] + alpha) * 6;
#pragma unroll
            for (int c = 0; c < 3; c++)
                u[c] = {(Real)src[2 * c], (Real)src[2 * c + 1]};
        }
#pragma unroll
        for (int c = 0; c < 3; c++) sU[eml][c][ipnt] = u[c];
    }
    __syncthreads();
    if (!active) return;
    const bool axial = s.elem_axial[e] != 0;
    if ((axial && i == 0) != AX) return;
    PointG<Real> G;
    load_point_g<Real>(sG, axial, i, j, G);
    PointOps<Real> g{}, q{};
    load_point_ops(s.elem_ops, e, ipnt, g, q);
    cpx<Real> eps[6];
    if (PRT && s.prt_idx[e] >= 0) {
        cpx<Real> g9[9];
        grad9_point<Real>(sU[eml], u, (Real)alpha, AX, i, j, G, g, g9);
        if (s.rot != nullptr) {
            const Real *r = s.rot + (size_t)e * 4 * NPEM + ipnt;
            rot9<cpx<Real>, Real>(g9, r[0], r[NPEM], r[2 * NPEM],
                                  r[3 * NPEM]);
        }
        prt_sph_to_und6<cpx<Real>, Real>(
            g9, s.prt + (size_t)s.prt_idx[e] * 8 * NPEM + ipnt, eps);
    } else {
        grad6_point<Real>(sU[eml], u, (Real)alpha, AX, i, j, G, g, eps);
        if (ROT) {
            const Real *r = s.rot + (size_t)e * 4 * NPEM + ipnt;
            rot_spz_rtz6<cpx<Real>, Real>(eps, r[0], r[NPEM], r[2 * NPEM],
                                          r[3 * NPEM]);
        }
    }
    const int nu1 = s.elem_nu1[e];
    cpx<Real> *dst = reinterpret_cast<cpx<Real> *>(a.scratch) +
                     a.scr_off[e] + (size_t)ipnt * scr_row(nu1) + alpha;
#pragma unroll
    for (int d = 0; d < 6; d++)
        dst[(size_t)d * NPEM * scr_row(nu1)] = eps[d];
}

// C: quadrature of the element-modes [em0, em1) from the scratch stress
// rows (rotated back from RTZ) into the element-local force buffer.
template <typename Real, bool ROT, bool PRT>
__global__ __launch_bounds__(CD_EMB_THREADS) void k_cd_quad(CardArgs<Real> a,
                                                         int em0, int em1) {
    __shared__ cpx<Real> sX[CD_EMB][3][NPEM];
    __shared__ cpx<Real> sY[CD_EMB][3][NPEM];
    __shared__ Real sG[2][NPEM];
    const StiffArgs<Real> &s = a.s;
    const int t = threadIdx.x;
    if (t < 2 * NPEM) sG[t / NPEM][t % NPEM] = s.gtab[t];
    __syncthreads();
    const int eml = t % CD_EMB;
    const int ipnt = t < CD_EMB * NPEM ? t / CD_EMB : 0;
    const int em = em0 + xcd_block() * CD_EMB + eml;
    const bool active = t < CD_EMB * NPEM && em < em1;
    const int i = ipnt / NPED, j = ipnt - (ipnt / NPED) * NPED;
    int e = 0, alpha = 0, nu1 = 0;
    bool axial = false;
    cpx<Real> sig[6];
    PointOps<Real> g{}, q{};
    const Real *xp = nullptr;
    if (active) {
        const EMItem it = s.em_list[em];
        e = it.elem;
        alpha = it.alpha;
        nu1 = s.elem_nu1[e];
        axial = s.elem_axial[e] != 0;
        const cpx<Real> *src = reinterpret_cast<const cpx<Real> *>(a.scratch) +
                               a.scr_off[e] + (size_t)ipnt * scr_row(nu1) +
                               alpha;
#pragma unroll
        for (int d = 0; d < 6; d++)
            sig[d] = src[(size_t)d * NPEM * scr_row(nu1)];
        if (PRT && s.prt_idx[e] >= 0)
            xp = s.prt + (size_t)s.prt_idx[e] * 8 * NPEM + ipnt;
        if (ROT && xp == nullptr) {
            const Real *r = s.rot + (size_t)e * 4 * NPEM + ipnt;
            rot_rtz_spz6<cpx<Real>, Real>(sig, r[0], r[NPEM], r[2 * NPEM],
                                          r[3 * NPEM]);
        }
        load_point_ops(s.elem_ops, e, ipnt, g, q);
    }
    PointG<Real> G;
    load_point_g<Real>(sG, axial, i, j, G);
    cpx<Real> F[3];
    if (PRT) {
        cpx<Real> L[3];
        if (active && xp != nullptr) {
            // back through the transpose, from RTZ, and computeQuad9
            // (SolidElement.cpp:145-159)
            cpx<Real> f9[9];
            prt_und_to_sph9<cpx<Real>, Real>(sig, xp + 4 * NPEM, f9);
            if (s.rot != nullptr) {
                const Real *r = s.rot + (size_t)e * 4 * NPEM + ipnt;
                rot9<cpx<Real>, Real>(f9, -r[0], r[NPEM], -r[2 * NPEM],
                                      r[3 * NPEM]);
            }
            quad9_prep<Real>(f9, (Real)alpha, axial && i == 0, q, sX[eml],
                             sY[eml], ipnt, L);
        } else if (active) {
            quad6_prep<Real>(sig, (Real)alpha, axial && i == 0, q, sX[eml],
                             sY[eml], ipnt, L);
        }
        __syncthreads();
        if (active) quad6_finish<Real>(sX[eml], sY[eml], i, j, G, L, F);
    } else {
        quad6_point<Real>(sig, (Real)alpha, axial && i == 0, q, sX[eml],
                          sY[eml], ipnt, i, j, G, active, F);
    }
    if (active && alpha < s.point_nu1[s.elem_points[e * NPEM + ipnt]]) {
        Real *dst = s.eb + ((size_t)s.elem_eb_off[e] + (size_t)ipnt * nu1 +
                            alpha) * 6;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            dst[2 * c] = F[c].re;
            dst[2 * c + 1] = F[c].im;
        }
    }
}

// B: stress of one GLL point of one element per block of three waves.
// Wave q packs strain pair q (rows 2q, 2q+1: (ss, pp), (zz, pz), (sz, sp))
// as x + i y into its own LDS buffer and inverse-transforms it; after a
// block barrier the three waves form the stress slice by slice from all
// six strains (Isotropic::strainToStress, Isotropic.hpp:155-172;
// TransverselyIsotropic::strainToStress, TransverselyIsotropic.hpp:161-
// 180; Anisotropic::strainToStress, Anisotropic.hpp:203-274) with the SLS
// memory update, write it back in place, and after a second barrier each
// wave forward-transforms its pair and unpacks it into the scratch rows.
// Between the transforms slice k sits at buffer position pos[k] (the
// digit reversal of the in-place FFT); the slice loop runs over slices so
// its global accesses stay consecutive.
template <typename Real, int NS, int MT>
__device__ __forceinline__ void cd_stress_slices(const CardArgs<Real> &a,
                                                 cpx<Real> *b0, cpx<Real> *b1,
                                                 cpx<Real> *b2,
                                                 const uint16_t *pos,
                                                 const SliceMat<Real> &M,
                                                 const double *abg, bool att,
                                                 int nr, int t) {
    // two slices in flight without attenuation; with it the memory
    // variables of one slice already fill the registers
    constexpr int UNR = NS != 0 ? 1 : 2;
#pragma unroll UNR
    for (int k = t; k < nr; k += CD_PT) {
        const int p = pad(pos[k]);
        const cpx<Real> z01 = b0[p], z23 = b1[p], z45 = b2[p];
        const Real e[6] = {z01.re, z01.im, z23.re, z23.im, z45.re, z45.im};
        Real sv[6];
        if constexpr (MT == 0) {
            const Real lam = M.p(0, k), mu = M.p(1, k);
            const Real mu2 = mu * (Real)2;
            const Real tl = (e[0] + e[1] + e[2]) * lam;
            sv[0] = tl + e[0] * mu2;
            sv[1] = tl + e[1] * mu2;
            sv[2] = tl + e[2] * mu2;
            sv[3] = e[3] * mu;
            sv[4] = e[4] * mu;
            sv[5] = e[5] * mu;
        } else if constexpr (MT == 1) {
            const Real A = M.p(0, k), C = M.p(1, k), F = M.p(2, k);
            const Real Lm = M.p(3, k), N = M.p(4, k);
            const Real mN2 = N * (Real)-2;
            const Real t3 = e[0] + e[1];
            const Real t4 = t3 * A + e[2] * F;
            sv[0] = t4 + e[1] * mN2;
            sv[1] = t4 + e[0] * mN2;
            sv[2] = t3 * F + e[2] * C;
            sv[3] = e[3] * Lm;
            sv[4] = e[4] * Lm;
            sv[5] = e[5] * N;
        } else {
            aniso_hooke<Real, Real>(e, [&](int c) { return M.p(c, k); }, sv);
        }
        if (NS != 0 && att) {
            const Real dl = M.dl[k], dm = M.dm[k];
            const Real dm2 = dm * (Real)2;
            const Real tr = e[0] + e[1] + e[2];
            double *mvk = M.mv + k;
            Real *dsk = M.ds + k;
            const int cs = (int)M.cs;
#pragma unroll
            for (int d = 0; d < 6; d++)
                att_update<Real, NS>(
                    mvk, dsk, cs, a.nsls, abg, d,
                    d < 3 ? tr * dl + e[d] * dm2 : e[d] * dm, sv[d]);
        }
        b0[p] = {sv[0], sv[1]};
        b1[p] = {sv[2], sv[3]};
        b2[p] = {sv[4], sv[5]};
    }
}

// B kernel.  MT = material type of the batch's elements (0 isotropic, 1
// TI, 2 anisotropic); NS = number of SLS (0: no attenuation, -1: taken
// from the arguments).  25 blocks per element of the batch list3d[w0 ..),
// one per GLL point.
template <typename Real, int NS, int MT>
__global__ __launch_bounds__(CD_PT, CardOcc<Real>::value) void k_cd_stress(
    CardArgs<Real> a, int w0) {
    __shared__ double abg[3 * CD_MAX_SLS];    // SLS alpha, beta, gamma
    const int t = threadIdx.x;
    const int lane = t & 63;
    // the wave index is uniform across the wave: keep it (and all that is
    // derived from it) in scalar registers
    const int q = __builtin_amdgcn_readfirstlane(t >> 6);
    const int w = w0 + (int)blockIdx.x / NPEM;
    const int pp = (int)blockIdx.x % NPEM;
    const int e = a.list3d[w];
    const int nu1 = a.s.elem_nu1[e];
    const int nr = a.elem_nr[e];
    const uint32_t *plan = a.plan + (size_t)nr * CARD_PLAN_W;
    FftLds<Real> L;
    L.carve(fft_lds, nr, CD_PW);
    fft_load_tables<Real>(L, a.twiddle, a.tw_off, a.slice_pos, a.pos_off,
                          plan, nr, t, CD_PT);
    if (NS != 0 && t < 3 * a.nsls) {
        const int c = t / a.nsls, l = t - c * a.nsls;
        abg[c * CD_MAX_SLS + l] = a.abg[t];
    }
    __syncthreads();
    const bool even = (nr & 1) == 0;
    const size_t rs = (size_t)NPEM * scr_row(nu1);
    cpx<Real> *r0 = reinterpret_cast<cpx<Real> *>(a.scratch) + a.scr_off[e] +
                    (size_t)pp * scr_row(nu1);
    cpx<Real> *b = L.buf(q);
    pack_pair<Real>(b, r0 + 2 * q * rs, r0 + (2 * q + 1) * rs, nr, nu1, even,
                    lane);
    // both transforms from one call site (the inverse, the stress, then
    // the forward transform), which halves the FFT's code
#pragma unroll 1
    for (int pass = 0; pass < 2; pass++) {
        wave_fft<Real>(b, plan, pass == 0, L.tw, lane);
        if (pass == 1) break;
        __syncthreads();
    {
        SliceMat<Real> M;
        M.ps = (size_t)NPEM * nr;
        M.par = a.prop3d + a.prop_off[e] + (size_t)pp * nr;
        const int ip4 = NS != 0 ? att_slot(pp, a.npa) : -1;
        const bool att = NS != 0 && ip4 >= 0;
        if (NS != 0 && att) {
            const int64_t sl0 = a.slice_off[e];
            const int npa = a.npa;
            M.cs = (size_t)npa * nr;
            M.dl = a.att3d + (size_t)sl0 * 2 * npa + (size_t)ip4 * nr;
            M.dm = M.dl + M.cs;
            M.mv = a.memvar + (size_t)sl0 * a.nsls * 6 * npa +
                   (size_t)ip4 * nr;
            M.ds = a.dstress + (size_t)sl0 * 6 * npa + (size_t)ip4 * nr;
        }
        cd_stress_slices<Real, NS, MT>(a, L.buf(0), L.buf(1), L.buf(2),
                                        L.pos, M, abg, att, nr, t);
    }
        __syncthreads();
    }
    unpack_pair<Real>(b, r0 + 2 * q * rs, r0 + (2 * q + 1) * rs, nr, nu1,
                      even, lane);
}

template <typename Real, int MT>
static void launch_stress_mt(hipStream_t st, const CardArgs<Real> &a, int w0,
                             int n, int nr_max) {
    const dim3 grid(n * NPEM), block(CD_PT);
    const size_t lds = fft_lds_bytes<Real>(nr_max, CD_PW);
    if (a.nsls == 0)
        hipLaunchKernelGGL((k_cd_stress<Real, 0, MT>), grid, block, lds, st,
                           a, w0);
    else if (a.nsls == 5)
        hipLaunchKernelGGL((k_cd_stress<Real, 5, MT>), grid, block, lds, st,
                           a, w0);
    else
        hipLaunchKernelGGL((k_cd_stress<Real, -1, MT>), grid, block, lds, st,
                           a, w0);
}

// The three launches of a batch (gradient, stress, quadrature); batches
// hold elements of one material type (list3d is grouped by type, each
// group heaviest first) whose strain and stress fit the scratch together.
template <typename Real>
void launch_cd_grad(hipStream_t st, const CardArgs<Real> &a,
                    const CardBatch &bt) {
    const int gem = (bt.em_end - bt.em_begin + CD_EMB - 1) / CD_EMB;
    if (gem <= 0) return;
    const bool rot = bt.mtype != 0 && a.s.rot != nullptr;
    const bool prt = a.s.prt_idx != nullptr;
#define AXS3D_CDG(ROT_, AX_, G_, B_, E_)                                      \
    do {                                                                      \
        if (prt)                                                              \
            hipLaunchKernelGGL((k_cd_grad<Real, ROT_, AX_, true>), dim3(G_),  \
                               dim3(CD_EMB_THREADS), 0, st, a, B_, E_);       \
        else                                                                  \
            hipLaunchKernelGGL((k_cd_grad<Real, ROT_, AX_, false>), dim3(G_), \
                               dim3(CD_EMB_THREADS), 0, st, a, B_, E_);       \
    } while (0)
    if (rot) AXS3D_CDG(true, false, gem, bt.em_begin, bt.em_end);
    else AXS3D_CDG(false, false, gem, bt.em_begin, bt.em_end);
    const int gax = (bt.ax_end - bt.ax_begin + CD_EMB - 1) / CD_EMB;
    if (gax <= 0) return;
    if (rot) AXS3D_CDG(true, true, gax, bt.ax_begin, bt.ax_end);
    else AXS3D_CDG(false, true, gax, bt.ax_begin, bt.ax_end);
#undef AXS3D_CDG
}

template <typename Real>
void launch_cd_stress(hipStream_t st, const CardArgs<Real> &a,
                      const CardBatch &bt) {
    const int n = bt.end - bt.begin;
    if (n <= 0) return;
    switch (bt.mtype) {
        case 0: launch_stress_mt<Real, 0>(st, a, bt.begin, n, bt.nr_max); break;
        case 1: launch_stress_mt<Real, 1>(st, a, bt.begin, n, bt.nr_max); break;
        default:
            launch_stress_mt<Real, 2>(st, a, bt.begin, n, bt.nr_max);
            break;
    }
}

template <typename Real>
void launch_cd_quad(hipStream_t st, const CardArgs<Real> &a,
                    const CardBatch &bt) {
    const int gem = (bt.em_end - bt.em_begin + CD_EMB - 1) / CD_EMB;
    if (gem <= 0) return;
    const bool rot = bt.mtype != 0 && a.s.rot != nullptr;
    const bool prt = a.s.prt_idx != nullptr;
    if (rot && prt)
        hipLaunchKernelGGL((k_cd_quad<Real, true, true>), dim3(gem),
                           dim3(CD_EMB_THREADS), 0, st, a, bt.em_begin,
                           bt.em_end);
    else if (rot)
        hipLaunchKernelGGL((k_cd_quad<Real, true, false>), dim3(gem),
                           dim3(CD_EMB_THREADS), 0, st, a, bt.em_begin,
                           bt.em_end);
    else if (prt)
        hipLaunchKernelGGL((k_cd_quad<Real, false, true>), dim3(gem),
                           dim3(CD_EMB_THREADS), 0, st, a, bt.em_begin,
                           bt.em_end);
    else
        hipLaunchKernelGGL((k_cd_quad<Real, false, false>), dim3(gem),
                           dim3(CD_EMB_THREADS), 0, st, a, bt.em_begin,
                           bt.em_end);
}

// ------------------------------------------------- fluid, 3D medium
// FluidElement::displToStiff, cardinal branch (src/core/element/
// FluidElement.cpp:101-148): computeGrad3 (GradientQuadrature.hpp:163-
// 175) of the displacement potential into the scratch rows [c][25][nu1]
// (c = dp/ds, i alpha p / s, dp/dz), C2R of the three series, K = 1 / rho
// per slice (Acoustic::strainToStress _3D_CD, Acoustic.hpp:80-92), R2C,
// computeQuad3 (GradientQuadrature.hpp:234-252) into the fluid
// element-local forces.  The same three-kernel split as the solids.
template <typename Real>
__global__ __launch_bounds__(CD_EMB_THREADS) void k_cdf_grad(
    FluidCardArgs<Real> a, int em0, int em1) {
    __shared__ cpx<Real> sP[CD_EMB][NPEM];
    __shared__ Real sG[2][NPEM];
    const FluidArgs<Real> &f = a.f;
    const int t = threadIdx.x;
    if (t < 2 * NPEM) sG[t / NPEM][t % NPEM] = f.gtab[t];
    const int eml = t % CD_EMB;
    const int ipnt = t < CD_EMB * NPEM ? t / CD_EMB : 0;
    const int em = em0 + xcd_block() * CD_EMB + eml;
    const bool active = t < CD_EMB * NPEM && em < em1;
    const int i = ipnt / NPED, j = ipnt - (ipnt / NPED) * NPED;
    int e = 0, alpha = 0;
    cpx<Real> p = {0, 0};
    if (active) {
        const EMItem it = f.em_list[em];
        e = it.elem;
        alpha = it.alpha;
        const int ip = f.elem_points[e * NPEM + ipnt];
        if (alpha < f.point_nu1[ip]) {
            const double *src = f.chi + (size_t)(f.point_off[ip] + alpha) * 2;
            p = {(Real)src[0], (Real)src[1]};
        }
        sP[eml][ipnt] = p;
    }
    __syncthreads();
    if (!active) return;
    const bool axial = f.elem_axial[e] != 0;
    PointG<Real> G;
    load_point_g<Real>(sG, axial, i, j, G);
    PointOps<Real> g{}, q{};
    load_point_ops(f.elem_ops, e, ipnt, g, q);
    cpx<Real> x = {0, 0}, y = {0, 0};
#pragma unroll
    for (int k = 0; k < NPED; k++) {
        x = cfma(sP[eml][k * NPED + j], G.gxi_col[k], x);
        y = cfma(sP[eml][i * NPED + k], G.g_col[k], y);
    }
    const Real ralpha = (Real)alpha;
    const cpx<Real> dpds = cadd(cmul(x, g.dzdeta), cmul(y, g.dzdxii));
    const cpx<Real> dpdz = cadd(cmul(x, g.dsdeta), cmul(y, g.dsdxii));
    const cpx<Real> ps = (axial && i == 0) ? ialpha(ralpha, dpds)
                                           : cmul(ialpha(ralpha, p), g.invs);
    const int nu1 = f.elem_nu1[e];
    cpx<Real> *dst = reinterpret_cast<cpx<Real> *>(a.scratch) + a.scr_off[e] +
                     (size_t)ipnt * scr_row(nu1) + alpha;
    const size_t rs = (size_t)NPEM * scr_row(nu1);
    dst[0] = dpds;
    dst[rs] = ps;
    dst[2 * rs] = dpdz;
}

// One fluid GLL point per block of two waves: wave 0 transforms the pair
// (dp/ds, p_phi), wave 1 dp/dz (y = 0); each scales its own buffer by K
// slice by slice and transforms it back.
template <typename Real>
__global__ __launch_bounds__(128) void k_cdf_stress(FluidCardArgs<Real> a,
                                                    int w0) {
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int q = __builtin_amdgcn_readfirstlane(t >> 6);
    const int w = w0 + (int)blockIdx.x / NPEM;
    const int pp = (int)blockIdx.x % NPEM;
    const int e = a.list3d[w];
    const int nr = a.elem_nr[e];
    const int nu1 = nr / 2 + 1;
    const uint32_t *plan = a.plan + (size_t)nr * CARD_PLAN_W;
    FftLds<Real> L;
    L.carve(fft_lds, nr, 2);
    fft_load_tables<Real>(L, a.twiddle, a.tw_off, a.slice_pos, a.pos_off,
                          plan, nr, t, 128);
    const cpx<Real> *tw = L.tw;
    const uint16_t *pos = L.pos;
    __syncthreads();
    const bool even = (nr & 1) == 0;
    const size_t rs = (size_t)NPEM * scr_row(nu1);
    cpx<Real> *r0 = reinterpret_cast<cpx<Real> *>(a.scratch) + a.scr_off[e] +
                    (size_t)pp * scr_row(nu1);
    cpx<Real> *b = L.buf(q);
    // wave 1: rows (2, zero); the zero row is read as row 2's imaginary
    // partner, so pack it from row 2 and clear the y part
    if (q == 0) {
        pack_pair<Real>(b, r0, r0 + rs, nr, nu1, even, lane);
    } else {
        const cpx<Real> *X = r0 + 2 * rs;
        const int nyq = even ? nu1 - 1 : -1;
        for (int k = lane; k < nr; k += 64) {
            const bool lo = k < nu1;
            const cpx<Real> x = X[lo ? k : nr - k];
            const bool real_only = k == 0 || k == nyq;
            const Real sc = k == nyq ? (Real)2 : (Real)1;
            const Real sim = real_only ? (Real)0 : (lo ? (Real)1 : (Real)-1);
            b[pad(k)] = {sc * x.re, sim * x.im};
        }
        wave_sync();
    }
    wave_fft<Real>(b, plan, true, tw, lane);
    const Real *K = a.k3d + a.k3d_off[e] + (size_t)pp * nr;
    for (int k = lane; k < nr; k += 64) {
        const int pk = pad(pos[k]);
        const Real kk = K[k];
        const cpx<Real> z = b[pk];
        b[pk] = {z.re * kk, z.im * kk};
    }
    wave_sync();
    wave_fft<Real>(b, plan, false, tw, lane);
    if (q == 0) {
        unpack_pair<Real>(b, r0, r0 + rs, nr, nu1, even, lane);
    } else {
        cpx<Real> *X = r0 + 2 * rs;
        const Real h0 = (Real)0.5 / (Real)nr;
        for (int k = lane; k < nu1; k += 64) {
            const cpx<Real> zk = b[pad(k)];
            const cpx<Real> zn = b[pad(k == 0 ? 0 : nr - k)];
            const Real h = (even && k == nu1 - 1) ? h0 * (Real)0.5 : h0;
            X[k] = {(zk.re + zn.re) * h, (zk.im - zn.im) * h};
        }
    }
}

template <typename Real>
__global__ __launch_bounds__(CD_EMB_THREADS) void k_cdf_quad(
    FluidCardArgs<Real> a, int em0, int em1) {
    __shared__ cpx<Real> sX[CD_EMB][NPEM];
    __shared__ cpx<Real> sY[CD_EMB][NPEM];
    __shared__ Real sG[2][NPEM];
    const FluidArgs<Real> &f = a.f;
    const int t = threadIdx.x;
    if (t < 2 * NPEM) sG[t / NPEM][t % NPEM] = f.gtab[t];
    const int eml = t % CD_EMB;
    const int ipnt = t < CD_EMB * NPEM ? t / CD_EMB : 0;
    const int em = em0 + xcd_block() * CD_EMB + eml;
    const bool active = t < CD_EMB * NPEM && em < em1;
    const int i = ipnt / NPED, j = ipnt - (ipnt / NPED) * NPED;
    int e = 0, alpha = 0, nu1 = 0;
    bool axial = false;
    cpx<Real> Lt = {0, 0};
    if (active) {
        const EMItem it = f.em_list[em];
        e = it.elem;
        alpha = it.alpha;
        nu1 = f.elem_nu1[e];
        axial = f.elem_axial[e] != 0;
        const cpx<Real> *src = reinterpret_cast<const cpx<Real> *>(a.scratch) +
                               a.scr_off[e] + (size_t)ipnt * scr_row(nu1) +
                               alpha;
        const size_t rs = (size_t)NPEM * scr_row(nu1);
        const cpx<Real> q0 = src[0], q1 = src[rs], q2 = src[2 * rs];
        PointOps<Real> g{}, q{};
        load_point_ops(f.elem_ops, e, ipnt, g, q);
        cpx<Real> X, Y;
        quad_xy(q0, q2, q, X, Y);
        const cpx<Real> bq = ialpha((Real)alpha, q1);
        Lt = cmul(bq, -q.invs);
        if (axial && i == 0) {
            X = csub(X, cmul(bq, q.dzdeta));
            Y = csub(Y, cmul(bq, q.dzdxii));
        }
        sX[eml][ipnt] = X;
        sY[eml][ipnt] = Y;
    }
    __syncthreads();
    if (!active) return;
    if (alpha >= f.point_nu1[f.elem_points[e * NPEM + ipnt]]) return;
    PointG<Real> G;
    load_point_g<Real>(sG, axial, i, j, G);
    cpx<Real> acc = {0, 0}, bsum = {0, 0};
#pragma unroll
    for (int k = 0; k < NPED; k++) {
        acc = cfma(sX[eml][k * NPED + j], G.gxi_row[k], acc);
        bsum = cfma(sY[eml][i * NPED + k], G.g_row[k], bsum);
    }
    const cpx<Real> F = cadd(cadd(acc, bsum), Lt);
    Real *dst = f.eb + ((size_t)f.elem_eb_off[e] + (size_t)ipnt * nu1 +
                        alpha) * 2;
    dst[0] = F.re;
    dst[1] = F.im;
}

template <typename Real>
void launch_cdf_batch(hipStream_t st, const FluidCardArgs<Real> &a,
                      const CardBatch &bt) {
    const int n = bt.end - bt.begin;
    const int gem = (bt.em_end - bt.em_begin + CD_EMB - 1) / CD_EMB;
    if (n <= 0 || gem <= 0) return;
    hipLaunchKernelGGL(k_cdf_grad<Real>, dim3(gem), dim3(CD_EMB_THREADS), 0,
                       st, a, bt.em_begin, bt.em_end);
    hipLaunchKernelGGL(k_cdf_stress<Real>, dim3(n * NPEM), dim3(128),
                       fft_lds_bytes<Real>(bt.nr_max, 2), st, a, bt.begin);
    hipLaunchKernelGGL(k_cdf_quad<Real>, dim3(gem), dim3(CD_EMB_THREADS), 0,
                       st, a, bt.em_begin, bt.em_end);
}

// Mass3D::computeAccel (src/core/point/mass/Mass3D.cpp:56-68) for the
// solid points with a 3D mass, after the point kernel has assembled their
// stiffness (axis and Nyquist rules applied): C2R, times 1/m slice by
// slice, R2C.  One block of two waves per point: wave 0 takes components
// (s, phi) packed as x + i y, wave 1 component z (y = 0).  With MARCH the
// result is the acceleration of NewmarkTimeScheme::update
// (NewmarkTimeScheme.hpp:48-59, FP64); without, it goes back to stiff.
template <typename Real, bool MARCH>
__global__ __launch_bounds__(128) void k_mass3d(Mass3DArgs<Real> a,
                                                Real *__restrict__ stiff) {
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int q = __builtin_amdgcn_readfirstlane(t >> 6);
    const int i = blockIdx.x;
    const int p = a.point[i];
    const int nr = a.point_nr[p];
    const int nu1 = nr / 2 + 1;
    const bool even = (nr & 1) == 0;
    const uint32_t *plan = a.plan + (size_t)nr * CARD_PLAN_W;
    FftLds<Real> L;
    L.carve(fft_lds, nr, 2);
    fft_load_tables<Real>(L, a.twiddle, a.tw_off, a.slice_pos, a.pos_off,
                          plan, nr, t, 128);
    const cpx<Real> *tw = L.tw;
    const uint16_t *pos = L.pos;
    __syncthreads();
    // component c of mode m of the point: stiff complex [off + m][3]
    cpx<Real> *S = reinterpret_cast<cpx<Real> *>(stiff) +
                   (size_t)a.point_off[p] * 3;
    const int cx = q == 0 ? 0 : 2;
    cpx<Real> *b = L.buf(q);
    // C2R input (SolverFFTW conventions, as pack_pair): Hermitian
    // extension of the modes, mean and Nyquist real, Nyquist doubled
    const int nyq = even ? nu1 - 1 : -1;
    for (int k = lane; k < nr; k += 64) {
        const bool lo = k < nu1;
        const int kk = lo ? k : nr - k;
        const cpx<Real> x = S[(size_t)kk * 3 + cx];
        const cpx<Real> y =
            q == 0 ? S[(size_t)kk * 3 + 1] : cpx<Real>{(Real)0, (Real)0};
        const bool real_only = k == 0 || k == nyq;
        const Real sc = k == nyq ? (Real)2 : (Real)1;
        const Real sim = real_only ? (Real)0 : (lo ? (Real)1 : (Real)-1);
        b[pad(k)] = {sc * x.re - sim * y.im, sim * x.im + sc * y.re};
    }
    wave_sync();
    wave_fft<Real>(b, plan, true, tw, lane);
    const Real *im = a.inv_mass + a.inv_off[i];
    const int64_t ko = a.k_off != nullptr ? a.k_off[i] : -1;
    if (ko >= 0) {
        // MassOceanLoad3D::computeAccel (MassOceanLoad3D.cpp:41-56): per
        // slice a = F / m - (F . k) k, with F = (s, phi) from wave 0's
        // buffer and z from wave 1's (block-uniform branch)
        __syncthreads();
        cpx<Real> *b0 = L.buf(0), *b1 = L.buf(1);
        const Real *kk = a.ocean_k + ko;
        for (int k = t; k < nr; k += 128) {
            const int pk = pad(pos[k]);
            const Real m = im[k];
            const Real ks = kk[3 * k], kp = kk[3 * k + 1], kz = kk[3 * k + 2];
            const cpx<Real> sp = b0[pk];
            const Real fz = b1[pk].re;
            const Real dot = sp.re * ks + sp.im * kp + fz * kz;
            b0[pk] = {sp.re * m - dot * ks, sp.im * m - dot * kp};
            b1[pk] = {fz * m - dot * kz, (Real)0};
        }
        __syncthreads();
    } else {
        for (int k = lane; k < nr; k += 64) {
            const int pk = pad(pos[k]);
            const Real m = im[k];
            const cpx<Real> z = b[pk];
            b[pk] = {z.re * m, z.im * m};
        }
        wave_sync();
    }
    wave_fft<Real>(b, plan, false, tw, lane);
    // R2C output (as unpack_pair): X into component cx, Y into 1 (wave 0)
    const Real h0 = (Real)0.5 / (Real)nr;
    bool bad = false, bad0 = false;
    const int ai = a.abc.idx != nullptr ? a.abc.idx[p] : -1;
    for (int k = lane; k < nu1; k += 64) {
        const cpx<Real> zk = b[pad(k)];
        const cpx<Real> zn = b[pad(k == 0 ? 0 : nr - k)];
        const Real h = k == nyq ? h0 * (Real)0.5 : h0;
        cpx<Real> v[2];
        v[0] = {(zk.re + zn.re) * h, (zk.im - zn.im) * h};
        v[1] = {(zk.im + zn.im) * h, (zn.re - zk.re) * h};
        const int nc = q == 0 ? 2 : 1;
        for (int c = 0; c < nc; c++) {
            const size_t m6 = ((size_t)a.point_off[p] + k) * 6 +
                              2 * (q == 0 ? c : 2);
            if (!MARCH) {
                stiff[m6] = v[c].re;
                stiff[m6 + 1] = v[c].im;
                continue;
            }
            const double hdt = .5 * a.dt, hdt2 = .5 * a.dt * a.dt;
            double an[2] = {(double)v[c].re, (double)v[c].im};
            if (ai >= 0) {
                // sponge (Sponge.hpp:74-96), after the mass
                const double *tt =
                    a.abc.sterm +
                    ((size_t)a.abc.term_off[ai] + k) * 6 + 2 * (q == 0 ? c : 2);
                an[0] -= tt[0];
                an[1] -= tt[1];
            }
#pragma unroll
            for (int ri = 0; ri < 2; ri++) {
                const double u0 = a.displ[m6 + ri];
                double vv, uu;
                if (a.scheme == 1) {
                    // SymplecticTimeScheme::update
                    vv = a.veloc[m6 + ri] + a.sym_pi * a.dt * an[ri];
                    uu = u0 + a.sym_kappa * a.dt * vv;
                } else {
                    vv = a.veloc[m6 + ri] +
                         hdt * (a.accel[m6 + ri] + an[ri]);
                    uu = u0 + (a.dt * vv + hdt2 * an[ri]);
                    a.accel[m6 + ri] = an[ri];
                }
                a.veloc[m6 + ri] = vv;
                a.displ[m6 + ri] = uu;
                bad |= !isfinite(uu);
                bad0 |= !isfinite(u0);
            }
        }
    }
    // bit 0: started from a non-finite displacement, bit 1: left one
    if (MARCH && (bad0 || bad))
        atomicOr(a.flag, (bad0 ? 1 : 0) | (bad ? 2 : 0));
}

template <typename Real>
void launch_mass3d(hipStream_t st, bool march, const Mass3DArgs<Real> &a,
                   Real *stiff) {
    if (a.n <= 0) return;
    const size_t lds = fft_lds_bytes<Real>(a.nr_max, 2);
    if (march)
        hipLaunchKernelGGL((k_mass3d<Real, true>), dim3(a.n), dim3(128), lds,
                           st, a, stiff);
    else
        hipLaunchKernelGGL((k_mass3d<Real, false>), dim3(a.n), dim3(128),
                           lds, st, a, stiff);
}

// C2R input of a pair of real series from FP64 Fourier columns (st
// doubles between modes, re/im interleaved): buf[k] = Xfull[k] + i
// Yfull[k] under SolverFFTW's C2R rules, as pack_pair; Y null: zero.
template <typename Real>
__device__ __forceinline__ void pack_fields(cpx<Real> *buf,
                                            const double *__restrict__ X,
                                            const double *__restrict__ Y,
                                            int st, int nr, int nu1,
                                            bool even, int lane) {
    const int nyq = even ? nu1 - 1 : -1;
    for (int k = lane; k < nr; k += 64) {
        const bool lo = k < nu1;
        const size_t kk = (size_t)(lo ? k : nr - k) * st;
        const cpx<Real> x = {(Real)X[kk], (Real)X[kk + 1]};
        const cpx<Real> y = Y != nullptr
                                ? cpx<Real>{(Real)Y[kk], (Real)Y[kk + 1]}
                                : cpx<Real>{(Real)0, (Real)0};
        const bool real_only = k == 0 || k == nyq;
        const Real sc = k == nyq ? (Real)2 : (Real)1;
        const Real sim = real_only ? (Real)0 : (lo ? (Real)1 : (Real)-1);
        buf[pad(k)] = {sc * x.re - sim * y.im, sim * x.im + sc * y.re};
    }
    wave_sync();
}

// R2C output of a pair (as unpack_pair), added to FP64 Fourier columns.
template <typename Real>
__device__ __forceinline__ void unpack_add(const cpx<Real> *buf,
                                           double *__restrict__ X,
                                           double *__restrict__ Y, int st,
                                           int nr, int nu1, bool even,
                                           int lane) {
    const Real h0 = (Real)0.5 / (Real)nr;
    for (int k = lane; k < nu1; k += 64) {
        const cpx<Real> zk = buf[pad(k)];
        const cpx<Real> zn = buf[pad(k == 0 ? 0 : nr - k)];
        const Real h = (even && k == nu1 - 1) ? h0 * (Real)0.5 : h0;
        const size_t m = (size_t)k * st;
        X[m] += (double)((zk.re + zn.re) * h);
        X[m + 1] += (double)((zk.im - zn.im) * h);
        if (Y != nullptr) {
            Y[m] += (double)((zk.im + zn.im) * h);
            Y[m + 1] += (double)((zn.re - zk.re) * h);
        }
    }
}

// Terms of the absorbing boundaries of one step (AbcArgs), one block of
// three waves per boundary point; NC = 3 (solid) or 1 (fluid).
//  1D parts, order by order: the ClaytonSolid1D force (ClaytonSolid1D.cpp:
//    16-29) or ClaytonFluid1D's v area / (rho vp) (ClaytonFluid1D.cpp:
//    15-22), and the sponge's 2 gamma v + gamma^2 u (Sponge.hpp:80-83).
//  3D Clayton (ClaytonSolid3D.cpp:43-67, ClaytonFluid3D.cpp:33-51): the
//    velocity to the slices (wave 0: s + i phi, wave 1: z; fluid: wave 0),
//    the operator slice by slice, back to orders.
//  3D sponge (Sponge.hpp:84-92): wave c transforms u_c + i v_c, slice by
//    slice w_c = 2 gamma v_c + gamma^2 u_c, back to orders (the point
//    kernel subtracts the term from the acceleration and drops the
//    imaginary part of its order 0, as the reference's C2R/R2C round trip
//    of the acceleration does).
template <typename Real, int NC>
__global__ __launch_bounds__(192) void k_abc(AbcArgs<Real> a) {
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int q = __builtin_amdgcn_readfirstlane(t >> 6);
    const int i = blockIdx.x;
    const int p = a.point[i];
    const int nr = a.point_nr[p];
    const int nu1 = nr / 2 + 1;
    const bool even = (nr & 1) == 0;
    const int st = 2 * NC;
    const double *U = a.displ + (size_t)a.point_off[p] * st;
    const double *V = a.veloc + (size_t)a.point_off[p] * st;
    double *CT = a.cterm + (size_t)a.term_off[i] * st;
    double *ST = a.sterm + (size_t)a.term_off[i] * st;
    const double *c1 = a.cl1 + (size_t)i * (NC == 3 ? 4 : 1);
    const double g1 = a.sp1[2 * i], g2 = a.sp1[2 * i + 1];
    for (int al = t; al < nu1; al += 192) {
        const size_t m = (size_t)al * st;
#pragma unroll
        for (int ri = 0; ri < 2; ri++) {
            if (NC == 3) {
                const double vs = V[m + ri], vp = V[m + 2 + ri],
                             vz = V[m + 4 + ri];
                CT[m + ri] = c1[0] * vs + c1[2] * vz;
                CT[m + 2 + ri] = c1[3] * vp;
                CT[m + 4 + ri] = c1[2] * vs + c1[1] * vz;
            } else {
                CT[m + ri] = c1[0] * V[m + ri];
            }
#pragma unroll
            for (int c = 0; c < NC; c++)
                ST[m + 2 * c + ri] =
                    g1 * V[m + 2 * c + ri] + g2 * U[m + 2 * c + ri];
        }
    }
    const int64_t c3 = a.cl3_off[i], s3 = a.sp3_off[i];
    if (c3 < 0 && s3 < 0) return;
    const uint32_t *plan = a.plan + (size_t)nr * CARD_PLAN_W;
    FftLds<Real> L;
    L.carve(fft_lds, nr, 3);
    fft_load_tables<Real>(L, a.twiddle, a.tw_off, a.slice_pos, a.pos_off,
                          plan, nr, t, 192);
    __syncthreads();
    cpx<Real> *b = L.buf(q);
    constexpr int NW = NC == 3 ? 2 : 1;   // waves of the real outputs
    if (c3 >= 0) {
        if (q < NW) {
            pack_fields<Real>(b, V + (q == 0 ? 0 : 4),
                              (NC == 3 && q == 0) ? V + 2 : nullptr, st, nr,
                              nu1, even, lane);
            wave_fft<Real>(b, plan, true, L.tw, lane);
        }
        __syncthreads();
        const Real *M = a.cl3 + (size_t)c3 * (NC == 3 ? 6 : 1);
        for (int k = t; k < nr; k += 192) {
            const int pk = pad(L.pos[k]);
            if (NC == 3) {
                const cpx<Real> z0 = L.buf(0)[pk], z1 = L.buf(1)[pk];
                const Real vs = z0.re, vp = z0.im, vz = z1.re;
                const Real *m = M + (size_t)k * 6;
                L.buf(0)[pk] = {m[0] * vs + m[3] * vp + m[4] * vz,
                                m[3] * vs + m[1] * vp + m[5] * vz};
                L.buf(1)[pk] = {m[4] * vs + m[5] * vp + m[2] * vz, (Real)0};
            } else {
                L.buf(0)[pk] = {M[k] * L.buf(0)[pk].re, (Real)0};
            }
        }
        __syncthreads();
        if (q < NW) {
            wave_fft<Real>(b, plan, false, L.tw, lane);
            unpack_add<Real>(b, CT + (q == 0 ? 0 : 4),
                             (NC == 3 && q == 0) ? CT + 2 : nullptr, st, nr,
                             nu1, even, lane);
        }
        __syncthreads();
    }
    if (s3 >= 0) {
        if (q < NC) {
            pack_fields<Real>(b, U + 2 * q, V + 2 * q, st, nr, nu1, even,
                              lane);
            wave_fft<Real>(b, plan, true, L.tw, lane);
        }
        __syncthreads();
        const Real *G = a.sp3 + (size_t)s3 * 2;
        for (int k = t; k < nr; k += 192) {
            const int pk = pad(L.pos[k]);
            const Real ga = G[2 * k], gb = G[2 * k + 1];
            Real w[3] = {(Real)0, (Real)0, (Real)0};
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const cpx<Real> z = L.buf(c)[pk];
                w[c] = ga * z.im + gb * z.re;
            }
            L.buf(0)[pk] = {w[0], NC == 3 ? w[1] : (Real)0};
            if (NC == 3) L.buf(1)[pk] = {w[2], (Real)0};
        }
        __syncthreads();
        if (q < NW) {
            wave_fft<Real>(b, plan, false, L.tw, lane);
            unpack_add<Real>(b, ST + (q == 0 ? 0 : 4),
                             (NC == 3 && q == 0) ? ST + 2 : nullptr, st, nr,
                             nu1, even, lane);
        }
    }
}

template <typename Real, int NC>
void launch_abc(hipStream_t st, const AbcArgs<Real> &a) {
    if (a.n <= 0) return;
    const size_t lds = fft_lds_bytes<Real>(std::max(a.nr_max, 1), 3);
    hipLaunchKernelGGL((k_abc<Real, NC>), dim3(a.n), dim3(192), lds, st, a);
}

// C2R input of one wave's buffer (SolverFFTW conventions, as pack_pair):
// the Hermitian extension of the modes X(k) + i Y(k), mean and Nyquist
// real, Nyquist doubled.
template <typename Real, typename GX, typename GY>
__device__ __forceinline__ void c2r_pack(cpx<Real> *b, int nr, int nu1,
                                         bool even, int lane, GX gx, GY gy) {
    const int nyq = even ? nu1 - 1 : -1;
    for (int k = lane; k < nr; k += 64) {
        const bool lo = k < nu1;
        const int kk = lo ? k : nr - k;
        const cpx<Real> x = gx(kk);
        const cpx<Real> y = gy(kk);
        const bool real_only = k == 0 || k == nyq;
        const Real sc = k == nyq ? (Real)2 : (Real)1;
        const Real sim = real_only ? (Real)0 : (lo ? (Real)1 : (Real)-1);
        b[pad(k)] = {sc * x.re - sim * y.im, sim * x.im + sc * y.re};
    }
}

// R2C output of one wave's buffer (as unpack_pair): the modes of the real
// series x (v0) and y (v1) packed as x + i y, 1 / Nr, Nyquist halved.
template <typename Real, typename F>
__device__ __forceinline__ void r2c_unpack(const cpx<Real> *b, int nr,
                                           int nu1, bool even, int lane,
                                           F put) {
    const int nyq = even ? nu1 - 1 : -1;
    const Real h0 = (Real)0.5 / (Real)nr;
    for (int k = lane; k < nu1; k += 64) {
        const cpx<Real> zk = b[pad(k)];
        const cpx<Real> zn = b[pad(k == 0 ? 0 : nr - k)];
        const Real h = k == nyq ? h0 * (Real)0.5 : h0;
        put(k, cpx<Real>{(zk.re + zn.re) * h, (zk.im - zn.im) * h},
            cpx<Real>{(zk.im + zn.im) * h, (zn.re - zk.re) * h});
    }
}

// Deferred fluid points, one block of two waves each (see FluidDeferArgs):
//   COUPLE  for a 3D pair (SolidFluidCoupling3D::coupleSolidToFluid,
//           SolidFluidCoupling3D.cpp:55-73): C2R of the solid displacement
//           (wave 0: s + i phi, wave 1: z), n . u per slice, R2C (wave 1)
//           added to the fluid stiffness; then coupleFluidToSolid (:76-94):
//           C2R of the fluid stiffness, times n / m_fluid per slice, R2C
//           (wave 0: s + i phi, wave 1: z) into csolid3d
//   AXIAL   fluid axis (orders >= 1 zero) and free surface (all zero)
//   MASS    Nyquist imaginary part zero, then Mass3D (C2R, times 1 / m per
//           slice, R2C) or Mass1D
//   MARCH   Newmark (FP64), fluid stiffness zeroed
template <typename Real, bool COUPLE, bool AXIAL, bool MASS, bool MARCH>
__global__ __launch_bounds__(128) void k_fluid_defer(FluidDeferArgs<Real> a) {
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int q = __builtin_amdgcn_readfirstlane(t >> 6);
    const int i = blockIdx.x;
    const int p = a.point[i];
    const int nr = a.point_nr[p];
    const int nu1 = nr / 2 + 1;
    const bool even = (nr & 1) == 0;
    const uint32_t *plan = a.plan + (size_t)nr * CARD_PLAN_W;
    const size_t foff = a.point_off[p];
    FftLds<Real> L;
    L.carve(fft_lds, nr, 2);
    fft_load_tables<Real>(L, a.twiddle, a.tw_off, a.slice_pos, a.pos_off,
                          plan, nr, t, 128);
    const cpx<Real> *tw = L.tw;
    const uint16_t *pos = L.pos;
    // the point's fluid stiffness (nu1 complex) after the buffers
    cpx<Real> *sf = reinterpret_cast<cpx<Real> *>(L.extra);
    {
        const cpx<Real> *fs = reinterpret_cast<const cpx<Real> *>(a.fstiff);
        for (int k = t; k < nu1; k += 128) sf[k] = fs[foff + k];
    }
    __syncthreads();
    const cpx<Real> zero = {(Real)0, (Real)0};
    cpx<Real> *b = L.buf(q);
    const int j = COUPLE ? a.pair3d[i] : -1;
    if (COUPLE && j >= 0) {
        const int sp = a.pair_solid[j];
        const double *u = a.solid_displ + (size_t)a.solid_off[sp] * 6;
        const int cx = q == 0 ? 0 : 2;
        c2r_pack<Real>(
            b, nr, nu1, even, lane,
            [&](int k) {
                return cpx<Real>{(Real)u[k * 6 + 2 * cx],
                                 (Real)u[k * 6 + 2 * cx + 1]};
            },
            [&](int k) {
                return q == 0 ? cpx<Real>{(Real)u[k * 6 + 2],
                                          (Real)u[k * 6 + 3]}
                              : zero;
            });
        wave_sync();
        wave_fft<Real>(b, plan, true, tw, lane);
        __syncthreads();
        const Real *nn = a.nu + a.n_off[j] * 3;
        for (int k = t; k < nr; k += 128) {
            const int pk = pad(pos[k]);
            const cpx<Real> sphi = L.buf(0)[pk];
            const Real f = nn[3 * k] * sphi.re + nn[3 * k + 1] * sphi.im +
                           nn[3 * k + 2] * L.buf(1)[pk].re;
            L.buf(1)[pk] = {f, (Real)0};
        }
        __syncthreads();
        if (q == 1) {
            wave_fft<Real>(b, plan, false, tw, lane);
            r2c_unpack<Real>(b, nr, nu1, even, lane,
                             [&](int k, cpx<Real> v0, cpx<Real>) {
                                 sf[k] = cadd(sf[k], v0);
                             });
        }
        __syncthreads();
        if (q == 0) {
            c2r_pack<Real>(b, nr, nu1, even, lane,
                           [&](int k) { return sf[k]; },
                           [&](int) { return zero; });
            wave_sync();
            wave_fft<Real>(b, plan, true, tw, lane);
        }
        __syncthreads();
        const Real *na = a.nam + a.n_off[j] * 3;
        for (int k = t; k < nr; k += 128) {
            const int pk = pad(pos[k]);
            const Real g = L.buf(0)[pk].re;
            L.buf(0)[pk] = {g * na[3 * k], g * na[3 * k + 1]};
            L.buf(1)[pk] = {g * na[3 * k + 2], (Real)0};
        }
        __syncthreads();
        wave_fft<Real>(b, plan, false, tw, lane);
        double *cs = a.csolid3d + (size_t)a.c3off[j] * 6;
        r2c_unpack<Real>(b, nr, nu1, even, lane,
                         [&](int k, cpx<Real> v0, cpx<Real> v1) {
                             if (q == 0) {
                                 cs[k * 6 + 0] = (double)v0.re;
                                 cs[k * 6 + 1] = (double)v0.im;
                                 cs[k * 6 + 2] = (double)v1.re;
                                 cs[k * 6 + 3] = (double)v1.im;
                             } else {
                                 cs[k * 6 + 4] = (double)v0.re;
                                 cs[k * 6 + 5] = (double)v0.im;
                             }
                         });
        __syncthreads();
    }
    if (AXIAL) {
        const uint8_t bc = a.point_bc[p];
        for (int k = t; k < nu1; k += 128)
            if (((bc & 1) && k >= 1) || (bc & 2)) sf[k] = zero;
        __syncthreads();
    }
    if (MASS) {
        if (even && t == 0) sf[nu1 - 1].im = (Real)0;
        __syncthreads();
        const int64_t mo = a.minv_off[i];
        if (mo >= 0) {
            if (q == 0) {
                c2r_pack<Real>(b, nr, nu1, even, lane,
                               [&](int k) { return sf[k]; },
                               [&](int) { return zero; });
                wave_sync();
                wave_fft<Real>(b, plan, true, tw, lane);
                const Real *m = a.minv + mo;
                for (int k = lane; k < nr; k += 64) {
                    const int pk = pad(pos[k]);
                    b[pk] = {b[pk].re * m[k], (Real)0};
                }
                wave_sync();
                wave_fft<Real>(b, plan, false, tw, lane);
                r2c_unpack<Real>(b, nr, nu1, even, lane,
                                 [&](int k, cpx<Real> v0, cpx<Real>) {
                                     sf[k] = v0;
                                 });
            }
        } else {
            const Real im = (Real)a.inv_mass[p];
            for (int k = t; k < nu1; k += 128)
                sf[k] = {sf[k].re * im, sf[k].im * im};
        }
        __syncthreads();
    }
    cpx<Real> *fs = reinterpret_cast<cpx<Real> *>(a.fstiff) + foff;
    if (!MARCH) {
        for (int k = t; k < nu1; k += 128) fs[k] = sf[k];
        return;
    }
    const double hdt = .5 * a.dt, hdt2 = .5 * a.dt * a.dt;
    bool bad = false, bad0 = false;
    const int ai = a.abc.idx != nullptr ? a.abc.idx[p] : -1;
    for (int k = t; k < nu1; k += 128) {
        const size_t m2 = (foff + k) * 2;
        double an[2] = {(double)sf[k].re, (double)sf[k].im};
        // sponge (Sponge.hpp:74-96), after the mass in the fused form;
        // the per-phase form has applied it in its own phase (the fluid
        // point kernel's SPONGE pass covers the deferred points too)
        if (MASS && ai >= 0)
            sponge_sub<1>(a.abc, ai, k, even ? nu1 - 1 : -1, an);
#pragma unroll
        for (int ri = 0; ri < 2; ri++) {
            const double u0 = a.chi[m2 + ri];
            double vv, uu;
            if (a.scheme == 1) {
                // SymplecticTimeScheme::update
                vv = a.chi_v[m2 + ri] + a.sym_pi * a.dt * an[ri];
                uu = u0 + a.sym_kappa * a.dt * vv;
            } else {
                vv = a.chi_v[m2 + ri] + hdt * (a.chi_a[m2 + ri] + an[ri]);
                uu = u0 + (a.dt * vv + hdt2 * an[ri]);
                a.chi_a[m2 + ri] = an[ri];
            }
            a.chi_v[m2 + ri] = vv;
            a.chi[m2 + ri] = uu;
            bad |= !isfinite(uu);
            bad0 |= !isfinite(u0);
        }
        fs[k] = zero;
    }
    // bit 0: started from a non-finite potential, bit 1: left one
    if (bad0 || bad) atomicOr(a.flag, (bad0 ? 1 : 0) | (bad ? 2 : 0));
}

template <typename Real>
void launch_fluid_defer(hipStream_t st, int phases,
                        const FluidDeferArgs<Real> &a) {
    if (a.n <= 0) return;
    // two series and the point's nu1 fluid stiffness values
    const size_t lds = fft_lds_bytes<Real>(
        a.nr_max, 2, (a.nr_max / 2 + 1) * (int)sizeof(cpx<Real>));
    const bool c = (phases & PHASE_COUPLE) != 0;
    const bool x = (phases & PHASE_AXIAL) != 0;
    const bool m = (phases & PHASE_MASS) != 0;
    const bool t = (phases & PHASE_MARCH) != 0;
#define AXS3D_DEFER(C_, X_, M_, T_)                                           \
    if (c == C_ && x == X_ && m == M_ && t == T_) {                           \
        hipLaunchKernelGGL((k_fluid_defer<Real, C_, X_, M_, T_>), dim3(a.n),  \
                           dim3(128), lds, st, a);                            \
        return;                                                               \
    }
    AXS3D_DEFER(true, false, false, false)
    AXS3D_DEFER(false, true, false, false)
    AXS3D_DEFER(false, false, true, false)
    AXS3D_DEFER(false, false, false, true)
    AXS3D_DEFER(false, true, true, true)
    AXS3D_DEFER(true, true, true, true)
#undef AXS3D_DEFER
}

// Adjoint sensitivity of 3D isotropic elements per slice (one block of
// three waves per GLL point of each element of the batch): the strain
// rows of psi (scratch, from k_cd_grad on psi) and of u (scrb, from
// k_cd_grad on u) to the slices by C2R, then with IF the integral factor
// and Nr slices
//   kern3d[slice][0][k] += scale IF / Nr tr e_psi tr e_u
//   kern3d[slice][1][k] += scale IF / Nr (2 sum_ii e_psi,ii e_u,ii +
//                                          sum_shear g_psi g_u)
// (Parseval: the Fourier-space pairing of strain and stress equals the
// mean over the slices of their product), in the layout of the host's
// per-slice lambda, mu ([slice][2][25], slices at slice_off[e]).
template <typename Real, int MT>
__global__ __launch_bounds__(CD_PT) void k_cd_sens(CardArgs<Real> a,
                                                   const Real *scrb, int w0,
                                                   const double *ifac,
                                                   double *kern3d,
                                                   double scale) {
    constexpr int NP = MT == 0 ? 2 : MT == 1 ? 5 : 21;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int q = __builtin_amdgcn_readfirstlane(t >> 6);
    const int w = w0 + (int)blockIdx.x / NPEM;
    const int pp = (int)blockIdx.x % NPEM;
    const int e = a.list3d[w];
    const int nu1 = a.s.elem_nu1[e];
    const int nr = a.elem_nr[e];
    const uint32_t *plan = a.plan + (size_t)nr * CARD_PLAN_W;
    FftLds<Real> L;
    L.carve(fft_lds, nr, 2 * CD_PW);     // psi's three series, then u's
    fft_load_tables<Real>(L, a.twiddle, a.tw_off, a.slice_pos, a.pos_off,
                          plan, nr, t, CD_PT);
    const cpx<Real> *tw = L.tw;
    const uint16_t *pos = L.pos;
    __syncthreads();
    const bool even = (nr & 1) == 0;
    const size_t rs = (size_t)NPEM * scr_row(nu1);
    const cpx<Real> *ra = reinterpret_cast<const cpx<Real> *>(a.scratch) +
                          a.scr_off[e] + (size_t)pp * scr_row(nu1);
    const cpx<Real> *rb = reinterpret_cast<const cpx<Real> *>(scrb) +
                          a.scr_off[e] + (size_t)pp * scr_row(nu1);
    pack_pair<Real>(L.buf(q), ra + 2 * q * rs, ra + (2 * q + 1) * rs, nr,
                    nu1, even, lane);
    wave_fft<Real>(L.buf(q), plan, true, tw, lane);
    pack_pair<Real>(L.buf(CD_PW + q), rb + 2 * q * rs,
                    rb + (2 * q + 1) * rs, nr, nu1, even, lane);
    wave_fft<Real>(L.buf(CD_PW + q), plan, true, tw, lane);
    __syncthreads();
    const double f = scale * ifac[(size_t)e * 6 * NPEM + 5 * NPEM + pp] /
                     (double)nr;
    // [slice][param][25] from the element's property offset (the strains
    // of TI and anisotropic batches are in RTZ already, as k_cd_grad
    // formed them)
    double *K = kern3d + a.prop_off[e];
    for (int k = t; k < nr; k += CD_PT) {
        const int pk = pad(pos[k]);
        double ep[6], eu[6];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const cpx<Real> zp = L.buf(c)[pk], zu = L.buf(CD_PW + c)[pk];
            ep[2 * c] = (double)zp.re;
            ep[2 * c + 1] = (double)zp.im;
            eu[2 * c] = (double)zu.re;
            eu[2 * c + 1] = (double)zu.im;
        }
        mat_density<MT>(ep, eu, [](double x, double y) { return x * y; },
                        [&](int c, double v) {
                            K[((size_t)k * NP + c) * NPEM + pp] += f * v;
                        });
    }
}

template <typename Real>
void launch_cd_sens(hipStream_t st, const CardArgs<Real> &a, const Real *scrb,
                    const CardBatch &bt, const double *ifac, double *kern3d,
                    double scale) {
    const int n = bt.end - bt.begin;
    if (n <= 0) return;
    const size_t lds = fft_lds_bytes<Real>(bt.nr_max, 2 * CD_PW);
    if (bt.mtype == 0)
        hipLaunchKernelGGL((k_cd_sens<Real, 0>), dim3(n * NPEM), dim3(CD_PT),
                           lds, st, a, scrb, bt.begin, ifac, kern3d, scale);
    else if (bt.mtype == 1)
        hipLaunchKernelGGL((k_cd_sens<Real, 1>), dim3(n * NPEM), dim3(CD_PT),
                           lds, st, a, scrb, bt.begin, ifac, kern3d, scale);
    else
        hipLaunchKernelGGL((k_cd_sens<Real, 2>), dim3(n * NPEM), dim3(CD_PT),
                           lds, st, a, scrb, bt.begin, ifac, kern3d, scale);
}

template <typename Real>
int card_blocks_per_cu(bool att, int nr_max) {
    int n = 0;
    const size_t lds = fft_lds_bytes<Real>(nr_max, CD_PW);
    const hipError_t r =
        att ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                  &n, k_cd_stress<Real, 5, 0>, CD_PT, lds)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                  &n, k_cd_stress<Real, 0, 0>, CD_PT, lds);
    if (r != hipSuccess) return 1;
    return n > 0 ? n : 1;
}

template int card_blocks_per_cu<float>(bool, int);
template int card_blocks_per_cu<double>(bool, int);
#define AXS3D_CD_INST(Real)                                                   \
    template void launch_cd_grad<Real>(hipStream_t, const CardArgs<Real> &, \
                                       const CardBatch &);                  \
    template void launch_cd_stress<Real>(hipStream_t,                       \
                                         const CardArgs<Real> &,            \
                                         const CardBatch &);                \
    template void launch_cd_quad<Real>(hipStream_t, const CardArgs<Real> &, \
                                       const CardBatch &);
AXS3D_CD_INST(float)
AXS3D_CD_INST(double)
template void launch_cdf_batch<float>(hipStream_t,
                                      const FluidCardArgs<float> &,
                                      const CardBatch &);
template void launch_cdf_batch<double>(hipStream_t,
                                       const FluidCardArgs<double> &,
                                       const CardBatch &);
template void launch_cd_sens<float>(hipStream_t, const CardArgs<float> &,
                                    const float *, const CardBatch &,
                                    const double *, double *, double);
template void launch_cd_sens<double>(hipStream_t, const CardArgs<double> &,
                                     const double *, const CardBatch &,
                                     const double *, double *, double);
template void launch_fluid_defer<float>(hipStream_t, int,
                                        const FluidDeferArgs<float> &);
template void launch_fluid_defer<double>(hipStream_t, int,
                                         const FluidDeferArgs<double> &);
template void launch_mass3d<float>(hipStream_t, bool,
                                   const Mass3DArgs<float> &, float *);
template void launch_mass3d<double>(hipStream_t, bool,
                                    const Mass3DArgs<double> &, double *);
template void launch_abc<float, 3>(hipStream_t, const AbcArgs<float> &);
template void launch_abc<float, 1>(hipStream_t, const AbcArgs<float> &);
template void launch_abc<double, 3>(hipStream_t, const AbcArgs<double> &);
template void launch_abc<double, 1>(hipStream_t, const AbcArgs<double> &);
#undef AXS3D_CD_INST

}  // namespace axs3d
