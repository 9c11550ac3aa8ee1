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
// Design: a persistent grid of blocks of five wave64s.  A block takes one
// element at a time from a global work counter (elements sorted by Nr,
// heaviest first) and keeps its strain, then its stress, in the Fourier
// domain in its own scratch slot, series-major [6*25][nu1max] complex.
// Inside an element every phase is split between the waves, which
// synchronise only among their own lanes; the block meets at two barriers
// per element (gradient done, stress done) and one for the work counter.
//  A. gradient: each wave takes two Fourier orders at a time (lanes 0-24
//     order a, lanes 25-49 order a+1, one lane per GLL point), exchanging
//     the displacement through a per-wave LDS tile.
//  B. five GLL points per wave.  The six strain series of a point are
//     packed two per complex array (x + i y) and inverse-transformed in
//     place in the wave's LDS buffer by a mixed-radix Stockham FFT (radices
//     4, 2, 3, 5, 7, 11, 13; a stage reads all its inputs into registers,
//     then writes its outputs; plans and twiddles are precomputed per Nr).
//     The first two arrays wait in registers while the third is
//     transformed; then stress slice by slice (3D lambda, mu; SLS memory
//     variables), forward transforms and unpack into the scratch.
//  C. quadrature, split like A, element-local forces to the buffer the
//     point kernel gathers.
#include "axs3d_common.hpp"
#include "axs3d_element.hpp"
#include "axs3d_launch.hpp"

namespace axs3d {

constexpr int CD_WAVES = 5;                        // waves per block
constexpr int CD_THREADS = CD_WAVES * 64;
constexpr int CD_MP = (CARD_NR_CAP + 63) / 64;     // slices per lane, max

template <typename T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

template <typename Real>
__device__ __forceinline__ cpx<Real> twid(const cpx<Real> *tw, int t,
                                          bool inv) {
    cpx<Real> w = tw[t];
    if (inv) w.im = -w.im;
    return w;
}

// One in-place Stockham stage of radix R on the series buf[0..n) held by
// one wave: out[(b/Ns) Ns R + b%Ns + q Ns] = sum_r W_R^{rq} W_n^{r (b%Ns)
// step} in[b + r n/R], b < n/R, step = n/(Ns R) (Govindaraju et al., SC
// 2008).  Forward: W = exp(-2 pi i / .); inverse (inv): conjugate.  tw[t] =
// exp(-2 pi i t / n) in LDS.  All lanes read all their inputs before any
// lane writes, so input and output share the buffer.  b % Ns comes from
// the magic multiplier mg = ceil(2^32 / Ns), exact for b Ns < 2^32.
template <typename Real, int R>
__device__ __forceinline__ void wstage(cpx<Real> *buf, int n, int Ns,
                                       uint32_t mg, int step, bool inv,
                                       const cpx<Real> *tw, int lane) {
    constexpr int M = (CARD_NR_CAP / R + 63) / 64;
    const int nb = n / R;
    cpx<Real> v[M][R];
#pragma unroll
    for (int m = 0; m < M; m++) {
        const int b = lane + 64 * m;
        if (b < nb) {
#pragma unroll
            for (int r = 0; r < R; r++) v[m][r] = buf[b + r * nb];
        }
    }
    wave_sync();
    const Real sg = inv ? (Real)1 : (Real)-1;
#pragma unroll
    for (int m = 0; m < M; m++) {
        const int b = lane + 64 * m;
        if (b < nb) {
            const int k =
                Ns == 1 ? 0 : b - (int)__umulhi((uint32_t)b, mg) * Ns;
            if (k > 0) {
#pragma unroll
                for (int r = 1; r < R; r++)
                    v[m][r] = cmulc(v[m][r], twid(tw, r * k * step, inv));
            }
            cpx<Real> *o = buf + (b - k) * R + k;
            const cpx<Real> *x = v[m];
            if constexpr (R == 2) {
                o[0] = cadd(x[0], x[1]);
                o[Ns] = csub(x[0], x[1]);
            } else if constexpr (R == 4) {
                const cpx<Real> s02 = cadd(x[0], x[2]), d02 = csub(x[0], x[2]);
                const cpx<Real> s13 = cadd(x[1], x[3]), d13 = csub(x[1], x[3]);
                // (-i) d13 forward, (+i) d13 inverse
                const cpx<Real> rd13 = {-sg * d13.im, sg * d13.re};
                o[0] = cadd(s02, s13);
                o[2 * Ns] = csub(s02, s13);
                o[Ns] = cadd(d02, rd13);
                o[3 * Ns] = csub(d02, rd13);
            } else if constexpr (R == 3) {
                const Real c1 = (Real)-0.5;
                const Real s1 = sg * (Real)0.86602540378443864676;
                const cpx<Real> tt = cadd(x[1], x[2]);
                const cpx<Real> d = csub(x[1], x[2]);
                const cpx<Real> mm = cfma(tt, c1, x[0]);
                const cpx<Real> id = {-s1 * d.im, s1 * d.re};
                o[0] = cadd(x[0], tt);
                o[Ns] = cadd(mm, id);
                o[2 * Ns] = csub(mm, id);
            } else if constexpr (R == 5) {
                const Real c1 = (Real)0.30901699437494742410;   // cos 2pi/5
                const Real c2 = (Real)-0.80901699437494742410;  // cos 4pi/5
                const Real s1 = sg * (Real)0.95105651629515357212;
                const Real s2 = sg * (Real)0.58778525229247312917;
                const cpx<Real> t1 = cadd(x[1], x[4]), d1 = csub(x[1], x[4]);
                const cpx<Real> t2 = cadd(x[2], x[3]), d2 = csub(x[2], x[3]);
                const cpx<Real> m1 =
                    cadd(x[0], cadd(cmul(t1, c1), cmul(t2, c2)));
                const cpx<Real> m2 =
                    cadd(x[0], cadd(cmul(t1, c2), cmul(t2, c1)));
                // i (s1 d1 + s2 d2) and i (s2 d1 - s1 d2)
                const cpx<Real> e1 = cadd(cmul(d1, s1), cmul(d2, s2));
                const cpx<Real> e2 = csub(cmul(d1, s2), cmul(d2, s1));
                const cpx<Real> ie1 = {-e1.im, e1.re}, ie2 = {-e2.im, e2.re};
                o[0] = cadd(x[0], cadd(t1, t2));
                o[Ns] = cadd(m1, ie1);
                o[4 * Ns] = csub(m1, ie1);
                o[2 * Ns] = cadd(m2, ie2);
                o[3 * Ns] = csub(m2, ie2);
            } else {
                // radices 7, 11, 13: direct DFT, one output at a time, the
                // kernel W_R^{rq} = tw[(rq mod R) n/R] read from LDS
#pragma unroll 1
                for (int q = 0; q < R; q++) {
                    cpx<Real> acc = x[0];
                    int rq = 0;
#pragma unroll
                    for (int r = 1; r < R; r++) {
                        rq += q;
                        if (rq >= R) rq -= R;
                        acc = cadd(acc, cmulc(x[r], twid(tw, rq * nb, inv)));
                    }
                    o[q * Ns] = acc;
                }
            }
        }
    }
    wave_sync();
}

// Whole FFT of buf[0..n) by one wave, stages from the plan.
template <typename Real>
__device__ __forceinline__ void wfft(cpx<Real> *buf, int n,
                                     const uint32_t *__restrict__ plan,
                                     bool inv, const cpx<Real> *tw, int lane) {
    const int nst = (int)plan[0];
    for (int s = 0; s < nst; s++) {
        const uint32_t w0 = plan[1 + 3 * s];
        const uint32_t mg = plan[2 + 3 * s];
        const int step = (int)plan[3 + 3 * s];
        const int R = (int)(w0 & 255u), Ns = (int)(w0 >> 8);
        switch (R) {
            case 4: wstage<Real, 4>(buf, n, Ns, mg, step, inv, tw, lane); break;
            case 2: wstage<Real, 2>(buf, n, Ns, mg, step, inv, tw, lane); break;
            case 3: wstage<Real, 3>(buf, n, Ns, mg, step, inv, tw, lane); break;
            case 5: wstage<Real, 5>(buf, n, Ns, mg, step, inv, tw, lane); break;
            case 7: wstage<Real, 7>(buf, n, Ns, mg, step, inv, tw, lane); break;
            case 11: wstage<Real, 11>(buf, n, Ns, mg, step, inv, tw, lane); break;
            default: wstage<Real, 13>(buf, n, Ns, mg, step, inv, tw, lane); break;
        }
    }
}

// C2R input of the strain pair (X, Y) = rows (2jp, 2jp+1) of point p:
// buf[k] = Xfull[k] + i Yfull[k], with Xfull the Hermitian extension
// under SolverFFTW's C2R rules (Nyquist doubled, Im of mean and Nyquist
// ignored).
template <typename Real>
__device__ __forceinline__ void pack_pair(cpx<Real> *buf,
                                          const cpx<Real> *__restrict__ X,
                                          const cpx<Real> *__restrict__ Y,
                                          int nr, int nu1, bool even,
                                          int lane) {
#pragma unroll 2
    for (int k = lane; k < nr; k += 64) {
        {
            cpx<Real> xf, yf;
            if (k == 0) {
                xf = {X[0].re, 0};
                yf = {Y[0].re, 0};
            } else if (k < nu1) {
                xf = X[k];
                yf = Y[k];
                if (even && k == nu1 - 1) {
                    xf = {(Real)2 * xf.re, 0};
                    yf = {(Real)2 * yf.re, 0};
                }
            } else {
                const cpx<Real> xs = X[nr - k], ys = Y[nr - k];
                xf = {xs.re, -xs.im};
                yf = {ys.re, -ys.im};
            }
            buf[k] = {xf.re - yf.im, xf.im + yf.re};
        }
    }
    wave_sync();
}

// R2C output: X_k = (Z_k + conj Z_{N-k}) / 2N, Y_k = (Z_k - conj Z_{N-k})
// / 2iN, Nyquist row halved; written to rows (2jp, 2jp+1).
template <typename Real>
__device__ __forceinline__ void unpack_pair(const cpx<Real> *buf,
                                            cpx<Real> *__restrict__ X,
                                            cpx<Real> *__restrict__ Y, int nr,
                                            int nu1, bool even, int lane) {
    const Real h0 = (Real)0.5 / (Real)nr;
#pragma unroll 2
    for (int k = lane; k < nu1; k += 64) {
        {
            const cpx<Real> zk = buf[k];
            const cpx<Real> zn = buf[k == 0 ? 0 : nr - k];
            const Real h = (even && k == nu1 - 1) ? h0 * (Real)0.5 : h0;
            X[k] = {(zk.re + zn.re) * h, (zk.im - zn.im) * h};
            Y[k] = {(zk.im + zn.im) * h, (zn.re - zk.re) * h};
        }
    }
    wave_sync();
}

// CG4 attenuation points (1,1),(1,3),(3,1),(3,3) (AttBuilder.cpp:141-147).
__device__ __forceinline__ int att_slot(int p, int npa) {
    if (npa == NPEM) return p;
    return p == 6 ? 0 : p == 8 ? 1 : p == 16 ? 2 : p == 18 ? 3 : -1;
}

// Per-slice material of one GLL point: pointers at slice 0 of element e;
// parameter / component c of slice k sits at [c * Nr + k] (material) or
// [c * npa * Nr + k] (attenuation).
template <typename Real>
struct SliceMat {
    const Real *par;             // this point's parameter 0; parameter c at
    size_t ps;                   // par + c * ps (ps = 25 Nr)
    const Real *dl, *dm;         // memory moduli of its attenuation slot
    Real *mv;                    // memory variables [nsls][6] of the slot
    Real *ds;                    // memory stress [6] of the slot
    size_t cs;                   // component step of mv, ds, dl/dm: npa Nr
    __device__ Real p(int c, int k) const { return par[c * ps + k]; }
};

// SLS memory of stress component d at slice k, when the point carries it
// (Attenuation::updateMem*, AttenuationFull/CG4::apply): sigma -= sum_l
// m_l; m_l = alpha_l m_l + beta_l ds_old + gamma_l ds_new; ds_old =
// ds_new, with ds = isotropic(dlambda, dmu)(eps) on the strain in the
// frame the stress is formed in.  Components are independent apart from
// the trace tr = eps_0 + eps_1 + eps_2.
template <typename Real, bool ATT>
__device__ __forceinline__ void att_comp(const CardArgs<Real> &a,
                                         const SliceMat<Real> &M, bool att,
                                         int d, int k, Real eps, Real tr,
                                         Real &sig) {
    if (!ATT || !att) return;
    const Real dl = M.dl[k], dm = M.dm[k];
    const Real dsn = d < 3 ? tr * dl + eps * (dm * (Real)2) : eps * dm;
    Real *ds = M.ds + (size_t)d * M.cs + k;
    const Real dso = *ds;
    const int nsls = a.nsls;
    Real *mv = M.mv + (size_t)d * M.cs + k;
    const size_t ls = 6 * M.cs;
    for (int l = 0; l < nsls; l++) sig -= mv[l * ls];
    for (int l = 0; l < nsls; l++) {
        const Real al = (Real)a.abg[l];
        const Real be = (Real)a.abg[nsls + l];
        const Real ga = (Real)a.abg[2 * nsls + l];
        Real &m = mv[l * ls];
        m = al * m + be * dso + ga * dsn;
    }
    *ds = dsn;
}

// Per-wave LDS tiles of phases A and C (two element-modes), aliased with
// the wave's FFT buffers.  Consecutive passes alternate between two
// copies (first index): the addresses then change from pass to pass and
// the compiler recomputes them instead of holding dozens of them in
// registers across the loop.
template <typename Real>
struct WaveTiles {
    cpx<Real> U[2][2][3][NPEM];
    cpx<Real> X[2][2][3][NPEM];
    cpx<Real> Y[2][2][3][NPEM];
};

// Two complex series of Nr per wave (FFT work buffers).
template <typename Real>
struct CdLds {
    cpx<Real> tw[CARD_NR_CAP];
    cpx<Real> wbuf[CD_WAVES][2][CARD_NR_CAP];
    Real G[2][NPEM];
    int work;
};
static_assert(sizeof(WaveTiles<float>) <= 2 * sizeof(cpx<float>) * CARD_NR_CAP,
              "phase A/C tiles must fit the wave buffers");
static_assert(sizeof(WaveTiles<double>) <=
                  2 * sizeof(cpx<double>) * CARD_NR_CAP,
              "phase A/C tiles must fit the wave buffers");

// Lane map of phases A and C: two Fourier orders x 25 GLL points per wave
// (lanes 0-24 order a0, 25-49 order a0+1), lanes 50-63 idle.
struct AcLane {
    bool ok;
    int hh, ipnt, i, j;
    __device__ explicit AcLane(int lane) {
        ok = lane < 2 * NPEM;
        hh = ok ? lane / NPEM : 0;
        ipnt = ok ? lane - hh * NPEM : 0;
        i = ipnt / NPED;
        j = ipnt - i * NPED;
    }
};


// sin, cos of theta and 2 theta of GLL point ipnt of element e for the
// RTZ rotation of TI and anisotropic media on spherical meshes
// (SolidElement.cpp:177-193), or false.
template <typename Real, int MT>
__device__ __forceinline__ bool rtz_of(const Real *rot, int e, int ipnt,
                                       Real (&r)[4]) {
    if (MT == 0 || rot == nullptr) return false;
    const Real *q = rot + (size_t)e * 4 * NPEM + ipnt;
#pragma unroll
    for (int c = 0; c < 4; c++) r[c] = q[c * NPEM];
    return true;
}

// A: gradient of element e, Fourier orders a0 = 2 wv, 2 wv + 2 CD_WAVES, ...
// into the scratch rows [d*25 + ipnt][alpha], rotated to RTZ where the
// medium works there.
template <typename Real, int MT>
__device__ __forceinline__ void cd_phase_grad(const StiffArgs<Real> *s,
                                           cpx<Real> *S, int nu1max,
                                           WaveTiles<Real> *T,
                                           const Real (*Gtab)[NPEM], int e,
                                           int nu1, int lane, int wv) {
    const AcLane ln(lane);
    const bool axial = s->elem_axial[e] != 0;
    const bool axial_row0 = axial && ln.i == 0;
    int nu1p = 0;
    uint32_t off = 0;
    PointOps<Real> g{}, q{};
    if (ln.ok) {
        const int p = s->elem_points[e * NPEM + ln.ipnt];
        nu1p = s->point_nu1[p];
        off = s->point_off[p];
        load_point_ops(s->elem_ops, e, ln.ipnt, g, q);
    }
    PointG<Real> G;
    load_point_g<Real>(Gtab, axial, ln.i, ln.j, G);
    Real rr[4];
    const bool rtz = rtz_of<Real, MT>(s->rot, e, ln.ipnt, rr);
    int par = 0;
#pragma unroll 1
    for (int a0 = 2 * wv; a0 < nu1; a0 += 2 * CD_WAVES, par ^= 1) {
        const int alpha = a0 + ln.hh;
        const bool active = ln.ok && alpha < nu1;
        cpx<Real> u[3] = {{0, 0}, {0, 0}, {0, 0}};
        if (active) {
            if (alpha < nu1p) {
                const double *src = s->displ + (size_t)(off + alpha) * 6;
#pragma unroll
                for (int c = 0; c < 3; c++)
                    u[c] = {(Real)src[2 * c], (Real)src[2 * c + 1]};
            }
#pragma unroll
            for (int c = 0; c < 3; c++) T->U[par][ln.hh][c][ln.ipnt] = u[c];
        }
        wave_sync();
        if (active) {
            cpx<Real> eps[6];
            grad6_point<Real>(T->U[par][ln.hh], u, (Real)alpha, axial_row0, ln.i,
                              ln.j, G, g, eps);
            if (rtz)
                rot_spz_rtz6<cpx<Real>, Real>(eps, rr[0], rr[1], rr[2],
                                              rr[3]);
#pragma unroll
            for (int d = 0; d < 6; d++)
                S[(size_t)(d * NPEM + ln.ipnt) * nu1max + alpha] = eps[d];
        }
        wave_sync();
    }
}

// B: the six series of GLL point pp in the wave's two LDS buffers.
// Isotropic and TI media (MT 0, 1): pair (2,3) into buffer 0, where the
// shear stress 3 is formed and eps_2 waits; pair (0,1) into buffer 1,
// which completes the normal stresses; both go back; then the shear pair
// (4,5).  Anisotropic media (MT 2) couple all six: pair (0,1) waits in
// the point's own scratch rows 0-1 (free once packed) while pairs (2,3)
// and (4,5) are in the buffers.  Isotropic::strainToStress
// (Isotropic.hpp:155-172), TransverselyIsotropic::strainToStress
// (TransverselyIsotropic.hpp:161-180), Anisotropic::strainToStress
// (Anisotropic.hpp:203-274).
template <typename Real, bool ATT, int MT>
__device__ __forceinline__ void cd_phase_point(const CardArgs<Real> *a,
                                               cpx<Real> *S, cpx<Real> *b0,
                                               cpx<Real> *b1,
                                               const cpx<Real> *tw, int e,
                                               int nr, int nu1, int pp,
                                               int lane) {
    const int nu1max = a->nu1max;
    const uint32_t *plan = a->plan + (size_t)nr * CARD_PLAN_W;
    const int64_t sl0 = a->slice_off[e];
    const bool even = (nr & 1) == 0;
    cpx<Real> *row[6];
#pragma unroll
    for (int d = 0; d < 6; d++) row[d] = S + (size_t)(d * NPEM + pp) * nu1max;
    SliceMat<Real> M;
    M.ps = (size_t)NPEM * nr;
    M.par = a->prop3d + a->prop_off[e] + (size_t)pp * nr;
    const int ip4 = ATT ? att_slot(pp, a->npa) : -1;
    const bool att = ATT && ip4 >= 0;
    if (ATT) {
        const int npa = a->npa;
        M.cs = (size_t)npa * nr;
        M.dl = a->att3d + (size_t)sl0 * 2 * npa + (size_t)ip4 * nr;
        M.dm = M.dl + M.cs;
        M.mv = a->memvar + (size_t)sl0 * a->nsls * 6 * npa + (size_t)ip4 * nr;
        M.ds = a->dstress + (size_t)sl0 * 6 * npa + (size_t)ip4 * nr;
    }
    if constexpr (MT != 2) {
        // (eps_2, eps_3) -> (eps_2, sigma_3) in b0
        pack_pair<Real>(b0, row[2], row[3], nr, nu1, even, lane);
        wfft<Real>(b0, nr, plan, true, tw, lane);
#pragma unroll 2
        for (int k = lane; k < nr; k += 64) {
            const cpx<Real> z = b0[k];
            // sigma_3 = mu eps_3 (isotropic) or L eps_3 (TI)
            Real s3 = z.im * M.p(MT == 0 ? 1 : 3, k);
            att_comp<Real, ATT>(*a, M, att, 3, k, z.im, 0, s3);
            b0[k] = {z.re, s3};
        }
        // (eps_0, eps_1) in b1: sigma_0, sigma_1 to b1, sigma_2 to b0
        pack_pair<Real>(b1, row[0], row[1], nr, nu1, even, lane);
        wfft<Real>(b1, nr, plan, true, tw, lane);
#pragma unroll 2
        for (int k = lane; k < nr; k += 64) {
            const cpx<Real> z = b1[k];
            const cpx<Real> y = b0[k];
            const Real e0 = z.re, e1 = z.im, e2 = y.re;
            Real s0, s1, s2;
            if constexpr (MT == 0) {
                const Real lam = M.p(0, k), mu = M.p(1, k);
                const Real mu2 = mu * (Real)2;
                const Real tl = (e0 + e1 + e2) * lam;
                s0 = tl + e0 * mu2;
                s1 = tl + e1 * mu2;
                s2 = tl + e2 * mu2;
            } else {
                const Real A = M.p(0, k), C = M.p(1, k), F = M.p(2, k);
                const Real mN2 = M.p(4, k) * (Real)-2;
                const Real t3 = e0 + e1;
                const Real t4 = t3 * A + e2 * F;
                s0 = t4 + e1 * mN2;
                s1 = t4 + e0 * mN2;
                s2 = t3 * F + e2 * C;
            }
            if (ATT && att) {
                const Real tr = e0 + e1 + e2;
                att_comp<Real, ATT>(*a, M, att, 0, k, e0, tr, s0);
                att_comp<Real, ATT>(*a, M, att, 1, k, e1, tr, s1);
                att_comp<Real, ATT>(*a, M, att, 2, k, e2, tr, s2);
            }
            b1[k] = {s0, s1};
            b0[k] = {s2, y.im};
        }
        wave_sync();
        wfft<Real>(b1, nr, plan, false, tw, lane);
        unpack_pair<Real>(b1, row[0], row[1], nr, nu1, even, lane);
        wfft<Real>(b0, nr, plan, false, tw, lane);
        unpack_pair<Real>(b0, row[2], row[3], nr, nu1, even, lane);
        // shears (eps_4, eps_5) in b0
        pack_pair<Real>(b0, row[4], row[5], nr, nu1, even, lane);
        wfft<Real>(b0, nr, plan, true, tw, lane);
#pragma unroll 2
        for (int k = lane; k < nr; k += 64) {
            const cpx<Real> z = b0[k];
            Real s4, s5;
            if constexpr (MT == 0) {
                const Real mu = M.p(1, k);
                s4 = z.re * mu;
                s5 = z.im * mu;
            } else {
                s4 = z.re * M.p(3, k);
                s5 = z.im * M.p(4, k);
            }
            att_comp<Real, ATT>(*a, M, att, 4, k, z.re, 0, s4);
            att_comp<Real, ATT>(*a, M, att, 5, k, z.im, 0, s5);
            b0[k] = {s4, s5};
        }
        wave_sync();
        wfft<Real>(b0, nr, plan, false, tw, lane);
        unpack_pair<Real>(b0, row[4], row[5], nr, nu1, even, lane);
    } else {
        // (eps_0, eps_1) waits in rows 0-1: slice k at row0[k] (k < nu1)
        // or row1[k - nu1] (nr - nu1 < nu1 <= nu1max)
        cpx<Real> *r0 = row[0], *r1 = row[1] - nu1;
        pack_pair<Real>(b0, row[0], row[1], nr, nu1, even, lane);
        wfft<Real>(b0, nr, plan, true, tw, lane);
        for (int k = lane; k < nr; k += 64) (k < nu1 ? r0 : r1)[k] = b0[k];
        wave_sync();
        pack_pair<Real>(b1, row[2], row[3], nr, nu1, even, lane);
        wfft<Real>(b1, nr, plan, true, tw, lane);
        pack_pair<Real>(b0, row[4], row[5], nr, nu1, even, lane);
        wfft<Real>(b0, nr, plan, true, tw, lane);
#pragma unroll 1
        for (int k = lane; k < nr; k += 64) {
            cpx<Real> *pk = (k < nu1 ? r0 : r1) + k;
            const cpx<Real> z01 = *pk, z23 = b1[k], z45 = b0[k];
            const Real ev[6] = {z01.re, z01.im, z23.re, z23.im, z45.re,
                                z45.im};
            Real sv[6];
            aniso_hooke<Real, Real>(ev, [&](int c) { return M.p(c, k); },
                                    sv);
            if (ATT && att) {
                const Real tr = ev[0] + ev[1] + ev[2];
#pragma unroll
                for (int d = 0; d < 6; d++)
                    att_comp<Real, ATT>(*a, M, att, d, k, ev[d], tr, sv[d]);
            }
            *pk = {sv[0], sv[1]};
            b1[k] = {sv[2], sv[3]};
            b0[k] = {sv[4], sv[5]};
        }
        wave_sync();
        wfft<Real>(b0, nr, plan, false, tw, lane);
        unpack_pair<Real>(b0, row[4], row[5], nr, nu1, even, lane);
        wfft<Real>(b1, nr, plan, false, tw, lane);
        unpack_pair<Real>(b1, row[2], row[3], nr, nu1, even, lane);
        for (int k = lane; k < nr; k += 64) b0[k] = (k < nu1 ? r0 : r1)[k];
        wave_sync();
        wfft<Real>(b0, nr, plan, false, tw, lane);
        unpack_pair<Real>(b0, row[0], row[1], nr, nu1, even, lane);
    }
}

// C: quadrature of element e from the scratch stress rows (rotated back
// from RTZ where needed) into the element-local force buffer.
template <typename Real, int MT>
__device__ __forceinline__ void cd_phase_quad(const StiffArgs<Real> *s,
                                           const cpx<Real> *S, int nu1max,
                                           WaveTiles<Real> *T,
                                           const Real (*Gtab)[NPEM], int e,
                                           int nu1, int lane, int wv) {
    const AcLane ln(lane);
    const bool axial = s->elem_axial[e] != 0;
    const bool axial_row0 = axial && ln.i == 0;
    int nu1p = 0;
    PointOps<Real> g{}, q{};
    if (ln.ok) {
        const int p = s->elem_points[e * NPEM + ln.ipnt];
        nu1p = s->point_nu1[p];
        load_point_ops(s->elem_ops, e, ln.ipnt, g, q);
    }
    PointG<Real> G;
    load_point_g<Real>(Gtab, axial, ln.i, ln.j, G);
    Real *eb = s->eb + (size_t)s->elem_eb_off[e] * 6;
    Real rr[4];
    const bool rtz = rtz_of<Real, MT>(s->rot, e, ln.ipnt, rr);
    int par = 0;
#pragma unroll 1
    for (int a0 = 2 * wv; a0 < nu1; a0 += 2 * CD_WAVES, par ^= 1) {
        const int alpha = a0 + ln.hh;
        const bool active = ln.ok && alpha < nu1;
        cpx<Real> Lt[3];
        if (active) {
            cpx<Real> sig[6];
#pragma unroll
            for (int d = 0; d < 6; d++)
                sig[d] = S[(size_t)(d * NPEM + ln.ipnt) * nu1max + alpha];
            if (rtz)
                rot_rtz_spz6<cpx<Real>, Real>(sig, rr[0], rr[1], rr[2],
                                              rr[3]);
            quad6_prep<Real>(sig, (Real)alpha, axial_row0, q,
                             T->X[par][ln.hh], T->Y[par][ln.hh], ln.ipnt, Lt);
        }
        wave_sync();
        if (active) {
            cpx<Real> F[3];
            quad6_finish<Real>(T->X[par][ln.hh], T->Y[par][ln.hh], ln.i, ln.j,
                               G, Lt, F);
            if (alpha < nu1p) {
                Real *dst = eb + ((size_t)ln.ipnt * nu1 + alpha) * 6;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    dst[2 * c] = F[c].re;
                    dst[2 * c + 1] = F[c].im;
                }
            }
        }
        wave_sync();
    }
}

// OCC = minimum waves per SIMD the register allocation must allow; picked
// per workload by measurement.  MT = material type of the elements of
// a.list3d (0 isotropic, 1 TI, 2 anisotropic).
template <typename Real, bool ATT, int OCC, int MT>
__global__ __launch_bounds__(CD_THREADS, OCC) void k_solid_stiff_cd(
    CardArgs<Real> a) {
    __shared__ CdLds<Real> L;
    const int t = threadIdx.x;
    const int lane = t & 63;
    // the wave index is uniform across the wave: keep it (and all that is
    // derived from it) in scalar registers
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
    if (t < 2 * NPEM) L.G[t / NPEM][t % NPEM] = a.s.gtab[t];
    const int nu1max = a.nu1max;
    cpx<Real> *S = reinterpret_cast<cpx<Real> *>(a.scratch) +
                   (size_t)blockIdx.x * 6 * NPEM * nu1max;
    cpx<Real> *b0 = L.wbuf[wv][0], *b1 = L.wbuf[wv][1];
    WaveTiles<Real> *T = reinterpret_cast<WaveTiles<Real> *>(b0);
    for (;;) {
        if (t == 0) L.work = atomicAdd(a.work_counter, 1);
        __syncthreads();
        const int w = L.work;
        if (w >= a.n3d) break;
        const int e = a.list3d[w];
        const int nu1 = a.s.elem_nu1[e];
        const int nr = a.elem_nr[e];
        {
            const cpx<Real> *twg = a.twiddle + a.tw_off[nr];
            for (int x = t; x < nr; x += CD_THREADS) L.tw[x] = twg[x];
        }
        cd_phase_grad<Real, MT>(&a.s, S, nu1max, T, L.G, e, nu1, lane, wv);
        __syncthreads();
        for (int pp = wv; pp < NPEM; pp += CD_WAVES)
            cd_phase_point<Real, ATT, MT>(&a, S, b0, b1, L.tw, e, nr, nu1,
                                          pp, lane);
        __syncthreads();
        cd_phase_quad<Real, MT>(&a.s, S, nu1max, T, L.G, e, nu1, lane, wv);
        // every wave is done with this element's scratch and with the
        // work slot before the next element is fetched
        __syncthreads();
    }
}

__global__ void k_reset_counter(int *c) {
    if (threadIdx.x < 3 && blockIdx.x == 0) c[threadIdx.x] = 0;
}

// Occupancy variants for isotropic media (the common case); TI and
// anisotropic media run the occupancy-2 variant.
template <typename Real, bool ATT, int MT>
static void launch_cd_mt(hipStream_t st, const CardArgs<Real> &a,
                         int n_slots, int occ) {
    if constexpr (MT != 0) {
        hipLaunchKernelGGL((k_solid_stiff_cd<Real, ATT, 2, MT>),
                           dim3(n_slots), dim3(CD_THREADS), 0, st, a);
    } else {
        switch (occ) {
            case 4:
                hipLaunchKernelGGL((k_solid_stiff_cd<Real, ATT, 4, MT>),
                                   dim3(n_slots), dim3(CD_THREADS), 0, st, a);
                break;
            case 3:
                hipLaunchKernelGGL((k_solid_stiff_cd<Real, ATT, 3, MT>),
                                   dim3(n_slots), dim3(CD_THREADS), 0, st, a);
                break;
            default:
                hipLaunchKernelGGL((k_solid_stiff_cd<Real, ATT, 2, MT>),
                                   dim3(n_slots), dim3(CD_THREADS), 0, st, a);
                break;
        }
    }
}

template <typename Real, bool ATT>
static void launch_cd_att(hipStream_t st, const CardArgs<Real> &a, int mt,
                          int n_slots, int occ) {
    if (mt == 0)
        launch_cd_mt<Real, ATT, 0>(st, a, n_slots, occ);
    else if (mt == 1)
        launch_cd_mt<Real, ATT, 1>(st, a, n_slots, occ);
    else
        launch_cd_mt<Real, ATT, 2>(st, a, n_slots, occ);
}

// One persistent launch per material type present; a.list3d holds the 3D
// elements grouped by type (each group heaviest first) with the group
// boundaries in a.type_begin, and a.work_counter one counter per type.
template <typename Real>
void launch_solid_stiff_cd(hipStream_t st, const CardArgs<Real> &a,
                           int n_slots, int occ) {
    if (a.n3d <= 0) return;
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, st,
                       a.work_counter);
    for (int mt = 0; mt < 3; mt++) {
        const int b = a.type_begin[mt], n = a.type_begin[mt + 1] - b;
        if (n <= 0) continue;
        CardArgs<Real> c = a;
        c.list3d = a.list3d + b;
        c.n3d = n;
        c.work_counter = a.work_counter + mt;
        if (a.nsls > 0)
            launch_cd_att<Real, true>(st, c, mt, n_slots, occ);
        else
            launch_cd_att<Real, false>(st, c, mt, n_slots, occ);
    }
}

template <typename Real, bool ATT, int OCC>
static int blocks_per_cu_v() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n, k_solid_stiff_cd<Real, ATT, OCC, 0>, CD_THREADS, 0) !=
        hipSuccess)
        return 1;
    return n > 0 ? n : 1;
}

template <typename Real>
int card_blocks_per_cu(bool att, int occ) {
    if (att) {
        if (occ == 4) return blocks_per_cu_v<Real, true, 4>();
        if (occ == 3) return blocks_per_cu_v<Real, true, 3>();
        return blocks_per_cu_v<Real, true, 2>();
    }
    if (occ == 4) return blocks_per_cu_v<Real, false, 4>();
    if (occ == 3) return blocks_per_cu_v<Real, false, 3>();
    return blocks_per_cu_v<Real, false, 2>();
}

template int card_blocks_per_cu<float>(bool, int);
template int card_blocks_per_cu<double>(bool, int);
template void launch_solid_stiff_cd<float>(hipStream_t,
                                           const CardArgs<float> &, int, int);
template void launch_solid_stiff_cd<double>(hipStream_t,
                                            const CardArgs<double> &, int,
                                            int);

}  // namespace axs3d
