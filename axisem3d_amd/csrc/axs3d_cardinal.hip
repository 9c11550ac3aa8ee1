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
// Design: a persistent grid, one block per scratch slot, pulls elements
// (heaviest first) from a global work counter.  For each element:
//  A. gradient: 250 threads = 10 Fourier orders x 25 GLL points per pass;
//     strain written to the slot's scratch, series-major [6*25][nu1].
//  B. per batch of GLL points: the 6 strain series of a point are packed
//     two per complex array (x + i y), inverse-transformed with an LDS
//     mixed-radix Stockham FFT (radices 4, 2, 3, 5, 7, 11, 13), turned into
//     stress slice by slice (3D lambda, mu; SLS memory variables in FP64),
//     forward-transformed and unpacked back into the scratch.
//  C. quadrature from the scratch, element-local forces to the buffer the
//     point kernel gathers.
#include "axs3d_common.hpp"
#include "axs3d_element.hpp"
#include "axs3d_launch.hpp"

namespace axs3d {

template <typename T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

__device__ __forceinline__ int next_radix(int m) {
    if (m % 4 == 0) return 4;
    if (m % 2 == 0) return 2;
    if (m % 3 == 0) return 3;
    if (m % 5 == 0) return 5;
    if (m % 7 == 0) return 7;
    if (m % 11 == 0) return 11;
    if (m % 13 == 0) return 13;
    return m;
}

// One Stockham stage of radix R over `nbatch` contiguous series of length
// n: out[(j/Ns)*Ns*R + j%Ns + q*Ns] = sum_r W_R^{rq} W_n^{r (j%Ns) n/(Ns R)}
// in[j + r n/R] for j < n/R (Govindaraju et al., SC 2008).  sign = -1
// forward, +1 inverse; tw[t] = exp(-2 pi i t / n).  Radices 2, 3, 4, 5
// use closed-form butterflies; 7, 11 and 13 form each output from the
// table (W_R^x = W_n^{x n/R}) and store it at once, keeping register use
// at 2R + O(1) values.
template <typename Real>
__device__ __forceinline__ cpx<Real> twid(const cpx<Real> *__restrict__ tw,
                                          int t, int sign) {
    cpx<Real> w = tw[t];
    if (sign > 0) w.im = -w.im;
    return w;
}

template <typename Real, int R>
__device__ __forceinline__ void fft_stage(const cpx<Real> *__restrict__ in,
                                          cpx<Real> *__restrict__ out, int n,
                                          int nbatch, int Ns, int sign,
                                          const cpx<Real> *__restrict__ tw) {
    const int nb = n / R;
    const int step = n / (Ns * R);
    const int total = nb * nbatch;
    const Real sg = (Real)sign;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int h = idx / nb;
        const int j = idx - h * nb;
        const cpx<Real> *a = in + (size_t)h * n;
        cpx<Real> *b = out + (size_t)h * n;
        const int k = j % Ns;
        const int idxD = (j / Ns) * Ns * R + k;
        if (R >= 7) {
            // y_q = sum_r x_r W_n^{r (k step + q n/R)}: stage twiddle and
            // DFT kernel from one table index, reduced incrementally
            // (k step + q n/R < n, so one subtraction keeps t < n)
            const int wstep = n / R;
#pragma unroll 1
            for (int q = 0; q < R; q++) {
                const int base = k * step + q * wstep;
                cpx<Real> acc = a[j];
                int t = 0;
#pragma unroll 1
                for (int r = 1; r < R; r++) {
                    t += base;
                    if (t >= n) t -= n;
                    acc = cadd(acc, cmulc(a[j + r * nb], twid(tw, t, sign)));
                }
                b[idxD + q * Ns] = acc;
            }
            continue;
        }
        cpx<Real> v[R < 7 ? R : 1];
#pragma unroll
        for (int r = 0; r < (R < 7 ? R : 1); r++) {
            cpx<Real> x = a[j + r * nb];
            if (r > 0 && k > 0) x = cmulc(x, twid(tw, r * k * step, sign));
            v[r] = x;
        }
        if (R == 2) {
            b[idxD] = cadd(v[0], v[1]);
            b[idxD + Ns] = csub(v[0], v[1]);
        } else if (R == 4) {
            const cpx<Real> s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
            const cpx<Real> s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
            // (-i) d13 forward, (+i) d13 inverse
            const cpx<Real> rd13 = {-sg * d13.im, sg * d13.re};
            b[idxD] = cadd(s02, s13);
            b[idxD + 2 * Ns] = csub(s02, s13);
            b[idxD + Ns] = cadd(d02, rd13);
            b[idxD + 3 * Ns] = csub(d02, rd13);
        } else if (R == 3) {
            const Real c1 = (Real)-0.5;
            const Real s1 = sg * (Real)0.86602540378443864676;
            const cpx<Real> t = cadd(v[1], v[2]);
            const cpx<Real> d = csub(v[1], v[2]);
            const cpx<Real> m = cfma(t, c1, v[0]);
            const cpx<Real> id = {-s1 * d.im, s1 * d.re};
            b[idxD] = cadd(v[0], t);
            b[idxD + Ns] = cadd(m, id);
            b[idxD + 2 * Ns] = csub(m, id);
        } else if (R == 5) {
            const Real c1 = (Real)0.30901699437494742410;   // cos 2pi/5
            const Real c2 = (Real)-0.80901699437494742410;  // cos 4pi/5
            const Real s1 = sg * (Real)0.95105651629515357212;
            const Real s2 = sg * (Real)0.58778525229247312917;
            const cpx<Real> t1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]);
            const cpx<Real> t2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
            const cpx<Real> m1 = cadd(v[0], cadd(cmul(t1, c1), cmul(t2, c2)));
            const cpx<Real> m2 = cadd(v[0], cadd(cmul(t1, c2), cmul(t2, c1)));
            // i (s1 d1 + s2 d2) and i (s2 d1 - s1 d2)
            const cpx<Real> e1 = cadd(cmul(d1, s1), cmul(d2, s2));
            const cpx<Real> e2 = csub(cmul(d1, s2), cmul(d2, s1));
            const cpx<Real> ie1 = {-e1.im, e1.re}, ie2 = {-e2.im, e2.re};
            b[idxD] = cadd(v[0], cadd(t1, t2));
            b[idxD + Ns] = cadd(m1, ie1);
            b[idxD + 4 * Ns] = csub(m1, ie1);
            b[idxD + 2 * Ns] = cadd(m2, ie2);
            b[idxD + 3 * Ns] = csub(m2, ie2);
        }
    }
}

// Batched FFT of `nbatch` series of length n held in LDS buffer a (b is the
// ping-pong partner).  Returns the buffer holding the result.  Block-wide:
// every thread must call it.
template <typename Real>
__device__ cpx<Real> *fft_batch(cpx<Real> *a, cpx<Real> *b, int n, int nbatch,
                                int sign, const cpx<Real> *__restrict__ tw) {
    int Ns = 1, rem = n;
    while (rem > 1) {
        const int R = next_radix(rem);
        switch (R) {
            case 4: fft_stage<Real, 4>(a, b, n, nbatch, Ns, sign, tw); break;
            case 2: fft_stage<Real, 2>(a, b, n, nbatch, Ns, sign, tw); break;
            case 3: fft_stage<Real, 3>(a, b, n, nbatch, Ns, sign, tw); break;
            case 5: fft_stage<Real, 5>(a, b, n, nbatch, Ns, sign, tw); break;
            case 7: fft_stage<Real, 7>(a, b, n, nbatch, Ns, sign, tw); break;
            case 11: fft_stage<Real, 11>(a, b, n, nbatch, Ns, sign, tw); break;
            default: fft_stage<Real, 13>(a, b, n, nbatch, Ns, sign, tw); break;
        }
        __syncthreads();
        cpx<Real> *t = a;
        a = b;
        b = t;
        Ns *= R;
        rem /= R;
    }
    return a;
}

// CG4 attenuation points (1,1),(1,3),(3,1),(3,3) (AttBuilder.cpp:141-147).
__device__ __forceinline__ int att_slot(int p, int npa) {
    if (npa == NPEM) return p;
    return p == 6 ? 0 : p == 8 ? 1 : p == 16 ? 2 : p == 18 ? 3 : -1;
}

template <typename Real>
struct CardLds {
    // phases A and C
    cpx<Real> sU[EMB][3][NPEM];
    cpx<Real> sX[EMB][3][NPEM];
    cpx<Real> sY[EMB][3][NPEM];
};

// OCC = minimum waves per SIMD the register allocation must allow (2, 3
// or 4 blocks of 256 threads per CU); picked per workload by measurement.
template <typename Real, bool ATT, int OCC>
__global__ __launch_bounds__(STIFF_BLOCK, OCC) void k_solid_stiff_cd(
    CardArgs<Real> a) {
    constexpr size_t fft_bytes = 2 * (size_t)CARD_FFT_CAP * sizeof(cpx<Real>);
    constexpr size_t ac_bytes = sizeof(CardLds<Real>);
    constexpr size_t lds_bytes = fft_bytes > ac_bytes ? fft_bytes : ac_bytes;
    __shared__ __align__(16) unsigned char smem[lds_bytes];
    __shared__ Real sG[2][NPEM];
    __shared__ int sWork;
    CardLds<Real> &L = *reinterpret_cast<CardLds<Real> *>(smem);
    cpx<Real> *buf0 = reinterpret_cast<cpx<Real> *>(smem);
    cpx<Real> *buf1 = buf0 + CARD_FFT_CAP;

    const StiffArgs<Real> &s = a.s;
    const int t = threadIdx.x;
    if (t < 2 * NPEM) sG[t / NPEM][t % NPEM] = s.gtab[t];
    const int nu1max = a.nu1max;
    cpx<Real> *S = reinterpret_cast<cpx<Real> *>(a.scratch) +
                   (size_t)blockIdx.x * 6 * NPEM * nu1max;
    // phase A/C thread map: 10 orders x 25 points, order fastest
    const bool lane_ok = t < EMB * NPEM;
    const int eml = t % EMB;
    const int ipnt = lane_ok ? t / EMB : 0;
    const int i = ipnt / NPED, j = ipnt - (ipnt / NPED) * NPED;

    for (;;) {
        if (t == 0) sWork = atomicAdd(a.work_counter, 1);
        __syncthreads();
        const int w = sWork;
        __syncthreads();
        if (w >= a.n3d) break;
        const int e = a.list3d[w];
        const int nu1 = s.elem_nu1[e];
        const int nr = a.elem_nr[e];
        const bool axial = s.elem_axial[e] != 0;
        const bool axial_row0 = axial && (i == 0);

        int nu1p = 0;
        uint32_t off = 0;
        if (lane_ok) {
            const int p = s.elem_points[e * NPEM + ipnt];
            nu1p = s.point_nu1[p];
            off = s.point_off[p];
        }

        // ---------------- A. gradient -> strain (Fourier) in scratch
        {
        PointG<Real> G;
        load_point_g<Real>(sG, axial, i, j, G);
        PointOps<Real> g{}, q{};
        if (lane_ok) load_point_ops(s.elem_ops, e, ipnt, g, q);
        for (int a0 = 0; a0 < nu1; a0 += EMB) {
            const int alpha = a0 + eml;
            const bool active = lane_ok && alpha < nu1;
            cpx<Real> u[3] = {{0, 0}, {0, 0}, {0, 0}};
            if (active) {
                if (alpha < nu1p) {
                    const double *src = s.displ + (size_t)(off + alpha) * 6;
#pragma unroll
                    for (int c = 0; c < 3; c++)
                        u[c] = {(Real)src[2 * c], (Real)src[2 * c + 1]};
                }
#pragma unroll
                for (int c = 0; c < 3; c++) L.sU[eml][c][ipnt] = u[c];
            }
            __syncthreads();
            if (active) {
                cpx<Real> eps[6];
                grad6_point<Real>(L.sU[eml], u, (Real)alpha, axial_row0, i, j,
                                  G, g, eps);
#pragma unroll
                for (int d = 0; d < 6; d++)
                    S[(size_t)(d * NPEM + ipnt) * nu1max + alpha] = eps[d];
            }
            __syncthreads();
        }
        }

        // ---------------- B. FFT -> stress (+ attenuation) -> FFT
        const cpx<Real> *tw = a.twiddle + a.tw_off[nr];
        const int np = min(NPEM, CARD_FFT_CAP / (3 * nr));
        const int64_t sl0 = a.slice_off[e];
        const Real inv_nr = (Real)1 / (Real)nr;
        const bool even = (nr & 1) == 0;
        for (int p0 = 0; p0 < NPEM; p0 += np) {
            const int npc = min(np, NPEM - p0);
            // pack: z = Xfull + i Yfull for dim pairs (0,1), (2,3), (4,5)
            for (int idx = t; idx < npc * 3 * nr; idx += blockDim.x) {
                const int qq = idx / (3 * nr);
                const int rem = idx - qq * 3 * nr;
                const int jp = rem / nr;
                const int k = rem - jp * nr;
                const int pp = p0 + qq;
                const cpx<Real> *X = S + (size_t)((2 * jp) * NPEM + pp) * nu1max;
                const cpx<Real> *Y = X + (size_t)NPEM * nu1max;
                cpx<Real> xf, yf;
                if (k == 0) {
                    xf = {X[0].re, 0};
                    yf = {Y[0].re, 0};
                } else if (k < nu1) {
                    if (even && k == nu1 - 1) {
                        xf = {(Real)2 * X[k].re, 0};
                        yf = {(Real)2 * Y[k].re, 0};
                    } else {
                        xf = X[k];
                        yf = Y[k];
                    }
                } else {
                    const cpx<Real> xs = X[nr - k], ys = Y[nr - k];
                    xf = {xs.re, -xs.im};
                    yf = {ys.re, -ys.im};
                }
                buf0[idx] = {xf.re - yf.im, xf.im + yf.re};
            }
            __syncthreads();
            cpx<Real> *res = fft_batch<Real>(buf0, buf1, nr, npc * 3, +1, tw);
            cpx<Real> *oth = res == buf0 ? buf1 : buf0;
            // stress slice by slice
            for (int idx = t; idx < npc * nr; idx += blockDim.x) {
                const int qq = idx / nr;
                const int k = idx - qq * nr;
                const int pp = p0 + qq;
                cpx<Real> *z = res + (size_t)qq * 3 * nr + k;
                Real eps[6], sig[6];
#pragma unroll
                for (int jp = 0; jp < 3; jp++) {
                    eps[2 * jp] = z[jp * nr].re;
                    eps[2 * jp + 1] = z[jp * nr].im;
                }
                const Real *lm = a.prop3d + ((size_t)(sl0 + k) * 2) * NPEM;
                const Real lam = lm[pp], mu = lm[NPEM + pp];
                const Real mu2 = mu * (Real)2;
                const Real tr = (eps[0] + eps[1] + eps[2]) * lam;
                sig[0] = tr + eps[0] * mu2;
                sig[1] = tr + eps[1] * mu2;
                sig[2] = tr + eps[2] * mu2;
                sig[3] = eps[3] * mu;
                sig[4] = eps[4] * mu;
                sig[5] = eps[5] * mu;
                if (ATT) {
                    const int ip4 = att_slot(pp, a.npa);
                    if (ip4 >= 0) {
                        const int npa = a.npa, nsls = a.nsls;
                        const size_t sl = (size_t)(sl0 + k);
                        double *mv = a.memvar + sl * nsls * 6 * npa + ip4;
                        Real *ds = a.dstress + sl * 6 * npa + ip4;
                        const Real *dlm = a.att3d + sl * 2 * npa + ip4;
                        const Real dl = dlm[0], dm = dlm[npa];
                        // stress -= sum of memory variables
                        double msum[6] = {0, 0, 0, 0, 0, 0};
                        for (int l = 0; l < nsls; l++)
#pragma unroll
                            for (int d = 0; d < 6; d++)
                                msum[d] += mv[(l * 6 + d) * npa];
#pragma unroll
                        for (int d = 0; d < 6; d++)
                            sig[d] = (Real)((double)sig[d] - msum[d]);
                        // new memory stress from this strain
                        const Real dm2 = dm * (Real)2;
                        const Real dtr = (eps[0] + eps[1] + eps[2]) * dl;
                        Real dsn[6];
                        dsn[0] = dtr + eps[0] * dm2;
                        dsn[1] = dtr + eps[1] * dm2;
                        dsn[2] = dtr + eps[2] * dm2;
                        dsn[3] = eps[3] * dm;
                        dsn[4] = eps[4] * dm;
                        dsn[5] = eps[5] * dm;
                        for (int l = 0; l < nsls; l++) {
                            const double al = a.abg[l];
                            const double be = a.abg[nsls + l];
                            const double ga = a.abg[2 * nsls + l];
#pragma unroll
                            for (int d = 0; d < 6; d++) {
                                double m = mv[(l * 6 + d) * npa];
                                m = al * m + be * (double)ds[d * npa];
                                m += ga * (double)dsn[d];
                                mv[(l * 6 + d) * npa] = m;
                            }
                        }
#pragma unroll
                        for (int d = 0; d < 6; d++) ds[d * npa] = dsn[d];
                    }
                }
#pragma unroll
                for (int jp = 0; jp < 3; jp++)
                    z[jp * nr] = {sig[2 * jp], sig[2 * jp + 1]};
            }
            __syncthreads();
            cpx<Real> *fres = fft_batch<Real>(res, oth, nr, npc * 3, -1, tw);
            // unpack X_k = (Z_k + conj Z_{N-k}) / 2, Y_k = (Z_k - conj Z_{N-k}) / 2i
            for (int idx = t; idx < npc * 3 * nu1; idx += blockDim.x) {
                const int qq = idx / (3 * nu1);
                const int rem = idx - qq * 3 * nu1;
                const int jp = rem / nu1;
                const int k = rem - jp * nu1;
                const int pp = p0 + qq;
                const cpx<Real> *z = fres + (size_t)(qq * 3 + jp) * nr;
                const cpx<Real> zk = z[k];
                const cpx<Real> zn = z[k == 0 ? 0 : nr - k];
                Real h = (Real)0.5 * inv_nr;
                if (even && k == nu1 - 1) h *= (Real)0.5;
                const cpx<Real> X = {(zk.re + zn.re) * h, (zk.im - zn.im) * h};
                const cpx<Real> Y = {(zk.im + zn.im) * h, (zn.re - zk.re) * h};
                S[(size_t)((2 * jp) * NPEM + pp) * nu1max + k] = X;
                S[(size_t)((2 * jp + 1) * NPEM + pp) * nu1max + k] = Y;
            }
            __syncthreads();
        }

        // ---------------- C. quadrature -> element-local forces
        {
        PointG<Real> G;
        load_point_g<Real>(sG, axial, i, j, G);
        PointOps<Real> g{}, q{};
        if (lane_ok) load_point_ops(s.elem_ops, e, ipnt, g, q);
        for (int a0 = 0; a0 < nu1; a0 += EMB) {
            const int alpha = a0 + eml;
            const bool active = lane_ok && alpha < nu1;
            cpx<Real> sig[6];
            if (active) {
#pragma unroll
                for (int d = 0; d < 6; d++)
                    sig[d] = S[(size_t)(d * NPEM + ipnt) * nu1max + alpha];
            }
            cpx<Real> F[3];
            quad6_point<Real>(sig, (Real)alpha, axial_row0, q,
                              L.sX[active ? eml : 0], L.sY[active ? eml : 0],
                              ipnt, i, j, G, active, F);
            if (active && alpha < nu1p) {
                Real *dst = s.eb + ((size_t)s.elem_eb_off[e] +
                                    (size_t)ipnt * nu1 + alpha) * 6;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    dst[2 * c] = F[c].re;
                    dst[2 * c + 1] = F[c].im;
                }
            }
            __syncthreads();
        }
        }
    }
}

__global__ void k_reset_counter(int *c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0;
}

template <typename Real, bool ATT>
static void launch_cd_att(hipStream_t st, const CardArgs<Real> &a, int n_slots,
                          int occ) {
    switch (occ) {
        case 4:
            hipLaunchKernelGGL((k_solid_stiff_cd<Real, ATT, 4>),
                               dim3(n_slots), dim3(STIFF_BLOCK), 0, st, a);
            break;
        case 3:
            hipLaunchKernelGGL((k_solid_stiff_cd<Real, ATT, 3>),
                               dim3(n_slots), dim3(STIFF_BLOCK), 0, st, a);
            break;
        default:
            hipLaunchKernelGGL((k_solid_stiff_cd<Real, ATT, 2>),
                               dim3(n_slots), dim3(STIFF_BLOCK), 0, st, a);
            break;
    }
}

template <typename Real>
void launch_solid_stiff_cd(hipStream_t st, const CardArgs<Real> &a,
                           int n_slots, int occ) {
    if (a.n3d <= 0) return;
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, st,
                       a.work_counter);
    if (a.nsls > 0)
        launch_cd_att<Real, true>(st, a, n_slots, occ);
    else
        launch_cd_att<Real, false>(st, a, n_slots, occ);
}

template <typename Real, bool ATT, int OCC>
static int blocks_per_cu_v() {
    int n = 0;
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &n, k_solid_stiff_cd<Real, ATT, OCC>, STIFF_BLOCK, 0);
    return n > 0 ? n : 1;
}

template <typename Real>
int card_blocks_per_cu(bool att, int occ) {
    if (att)
        return occ == 4 ? blocks_per_cu_v<Real, true, 4>()
             : occ == 3 ? blocks_per_cu_v<Real, true, 3>()
                        : blocks_per_cu_v<Real, true, 2>();
    return occ == 4 ? blocks_per_cu_v<Real, false, 4>()
         : occ == 3 ? blocks_per_cu_v<Real, false, 3>()
                    : blocks_per_cu_v<Real, false, 2>();
}

template int card_blocks_per_cu<float>(bool, int);
template int card_blocks_per_cu<double>(bool, int);
template void launch_solid_stiff_cd<float>(hipStream_t,
                                           const CardArgs<float> &, int, int);
template void launch_solid_stiff_cd<double>(hipStream_t,
                                            const CardArgs<double> &, int,
                                            int);

}  // namespace axs3d
