// axs3d_fft.hpp -- the azimuthal FFT of the cardinal path: one complex
// series of Nr per wave, in place in LDS, in passes of radix 2 to 16.
//
// Conventions (SolverFFTW, src/core/fft/SolverFFTW.hpp:101-168): the
// inverse transform (Fourier orders to slices, C2R side) uses exp(+2 pi i
// jk / Nr), the forward (R2C side) exp(-2 pi i jk / Nr); both unscaled
// here, the 1/Nr and the Nyquist rules live in the callers' pack and
// unpack.
//
// Algorithm.  Nr = R_0 R_1 ... R_{P-1}, each pass radix R_p in 2..16 (a
// plan per Nr, chosen on the host by an instruction-count model).  Pass p
// has span L_p = R_p m_p (L_0 = Nr, L_{p+1} = m_p) and Nr / R_p
// butterflies; butterfly b takes the elements base + r m_p, r < R_p, with
// base = (b / m_p) L_p + j, j = b % m_p, and writes its outputs back to the
// same elements, so the lanes of a wave never touch each other's data
// within a pass.
//  * inverse (decimation in frequency, passes in plan order): R-point DFT,
//    then output q times w^{jq}, w = exp(+2 pi i / L_p); slice k ends at
//    the mixed-radix digit reversal pos[k] of the pass radices;
//  * forward (decimation in time, passes in reverse order, input in that
//    digit-reversed order): input r times w^{jr}, w = exp(-2 pi i / L_p),
//    then the DFT; the output is in natural order.
// The R-point DFTs run in registers with compile-time coefficients: the
// primes 2, 3, 5, 7, 11, 13 and 4, 8 directly, the other composites by
// the prime-factor algorithm (coprime factors, no twiddles: 6, 10, 12, 14,
// 15) or Cooley-Tukey with constant twiddles (9, 16).  The pass twiddles
// come from a table of every power: for pass p, entry j (R_p - 1) + q - 1
// = exp(-2 pi i j q / L_p), j < m_p, 1 <= q < R_p; the tables of all passes
// hold Nr - 1 entries together (sum_p m_p (R_p - 1) telescopes).  Against
// per-butterfly powers of one twiddle this trades (R - 2) complex products
// for R - 1 LDS reads, and the higher radices halve the passes of the
// radix-2..8 design: fewer address computations and LDS round trips per
// element.
//
// Everything here is __host__ __device__ so that the host can run the same
// passes lane by lane (fft_host_run) and check the plans against a DFT.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <utility>
#include <vector>

#include "axs3d_common.hpp"

namespace axs3d {

// LDS index of element i of a padded series: one complex of padding every
// 16, so that strided accesses of the late passes spread over the banks.
__host__ __device__ __forceinline__ int pad(int i) { return i + (i >> 4); }

// Padded length of a series of n elements.
__host__ __device__ __forceinline__ int padded(int n) {
    return n + (n >> 4) + 1;
}

// ------------------------------------------------- compile-time helpers
namespace fftc {

constexpr double kPi = 3.14159265358979323846264338327950288;

// Taylor series on |x| <= pi/4 (20 terms: below 1e-17)
constexpr double sin_small(double x) {
    double x2 = x * x, t = x, s = x;
    for (int k = 1; k < 20; k++) {
        t *= -x2 / (double)((2 * k) * (2 * k + 1));
        s += t;
    }
    return s;
}
constexpr double cos_small(double x) {
    double x2 = x * x, t = 1.0, s = 1.0;
    for (int k = 1; k < 20; k++) {
        t *= -x2 / (double)((2 * k - 1) * (2 * k));
        s += t;
    }
    return s;
}

// cos and sin of 2 pi e / n, reduced exactly to an octant: theta = o pi/4
// + phi with 8 e = o n + r, phi = (pi/4) r / n
struct CS {
    double c, s;
};
constexpr CS cs2pi(long e, long n) {
    e %= n;
    if (e < 0) e += n;
    const long o = (8 * e) / n;
    const long r = 8 * e - o * n;
    double c = 0, s = 0;
    int q = 0;
    if (o % 2 == 0) {
        const double phi = (kPi / 4) * (double)r / (double)n;
        c = cos_small(phi);
        s = sin_small(phi);
        q = (int)(o / 2);
    } else {
        // theta = ((o + 1) / 2) pi/2 - psi, psi = (pi/4) (n - r) / n
        const double psi = (kPi / 4) * (double)(n - r) / (double)n;
        c = cos_small(psi);
        s = -sin_small(psi);
        q = (int)((o + 1) / 2);
    }
    switch (q & 3) {
        case 0: return {c, s};
        case 1: return {-s, c};
        case 2: return {-c, -s};
        default: return {s, -c};
    }
}

template <int... Is, typename F>
__host__ __device__ __forceinline__ void sfor_impl(
    std::integer_sequence<int, Is...>, F &&f) {
    (f(std::integral_constant<int, Is>{}), ...);
}
// f(integral_constant<int, i>) for i = 0 .. N-1, unrolled at compile time
template <int N, typename F>
__host__ __device__ __forceinline__ void sfor(F &&f) {
    sfor_impl(std::make_integer_sequence<int, N>{}, f);
}

// unique k in [0, A B) with k % A == k1 and k % B == k2 (A, B coprime)
constexpr int crt(int A, int B, int k1, int k2) {
    for (int k = 0; k < A * B; k++)
        if (k % A == k1 && k % B == k2) return k;
    return -1;
}

}  // namespace fftc

// v times W_N^E, W = exp(sg 2 pi i / N), sg = +1 (INV) or -1; exact for
// the multiples of a quarter turn, two products for odd eighths.
template <typename Real, int N, int E, bool INV>
__host__ __device__ __forceinline__ cpx<Real> twc(cpx<Real> v) {
    constexpr int e = ((E % N) + N) % N;
    if constexpr (e == 0) {
        return v;
    } else if constexpr ((4 * e) % N == 0) {
        constexpr int quarter = (4 * e) / N;          // 1, 2, 3
        constexpr int qs = INV ? quarter : (4 - quarter) % 4;
        if constexpr (qs == 1) return {-v.im, v.re};  // i v
        else if constexpr (qs == 2) return {-v.re, -v.im};
        else return {v.im, -v.re};                    // -i v
    } else if constexpr ((8 * e) % N == 0) {
        constexpr fftc::CS w = fftc::cs2pi(INV ? e : -e, N);
        constexpr Real h = (Real)0.70710678118654752440;
        // w = (+-1 +- i) h: (a + ib)(cr + i ci) with |cr| = |ci| = h
        constexpr Real cr = w.c > 0 ? (Real)1 : (Real)-1;
        constexpr Real ci = w.s > 0 ? (Real)1 : (Real)-1;
        return {(v.re * cr - v.im * ci) * h, (v.re * ci + v.im * cr) * h};
    } else {
        constexpr fftc::CS w = fftc::cs2pi(INV ? e : -e, N);
        constexpr Real c = (Real)w.c, s = (Real)w.s;
        return {v.re * c - v.im * s, v.re * s + v.im * c};
    }
}

// In-register R-point DFT, natural order in and out, W = exp(sg 2 pi i /
// R), unscaled.
template <typename Real, int R, bool INV>
struct Dft;

template <typename Real, bool INV>
struct Dft<Real, 1, INV> {
    __host__ __device__ static __forceinline__ void run(cpx<Real> (&)[1]) {}
};

template <typename Real, bool INV>
struct Dft<Real, 2, INV> {
    __host__ __device__ static __forceinline__ void run(cpx<Real> (&x)[2]) {
        const cpx<Real> a = x[0];
        x[0] = cadd(a, x[1]);
        x[1] = csub(a, x[1]);
    }
};

template <typename Real, bool INV>
struct Dft<Real, 3, INV> {
    __host__ __device__ static __forceinline__ void run(cpx<Real> (&x)[3]) {
        constexpr Real c1 = (Real)-0.5;
        constexpr Real s1 = (INV ? (Real)1 : (Real)-1) *
                            (Real)0.86602540378443864676;
        const cpx<Real> t = cadd(x[1], x[2]);
        const cpx<Real> d = csub(x[1], x[2]);
        const cpx<Real> m = cfma(t, c1, x[0]);
        const cpx<Real> id = {-s1 * d.im, s1 * d.re};
        x[0] = cadd(x[0], t);
        x[1] = cadd(m, id);
        x[2] = csub(m, id);
    }
};

template <typename Real, bool INV>
struct Dft<Real, 4, INV> {
    __host__ __device__ static __forceinline__ void run(cpx<Real> (&x)[4]) {
        const cpx<Real> s02 = cadd(x[0], x[2]), d02 = csub(x[0], x[2]);
        const cpx<Real> s13 = cadd(x[1], x[3]), d13 = csub(x[1], x[3]);
        // sg i d13
        const cpx<Real> rd = INV ? cpx<Real>{-d13.im, d13.re}
                                 : cpx<Real>{d13.im, -d13.re};
        x[0] = cadd(s02, s13);
        x[2] = csub(s02, s13);
        x[1] = cadd(d02, rd);
        x[3] = csub(d02, rd);
    }
};

template <typename Real, bool INV>
struct Dft<Real, 5, INV> {
    __host__ __device__ static __forceinline__ void run(cpx<Real> (&x)[5]) {
        constexpr Real sg = INV ? (Real)1 : (Real)-1;
        constexpr Real c1 = (Real)0.30901699437494742410;   // cos 2pi/5
        constexpr Real c2 = (Real)-0.80901699437494742410;  // cos 4pi/5
        constexpr Real s1 = sg * (Real)0.95105651629515357212;
        constexpr Real s2 = sg * (Real)0.58778525229247312917;
        const cpx<Real> t1 = cadd(x[1], x[4]), d1 = csub(x[1], x[4]);
        const cpx<Real> t2 = cadd(x[2], x[3]), d2 = csub(x[2], x[3]);
        const cpx<Real> m1 = cadd(x[0], cadd(cmul(t1, c1), cmul(t2, c2)));
        const cpx<Real> m2 = cadd(x[0], cadd(cmul(t1, c2), cmul(t2, c1)));
        const cpx<Real> e1 = cadd(cmul(d1, s1), cmul(d2, s2));
        const cpx<Real> e2 = csub(cmul(d1, s2), cmul(d2, s1));
        const cpx<Real> ie1 = {-e1.im, e1.re}, ie2 = {-e2.im, e2.re};
        x[0] = cadd(x[0], cadd(t1, t2));
        x[1] = cadd(m1, ie1);
        x[4] = csub(m1, ie1);
        x[2] = cadd(m2, ie2);
        x[3] = csub(m2, ie2);
    }
};

// odd primes 7, 11, 13: with t_r = x_r + x_{R-r}, d_r = x_r - x_{R-r},
// X_q = x_0 + sum_r cos(2 pi rq/R) t_r + sg i sum_r sin(2 pi rq/R) d_r
// and X_{R-q} the same with -sg
template <typename Real, int R, bool INV>
struct DftOddPrime {
    __host__ __device__ static __forceinline__ void run(cpx<Real> (&x)[R]) {
        constexpr int H = (R - 1) / 2;
        cpx<Real> t[H], d[H];
        fftc::sfor<H>([&](auto I) {
            constexpr int r = decltype(I)::value + 1;
            t[r - 1] = cadd(x[r], x[R - r]);
            d[r - 1] = csub(x[r], x[R - r]);
        });
        const cpx<Real> x0 = x[0];
        cpx<Real> sum = x0;
        fftc::sfor<H>([&](auto I) { sum = cadd(sum, t[decltype(I)::value]); });
        x[0] = sum;
        fftc::sfor<H>([&](auto Q) {
            constexpr int q = decltype(Q)::value + 1;
            cpx<Real> A = x0, B = {(Real)0, (Real)0};
            fftc::sfor<H>([&](auto Rr) {
                constexpr int r = decltype(Rr)::value + 1;
                constexpr fftc::CS w = fftc::cs2pi((long)r * q, R);
                constexpr Real c = (Real)w.c;
                constexpr Real s = (Real)w.s;
                A = cfma(t[r - 1], c, A);
                B = cfma(d[r - 1], s, B);
            });
            const cpx<Real> iB = INV ? cpx<Real>{-B.im, B.re}
                                     : cpx<Real>{B.im, -B.re};
            x[q] = cadd(A, iB);
            x[R - q] = csub(A, iB);
        });
    }
};
template <typename Real, bool INV>
struct Dft<Real, 7, INV> : DftOddPrime<Real, 7, INV> {};
template <typename Real, bool INV>
struct Dft<Real, 11, INV> : DftOddPrime<Real, 11, INV> {};
template <typename Real, bool INV>
struct Dft<Real, 13, INV> : DftOddPrime<Real, 13, INV> {};

// prime-factor algorithm, N = A B with A, B coprime: input index (B n1 + A
// n2) mod N, a 2D DFT without twiddles, output at the CRT index of (k1, k2)
template <typename Real, int A, int B, bool INV>
struct DftPfa {
    static constexpr int N = A * B;
    __host__ __device__ static __forceinline__ void run(cpx<Real> (&x)[N]) {
        cpx<Real> y[A][B];
        fftc::sfor<B>([&](auto I2) {
            constexpr int n2 = decltype(I2)::value;
            cpx<Real> t[A];
            fftc::sfor<A>([&](auto I1) {
                constexpr int n1 = decltype(I1)::value;
                t[n1] = x[(B * n1 + A * n2) % N];
            });
            Dft<Real, A, INV>::run(t);
            fftc::sfor<A>([&](auto K1) {
                y[decltype(K1)::value][n2] = t[decltype(K1)::value];
            });
        });
        fftc::sfor<A>([&](auto K1) {
            constexpr int k1 = decltype(K1)::value;
            cpx<Real> t[B];
            fftc::sfor<B>([&](auto I2) {
                t[decltype(I2)::value] = y[k1][decltype(I2)::value];
            });
            Dft<Real, B, INV>::run(t);
            fftc::sfor<B>([&](auto K2) {
                constexpr int k2 = decltype(K2)::value;
                x[fftc::crt(A, B, k1, k2)] = t[k2];
            });
        });
    }
};

// Cooley-Tukey, N = A B: X[k1 + A k2] = sum_n2 W_B^{n2 k2} W_N^{n2 k1}
// sum_n1 x[B n1 + n2] W_A^{n1 k1}
template <typename Real, int A, int B, bool INV>
struct DftCt {
    static constexpr int N = A * B;
    __host__ __device__ static __forceinline__ void run(cpx<Real> (&x)[N]) {
        cpx<Real> y[B][A];
        fftc::sfor<B>([&](auto I2) {
            constexpr int n2 = decltype(I2)::value;
            cpx<Real> t[A];
            fftc::sfor<A>([&](auto I1) {
                t[decltype(I1)::value] = x[B * decltype(I1)::value + n2];
            });
            Dft<Real, A, INV>::run(t);
            fftc::sfor<A>([&](auto K1) {
                constexpr int k1 = decltype(K1)::value;
                y[n2][k1] = twc<Real, N, n2 * k1, INV>(t[k1]);
            });
        });
        fftc::sfor<A>([&](auto K1) {
            constexpr int k1 = decltype(K1)::value;
            cpx<Real> t[B];
            fftc::sfor<B>([&](auto I2) {
                t[decltype(I2)::value] = y[decltype(I2)::value][k1];
            });
            Dft<Real, B, INV>::run(t);
            fftc::sfor<B>([&](auto K2) {
                constexpr int k2 = decltype(K2)::value;
                x[k1 + A * k2] = t[k2];
            });
        });
    }
};

template <typename Real, bool INV>
struct Dft<Real, 8, INV> : DftCt<Real, 2, 4, INV> {};
template <typename Real, bool INV>
struct Dft<Real, 6, INV> : DftPfa<Real, 2, 3, INV> {};
template <typename Real, bool INV>
struct Dft<Real, 9, INV> : DftCt<Real, 3, 3, INV> {};
template <typename Real, bool INV>
struct Dft<Real, 10, INV> : DftPfa<Real, 2, 5, INV> {};
template <typename Real, bool INV>
struct Dft<Real, 12, INV> : DftPfa<Real, 4, 3, INV> {};
template <typename Real, bool INV>
struct Dft<Real, 14, INV> : DftPfa<Real, 2, 7, INV> {};
template <typename Real, bool INV>
struct Dft<Real, 15, INV> : DftPfa<Real, 3, 5, INV> {};
template <typename Real, bool INV>
struct Dft<Real, 16, INV> : DftCt<Real, 4, 4, INV> {};

// (a * b) for a, b < 2^24: the full-rate 24-bit multiply
__host__ __device__ __forceinline__ int mul24(int a, int b) {
    return (int)(((uint32_t)a & 0xffffffu) * ((uint32_t)b & 0xffffffu));
}
// floor(a / m) from the magic multiplier mg = ceil(2^32 / m), exact for
// a m < 2^32
__host__ __device__ __forceinline__ int divm(int a, uint32_t mg) {
    return (int)(((uint64_t)(uint32_t)a * (uint64_t)mg) >> 32);
}

// Plan words per Nr (FFT_PLAN_W of them): word 0 = number of passes |
// (table length << 8); per pass p, in decimation-in-frequency order:
// [1 + 3p] = radix | (m << 8), [2 + 3p] = ceil(2^32 / m) (0 for m = 1),
// [3 + 3p] = offset of the pass's twiddles | (butterflies << 16).
constexpr int FFT_MAX_PASSES = 12;
constexpr int FFT_PLAN_W = 1 + 3 * FFT_MAX_PASSES;
// Largest Nr (the longest lucky number Nr <= 2048 is 2048 itself).
constexpr int FFT_NR_CAP = 2048;

// One pass of radix R of the series buf by lanes lane, lane + nlanes, ...
template <typename Real, int R, bool INV>
__host__ __device__ __forceinline__ void fft_pass(
    cpx<Real> *buf, int nb, int m, uint32_t mg,
    const cpx<Real> *__restrict__ tws, int lane, int nlanes) {
    const int L = m * R;
    for (int b = lane; b < nb; b += nlanes) {
        const int blk = m == 1 ? b : divm(b, mg);
        const int j = b - mul24(blk, m);
        const int base = mul24(blk, L) + j;
        int ad[R];
        cpx<Real> v[R];
        fftc::sfor<R>([&](auto I) {
            constexpr int r = decltype(I)::value;
            ad[r] = pad(base + mul24(r, m));
            v[r] = buf[ad[r]];
        });
        const cpx<Real> *w = tws + mul24(j, R - 1);
        if (!INV && m > 1) {
            fftc::sfor<R - 1>([&](auto I) {
                constexpr int r = decltype(I)::value + 1;
                const cpx<Real> t = w[r - 1];
                const cpx<Real> a = v[r];
                v[r] = {a.re * t.re - a.im * t.im, a.re * t.im + a.im * t.re};
            });
        }
        Dft<Real, R, INV>::run(v);
        if (INV && m > 1) {
            // times conj(table): exp(+2 pi i j q / L)
            fftc::sfor<R - 1>([&](auto I) {
                constexpr int q = decltype(I)::value + 1;
                const cpx<Real> t = w[q - 1];
                const cpx<Real> a = v[q];
                v[q] = {a.re * t.re + a.im * t.im, a.im * t.re - a.re * t.im};
            });
        }
        fftc::sfor<R>([&](auto I) {
            constexpr int r = decltype(I)::value;
            buf[ad[r]] = v[r];
        });
    }
}

// Pass p of the plan (radix switch).
template <typename Real, bool INV>
__host__ __device__ __forceinline__ void fft_pass_p(
    cpx<Real> *buf, const uint32_t *__restrict__ plan,
    const cpx<Real> *__restrict__ tw, int p, int lane, int nlanes) {
    const uint32_t w0 = plan[1 + 3 * p];
    const uint32_t mg = plan[2 + 3 * p];
    const uint32_t w2 = plan[3 + 3 * p];
    const int R = (int)(w0 & 255u), m = (int)(w0 >> 8);
    const int nb = (int)(w2 >> 16);
    const cpx<Real> *tws = tw + (w2 & 0xffffu);
    switch (R) {
#define AXS3D_FFT_CASE(RR)                                                   \
    case RR:                                                                 \
        fft_pass<Real, RR, INV>(buf, nb, m, mg, tws, lane, nlanes);          \
        break;
        AXS3D_FFT_CASE(2)
        AXS3D_FFT_CASE(3)
        AXS3D_FFT_CASE(4)
        AXS3D_FFT_CASE(5)
        AXS3D_FFT_CASE(6)
        AXS3D_FFT_CASE(7)
        AXS3D_FFT_CASE(8)
        AXS3D_FFT_CASE(9)
        AXS3D_FFT_CASE(10)
        AXS3D_FFT_CASE(11)
        AXS3D_FFT_CASE(12)
        AXS3D_FFT_CASE(13)
        AXS3D_FFT_CASE(14)
        AXS3D_FFT_CASE(15)
        AXS3D_FFT_CASE(16)
#undef AXS3D_FFT_CASE
        default: break;
    }
}

// Whole transform of buf by one wave (lane = lane index in the wave).
// The inverse takes the natural order and leaves slice k at pos[k]; the
// forward takes that order and leaves the natural one.
template <typename Real>
__device__ __forceinline__ void wave_fft(cpx<Real> *buf,
                                         const uint32_t *__restrict__ plan,
                                         bool inv, const cpx<Real> *tw,
                                         int lane) {
    const int np = (int)(plan[0] & 255u);
    if (inv) {
        for (int p = 0; p < np; p++) {
            fft_pass_p<Real, true>(buf, plan, tw, p, lane, 64);
            wave_sync();
        }
    } else {
        for (int p = np - 1; p >= 0; p--) {
            fft_pass_p<Real, false>(buf, plan, tw, p, lane, 64);
            wave_sync();
        }
    }
}

// ---------------------------------------------------------- host side
// The same passes on the host, every lane of a pass before the next pass
// (the lanes of a pass touch disjoint elements, so the order within a
// pass does not matter).
template <typename Real>
inline void fft_host_run(cpx<Real> *buf, const uint32_t *plan, bool inv,
                         const cpx<Real> *tw) {
    const int np = (int)(plan[0] & 255u);
    for (int i = 0; i < np; i++) {
        const int p = inv ? i : np - 1 - i;
        for (int lane = 0; lane < 64; lane++) {
            if (inv)
                fft_pass_p<Real, true>(buf, plan, tw, p, lane, 64);
            else
                fft_pass_p<Real, false>(buf, plan, tw, p, lane, 64);
        }
    }
}

// Model of the VALU instructions of one butterfly of radix R (addresses,
// twiddle products, the DFT) plus its LDS accesses; used to choose the
// pass radices.
inline double fft_butterfly_cost(int R) {
    static const double dft[17] = {0,  0,  4,  12, 16, 34, 40, 66, 56,
                                   84, 80, 140, 100, 172, 150, 150, 150};
    return 4.0 * R + 4.0 * (R - 1) + dft[R] + 3.0 * R + 10.0;
}

// Fixed cost of a pass in the same units (loop set-up, the wave's
// barrier, the LDS latency that a pass of few butterflies cannot hide).
constexpr double kFftPassCost = 40.0;

// Pass radices of Nr (prime factors <= 13), each <= max_radix (or a
// prime), minimising
// the modelled cost sum_p ceil((Nr / R_p) / 64) cost(R_p) (a pass has Nr /
// R_p butterflies whatever its place); descending order.  Empty if Nr has
// a prime factor above 13.
inline std::vector<int> fft_choose_radices(int nr, int max_radix) {
    std::vector<int> divs;
    for (int d = 1; d <= nr; d++)
        if (nr % d == 0) divs.push_back(d);
    // best[d] = min cost of splitting d into passes, choice[d] = radix
    std::vector<double> best(nr + 1, 1e300);
    std::vector<int> choice(nr + 1, 0);
    best[1] = 0.0;
    for (int d : divs) {
        if (d == 1) continue;
        for (int R = 2; R <= 16; R++) {
            // a prime factor is a pass of its own whatever the limit
            const bool prime = R == 2 || R == 3 || R == 5 || R == 7 ||
                               R == 11 || R == 13;
            if (R > max_radix && !prime) continue;
            if (d % R != 0 || best[d / R] >= 1e299) continue;
            const int nb = nr / R;
            const double c = best[d / R] + kFftPassCost +
                             (double)((nb + 63) / 64) * fft_butterfly_cost(R);
            if (c < best[d]) {
                best[d] = c;
                choice[d] = R;
            }
        }
    }
    std::vector<int> rad;
    if (best[nr] >= 1e299) return rad;
    for (int d = nr; d > 1; d /= choice[d]) rad.push_back(choice[d]);
    std::sort(rad.begin(), rad.end(), std::greater<int>());
    return rad;
}

// Plan words (FFT_PLAN_W), the twiddle table (Nr - 1 complex, FP64,
// interleaved) and the slice positions (Nr) of the given pass radices.
inline void fft_build_plan(int nr, const std::vector<int> &rad,
                           uint32_t *words, std::vector<double> &tw,
                           std::vector<uint16_t> &pos) {
    for (int i = 0; i < FFT_PLAN_W; i++) words[i] = 0;
    tw.clear();
    int span = nr, ntw = 0;
    for (size_t p = 0; p < rad.size(); p++) {
        const int R = rad[p], m = span / R;
        words[1 + 3 * p] = (uint32_t)R | ((uint32_t)m << 8);
        words[2 + 3 * p] =
            m == 1 ? 0u : (uint32_t)((((uint64_t)1 << 32) + m - 1) / m);
        words[3 + 3 * p] = (uint32_t)ntw | ((uint32_t)(nr / R) << 16);
        for (int j = 0; j < m; j++)
            for (int q = 1; q < R; q++) {
                const fftc::CS w = fftc::cs2pi(-(long)j * q, span);
                tw.push_back(w.c);
                tw.push_back(w.s);
            }
        ntw += m * (R - 1);
        span = m;
    }
    words[0] = (uint32_t)rad.size() | ((uint32_t)ntw << 8);
    pos.resize(nr);
    for (int sl = 0; sl < nr; sl++) {
        int rest = sl, ps = 0, sp = nr;
        for (int R : rad) {
            sp /= R;
            ps += (rest % R) * sp;
            rest /= R;
        }
        pos[sl] = (uint16_t)ps;
    }
}

// Bytes of the dynamic LDS of a block that holds the twiddle table, the
// slice positions and nbuf padded series for an Nr of at most nr_max,
// plus extra bytes at the end; the layout carved by FftLds.
template <typename Real>
__host__ __device__ __forceinline__ int fft_lds_tw_bytes(int nr) {
    return ((nr * (int)sizeof(cpx<Real>)) + 15) & ~15;
}
__host__ __device__ __forceinline__ int fft_lds_pos_bytes(int nr) {
    return ((nr * 2) + 15) & ~15;
}
template <typename Real>
__host__ __device__ __forceinline__ int fft_lds_buf_bytes(int nr) {
    return ((padded(nr) * (int)sizeof(cpx<Real>)) + 15) & ~15;
}
template <typename Real>
inline size_t fft_lds_bytes(int nr_max, int nbuf, int extra = 0) {
    return (size_t)fft_lds_tw_bytes<Real>(nr_max) + fft_lds_pos_bytes(nr_max) +
           (size_t)nbuf * fft_lds_buf_bytes<Real>(nr_max) + (size_t)extra;
}

// The block's view of that layout for its own Nr (blocks of one launch
// may have smaller Nr than the launch's maximum; each carves its own).
template <typename Real>
struct FftLds {
    cpx<Real> *tw;
    uint16_t *pos;
    unsigned char *bufs;
    int buf_bytes;
    unsigned char *extra;
    __device__ cpx<Real> *buf(int i) const {
        return reinterpret_cast<cpx<Real> *>(bufs + (size_t)i * buf_bytes);
    }
    __device__ void carve(unsigned char *lds, int nr, int nbuf) {
        tw = reinterpret_cast<cpx<Real> *>(lds);
        pos = reinterpret_cast<uint16_t *>(lds + fft_lds_tw_bytes<Real>(nr));
        bufs = lds + fft_lds_tw_bytes<Real>(nr) + fft_lds_pos_bytes(nr);
        buf_bytes = fft_lds_buf_bytes<Real>(nr);
        extra = bufs + (size_t)nbuf * buf_bytes;
    }
};

}  // namespace axs3d
