// axs3d_kernels.hip -- CDNA4 (gfx950) kernels of the stiffness path.
//
// Fourier-path solid element stiffness (1D material, no PRT): one thread
// per (element, Fourier order alpha, GLL point); ten element-modes per
// 256-thread block (four wave64s).  The element-mode displacement tile and
// the quadrature pre-products are staged in LDS so the two 5x5 GLL
// contractions (gradient and quadrature) read neighbours from LDS instead
// of HBM.  Assembly subtracts into the point stiffness with hardware float
// atomics (global_atomic_add_f32/f64 at L2).
//
// Reference semantics (src = SOLVER/src):
//   gather      SolidPoint::scatterDisplToElement  src/core/point/SolidPoint.hpp:82-98
//   gradient    GradientQuadrature::computeGrad6   src/core/element/grad/GradientQuadrature.hpp:207-231,405-459
//   stress      Isotropic::strainToStress          src/core/element/material/elastic/anisotropy/Isotropic.hpp:155-172
//   quadrature  GradientQuadrature::computeQuad6   GradientQuadrature.hpp:294-324,463-500
//   scatter     SolidPoint::gatherStiffFromElement src/core/point/SolidPoint.hpp:101-109
// The axial L'Hospital terms of computeQuad_1overS_addTo (row 0 of an
// axial element) are folded into the row-0 quadrature pre-products X and Y:
// F = G_xi X + Y G^T makes G_xi(:,0) X(0,:) and Y(0,:) G^T exactly the
// two correction products of GradientQuadrature.hpp:493-497.
#include "axs3d_common.hpp"
#include "axs3d_launch.hpp"

namespace axs3d {

constexpr int EMB = 10;          // element-modes per block
constexpr int STIFF_BLOCK = 256;

template <typename Real>
struct QuadOps {
    Real dsdxii, dsdeta, dzdxii, dzdeta, invs;
};

// Quadrature of one component pair (dF/dS, dF/dZ) at this point:
// X = dFdS*dzdeta + dFdZ*dsdeta, Y = dFdS*dzdxii + dFdZ*dsdxii
// (computeQuad_dSdZ, GradientQuadrature.hpp:471-474).
template <typename Real>
__device__ __forceinline__ void quad_xy(cpx<Real> dfds, cpx<Real> dfdz,
                                        const QuadOps<Real> &q, cpx<Real> &X,
                                        cpx<Real> &Y) {
    X = cadd(cmul(dfds, q.dzdeta), cmul(dfdz, q.dsdeta));
    Y = cadd(cmul(dfds, q.dzdxii), cmul(dfdz, q.dsdxii));
}

// Shared quadrature tail: from the six stress components at this point,
// form X, Y (with the axial fold) and the local 1/s term L, exchange X, Y
// through LDS and contract with G.  Returns F for the three components.
template <typename Real>
__device__ __forceinline__ void quad6_point(
    const cpx<Real> (&sig)[6], Real alpha, bool axial_row0,
    const QuadOps<Real> &q, cpx<Real> (*sX)[NPEM], cpx<Real> (*sY)[NPEM],
    int ipnt, int i, int j, const Real *gxi_row, const Real *g_row,
    bool active, cpx<Real> (&F)[3]) {
    cpx<Real> X[3], Y[3], L[3];
    if (active) {
        // (dFdS, dFdZ) per output component: (0,4), (5,3), (4,2)
        quad_xy(sig[0], sig[4], q, X[0], Y[0]);
        quad_xy(sig[5], sig[3], q, X[1], Y[1]);
        quad_xy(sig[4], sig[2], q, X[2], Y[2]);
        // lower-order terms (GradientQuadrature.hpp:311-318)
        cpx<Real> b[3];
        b[0] = csub(ialpha(alpha, sig[5]), sig[1]);
        b[1] = cadd(ialpha(alpha, sig[1]), sig[5]);
        b[2] = ialpha(alpha, sig[3]);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            L[c] = cmul(b[c], -q.invs);
            if (axial_row0) {
                X[c] = csub(X[c], cmul(b[c], q.dzdeta));
                Y[c] = csub(Y[c], cmul(b[c], q.dzdxii));
            }
            sX[c][ipnt] = X[c];
            sY[c][ipnt] = Y[c];
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            cpx<Real> a = {0, 0}, bsum = {0, 0};
#pragma unroll
            for (int k = 0; k < NPED; k++) {
                a = cfma(sX[c][k * NPED + j], gxi_row[k], a);
                bsum = cfma(sY[c][i * NPED + k], g_row[k], bsum);
            }
            F[c] = cadd(cadd(a, bsum), L[c]);
        }
    }
}

template <typename Real>
__device__ __forceinline__ void atomic_sub_cpx(Real *dst, cpx<Real> v) {
    unsafeAtomicAdd(dst, -v.re);
    unsafeAtomicAdd(dst + 1, -v.im);
}

template <typename Real>
__global__ __launch_bounds__(STIFF_BLOCK) void k_solid_stiff_fr(
    int n_em, const EMItem *__restrict__ em_list,
    const int32_t *__restrict__ elem_points,
    const uint8_t *__restrict__ elem_axial,
    const Real *__restrict__ elem_ops, const Real *__restrict__ elem_mat,
    const uint32_t *__restrict__ point_off,
    const int32_t *__restrict__ point_nu1, const double *__restrict__ displ,
    Real *__restrict__ stiff, const Real *__restrict__ gtab) {
    __shared__ cpx<Real> sU[EMB][3][NPEM];
    __shared__ cpx<Real> sX[EMB][3][NPEM];
    __shared__ cpx<Real> sY[EMB][3][NPEM];
    __shared__ Real sG[2][NPEM];

    const int t = threadIdx.x;
    if (t < 2 * NPEM) sG[t / NPEM][t % NPEM] = gtab[t];
    const int eml = t / NPEM;
    const int ipnt = t - eml * NPEM;
    const int em = blockIdx.x * EMB + eml;
    const bool active = (eml < EMB) && (em < n_em);
    const int i = ipnt / NPED, j = ipnt - (ipnt / NPED) * NPED;

    int e = 0, alpha = 0, nu1p = 0;
    uint32_t off = 0;
    bool axial = false;
    cpx<Real> u[3] = {{0, 0}, {0, 0}, {0, 0}};
    if (active) {
        const EMItem it = em_list[em];
        e = it.elem;
        alpha = it.alpha;
        const int p = elem_points[e * NPEM + ipnt];
        nu1p = point_nu1[p];
        off = point_off[p];
        axial = elem_axial[e] != 0;
        if (alpha < nu1p) {
            const double *src = displ + (size_t)(off + alpha) * 6;
#pragma unroll
            for (int c = 0; c < 3; c++)
                u[c] = {(Real)src[2 * c], (Real)src[2 * c + 1]};
        }
#pragma unroll
        for (int c = 0; c < 3; c++) sU[eml][c][ipnt] = u[c];
    }
    __syncthreads();

    Real gxi_col[NPED], g_col[NPED], gxi_row[NPED], g_row[NPED];
    const Real *Gxi = sG[axial ? 1 : 0];
#pragma unroll
    for (int k = 0; k < NPED; k++) {
        gxi_col[k] = Gxi[k * NPED + i];   // GT_xi(i,k) = G_xi(k,i)
        g_col[k] = sG[0][k * NPED + j];   // G(k,j)
        gxi_row[k] = Gxi[i * NPED + k];   // G_xi(i,k)
        g_row[k] = sG[0][j * NPED + k];   // GT(k,j) = G(j,k)
    }

    cpx<Real> sig[6];
    QuadOps<Real> qi{};
    const Real ralpha = (Real)alpha;
    const bool axial_row0 = axial && (i == 0);
    if (active) {
        const Real *ops = elem_ops + (size_t)e * ELEM_OPS * NPEM + ipnt;
        const Real dsdxii = ops[OP_DSDXII * NPEM];
        const Real dsdeta = ops[OP_DSDETA * NPEM];
        const Real dzdxii = ops[OP_DZDXII * NPEM];
        const Real dzdeta = ops[OP_DZDETA * NPEM];
        const Real invs = ops[OP_INVS * NPEM];
        qi.dsdxii = ops[OP_DSDXII_IF * NPEM];
        qi.dsdeta = ops[OP_DSDETA_IF * NPEM];
        qi.dzdxii = ops[OP_DZDXII_IF * NPEM];
        qi.dzdeta = ops[OP_DZDETA_IF * NPEM];
        qi.invs = ops[OP_INVS_IF * NPEM];

        // higher-order terms (computeGrad_dSdZ, GradientQuadrature.hpp:405-418)
        cpx<Real> dS[3], dZ[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            cpx<Real> x = {0, 0}, y = {0, 0};
#pragma unroll
            for (int k = 0; k < NPED; k++) {
                x = cfma(sU[eml][c][k * NPED + j], gxi_col[k], x);
                y = cfma(sU[eml][c][i * NPED + k], g_col[k], y);
            }
            dS[c] = cadd(cmul(x, dzdeta), cmul(y, dzdxii));
            dZ[c] = cadd(cmul(x, dsdeta), cmul(y, dsdxii));
        }
        // Voigt strain [ss, pp, zz, pz, sz, sp] (computeGrad6, :207-231)
        cpx<Real> eps[6];
        eps[0] = dS[0];
        eps[4] = cadd(dZ[0], dS[2]);
        eps[5] = dS[1];
        eps[3] = dZ[1];
        eps[2] = dZ[2];
        // lower-order terms: (i alpha u0 - u1, i alpha u1 + u0, i alpha u2)/s
        if (!axial_row0) {
            const cpx<Real> a0 = csub(ialpha(ralpha, u[0]), u[1]);
            const cpx<Real> a1 = cadd(ialpha(ralpha, u[1]), u[0]);
            const cpx<Real> a2 = ialpha(ralpha, u[2]);
            eps[5] = cfma(a0, invs, eps[5]);
            eps[1] = cmul(a1, invs);
            eps[3] = cfma(a2, invs, eps[3]);
        } else {
            // axis: d/ds of the lower-order numerators (L'Hospital),
            // computeGrad_1overS_setTo/addTo row 0, :426-437
            cpx<Real> xa[3] = {{0, 0}, {0, 0}, {0, 0}};
            cpx<Real> ya[3] = {{0, 0}, {0, 0}, {0, 0}};
#pragma unroll
            for (int k = 0; k < NPED; k++) {
                const cpx<Real> uc0 = sU[eml][0][k * NPED + j];
                const cpx<Real> uc1 = sU[eml][1][k * NPED + j];
                const cpx<Real> uc2 = sU[eml][2][k * NPED + j];
                const cpx<Real> ur0 = sU[eml][0][k];
                const cpx<Real> ur1 = sU[eml][1][k];
                const cpx<Real> ur2 = sU[eml][2][k];
                xa[0] = cfma(csub(ialpha(ralpha, uc0), uc1), gxi_col[k], xa[0]);
                xa[1] = cfma(cadd(ialpha(ralpha, uc1), uc0), gxi_col[k], xa[1]);
                xa[2] = cfma(ialpha(ralpha, uc2), gxi_col[k], xa[2]);
                ya[0] = cfma(csub(ialpha(ralpha, ur0), ur1), g_col[k], ya[0]);
                ya[1] = cfma(cadd(ialpha(ralpha, ur1), ur0), g_col[k], ya[1]);
                ya[2] = cfma(ialpha(ralpha, ur2), g_col[k], ya[2]);
            }
            eps[5] = cadd(eps[5], cadd(cmul(xa[0], dzdeta), cmul(ya[0], dzdxii)));
            eps[1] = cadd(cmul(xa[1], dzdeta), cmul(ya[1], dzdxii));
            eps[3] = cadd(eps[3], cadd(cmul(xa[2], dzdeta), cmul(ya[2], dzdxii)));
        }
        // isotropic Hooke (Isotropic.hpp:155-172)
        const Real *mat = elem_mat + (size_t)e * 2 * NPEM + ipnt;
        const Real lam = mat[0];
        const Real mu = mat[NPEM];
        const Real mu2 = mu * (Real)2;
        const cpx<Real> tr = cmul(cadd(cadd(eps[0], eps[1]), eps[2]), lam);
        sig[0] = cfma(eps[0], mu2, tr);
        sig[1] = cfma(eps[1], mu2, tr);
        sig[2] = cfma(eps[2], mu2, tr);
        sig[3] = cmul(eps[3], mu);
        sig[4] = cmul(eps[4], mu);
        sig[5] = cmul(eps[5], mu);
    }

    cpx<Real> F[3];
    quad6_point<Real>(sig, ralpha, axial_row0, qi, sX[active ? eml : 0],
                      sY[active ? eml : 0], ipnt, i, j, gxi_row, g_row,
                      active, F);
    if (active && alpha < nu1p) {
        Real *dst = stiff + (size_t)(off + alpha) * 6;
#pragma unroll
        for (int c = 0; c < 3; c++) atomic_sub_cpx(dst + 2 * c, F[c]);
    }
}

// Moment source: stress = -pattern * stf (SolidElement.cpp:279-280), then
// computeQuad6_NoIntegration (GradientQuadrature.hpp:363-390) and the
// subtracting scatter.  One thread per (alpha < nu1c, GLL point).
template <typename Real>
__global__ __launch_bounds__(STIFF_BLOCK) void k_moment_source(
    int elem, int nu1c, const Real *__restrict__ pattern, const double *stf,
    int tstep, const int32_t *__restrict__ elem_points,
    const uint8_t *__restrict__ elem_axial, const Real *__restrict__ elem_ops,
    const uint32_t *__restrict__ point_off,
    const int32_t *__restrict__ point_nu1, Real *__restrict__ stiff,
    const Real *__restrict__ gtab) {
    __shared__ cpx<Real> sX[EMB][3][NPEM];
    __shared__ cpx<Real> sY[EMB][3][NPEM];
    __shared__ Real sG[2][NPEM];
    const int t = threadIdx.x;
    if (t < 2 * NPEM) sG[t / NPEM][t % NPEM] = gtab[t];
    __syncthreads();
    const int eml = t / NPEM;
    const int ipnt = t - eml * NPEM;
    const int alpha = blockIdx.x * EMB + eml;
    const bool active = (eml < EMB) && (alpha < nu1c);
    const int i = ipnt / NPED, j = ipnt - (ipnt / NPED) * NPED;
    const bool axial = elem_axial[elem] != 0;
    const Real *Gxi = sG[axial ? 1 : 0];
    Real gxi_row[NPED], g_row[NPED];
#pragma unroll
    for (int k = 0; k < NPED; k++) {
        gxi_row[k] = Gxi[i * NPED + k];
        g_row[k] = sG[0][j * NPED + k];
    }
    QuadOps<Real> q{};
    cpx<Real> sig[6];
    const Real sv = (Real)stf[tstep];
    if (active) {
        const Real *ops = elem_ops + (size_t)elem * ELEM_OPS * NPEM + ipnt;
        q.dsdxii = ops[OP_DSDXII * NPEM];
        q.dsdeta = ops[OP_DSDETA * NPEM];
        q.dzdxii = ops[OP_DZDXII * NPEM];
        q.dzdeta = ops[OP_DZDETA * NPEM];
        q.invs = ops[OP_INVS * NPEM];
        const Real *pat = pattern + ((size_t)alpha * 6 * NPEM + ipnt) * 2;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const Real re = pat[k * NPEM * 2] * sv;
            const Real im = pat[k * NPEM * 2 + 1] * sv;
            sig[k] = {-re, -im};
        }
    }
    cpx<Real> F[3];
    quad6_point<Real>(sig, (Real)alpha, axial && i == 0, q,
                      sX[active ? eml : 0], sY[active ? eml : 0], ipnt, i, j,
                      gxi_row, g_row, active, F);
    if (active) {
        const int p = elem_points[elem * NPEM + ipnt];
        if (alpha < point_nu1[p]) {
            Real *dst = stiff + (size_t)(point_off[p] + alpha) * 6;
#pragma unroll
            for (int c = 0; c < 3; c++) atomic_sub_cpx(dst + 2 * c, F[c]);
        }
    }
}

// Point update, one thread per (point, mode).  Flags select the phases:
//   AXIAL  AxialBoundary::apply, solid (AxialBoundary.cpp:21-42)
//   MASS   Nyquist masking + Mass1D (SolidPoint.cpp:28-36, Mass1D.hpp:31)
//   MARCH  NewmarkTimeScheme::update (NewmarkTimeScheme.hpp:48-59)
// With MASS and MARCH fused the acceleration is formed in FP64 from the
// stiffness (FP32 or FP64) and never stored back into it.
template <typename Real, bool AXIAL, bool MASS, bool MARCH>
__global__ __launch_bounds__(256) void k_point(
    int64_t n_modes, const int32_t *__restrict__ mode_point,
    const uint32_t *__restrict__ point_off,
    const int32_t *__restrict__ point_nr,
    const uint8_t *__restrict__ point_axial,
    const double *__restrict__ inv_mass, Real *__restrict__ stiff,
    double *__restrict__ displ, double *__restrict__ veloc,
    double *__restrict__ accel, double dt, double half_dt,
    double half_dt_dt, int *__restrict__ flag) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_modes) return;
    const int p = mode_point[k];
    const int alpha = (int)(k - (int64_t)point_off[p]);
    Real *sp = stiff + (size_t)k * 6;
    Real s[6];
#pragma unroll
    for (int q = 0; q < 6; q++) s[q] = sp[q];
    if (AXIAL && point_axial[p]) {
        if (alpha == 0) {
            s[0] = s[1] = s[2] = s[3] = 0;
        } else if (alpha == 1) {
            // s0 <- (s0 - i s1) / 2 ; s1 <- i s0 ; s2 <- 0
            const Real h = (Real)0.5;
            const Real r0 = (s[0] + s[3]) * h;
            const Real i0 = (s[1] - s[2]) * h;
            s[0] = r0;
            s[1] = i0;
            s[2] = -i0;
            s[3] = r0;
            s[4] = s[5] = 0;
        } else {
#pragma unroll
            for (int q = 0; q < 6; q++) s[q] = 0;
        }
    }
    if (MASS) {
        const int nr = point_nr[p];
        if ((nr & 1) == 0 && alpha == nr / 2) s[1] = s[3] = s[5] = 0;
    }
    if (MASS && !MARCH) {
        const Real im = (Real)inv_mass[p];
#pragma unroll
        for (int q = 0; q < 6; q++) s[q] *= im;
    }
    if (!MARCH) {
#pragma unroll
        for (int q = 0; q < 6; q++) sp[q] = s[q];
        return;
    }
    double a_new[6];
    if (MASS) {
        const double im = inv_mass[p];
#pragma unroll
        for (int q = 0; q < 6; q++) a_new[q] = (double)s[q] * im;
    } else {
#pragma unroll
        for (int q = 0; q < 6; q++) a_new[q] = (double)s[q];
    }
    double *up = displ + (size_t)k * 6;
    double *vp = veloc + (size_t)k * 6;
    double *ap = accel + (size_t)k * 6;
    bool bad = false;
#pragma unroll
    for (int q = 0; q < 6; q++) {
        double v = vp[q] + half_dt * (ap[q] + a_new[q]);
        double u = up[q] + (dt * v + half_dt_dt * a_new[q]);
        vp[q] = v;
        ap[q] = a_new[q];
        up[q] = u;
        bad |= !isfinite(u);
        sp[q] = 0;
    }
    if (bad) atomicOr(flag, 1);
}

// Receivers: one wave64 per station.  In-plane weighted sum over the
// non-zero weights, then Re sum_alpha c_alpha x_alpha e^{i alpha phi}
// with c_0 = 1 and c_alpha = 2 (Station.hpp:53-73, eigen_tools.hpp:79-106).
__global__ __launch_bounds__(64) void k_record(
    int nrec, const int32_t *__restrict__ rec_elem,
    const double *__restrict__ rec_w, const double *__restrict__ rec_phi,
    const int32_t *__restrict__ elem_points,
    const int32_t *__restrict__ elem_nu1,
    const uint32_t *__restrict__ point_off,
    const int32_t *__restrict__ point_nu1, const double *__restrict__ displ,
    double *__restrict__ out) {
    const int r = blockIdx.x;
    if (r >= nrec) return;
    const int lane = threadIdx.x;
    const int e = rec_elem[r];
    const int nu1 = elem_nu1[e];
    const double phi = rec_phi[r];
    double acc[3] = {0, 0, 0};
    for (int alpha = lane; alpha < nu1; alpha += 64) {
        double xr[3] = {0, 0, 0}, xi[3] = {0, 0, 0};
        for (int p = 0; p < NPEM; p++) {
            const double w = rec_w[r * NPEM + p];
            if (fabs(w) < 2.220446049250313e-16) continue;
            const int ip = elem_points[e * NPEM + p];
            if (alpha >= point_nu1[ip]) continue;
            const double *u = displ + (size_t)(point_off[ip] + alpha) * 6;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                xr[c] += u[2 * c] * w;
                xi[c] += u[2 * c + 1] * w;
            }
        }
        double sn, cs;
        sincos((double)alpha * phi, &sn, &cs);
        const double fac = alpha == 0 ? 1.0 : 2.0;
#pragma unroll
        for (int c = 0; c < 3; c++) acc[c] += fac * (xr[c] * cs - xi[c] * sn);
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        double v = acc[c];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        acc[c] = v;
    }
    if (lane == 0) {
#pragma unroll
        for (int c = 0; c < 3; c++) out[r * 3 + c] = acc[c];
    }
}

// ------------------------------------------------------------ launchers
template <typename Real>
void launch_solid_stiff_fr(hipStream_t st, int n_em, const EMItem *em_list,
                           const int32_t *elem_points,
                           const uint8_t *elem_axial, const Real *elem_ops,
                           const Real *elem_mat, const uint32_t *point_off,
                           const int32_t *point_nu1, const double *displ,
                           Real *stiff, const Real *gtab) {
    if (n_em <= 0) return;
    const int grid = (n_em + EMB - 1) / EMB;
    hipLaunchKernelGGL(k_solid_stiff_fr<Real>, dim3(grid), dim3(STIFF_BLOCK),
                       0, st, n_em, em_list, elem_points, elem_axial,
                       elem_ops, elem_mat, point_off, point_nu1, displ,
                       stiff, gtab);
}

template <typename Real>
void launch_moment_source(hipStream_t st, int elem, int nu1c,
                          const Real *pattern, const double *stf, int tstep,
                          const int32_t *elem_points,
                          const uint8_t *elem_axial, const Real *elem_ops,
                          const uint32_t *point_off, const int32_t *point_nu1,
                          Real *stiff, const Real *gtab) {
    if (nu1c <= 0) return;
    const int grid = (nu1c + EMB - 1) / EMB;
    hipLaunchKernelGGL(k_moment_source<Real>, dim3(grid), dim3(STIFF_BLOCK),
                       0, st, elem, nu1c, pattern, stf, tstep, elem_points,
                       elem_axial, elem_ops, point_off, point_nu1, stiff,
                       gtab);
}

template <typename Real>
void launch_point(hipStream_t st, int phases, int64_t n_modes,
                  const int32_t *mode_point, const uint32_t *point_off,
                  const int32_t *point_nr, const uint8_t *point_axial,
                  const double *inv_mass, Real *stiff, double *displ,
                  double *veloc, double *accel, double dt, int *flag) {
    if (n_modes <= 0) return;
    const int block = 256;
    const int64_t grid = (n_modes + block - 1) / block;
    const double half_dt = .5 * dt;
    const double half_dt_dt = .5 * dt * dt;
#define AXS3D_POINT_CASE(A, M, T)                                            \
    case (A) | ((M) << 1) | ((T) << 2):                                      \
        hipLaunchKernelGGL((k_point<Real, A, M, T>), dim3(grid), dim3(block), \
                           0, st, n_modes, mode_point, point_off, point_nr,  \
                           point_axial, inv_mass, stiff, displ, veloc, accel, \
                           dt, half_dt, half_dt_dt, flag);                    \
        break;
    switch (phases) {
        AXS3D_POINT_CASE(1, 0, 0)
        AXS3D_POINT_CASE(0, 1, 0)
        AXS3D_POINT_CASE(0, 0, 1)
        AXS3D_POINT_CASE(1, 1, 1)
        default:
            break;
    }
#undef AXS3D_POINT_CASE
}

void launch_record(hipStream_t st, int nrec, const int32_t *rec_elem,
                   const double *rec_w, const double *rec_phi,
                   const int32_t *elem_points, const int32_t *elem_nu1,
                   const uint32_t *point_off, const int32_t *point_nu1,
                   const double *displ, double *out) {
    if (nrec <= 0) return;
    hipLaunchKernelGGL(k_record, dim3(nrec), dim3(64), 0, st, nrec, rec_elem,
                       rec_w, rec_phi, elem_points, elem_nu1, point_off,
                       point_nu1, displ, out);
}

#define AXS3D_INSTANTIATE(Real)                                               \
    template void launch_solid_stiff_fr<Real>(                                \
        hipStream_t, int, const EMItem *, const int32_t *, const uint8_t *,   \
        const Real *, const Real *, const uint32_t *, const int32_t *,        \
        const double *, Real *, const Real *);                                \
    template void launch_moment_source<Real>(                                 \
        hipStream_t, int, int, const Real *, const double *, int,             \
        const int32_t *, const uint8_t *, const Real *, const uint32_t *,     \
        const int32_t *, Real *, const Real *);                               \
    template void launch_point<Real>(                                         \
        hipStream_t, int, int64_t, const int32_t *, const uint32_t *,         \
        const int32_t *, const uint8_t *, const double *, Real *, double *,   \
        double *, double *, double, int *);

AXS3D_INSTANTIATE(float)
AXS3D_INSTANTIATE(double)

}  // namespace axs3d
