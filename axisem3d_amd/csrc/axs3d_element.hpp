// axs3d_element.hpp -- per-GLL-point gradient and quadrature device code
// shared by the Fourier-path and cardinal-path element kernels.
//
// One thread owns one (element, Fourier order alpha, GLL point); the 25
// threads of an element-mode exchange the displacement (gradient) and the
// quadrature pre-products X, Y through LDS.  Reference (src = SOLVER/src):
//   computeGrad6   src/core/element/grad/GradientQuadrature.hpp:207-231,405-459
//   computeQuad6   GradientQuadrature.hpp:294-324,463-500
#pragma once

#include "axs3d_common.hpp"

namespace axs3d {

constexpr int EMB = 10;          // element-modes per 256-thread block
constexpr int STIFF_BLOCK = 256;

template <typename Real>
struct PointOps {                // inverse-Jacobian entries and 1/s
    Real dsdxii, dsdeta, dzdxii, dzdeta, invs;
};

// Loads the plain (gradient) and integral-factor-scaled (quadrature)
// operators of point ipnt of element e, device layout [e][ELEM_OPS][25].
template <typename Real>
__device__ __forceinline__ void load_point_ops(const Real *__restrict__ ops_all,
                                               int e, int ipnt,
                                               PointOps<Real> &g,
                                               PointOps<Real> &q) {
    const Real *ops = ops_all + (size_t)e * ELEM_OPS * NPEM + ipnt;
    g.dsdxii = ops[OP_DSDXII * NPEM];
    g.dsdeta = ops[OP_DSDETA * NPEM];
    g.dzdxii = ops[OP_DZDXII * NPEM];
    g.dzdeta = ops[OP_DZDETA * NPEM];
    g.invs = ops[OP_INVS * NPEM];
    q.dsdxii = ops[OP_DSDXII_IF * NPEM];
    q.dsdeta = ops[OP_DSDETA_IF * NPEM];
    q.dzdxii = ops[OP_DZDXII_IF * NPEM];
    q.dzdeta = ops[OP_DZDETA_IF * NPEM];
    q.invs = ops[OP_INVS_IF * NPEM];
}

// The four 5-vectors of G this point needs: GT_xi(i,.), G(.,j) for the
// gradient and G_xi(i,.), GT(.,j) for the quadrature.  sG = [G_GLL, G_GLJ].
template <typename Real>
struct PointG {
    Real gxi_col[NPED], g_col[NPED], gxi_row[NPED], g_row[NPED];
};

template <typename Real>
__device__ __forceinline__ void load_point_g(const Real (*sG)[NPEM],
                                             bool axial, int i, int j,
                                             PointG<Real> &G) {
    const Real *Gxi = sG[axial ? 1 : 0];
#pragma unroll
    for (int k = 0; k < NPED; k++) {
        G.gxi_col[k] = Gxi[k * NPED + i];   // GT_xi(i,k) = G_xi(k,i)
        G.g_col[k] = sG[0][k * NPED + j];   // G(k,j)
        G.gxi_row[k] = Gxi[i * NPED + k];   // G_xi(i,k)
        G.g_row[k] = sG[0][j * NPED + k];   // GT(k,j) = G(j,k)
    }
}

// computeGrad6 at point (i, j) of one element-mode whose displacement
// tile sU[3][25] is in LDS; u = this point's own displacement.  Voigt
// strain [ss, pp, zz, pz, sz, sp] with engineering shears.
template <typename Real>
__device__ __forceinline__ void grad6_point(const cpx<Real> (*sU)[NPEM],
                                            const cpx<Real> (&u)[3],
                                            Real ralpha, bool axial_row0,
                                            int i, int j,
                                            const PointG<Real> &G,
                                            const PointOps<Real> &g,
                                            cpx<Real> (&eps)[6]) {
    // higher-order terms (computeGrad_dSdZ, GradientQuadrature.hpp:405-418)
    cpx<Real> dS[3], dZ[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        cpx<Real> x = {0, 0}, y = {0, 0};
#pragma unroll
        for (int k = 0; k < NPED; k++) {
            x = cfma(sU[c][k * NPED + j], G.gxi_col[k], x);
            y = cfma(sU[c][i * NPED + k], G.g_col[k], y);
        }
        dS[c] = cadd(cmul(x, g.dzdeta), cmul(y, g.dzdxii));
        dZ[c] = cadd(cmul(x, g.dsdeta), cmul(y, g.dsdxii));
    }
    eps[0] = dS[0];
    eps[4] = cadd(dZ[0], dS[2]);
    eps[5] = dS[1];
    eps[3] = dZ[1];
    eps[2] = dZ[2];
    // lower-order terms: (i alpha u0 - u1, i alpha u1 + u0, i alpha u2) / s
    if (!axial_row0) {
        const cpx<Real> a0 = csub(ialpha(ralpha, u[0]), u[1]);
        const cpx<Real> a1 = cadd(ialpha(ralpha, u[1]), u[0]);
        const cpx<Real> a2 = ialpha(ralpha, u[2]);
        eps[5] = cfma(a0, g.invs, eps[5]);
        eps[1] = cmul(a1, g.invs);
        eps[3] = cfma(a2, g.invs, eps[3]);
    } else {
        // axis: d/ds of the lower-order numerators (L'Hospital),
        // computeGrad_1overS_setTo/addTo row 0, :426-437
        cpx<Real> xa[3] = {{0, 0}, {0, 0}, {0, 0}};
        cpx<Real> ya[3] = {{0, 0}, {0, 0}, {0, 0}};
#pragma unroll
        for (int k = 0; k < NPED; k++) {
            const cpx<Real> uc0 = sU[0][k * NPED + j];
            const cpx<Real> uc1 = sU[1][k * NPED + j];
            const cpx<Real> uc2 = sU[2][k * NPED + j];
            const cpx<Real> ur0 = sU[0][k];
            const cpx<Real> ur1 = sU[1][k];
            const cpx<Real> ur2 = sU[2][k];
            xa[0] = cfma(csub(ialpha(ralpha, uc0), uc1), G.gxi_col[k], xa[0]);
            xa[1] = cfma(cadd(ialpha(ralpha, uc1), uc0), G.gxi_col[k], xa[1]);
            xa[2] = cfma(ialpha(ralpha, uc2), G.gxi_col[k], xa[2]);
            ya[0] = cfma(csub(ialpha(ralpha, ur0), ur1), G.g_col[k], ya[0]);
            ya[1] = cfma(cadd(ialpha(ralpha, ur1), ur0), G.g_col[k], ya[1]);
            ya[2] = cfma(ialpha(ralpha, ur2), G.g_col[k], ya[2]);
        }
        eps[5] = cadd(eps[5], cadd(cmul(xa[0], g.dzdeta), cmul(ya[0], g.dzdxii)));
        eps[1] = cadd(cmul(xa[1], g.dzdeta), cmul(ya[1], g.dzdxii));
        eps[3] = cadd(eps[3], cadd(cmul(xa[2], g.dzdeta), cmul(ya[2], g.dzdxii)));
    }
}

// computeGrad9 at point (i, j) (GradientQuadrature.hpp:181-201): the full
// gradient g[3 c + d] = d u_c / d x_d, x = (s, phi, z), for the particle
// relabelling path; the same contractions as grad6_point with the s-z and
// phi pairs kept apart.
template <typename Real>
__device__ __forceinline__ void grad9_point(const cpx<Real> (*sU)[NPEM],
                                            const cpx<Real> (&u)[3],
                                            Real ralpha, bool axial_row0,
                                            int i, int j,
                                            const PointG<Real> &G,
                                            const PointOps<Real> &g,
                                            cpx<Real> (&g9)[9]) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
        cpx<Real> x = {0, 0}, y = {0, 0};
#pragma unroll
        for (int k = 0; k < NPED; k++) {
            x = cfma(sU[c][k * NPED + j], G.gxi_col[k], x);
            y = cfma(sU[c][i * NPED + k], G.g_col[k], y);
        }
        g9[3 * c] = cadd(cmul(x, g.dzdeta), cmul(y, g.dzdxii));
        g9[3 * c + 2] = cadd(cmul(x, g.dsdeta), cmul(y, g.dsdxii));
    }
    if (!axial_row0) {
        g9[1] = cmul(csub(ialpha(ralpha, u[0]), u[1]), g.invs);
        g9[4] = cmul(cadd(ialpha(ralpha, u[1]), u[0]), g.invs);
        g9[7] = cmul(ialpha(ralpha, u[2]), g.invs);
    } else {
        cpx<Real> xa[3] = {{0, 0}, {0, 0}, {0, 0}};
        cpx<Real> ya[3] = {{0, 0}, {0, 0}, {0, 0}};
#pragma unroll
        for (int k = 0; k < NPED; k++) {
            const cpx<Real> uc0 = sU[0][k * NPED + j];
            const cpx<Real> uc1 = sU[1][k * NPED + j];
            const cpx<Real> uc2 = sU[2][k * NPED + j];
            const cpx<Real> ur0 = sU[0][k];
            const cpx<Real> ur1 = sU[1][k];
            const cpx<Real> ur2 = sU[2][k];
            xa[0] = cfma(csub(ialpha(ralpha, uc0), uc1), G.gxi_col[k], xa[0]);
            xa[1] = cfma(cadd(ialpha(ralpha, uc1), uc0), G.gxi_col[k], xa[1]);
            xa[2] = cfma(ialpha(ralpha, uc2), G.gxi_col[k], xa[2]);
            ya[0] = cfma(csub(ialpha(ralpha, ur0), ur1), G.g_col[k], ya[0]);
            ya[1] = cfma(cadd(ialpha(ralpha, ur1), ur0), G.g_col[k], ya[1]);
            ya[2] = cfma(ialpha(ralpha, ur2), G.g_col[k], ya[2]);
        }
#pragma unroll
        for (int c = 0; c < 3; c++)
            g9[3 * c + 1] = cadd(cmul(xa[c], g.dzdeta), cmul(ya[c], g.dzdxii));
    }
}

// A gradient-shaped tensor g -> Q g Q^T, Q the rotation by the colatitude
// in the (s, z) plane (CoordTransformSpherical::transformSPZ_RTZ9,
// CoordTransformSpherical.hpp:102-141); with -sin it is the way back
// (transformRTZ_SPZ9, :144-184).
template <typename T, typename Real>
__device__ __forceinline__ void rot9(T (&g)[9], Real s1, Real c1, Real s2,
                                     Real c2) {
    const T s08 = g[0] + g[8], d08 = g[0] - g[8];
    const T s26 = g[2] + g[6], d26 = g[2] - g[6];
    g[0] = (s08 + d08 * c2 - s26 * s2) * (Real)0.5;
    g[2] = (d26 + s26 * c2 + d08 * s2) * (Real)0.5;
    g[6] = g[2] - d26;
    g[8] = s08 - g[0];
    const T g1 = g[1], g3 = g[3];
    g[1] = g1 * c1 - g[7] * s1;
    g[7] = g[7] * c1 + g1 * s1;
    g[3] = g3 * c1 - g[5] * s1;
    g[5] = g[5] * c1 + g3 * s1;
}

// Particle relabelling (PRT::sphericalToUndulated6, PRT.hpp:327-340): the
// undulated Voigt strain (engineering shears) of the spherical gradient
// g with the inverse Jacobian X0..X3 at x[0], x[NPEM], ...
template <typename T, typename Real>
__device__ __forceinline__ void prt_sph_to_und6(const T (&g)[9],
                                                const Real *x, T (&e)[6]) {
    const Real X0 = x[0], X1 = x[NPEM], X2 = x[2 * NPEM], X3 = x[3 * NPEM];
    e[0] = g[0] * X0 + g[2] * X1;
    e[1] = g[4] * X0 + g[5] * X2;
    e[2] = g[8] * X3;
    e[3] = g[7] * X0 + g[8] * X2 + g[5] * X3;
    e[4] = g[6] * X0 + g[8] * X1 + g[2] * X3;
    e[5] = (g[3] + g[1]) * X0 + g[5] * X1 + g[2] * X2;
}

// PRT::undulatedToSpherical6 (PRT.hpp:374-389): the spherical tensor of
// the undulated stress s with the scaled inverse Jacobian X_k |J| at
// xj[0], xj[NPEM], ...; the transpose of prt_sph_to_und6.
template <typename T, typename Real>
__device__ __forceinline__ void prt_und_to_sph9(const T (&s)[6],
                                                const Real *xj, T (&f)[9]) {
    const Real Y0 = xj[0], Y1 = xj[NPEM], Y2 = xj[2 * NPEM];
    const Real Y3 = xj[3 * NPEM];
    f[0] = s[0] * Y0;
    f[1] = s[5] * Y0;
    f[2] = s[0] * Y1 + s[4] * Y3 + s[5] * Y2;
    f[3] = f[1];
    f[4] = s[1] * Y0;
    f[5] = s[1] * Y2 + s[3] * Y3 + s[5] * Y1;
    f[6] = s[4] * Y0;
    f[7] = s[3] * Y0;
    f[8] = s[2] * Y3 + s[3] * Y2 + s[4] * Y1;
}

// Isotropic Hooke's law (Isotropic::strainToStress, src/core/element/
// material/elastic/anisotropy/Isotropic.hpp:155-172).
template <typename Real>
__device__ __forceinline__ void iso_hooke(const cpx<Real> (&eps)[6], Real lam,
                                          Real mu, cpx<Real> (&sig)[6]) {
    const Real mu2 = mu * (Real)2;
    const cpx<Real> tr = cmul(cadd(cadd(eps[0], eps[1]), eps[2]), lam);
    sig[0] = cfma(eps[0], mu2, tr);
    sig[1] = cfma(eps[1], mu2, tr);
    sig[2] = cfma(eps[2], mu2, tr);
    sig[3] = cmul(eps[3], mu);
    sig[4] = cmul(eps[4], mu);
    sig[5] = cmul(eps[5], mu);
}

// Transversely isotropic Hooke's law in RTZ, radial symmetry axis
// (TransverselyIsotropic::strainToStress, src/core/element/material/
// elastic/anisotropy/TransverselyIsotropic.hpp:161-180), in the
// reference's operation order.  T = cpx<Real> or Real.
template <typename T, typename Real>
__device__ __forceinline__ void ti_hooke(const T (&e)[6], Real A, Real C,
                                         Real F, Real L, Real N, T (&s)[6]) {
    const Real mN2 = N * (Real)-2;
    const T s3 = e[0] + e[1];
    const T s4 = s3 * A + e[2] * F;
    s[0] = s4 + e[1] * mN2;
    s[1] = s4 + e[0] * mN2;
    s[2] = s3 * F + e[2] * C;
    s[3] = e[3] * L;
    s[4] = e[4] * L;
    s[5] = e[5] * N;
}

// Full anisotropy (Anisotropic::strainToStress, Anisotropic.hpp:203-274):
// s_i = sum_j C_ij e_j, j in order; c = C11 C12 .. C16 C22 .. C66 read
// through idx(k) (the k-th of the 21).
template <typename T, typename Real, typename Get>
__device__ __forceinline__ void aniso_hooke(const T (&e)[6], Get c,
                                            T (&s)[6]) {
    // row i, column j -> index in the 21 upper-triangle entries
    constexpr int ij[6][6] = {{0, 1, 2, 3, 4, 5},      {1, 6, 7, 8, 9, 10},
                              {2, 7, 11, 12, 13, 14},  {3, 8, 12, 15, 16, 17},
                              {4, 9, 13, 16, 18, 19},  {5, 10, 14, 17, 19, 20}};
#pragma unroll
    for (int i = 0; i < 6; i++) {
        T acc = e[0] * (Real)c(ij[i][0]);
#pragma unroll
        for (int j = 1; j < 6; j++) acc = acc + e[j] * (Real)c(ij[i][j]);
        s[i] = acc;
    }
}

// Densities of the bilinear form psi . sigma_theta(u) with respect to the
// parameters theta of material type MT (0: lambda, mu; 1: A, C, F, L, N;
// 2: the 21 Cij in the order of aniso_hooke), for Voigt strains with
// engineering shears: d(psi . C u)/d theta.  pr(x, y) is the pairing of
// two strain components (x y for slices, Re conj(x) y for orders);
// add(c, v) receives the density of parameter c.
template <int MT, typename T, typename Pair, typename Add>
__device__ __forceinline__ void mat_density(const T (&p)[6], const T (&u)[6],
                                            Pair pr, Add add) {
    if constexpr (MT == 0) {
        const T tp = p[0] + p[1] + p[2], tu = u[0] + u[1] + u[2];
        add(0, pr(tp, tu));
        add(1, 2. * (pr(p[0], u[0]) + pr(p[1], u[1]) + pr(p[2], u[2])) +
                   pr(p[3], u[3]) + pr(p[4], u[4]) + pr(p[5], u[5]));
    } else if constexpr (MT == 1) {
        const T p01 = p[0] + p[1], u01 = u[0] + u[1];
        add(0, pr(p01, u01));
        add(1, pr(p[2], u[2]));
        add(2, pr(p01, u[2]) + pr(p[2], u01));
        add(3, pr(p[3], u[3]) + pr(p[4], u[4]));
        add(4, -2. * (pr(p[0], u[1]) + pr(p[1], u[0])) + pr(p[5], u[5]));
    } else {
        int c = 0;
#pragma unroll
        for (int i = 0; i < 6; i++)
#pragma unroll
            for (int j = i; j < 6; j++, c++)
                add(c, i == j ? pr(p[i], u[i])
                              : pr(p[i], u[j]) + pr(p[j], u[i]));
    }
}

// (s, phi, z) -> (R, T, Z) for a Voigt strain with engineering shears and
// back for a stress (CoordTransformSpherical::transformSPZ_RTZ6 /
// transformRTZ_SPZ6, src/core/element/crd/CoordTransformSpherical.hpp:
// 186-249); sin, cos of theta and 2 theta of this GLL point.
template <typename T, typename Real>
__device__ __forceinline__ void rot_spz_rtz6(T (&e)[6], Real s1, Real c1,
                                             Real s2, Real c2) {
    const T s02 = e[0] + e[2];
    const T d02 = e[0] - e[2];
    e[0] = (s02 + d02 * c2 - e[4] * s2) * (Real)0.5;
    e[2] = s02 - e[0];
    e[4] = e[4] * c2 + d02 * s2;
    const T e3 = e[3];
    e[3] = e3 * c1 + e[5] * s1;
    e[5] = e[5] * c1 - e3 * s1;
}

template <typename T, typename Real>
__device__ __forceinline__ void rot_rtz_spz6(T (&s)[6], Real s1, Real c1,
                                             Real s2, Real c2) {
    const T s02 = s[0] + s[2];
    const T d02 = s[0] - s[2];
    s[0] = (s02 + d02 * c2) * (Real)0.5 + s[4] * s2;
    s[2] = s02 - s[0];
    s[4] = s[4] * c2 - d02 * s2 * (Real)0.5;
    const T s3 = s[3];
    s[3] = s3 * c1 - s[5] * s1;
    s[5] = s[5] * c1 + s3 * s1;
}

// X = dFdS*dzdeta + dFdZ*dsdeta, Y = dFdS*dzdxii + dFdZ*dsdxii
// (computeQuad_dSdZ, GradientQuadrature.hpp:471-474).
template <typename Real>
__device__ __forceinline__ void quad_xy(cpx<Real> dfds, cpx<Real> dfdz,
                                        const PointOps<Real> &q, cpx<Real> &X,
                                        cpx<Real> &Y) {
    X = cadd(cmul(dfds, q.dzdeta), cmul(dfdz, q.dsdeta));
    Y = cadd(cmul(dfds, q.dzdxii), cmul(dfdz, q.dsdxii));
}

// Quadrature at this point, first half: forms X, Y from the six stress
// components (with the axial L'Hospital corrections of
// computeQuad_1overS_addTo, :493-497, folded into row 0 of X and Y), puts
// them in the LDS tiles sX, sY, and returns the local 1/s term L
// (GradientQuadrature.hpp:311-318).
template <typename Real>
__device__ __forceinline__ void quad_prep_core(
    const cpx<Real> (&X)[3], const cpx<Real> (&Y)[3], const cpx<Real> (&b)[3],
    bool axial_row0, const PointOps<Real> &q, cpx<Real> (*sX)[NPEM],
    cpx<Real> (*sY)[NPEM], int ipnt, cpx<Real> (&L)[3]);

template <typename Real>
__device__ __forceinline__ void quad6_prep(
    const cpx<Real> (&sig)[6], Real alpha, bool axial_row0,
    const PointOps<Real> &q, cpx<Real> (*sX)[NPEM], cpx<Real> (*sY)[NPEM],
    int ipnt, cpx<Real> (&L)[3]) {
    cpx<Real> X[3], Y[3];
    // (dFdS, dFdZ) per output component: (0,4), (5,3), (4,2)
    quad_xy(sig[0], sig[4], q, X[0], Y[0]);
    quad_xy(sig[5], sig[3], q, X[1], Y[1]);
    quad_xy(sig[4], sig[2], q, X[2], Y[2]);
    cpx<Real> b[3];
    b[0] = csub(ialpha(alpha, sig[5]), sig[1]);
    b[1] = cadd(ialpha(alpha, sig[1]), sig[5]);
    b[2] = ialpha(alpha, sig[3]);
    quad_prep_core<Real>(X, Y, b, axial_row0, q, sX, sY, ipnt, L);
}

// The same first half for a 9-component tensor f (computeQuad9,
// GradientQuadrature.hpp:258-288): (dFdS, dFdZ) = (f0, f2), (f3, f5),
// (f6, f8); lower-order numerators i alpha f1 - f4, i alpha f4 + f1,
// i alpha f7.
template <typename Real>
__device__ __forceinline__ void quad9_prep(
    const cpx<Real> (&f)[9], Real alpha, bool axial_row0,
    const PointOps<Real> &q, cpx<Real> (*sX)[NPEM], cpx<Real> (*sY)[NPEM],
    int ipnt, cpx<Real> (&L)[3]) {
    cpx<Real> X[3], Y[3];
#pragma unroll
    for (int c = 0; c < 3; c++) quad_xy(f[3 * c], f[3 * c + 2], q, X[c], Y[c]);
    cpx<Real> b[3];
    b[0] = csub(ialpha(alpha, f[1]), f[4]);
    b[1] = cadd(ialpha(alpha, f[4]), f[1]);
    b[2] = ialpha(alpha, f[7]);
    quad_prep_core<Real>(X, Y, b, axial_row0, q, sX, sY, ipnt, L);
}

template <typename Real>
__device__ __forceinline__ void quad_prep_core(
    const cpx<Real> (&Xin)[3], const cpx<Real> (&Yin)[3],
    const cpx<Real> (&b)[3], bool axial_row0, const PointOps<Real> &q,
    cpx<Real> (*sX)[NPEM], cpx<Real> (*sY)[NPEM], int ipnt,
    cpx<Real> (&L)[3]) {
    cpx<Real> X[3] = {Xin[0], Xin[1], Xin[2]};
    cpx<Real> Y[3] = {Yin[0], Yin[1], Yin[2]};
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

// Second half, once the element-mode's X, Y tiles are complete:
// F = G_xi X + Y G^T + L at point (i, j).
template <typename Real>
__device__ __forceinline__ void quad6_finish(const cpx<Real> (*sX)[NPEM],
                                             const cpx<Real> (*sY)[NPEM],
                                             int i, int j,
                                             const PointG<Real> &G,
                                             const cpx<Real> (&L)[3],
                                             cpx<Real> (&F)[3]) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
        cpx<Real> a = {0, 0}, bsum = {0, 0};
#pragma unroll
        for (int k = 0; k < NPED; k++) {
            a = cfma(sX[c][k * NPED + j], G.gxi_row[k], a);
            bsum = cfma(sY[c][i * NPED + k], G.g_row[k], bsum);
        }
        F[c] = cadd(cadd(a, bsum), L[c]);
    }
}

// Both halves with a block barrier between them: every thread of the block
// must call it.
template <typename Real>
__device__ __forceinline__ void quad6_point(
    const cpx<Real> (&sig)[6], Real alpha, bool axial_row0,
    const PointOps<Real> &q, cpx<Real> (*sX)[NPEM], cpx<Real> (*sY)[NPEM],
    int ipnt, int i, int j, const PointG<Real> &G, bool active,
    cpx<Real> (&F)[3]) {
    cpx<Real> L[3];
    if (active) quad6_prep(sig, alpha, axial_row0, q, sX, sY, ipnt, L);
    __syncthreads();
    if (active) quad6_finish<Real>(sX, sY, i, j, G, L, F);
}

}  // namespace axs3d
