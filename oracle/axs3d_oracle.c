/*
 * axs3d_oracle.c -- FP64 CPU restatement of AxiSEM3D's per-time-step
 * stiffness path.  TEST INFRASTRUCTURE ONLY: this library is the checker
 * that the HIP path is compared against, and the CPU baseline timed beside
 * it.  It is never linked into, loaded by or called from the product path.
 *
 * Every function follows the reference file:line named in its comment.
 * Paths are relative to the reference repository root; src/ abbreviates
 * SOLVER/src/.  All arithmetic is double precision ("Real = double", the
 * reference built with -Ddouble=true).  Operation order inside each stage
 * follows the reference so that FP64 results agree to rounding.
 *
 * Layouts (shared with the HIP library's C ABI):
 *  - point fields: complex double [point_offset[p] + alpha][3];
 *  - element operands gq[e][6][25] = dsdxii, dsdeta, dzdxii, dzdeta,
 *    inv_s, ifact (the inverse-Jacobian entries keep the reference's
 *    Jacobian-entry names, src/preloop/se_model/quad/Quad.hpp:166-178);
 *  - element-local fields [alpha][dim][25], ipnt = ipol * 5 + jpol.
 */
#include <complex.h>
#include <math.h>
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "axs3d_oracle.h"

#define P 5
#define N 25

typedef double complex cplx;

/* G tables: G[i][j] = l_i'(x_j), set by oracle_set_gmat
 * (GradientQuadrature::setGMat, src/core/element/grad/
 * GradientQuadrature.hpp:510-516). */
static double sG_GLL[P][P], sG_GLJ[P][P];

void oracle_set_gmat(const double *g_gll, const double *g_glj) {
    for (int i = 0; i < P; i++)
        for (int j = 0; j < P; j++) {
            sG_GLL[i][j] = g_gll[i * P + j];
            sG_GLJ[i][j] = g_glj[i * P + j];
        }
}

/* ------------------------------------------------------------------ */
/* GradientQuadrature restatement (GradientQuadrature.hpp:163-500)     */
/* ------------------------------------------------------------------ */
typedef struct {
    const double *dsdxii, *dsdeta, *dzdxii, *dzdeta, *inv_s, *ifact;
    int axial;
} gq_t;

static gq_t gq_of(const double *gq_e, int axial) {
    gq_t g = {gq_e, gq_e + N, gq_e + 2 * N, gq_e + 3 * N, gq_e + 4 * N,
              gq_e + 5 * N, axial};
    return g;
}

/* computeGrad_dSdZ, GradientQuadrature.hpp:405-418:
 * X = GT_xi * F, Y = F * G_GLL; dFdS = X.*dzdeta + Y.*dzdxii;
 * dFdZ = X.*dsdeta + Y.*dsdxii. */
static void grad_dsdz(const gq_t *g, const cplx *F, cplx *dFdS, cplx *dFdZ) {
    const double (*Gxi)[P] = g->axial ? sG_GLJ : sG_GLL;
    for (int i = 0; i < P; i++)
        for (int j = 0; j < P; j++) {
            cplx x = 0, y = 0;
            for (int k = 0; k < P; k++) {
                x += Gxi[k][i] * F[k * P + j];     /* GT(i,k) = G(k,i) */
                y += F[i * P + k] * sG_GLL[k][j];
            }
            int p = i * P + j;
            dFdS[p] = x * g->dzdeta[p] + y * g->dzdxii[p];
            dFdZ[p] = x * g->dsdeta[p] + y * g->dsdxii[p];
        }
}

/* computeGrad_1overS_setTo/addTo, GradientQuadrature.hpp:422-459:
 * AoS (=|+=) A .* inv_s; axial row 0 gets the L'Hospital derivative. */
static void grad_1overs(const gq_t *g, const cplx *A, cplx *AoS, int add) {
    const double (*Gxi)[P] = g->axial ? sG_GLJ : sG_GLL;
    for (int p = 0; p < N; p++) {
        cplx v = A[p] * g->inv_s[p];
        AoS[p] = add ? AoS[p] + v : v;
    }
    if (g->axial) {
        for (int j = 0; j < P; j++) {
            cplx x = 0, y = 0;
            for (int k = 0; k < P; k++) {
                x += Gxi[k][0] * A[k * P + j];
                y += A[0 * P + k] * sG_GLL[k][j];
            }
            cplx v = x * g->dzdeta[j] + y * g->dzdxii[j];
            AoS[j] = add ? AoS[j] + v : v;
        }
    }
}

/* computeGrad6, GradientQuadrature.hpp:207-231.  u[3][25] -> e[6][25],
 * Voigt [ss, pp, zz, pz, sz, sp] with engineering shears. */
void oracle_grad6(const double *gq_e, int axial, int alpha, const cplx *u,
                  cplx *e) {
    gq_t g = gq_of(gq_e, axial);
    grad_dsdz(&g, u + 0 * N, e + 0 * N, e + 4 * N);
    grad_dsdz(&g, u + 1 * N, e + 5 * N, e + 3 * N);
    grad_dsdz(&g, u + 2 * N, e + 1 * N, e + 2 * N);
    for (int p = 0; p < N; p++) e[4 * N + p] += e[1 * N + p];
    cplx sA[3][N];
    cplx ai = (double)alpha * I;
    for (int p = 0; p < N; p++) {
        sA[0][p] = ai * u[0 * N + p] - u[1 * N + p];
        sA[1][p] = ai * u[1 * N + p] + u[0 * N + p];
        sA[2][p] = ai * u[2 * N + p];
    }
    grad_1overs(&g, sA[0], e + 5 * N, 1);
    grad_1overs(&g, sA[1], e + 1 * N, 0);
    grad_1overs(&g, sA[2], e + 3 * N, 1);
}

/* computeGrad9, GradientQuadrature.hpp:181-201.  u[3][25] -> g[9][25],
 * g[3 i + j] = d u_i / d x_j for x = (s, phi, z). */
void oracle_grad9(const double *gq_e, int axial, int alpha, const cplx *u,
                  cplx *g9) {
    gq_t g = gq_of(gq_e, axial);
    for (int c = 0; c < 3; c++)
        grad_dsdz(&g, u + c * N, g9 + (3 * c) * N, g9 + (3 * c + 2) * N);
    cplx sA[3][N];
    cplx ai = (double)alpha * I;
    for (int p = 0; p < N; p++) {
        sA[0][p] = ai * u[0 * N + p] - u[1 * N + p];
        sA[1][p] = ai * u[1 * N + p] + u[0 * N + p];
        sA[2][p] = ai * u[2 * N + p];
    }
    for (int c = 0; c < 3; c++) grad_1overs(&g, sA[c], g9 + (3 * c + 1) * N, 0);
}

/* computeGrad3, GradientQuadrature.hpp:163-175: p[25] -> p_i[3][25] =
 * (dp/ds, i alpha p / s, dp/dz). */
void oracle_grad3(const double *gq_e, int axial, int alpha, const cplx *p,
                  cplx *pi) {
    gq_t g = gq_of(gq_e, axial);
    grad_dsdz(&g, p, pi + 0 * N, pi + 2 * N);
    cplx sA[N];
    cplx ai = (double)alpha * I;
    for (int k = 0; k < N; k++) sA[k] = ai * p[k];
    grad_1overs(&g, sA, pi + 1 * N, 0);
}

/* computeQuad_dSdZ, GradientQuadrature.hpp:463-478:
 * X = dFdS.*dzdeta + dFdZ.*dsdeta; Y = dFdS.*dzdxii + dFdZ.*dsdxii;
 * F = G_xi * X + Y * GT_GLL. */
static void quad_dsdz(int axial, const cplx *dFdS, const cplx *dFdZ,
                      cplx *F, const double *dsdxii, const double *dsdeta,
                      const double *dzdxii, const double *dzdeta) {
    const double (*Gxi)[P] = axial ? sG_GLJ : sG_GLL;
    cplx X[N], Y[N];
    for (int p = 0; p < N; p++) {
        X[p] = dFdS[p] * dzdeta[p] + dFdZ[p] * dsdeta[p];
        Y[p] = dFdS[p] * dzdxii[p] + dFdZ[p] * dsdxii[p];
    }
    for (int i = 0; i < P; i++)
        for (int j = 0; j < P; j++) {
            cplx a = 0, b = 0;
            for (int k = 0; k < P; k++) {
                a += Gxi[i][k] * X[k * P + j];
                b += Y[i * P + k] * sG_GLL[j][k];   /* GT(k,j) = G(j,k) */
            }
            F[i * P + j] = a + b;
        }
}

/* computeQuad_1overS_addTo, GradientQuadrature.hpp:482-500 */
static void quad_1overs(int axial, const cplx *A, cplx *AoS,
                        const double *oneovers, const double *dzdxii,
                        const double *dzdeta) {
    const double (*Gxi)[P] = axial ? sG_GLJ : sG_GLL;
    for (int p = 0; p < N; p++) AoS[p] -= A[p] * oneovers[p];
    if (axial) {
        cplx X0[P], Y0[P];
        for (int j = 0; j < P; j++) {
            X0[j] = A[j] * dzdeta[j];
            Y0[j] = A[j] * dzdxii[j];
        }
        for (int i = 0; i < P; i++)
            for (int j = 0; j < P; j++) AoS[i * P + j] -= Gxi[i][0] * X0[j];
        for (int j = 0; j < P; j++) {
            cplx s = 0;
            for (int k = 0; k < P; k++) s += Y0[k] * sG_GLL[j][k];
            AoS[0 * P + j] -= s;
        }
    }
}

static void quad6_ops(int axial, int alpha, const cplx *sij, cplx *fi,
                      const double *ds_xi, const double *ds_eta,
                      const double *dz_xi, const double *dz_eta,
                      const double *ovs) {
    quad_dsdz(axial, sij + 0 * N, sij + 4 * N, fi + 0 * N, ds_xi, ds_eta,
              dz_xi, dz_eta);
    quad_dsdz(axial, sij + 5 * N, sij + 3 * N, fi + 1 * N, ds_xi, ds_eta,
              dz_xi, dz_eta);
    quad_dsdz(axial, sij + 4 * N, sij + 2 * N, fi + 2 * N, ds_xi, ds_eta,
              dz_xi, dz_eta);
    cplx sA[3][N];
    cplx ai = (double)alpha * I;
    for (int p = 0; p < N; p++) {
        sA[0][p] = ai * sij[5 * N + p] - sij[1 * N + p];
        sA[1][p] = ai * sij[1 * N + p] + sij[5 * N + p];
        sA[2][p] = ai * sij[3 * N + p];
    }
    for (int c = 0; c < 3; c++)
        quad_1overs(axial, sA[c], fi + c * N, ovs, dz_xi, dz_eta);
}

/* computeQuad6, GradientQuadrature.hpp:294-324: operators scaled by the
 * integral factor (precomputed in FP64, GradientQuadrature.hpp:127-137). */
void oracle_quad6(const double *gq_e, int axial, int alpha, const cplx *sij,
                  cplx *fi) {
    gq_t g = gq_of(gq_e, axial);
    double a[N], b[N], c[N], d[N], o[N];
    for (int p = 0; p < N; p++) {
        a[p] = g.dsdxii[p] * g.ifact[p];
        b[p] = g.dsdeta[p] * g.ifact[p];
        c[p] = g.dzdxii[p] * g.ifact[p];
        d[p] = g.dzdeta[p] * g.ifact[p];
        o[p] = g.inv_s[p] * g.ifact[p];
    }
    quad6_ops(axial, alpha, sij, fi, a, b, c, d, o);
}

/* computeQuad9, GradientQuadrature.hpp:258-288: f[9][25] -> fi[3][25],
 * operators scaled by the integral factor. */
static void quad9_ops(int axial, int alpha, const cplx *f, cplx *fi,
                      const double *a, const double *b, const double *c,
                      const double *d, const double *o) {
    for (int k = 0; k < 3; k++)
        quad_dsdz(axial, f + (3 * k) * N, f + (3 * k + 2) * N, fi + k * N,
                  a, b, c, d);
    cplx sA[3][N];
    cplx ai = (double)alpha * I;
    for (int p = 0; p < N; p++) {
        sA[0][p] = ai * f[1 * N + p] - f[4 * N + p];
        sA[1][p] = ai * f[4 * N + p] + f[1 * N + p];
        sA[2][p] = ai * f[7 * N + p];
    }
    for (int k = 0; k < 3; k++) quad_1overs(axial, sA[k], fi + k * N, o, c, d);
}

void oracle_quad9(const double *gq_e, int axial, int alpha, const cplx *f,
                  cplx *fi) {
    gq_t g = gq_of(gq_e, axial);
    double a[N], b[N], c[N], d[N], o[N];
    for (int p = 0; p < N; p++) {
        a[p] = g.dsdxii[p] * g.ifact[p];
        b[p] = g.dsdeta[p] * g.ifact[p];
        c[p] = g.dzdxii[p] * g.ifact[p];
        d[p] = g.dzdeta[p] * g.ifact[p];
        o[p] = g.inv_s[p] * g.ifact[p];
    }
    quad9_ops(axial, alpha, f, fi, a, b, c, d, o);
}

/* computeQuad9_NoIntegration, GradientQuadrature.hpp:330-360 */
static void quad9_noint(const double *gq_e, int axial, int alpha,
                        const cplx *f, cplx *fi) {
    gq_t g = gq_of(gq_e, axial);
    quad9_ops(axial, alpha, f, fi, g.dsdxii, g.dsdeta, g.dzdxii, g.dzdeta,
              g.inv_s);
}

/* computeQuad3, GradientQuadrature.hpp:234-252: q_i[3][25] -> q[25]. */
void oracle_quad3(const double *gq_e, int axial, int alpha, const cplx *qi,
                  cplx *q) {
    gq_t g = gq_of(gq_e, axial);
    double a[N], b[N], c[N], d[N], o[N];
    for (int p = 0; p < N; p++) {
        a[p] = g.dsdxii[p] * g.ifact[p];
        b[p] = g.dsdeta[p] * g.ifact[p];
        c[p] = g.dzdxii[p] * g.ifact[p];
        d[p] = g.dzdeta[p] * g.ifact[p];
        o[p] = g.inv_s[p] * g.ifact[p];
    }
    quad_dsdz(axial, qi + 0 * N, qi + 2 * N, q, a, b, c, d);
    cplx sA[N];
    cplx ai = (double)alpha * I;
    for (int p = 0; p < N; p++) sA[p] = ai * qi[1 * N + p];
    quad_1overs(axial, sA, q, o, c, d);
}

/* computeQuad6_NoIntegration, GradientQuadrature.hpp:363-390 */
void oracle_quad6_noint(const double *gq_e, int axial, int alpha,
                        const cplx *sij, cplx *fi) {
    gq_t g = gq_of(gq_e, axial);
    quad6_ops(axial, alpha, sij, fi, g.dsdxii, g.dsdeta, g.dzdxii, g.dzdeta,
              g.inv_s);
}

/* Isotropic::strainToStress, src/core/element/material/elastic/anisotropy/
 * Isotropic.hpp:155-172 (block 3 holds lambda * tr(e) first). */
void oracle_iso_stress(const double *lam, const double *mu, const cplx *e,
                       cplx *s) {
    for (int p = 0; p < N; p++) {
        cplx tr = (e[0 * N + p] + e[1 * N + p] + e[2 * N + p]) * lam[p];
        double mu2 = mu[p] * 2.;
        s[0 * N + p] = tr + e[0 * N + p] * mu2;
        s[1 * N + p] = tr + e[1 * N + p] * mu2;
        s[2 * N + p] = tr + e[2 * N + p] * mu2;
        s[3 * N + p] = e[3 * N + p] * mu[p];
        s[4 * N + p] = e[4 * N + p] * mu[p];
        s[5 * N + p] = e[5 * N + p] * mu[p];
    }
}


/* TransverselyIsotropic::strainToStress, src/core/element/material/
 * elastic/anisotropy/TransverselyIsotropic.hpp:161-180, RTZ Voigt with
 * engineering shears; par = [A, C, F, L, N][25]; T = cplx (Fourier) or
 * double (cardinal).  Blocks 3 and 4 of the stress hold e0 + e1 and
 * A (e0 + e1) + F e2 first, as in the reference. */
#define TI_STRESS(e, s, par, p)                                              \
    do {                                                                     \
        const double A_ = (par)[0 * N + (p)], C_ = (par)[1 * N + (p)];       \
        const double F_ = (par)[2 * N + (p)], L_ = (par)[3 * N + (p)];       \
        const double N_ = (par)[4 * N + (p)], mN2_ = N_ * -2.;               \
        (s)[3 * N + (p)] = (e)[0 * N + (p)] + (e)[1 * N + (p)];              \
        (s)[4 * N + (p)] = (s)[3 * N + (p)] * A_ + (e)[2 * N + (p)] * F_;    \
        (s)[0 * N + (p)] = (s)[4 * N + (p)] + (e)[1 * N + (p)] * mN2_;       \
        (s)[1 * N + (p)] = (s)[4 * N + (p)] + (e)[0 * N + (p)] * mN2_;       \
        (s)[2 * N + (p)] = (s)[3 * N + (p)] * F_ + (e)[2 * N + (p)] * C_;    \
        (s)[3 * N + (p)] = (e)[3 * N + (p)] * L_;                            \
        (s)[4 * N + (p)] = (e)[4 * N + (p)] * L_;                            \
        (s)[5 * N + (p)] = (e)[5 * N + (p)] * N_;                            \
    } while (0)

/* Anisotropic::strainToStress, src/core/element/material/elastic/
 * anisotropy/Anisotropic.hpp:203-274: s_i = sum_j C_ij e_j in order
 * j = 0..5; par = the 21 upper-triangle Cij (C11 C12 .. C16 C22 .. C66),
 * each [25]. */
static const int kCijIndex[6][6] = {
    {0, 1, 2, 3, 4, 5},      {1, 6, 7, 8, 9, 10},    {2, 7, 11, 12, 13, 14},
    {3, 8, 12, 15, 16, 17},  {4, 9, 13, 16, 18, 19}, {5, 10, 14, 17, 19, 20}};
#define ANISO_STRESS(T, e, s, par, p)                                        \
    do {                                                                     \
        for (int i_ = 0; i_ < 6; i_++) {                                     \
            T acc_ = (e)[0 * N + (p)] * (par)[kCijIndex[i_][0] * N + (p)];   \
            for (int j_ = 1; j_ < 6; j_++)                                   \
                acc_ += (e)[j_ * N + (p)] * (par)[kCijIndex[i_][j_] * N + (p)]; \
            (s)[i_ * N + (p)] = acc_;                                        \
        }                                                                    \
    } while (0)

/* Stress of one Fourier order (FR) or one slice (CD) for material type
 * mt (0 isotropic [lambda, mu], 1 TI [A, C, F, L, N], 2 anisotropic [21
 * Cij]); par = that type's parameters, each [25]. */
static void stress_fr(int mt, const double *par, const cplx *e, cplx *s) {
    if (mt == 0) {
        oracle_iso_stress(par, par + N, e, s);
    } else if (mt == 1) {
        for (int p = 0; p < N; p++) TI_STRESS(e, s, par, p);
    } else {
        for (int p = 0; p < N; p++) ANISO_STRESS(cplx, e, s, par, p);
    }
}

static void stress_cd(int mt, const double *par, const double *e, double *s) {
    if (mt == 0) {
        const double *lam = par, *mu = par + N;
        for (int p = 0; p < N; p++) {
            double tr = (e[0 * N + p] + e[1 * N + p] + e[2 * N + p]) * lam[p];
            double mu2 = mu[p] * 2.;
            s[0 * N + p] = tr + e[0 * N + p] * mu2;
            s[1 * N + p] = tr + e[1 * N + p] * mu2;
            s[2 * N + p] = tr + e[2 * N + p] * mu2;
            s[3 * N + p] = e[3 * N + p] * mu[p];
            s[4 * N + p] = e[4 * N + p] * mu[p];
            s[5 * N + p] = e[5 * N + p] * mu[p];
        }
    } else if (mt == 1) {
        for (int p = 0; p < N; p++) TI_STRESS(e, s, par, p);
    } else {
        for (int p = 0; p < N; p++) ANISO_STRESS(double, e, s, par, p);
    }
}

int oracle_mat_nparams(int mt) { return mt == 0 ? 2 : (mt == 1 ? 5 : 21); }

/* Test hook: stress of one Fourier order for material type mt. */
void oracle_material_stress(int mt, const double *par, const cplx *e,
                            cplx *s) {
    stress_fr(mt, par, e, s);
}

/* CoordTransformSpherical::transformSPZ_RTZ6 (strain, engineering shears)
 * and transformRTZ_SPZ6 (stress), src/core/element/crd/
 * CoordTransformSpherical.hpp:186-249, per GLL point with its colatitude
 * theta (Element::createCoordTransform, src/core/element/Element.cpp:
 * 133-151). */
void oracle_spz_to_rtz6(const double *theta, cplx *e) {
    for (int p = 0; p < N; p++) {
        double s1 = sin(theta[p]), c1 = cos(theta[p]);
        double s2 = sin(2. * theta[p]), c2 = cos(2. * theta[p]);
        cplx s02 = e[0 * N + p] + e[2 * N + p];
        cplx d02 = e[0 * N + p] - e[2 * N + p];
        e[0 * N + p] = (s02 + d02 * c2 - e[4 * N + p] * s2) * .5;
        e[2 * N + p] = s02 - e[0 * N + p];
        e[4 * N + p] = e[4 * N + p] * c2 + d02 * s2;
        cplx e3 = e[3 * N + p];
        e[3 * N + p] = e3 * c1 + e[5 * N + p] * s1;
        e[5 * N + p] = e[5 * N + p] * c1 - e3 * s1;
    }
}

void oracle_rtz_to_spz6(const double *theta, cplx *s) {
    for (int p = 0; p < N; p++) {
        double s1 = sin(theta[p]), c1 = cos(theta[p]);
        double s2 = sin(2. * theta[p]), c2 = cos(2. * theta[p]);
        cplx s02 = s[0 * N + p] + s[2 * N + p];
        cplx d02 = s[0 * N + p] - s[2 * N + p];
        s[0 * N + p] = (s02 + d02 * c2) * .5 + s[4 * N + p] * s2;
        s[2 * N + p] = s02 - s[0 * N + p];
        s[4 * N + p] = s[4 * N + p] * c2 - d02 * s2 * .5;
        cplx s3 = s[3 * N + p];
        s[3 * N + p] = s3 * c1 - s[5 * N + p] * s1;
        s[5 * N + p] = s[5 * N + p] * c1 + s3 * s1;
    }
}

/* transformSPZ_RTZ9 / transformRTZ_SPZ9 of a gradient-shaped tensor g[9]
 * (CoordTransformSpherical.hpp:102-184): g -> Q g Q^T with Q the rotation
 * by the colatitude in the (s, z) plane, and back. */
static void rot9(const double *theta, cplx *g, double sgn) {
    for (int p = 0; p < N; p++) {
        double s1 = sgn * sin(theta[p]), c1 = cos(theta[p]);
        double s2 = sgn * sin(2. * theta[p]), c2 = cos(2. * theta[p]);
        cplx *g0 = g + 0 * N + p, *g2 = g + 2 * N + p;
        cplx *g6 = g + 6 * N + p, *g8 = g + 8 * N + p;
        cplx s08 = *g0 + *g8, d08 = *g0 - *g8;
        cplx s26 = *g2 + *g6, d26 = *g2 - *g6;
        *g0 = (s08 + d08 * c2 - s26 * s2) * .5;
        *g2 = (d26 + s26 * c2 + d08 * s2) * .5;
        *g6 = *g2 - d26;
        *g8 = s08 - *g0;
        cplx g1 = g[1 * N + p], g3 = g[3 * N + p];
        g[1 * N + p] = g1 * c1 - g[7 * N + p] * s1;
        g[7 * N + p] = g[7 * N + p] * c1 + g1 * s1;
        g[3 * N + p] = g3 * c1 - g[5 * N + p] * s1;
        g[5 * N + p] = g[5 * N + p] * c1 + g3 * s1;
    }
}

void oracle_spz_to_rtz9(const double *theta, cplx *g) { rot9(theta, g, 1.); }
void oracle_rtz_to_spz9(const double *theta, cplx *g) { rot9(theta, g, -1.); }

/* PRT::sphericalToUndulated6 (PRT.hpp:327-340): the undulated Voigt strain
 * (engineering shears) from the spherical gradient g9 with the inverse
 * Jacobian x = [X0, X1, X2, X3] (each [25]). */
#define PRT_S2U(T, x, g, e)                                                   \
    for (int p = 0; p < N; p++) {                                             \
        double X0 = x[p], X1 = x[N + p], X2 = x[2 * N + p], X3 = x[3 * N + p];\
        T G0 = g[0 * N + p], G1 = g[1 * N + p], G2 = g[2 * N + p];            \
        T G3 = g[3 * N + p], G4 = g[4 * N + p], G5 = g[5 * N + p];            \
        T G6 = g[6 * N + p], G7 = g[7 * N + p], G8 = g[8 * N + p];            \
        e[0 * N + p] = G0 * X0 + G2 * X1;                                     \
        e[1 * N + p] = G4 * X0 + G5 * X2;                                     \
        e[2 * N + p] = G8 * X3;                                               \
        e[3 * N + p] = G7 * X0 + G8 * X2 + G5 * X3;                           \
        e[4 * N + p] = G6 * X0 + G8 * X1 + G2 * X3;                           \
        e[5 * N + p] = (G3 + G1) * X0 + G5 * X1 + G2 * X2;                    \
    }

/* PRT::undulatedToSpherical6 (PRT.hpp:374-389): the spherical tensor f9
 * of the undulated stress s6 with the scaled inverse Jacobian
 * xj = [X0 XJ, X1 XJ, X2 XJ, X3 XJ]; f[3] = f[1]. */
#define PRT_U2S(T, xj, s, f)                                                  \
    for (int p = 0; p < N; p++) {                                             \
        double Y0 = xj[p], Y1 = xj[N + p], Y2 = xj[2 * N + p];                \
        double Y3 = xj[3 * N + p];                                            \
        T S0 = s[0 * N + p], S1 = s[1 * N + p], S2 = s[2 * N + p];            \
        T S3 = s[3 * N + p], S4 = s[4 * N + p], S5 = s[5 * N + p];            \
        f[0 * N + p] = S0 * Y0;                                               \
        f[1 * N + p] = S5 * Y0;                                               \
        f[2 * N + p] = S0 * Y1 + S4 * Y3 + S5 * Y2;                           \
        f[3 * N + p] = f[1 * N + p];                                          \
        f[4 * N + p] = S1 * Y0;                                               \
        f[5 * N + p] = S1 * Y2 + S3 * Y3 + S5 * Y1;                           \
        f[6 * N + p] = S4 * Y0;                                               \
        f[7 * N + p] = S3 * Y0;                                               \
        f[8 * N + p] = S2 * Y3 + S3 * Y2 + S4 * Y1;                           \
    }

void oracle_prt_sph_to_und6(const double *x, const cplx *g9, cplx *e6) {
    PRT_S2U(cplx, x, g9, e6)
}

void oracle_prt_und_to_sph9(const double *x, const cplx *s6, cplx *f9) {
    PRT_U2S(cplx, x, s6, f9)
}

/* The scaled inverse Jacobian X_k XJ of a PRT operand set [5][25]
 * (the PRT constructor, PRT.hpp:57-80). */
static void prt_scaled(const double *x5, double *xj) {
    for (int k = 0; k < 4; k++)
        for (int p = 0; p < N; p++) xj[k * N + p] = x5[k * N + p] * x5[4 * N + p];
}

/* Material type of element e (Material::createElastic, src/preloop/
 * se_model/quad/physical/Material.cpp:366-379). */
static int elem_mt(const oracle_domain *d, int e) {
    return d->elem_mtype ? d->elem_mtype[e] : 0;
}

/* Elastic::inRTZ (TI and anisotropic) in a spherical mesh. */
static int elem_rtz(const oracle_domain *d, int e) {
    return d->elem_theta && elem_mt(d, e) != 0;
}

/* 1D parameters of element e. */
static const double *elem_par1d(const oracle_domain *d, int e) {
    if (d->elem_mtype) return d->mat1d + d->mat1d_off[e];
    return d->elem_iso + (int64_t)e * 2 * N;
}

/* 3D parameters of slice k of element e, [np][25]. */
static const double *elem_par3d(const oracle_domain *d, int e, int k) {
    int mt = elem_mt(d, e);
    if (d->elem_mtype)
        return d->prop3d + d->prop_off[e] +
               (int64_t)k * oracle_mat_nparams(mt) * N;
    return d->prop3d + (d->elem_slice_off[e] + k) * 2 * N;
}

/* ------------------------------------------------------------------ */
/* FFT (SolverFFTW conventions, src/core/fft/SolverFFTW.hpp:101-168)   */
/* ------------------------------------------------------------------ */
static int smallest_prime_factor(int n) {
    for (int p = 2; p * p <= n; p++)
        if (n % p == 0) return p;
    return n;
}

/* out[k] = sum_j in[j*is] W^{jk}, W = exp(sign 2 pi i / n) read from the
 * table of the top-level length N: W_n^x = tw[(x * L) % N], L = N / n.
 * Mixed radix, decimation in time on the smallest prime factor. */
static void fft_rec(const cplx *in, cplx *out, int n, int is,
                    const cplx *tw, int nt) {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    int p = smallest_prime_factor(n), m = n / p, L = nt / n;
    for (int r = 0; r < p; r++)
        fft_rec(in + (int64_t)r * is, out + (int64_t)r * m, m, is * p, tw, nt);
    cplx tloc[64], wloc[64];
    cplx *t = p <= 64 ? tloc : (cplx *)malloc(sizeof(cplx) * p);
    cplx *wp = p <= 64 ? wloc : (cplx *)malloc(sizeof(cplx) * p);
    for (int x = 0; x < p; x++) wp[x] = tw[(int64_t)x * m * L];
    for (int k = 0; k < m; k++) {
        /* r k L < p m L = nt: no reduction needed */
        for (int r = 0; r < p; r++)
            t[r] = out[r * m + k] * tw[(int64_t)r * k * L];
        for (int q = 0; q < p; q++) {
            cplx acc = t[0];
            int x = 0;
            for (int r = 1; r < p; r++) {
                x += q;
                if (x >= p) x -= p;
                acc += t[r] * wp[x];
            }
            out[q * m + k] = acc;
        }
    }
    if (wp != wloc) free(wp);
    if (t != tloc) free(t);
}

/* Twiddle tables exp(sign 2 pi i j / n), cached per n (n < 4096). */
static cplx *tw_cache[2][4096];

static const cplx *twiddles(int n, int sign) {
    int sl = sign > 0;
    if (n < 4096 && tw_cache[sl][n]) return tw_cache[sl][n];
    cplx *tw = (cplx *)malloc(sizeof(cplx) * n);
    for (int j = 0; j < n; j++)
        tw[j] = cexp(sign * 2. * M_PI * I * (double)j / (double)n);
    if (n < 4096) tw_cache[sl][n] = tw;
    return tw;
}

/* out = sum_j in[j] exp(sign 2 pi i jk/n); sign = +1 uses the conjugate
 * table (inverse transform). */
static void cfft(const cplx *in, cplx *out, int n, int sign) {
    const cplx *tw = twiddles(n, sign);
    fft_rec(in, out, n, 1, tw, n);
    if (n >= 4096) free((void *)tw);
}

/* computeR2C: X_k = (1/N) sum_n x_n e^{-2 pi i k n / N}, k <= N/2, Nyquist
 * (N even) halved (SolverFFTW.hpp:101-116). */
void oracle_r2c(const double *r, int nr, cplx *c) {
    cplx loc[4096];
    cplx *a = nr <= 2048 ? loc : (cplx *)malloc(sizeof(cplx) * nr * 2);
    for (int k = 0; k < nr; k++) a[k] = r[k] * (1. / nr);
    cfft(a, a + nr, nr, -1);
    int nc = nr / 2 + 1;
    for (int k = 0; k < nc; k++) c[k] = a[nr + k];
    if (nr % 2 == 0) c[nc - 1] *= .5;
    if (a != loc) free(a);
}

/* computeC2R: Nyquist (N even) doubled, then FFTW's unnormalised c2r on the
 * Hermitian extension, which takes only the real parts of X_0 and X_{N/2}
 * (SolverFFTW.hpp:138-153). */
void oracle_c2r(const cplx *c, int nr, double *r) {
    cplx loc[4096];
    cplx *a = nr <= 2048 ? loc : (cplx *)malloc(sizeof(cplx) * nr * 2);
    int nc = nr / 2 + 1;
    for (int k = 0; k < nr; k++) a[k] = 0;
    a[0] = creal(c[0]);
    for (int k = 1; k < nc; k++) {
        if (nr % 2 == 0 && k == nc - 1) {
            a[k] = 2. * creal(c[k]);
        } else {
            a[k] = c[k];
            a[nr - k] = conj(c[k]);
        }
    }
    cfft(a, a + nr, nr, +1);
    for (int k = 0; k < nr; k++) r[k] = creal(a[nr + k]);
    if (a != loc) free(a);
}

/* ------------------------------------------------------------------ */
/* Attenuation (src/core/element/material/elastic/attenuation)        */
/* ------------------------------------------------------------------ */
/* CG4 points (1,1),(1,3),(3,1),(3,3) (AttenuationCG4.hpp:203-212,
 * AttBuilder.cpp:141-147). */
static const int kCG4[4] = {6, 8, 16, 18};

typedef struct {
    int n;          /* rows: nu1 (FR) or Nr (CD) */
    int fr;         /* 1: complex Fourier rows   */
    double *dstress;      /* [n][6][npa] x (fr ? 2 : 1) */
    double *memvar;       /* [nsls][n][6][npa] x (fr ? 2 : 1) */
} att_elem;

static int att_npa(const oracle_domain *d) { return d->att_cg4 ? 4 : 25; }

static int elem_is3d(const oracle_domain *d, int e) {
    return d->elem_kind && d->elem_kind[e] == 1;
}

void *oracle_att_alloc(const oracle_domain *d) {
    if (d->nsls <= 0) return NULL;
    att_elem *st = (att_elem *)calloc(d->n_elem, sizeof(att_elem));
    int npa = att_npa(d);
    for (int e = 0; e < d->n_elem; e++) {
        int nr = d->elem_nr[e];
        st[e].fr = !elem_is3d(d, e);
        st[e].n = st[e].fr ? nr / 2 + 1 : nr;
        int w = st[e].fr ? 2 : 1;
        size_t sz = (size_t)st[e].n * 6 * npa * w;
        st[e].dstress = (double *)calloc(sz, sizeof(double));
        st[e].memvar = (double *)calloc(sz * d->nsls, sizeof(double));
    }
    return st;
}

void oracle_att_free(void *state, int n_elem) {
    att_elem *st = (att_elem *)state;
    if (!st) return;
    for (int e = 0; e < n_elem; e++) {
        free(st[e].dstress);
        free(st[e].memvar);
    }
    free(st);
}

/* AttenuationFull/CG4::apply on one row x (one Fourier order or one slice):
 * stress -= sum_l m_l; m_l = alpha_l m_l + beta_l dS; dS = iso(dlam,dmu)(e);
 * m_l += gamma_l dS  (AttenuationFull.hpp:192-207, AttenuationCG4.hpp:312-330,
 * Attenuation.hpp:92-140).  T = cplx (FR) or double (CD) via macro. */
#define ATT_ROW(T, strain, stress, dS, mv, dl, dm)                           \
    do {                                                                    \
        const double *al = d->att_abg, *be = al + d->nsls,                  \
                     *ga = be + d->nsls;                                    \
        int npa = att_npa(d);                                               \
        for (int l = 0; l < d->nsls; l++)                                   \
            for (int dd = 0; dd < 6; dd++)                                  \
                for (int q = 0; q < npa; q++) {                             \
                    int pt = d->att_cg4 ? kCG4[q] : q;                      \
                    stress[dd * N + pt] -= mv[(l * rows + x) * 6 * npa +    \
                                              dd * npa + q];                \
                }                                                           \
        for (int l = 0; l < d->nsls; l++)                                   \
            for (int k = 0; k < 6 * npa; k++) {                             \
                T *m = &mv[(l * rows + x) * 6 * npa + k];                   \
                *m = al[l] * *m + be[l] * dS[x * 6 * npa + k];              \
            }                                                               \
        for (int q = 0; q < npa; q++) {                                     \
            int pt = d->att_cg4 ? kCG4[q] : q;                              \
            T tr = (strain[0 * N + pt] + strain[1 * N + pt] +               \
                    strain[2 * N + pt]) * dl[q];                            \
            double dm2 = dm[q] * 2.;                                        \
            T *o = &dS[x * 6 * npa];                                        \
            o[0 * npa + q] = tr + strain[0 * N + pt] * dm2;                 \
            o[1 * npa + q] = tr + strain[1 * N + pt] * dm2;                 \
            o[2 * npa + q] = tr + strain[2 * N + pt] * dm2;                 \
            o[3 * npa + q] = strain[3 * N + pt] * dm[q];                    \
            o[4 * npa + q] = strain[4 * N + pt] * dm[q];                    \
            o[5 * npa + q] = strain[5 * N + pt] * dm[q];                    \
        }                                                                   \
        for (int l = 0; l < d->nsls; l++)                                   \
            for (int k = 0; k < 6 * npa; k++)                               \
                mv[(l * rows + x) * 6 * npa + k] += ga[l] * dS[x * 6 * npa + k]; \
    } while (0)

static void att_fr(const oracle_domain *d, int e, int alpha, const cplx *strain,
                   cplx *stress) {
    att_elem *st = &((att_elem *)d->att_state)[e];
    int npa = att_npa(d), rows = st->n, x = alpha;
    cplx *dS = (cplx *)st->dstress, *mv = (cplx *)st->memvar;
    const double *dl = d->att1d + (int64_t)e * 2 * npa, *dm = dl + npa;
    ATT_ROW(cplx, strain, stress, dS, mv, dl, dm);
}

static void att_cd(const oracle_domain *d, int e, int slice,
                   const double *strain, double *stress) {
    att_elem *st = &((att_elem *)d->att_state)[e];
    int npa = att_npa(d), rows = st->n, x = slice;
    double *dS = st->dstress, *mv = st->memvar;
    const double *dl = d->att3d +
                       (d->elem_slice_off[e] + slice) * 2 * (int64_t)npa;
    const double *dm = dl + npa;
    ATT_ROW(double, strain, stress, dS, mv, dl, dm);
}

/* Test hook: attenuation of one Fourier order of a 1D element (Fourier
 * path); strain, stress complex [6][25]. */
void oracle_att_apply_fr(const oracle_domain *d, int e, int alpha,
                         const cplx *strain, cplx *stress) {
    att_fr(d, e, alpha, strain, stress);
}

/* Test hook: attenuation of one slice of a 3D element (cardinal path). */
void oracle_att_apply_cd(const oracle_domain *d, int e, int slice,
                         const double *strain, double *stress) {
    att_cd(d, e, slice, strain, stress);
}

/* ------------------------------------------------------------------ */
/* Domain / time loop                                                  */
/* ------------------------------------------------------------------ */
static int64_t *make_offsets(int n, const int32_t *nr) {
    int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
    off[0] = 0;
    for (int i = 0; i < n; i++) off[i + 1] = off[i] + nr[i] / 2 + 1;
    return off;
}

/* SolidElement::computeStiff (SolidElement.cpp:213-222), no PRT:
 * collectDisplFromPoints (SolidElement.cpp:102-109, zero-padded,
 * SolidPoint.hpp:82-98) -> displToStiff (SolidElement.cpp:160-198) ->
 * addStiffToPoints (subtract, SolidPoint.hpp:101-109).
 * 1D material (Fourier path): computeGrad6 -> strainToStress_FR per mode
 * (+ attenuation per mode) -> computeQuad6.
 * 3D material (cardinal path): computeGrad6 -> gFFT_N6.computeC2R ->
 * strainToStress<_3D_CD> per slice (+ attenuation per slice) ->
 * gFFT_N6.computeR2C -> computeQuad6.  TI and anisotropic media work in
 * RTZ on spherical meshes: transformSPZ_RTZ6 after the gradient and
 * transformRTZ_SPZ6 before the quadrature (SolidElement.cpp:160-196). */
static void element_forces(const oracle_domain *d, int e, const cplx *displ,
                           cplx *fout);
static void element_scatter(const oracle_domain *d, int e, const cplx *fout,
                            cplx *stiff);

void oracle_solid_element_stiff(const oracle_domain *d, int e,
                                const cplx *displ, cplx *stiff) {
    int nu1 = d->elem_nr[e] / 2 + 1;
    cplx *f = (cplx *)malloc(sizeof(cplx) * nu1 * 3 * N);
    element_forces(d, e, displ, f);
    element_scatter(d, e, f, stiff);
    free(f);
}

/* The forces of an element with a particle relabelling transformation
 * (SolidElement::displToStiff, SolidElement.cpp:115-159): the 9-component
 * spherical gradient, rotated to RTZ on spherical meshes, to the
 * undulated strain per Fourier order (1D material) or per slice after
 * the C2R of all nine series (3D material), the stress, back to the
 * spherical tensor and its R2C, rotated back, and computeQuad9. */
static void element_forces_prt(const oracle_domain *d, int e,
                               const cplx *displ, cplx *fout) {
    const int32_t *pts = d->elem_points + (int64_t)e * N;
    int nr = d->elem_nr[e];
    int nu1 = nr / 2 + 1;
    const double *gq_e = d->elem_gq + (int64_t)e * 6 * N;
    int axial = d->elem_axial[e];
    int is3d = elem_is3d(d, e);
    int mt = elem_mt(d, e);
    const double *x5 = d->prt_data + d->prt_off[e];
    const double *theta = d->elem_theta ? d->elem_theta + (int64_t)e * N
                                        : NULL;
    double xj[4 * N];
    prt_scaled(x5, xj);
    cplx *g9 = (cplx *)malloc(sizeof(cplx) * nu1 * 9 * N);
    cplx *f9 = (cplx *)malloc(sizeof(cplx) * nu1 * 9 * N);
    cplx u[3 * N];
    for (int alpha = 0; alpha < nu1; alpha++) {
        for (int p = 0; p < N; p++) {
            int ip = pts[p];
            int pnu1 = d->point_nr[ip] / 2 + 1;
            for (int c = 0; c < 3; c++)
                u[c * N + p] = alpha < pnu1
                    ? displ[(d->offsets[ip] + alpha) * 3 + c] : 0.;
        }
        cplx *ga = g9 + (int64_t)alpha * 9 * N;
        oracle_grad9(gq_e, axial, alpha, u, ga);
        if (theta) oracle_spz_to_rtz9(theta, ga);
    }
    if (!is3d) {
        const double *par = elem_par1d(d, e);
        cplx e6[6 * N], s6[6 * N];
        for (int alpha = 0; alpha < nu1; alpha++) {
            const cplx *ga = g9 + (int64_t)alpha * 9 * N;
            PRT_S2U(cplx, x5, ga, e6)
            stress_fr(mt, par, e6, s6);
            if (d->nsls > 0 && d->att1d) att_fr(d, e, alpha, e6, s6);
            cplx *fa = f9 + (int64_t)alpha * 9 * N;
            PRT_U2S(cplx, xj, s6, fa)
        }
    } else {
        double *gcd = (double *)malloc(sizeof(double) * nr * 9 * N);
        double *fcd = (double *)malloc(sizeof(double) * nr * 9 * N);
        cplx *ser = (cplx *)malloc(sizeof(cplx) * nu1);
        double *rser = (double *)malloc(sizeof(double) * nr);
        for (int col = 0; col < 9 * N; col++) {
            for (int a = 0; a < nu1; a++) ser[a] = g9[(int64_t)a * 9 * N + col];
            oracle_c2r(ser, nr, rser);
            for (int k = 0; k < nr; k++) gcd[(int64_t)k * 9 * N + col] = rser[k];
        }
        double e6[6 * N], s6[6 * N];
        for (int k = 0; k < nr; k++) {
            const double *gk = gcd + (int64_t)k * 9 * N;
            double *fk = fcd + (int64_t)k * 9 * N;
            PRT_S2U(double, x5, gk, e6)
            stress_cd(mt, elem_par3d(d, e, k), e6, s6);
            if (d->nsls > 0 && d->att3d) att_cd(d, e, k, e6, s6);
            PRT_U2S(double, xj, s6, fk)
        }
        for (int col = 0; col < 9 * N; col++) {
            for (int k = 0; k < nr; k++) rser[k] = fcd[(int64_t)k * 9 * N + col];
            oracle_r2c(rser, nr, ser);
            for (int a = 0; a < nu1; a++) f9[(int64_t)a * 9 * N + col] = ser[a];
        }
        free(gcd);
        free(fcd);
        free(ser);
        free(rser);
    }
    for (int alpha = 0; alpha < nu1; alpha++) {
        cplx *fa = f9 + (int64_t)alpha * 9 * N;
        if (theta) oracle_rtz_to_spz9(theta, fa);
        oracle_quad9(gq_e, axial, alpha, fa, fout + (int64_t)alpha * 3 * N);
    }
    free(g9);
    free(f9);
}

/* The element's forces, fout[alpha][c][p] (before the subtraction). */
static void element_forces(const oracle_domain *d, int e, const cplx *displ,
                           cplx *fout) {
    if (d->prt_off && d->prt_off[e] >= 0) {
        element_forces_prt(d, e, displ, fout);
        return;
    }
    const int32_t *pts = d->elem_points + (int64_t)e * N;
    int nr = d->elem_nr[e];
    int nu1 = nr / 2 + 1;
    const double *gq_e = d->elem_gq + (int64_t)e * 6 * N;
    int axial = d->elem_axial[e];
    int is3d = elem_is3d(d, e);
    int mt = elem_mt(d, e), rtz = elem_rtz(d, e);
    cplx *eps = (cplx *)malloc(sizeof(cplx) * nu1 * 6 * N);
    cplx *sig = (cplx *)malloc(sizeof(cplx) * nu1 * 6 * N);
    cplx u[3 * N];
    for (int alpha = 0; alpha < nu1; alpha++) {
        for (int p = 0; p < N; p++) {
            int ip = pts[p];
            int pnu1 = d->point_nr[ip] / 2 + 1;
            for (int c = 0; c < 3; c++)
                u[c * N + p] = alpha < pnu1
                    ? displ[(d->offsets[ip] + alpha) * 3 + c] : 0.;
        }
        oracle_grad6(gq_e, axial, alpha, u, eps + (int64_t)alpha * 6 * N);
        if (rtz)
            oracle_spz_to_rtz6(d->elem_theta + (int64_t)e * N,
                               eps + (int64_t)alpha * 6 * N);
    }
    if (!is3d) {
        const double *par = elem_par1d(d, e);
        for (int alpha = 0; alpha < nu1; alpha++) {
            stress_fr(mt, par, eps + (int64_t)alpha * 6 * N,
                      sig + (int64_t)alpha * 6 * N);
            if (d->nsls > 0 && d->att1d)
                att_fr(d, e, alpha, eps + (int64_t)alpha * 6 * N,
                       sig + (int64_t)alpha * 6 * N);
        }
    } else {
        double *ecd = (double *)malloc(sizeof(double) * nr * 6 * N);
        double *scd = (double *)malloc(sizeof(double) * nr * 6 * N);
        cplx *ser = (cplx *)malloc(sizeof(cplx) * nu1);
        double *rser = (double *)malloc(sizeof(double) * nr);
        for (int col = 0; col < 6 * N; col++) {
            for (int a = 0; a < nu1; a++) ser[a] = eps[(int64_t)a * 6 * N + col];
            oracle_c2r(ser, nr, rser);
            for (int k = 0; k < nr; k++) ecd[(int64_t)k * 6 * N + col] = rser[k];
        }
        for (int k = 0; k < nr; k++) {
            const double *ek = ecd + (int64_t)k * 6 * N;
            double *sk = scd + (int64_t)k * 6 * N;
            stress_cd(mt, elem_par3d(d, e, k), ek, sk);
            if (d->nsls > 0 && d->att3d) att_cd(d, e, k, ek, sk);
        }
        for (int col = 0; col < 6 * N; col++) {
            for (int k = 0; k < nr; k++) rser[k] = scd[(int64_t)k * 6 * N + col];
            oracle_r2c(rser, nr, ser);
            for (int a = 0; a < nu1; a++) sig[(int64_t)a * 6 * N + col] = ser[a];
        }
        free(ecd);
        free(scd);
        free(ser);
        free(rser);
    }
    for (int alpha = 0; alpha < nu1; alpha++) {
        if (rtz)
            oracle_rtz_to_spz6(d->elem_theta + (int64_t)e * N,
                               sig + (int64_t)alpha * 6 * N);
        oracle_quad6(gq_e, axial, alpha, sig + (int64_t)alpha * 6 * N,
                     fout + (int64_t)alpha * 3 * N);
    }
    free(eps);
    free(sig);
}

/* addStiffToPoints: point stiff -= element force, over the point's own
 * orders (SolidPoint.hpp:101-109). */
static void element_scatter(const oracle_domain *d, int e, const cplx *fout,
                            cplx *stiff) {
    const int32_t *pts = d->elem_points + (int64_t)e * N;
    int nu1 = d->elem_nr[e] / 2 + 1;
    for (int alpha = 0; alpha < nu1; alpha++) {
        const cplx *f = fout + (int64_t)alpha * 3 * N;
        for (int p = 0; p < N; p++) {
            int ip = pts[p];
            int pnu1 = d->point_nr[ip] / 2 + 1;
            if (alpha < pnu1)
                for (int c = 0; c < 3; c++)
                    stiff[(d->offsets[ip] + alpha) * 3 + c] -= f[c * N + p];
        }
    }
}

/* Stiffness of all solid elements, subtracted into stiff in ascending
 * element order (Domain::computeStiffness).  The element forces are
 * computed by the OpenMP threads in any order, then scattered by one
 * thread in element order, so the result does not depend on the thread
 * count. */
static void solid_stiffness(const oracle_domain *d, const cplx *displ,
                            cplx *stiff) {
    int ne = d->n_elem;
    int64_t *foff = (int64_t *)malloc(sizeof(int64_t) * (ne + 1));
    foff[0] = 0;
    for (int e = 0; e < ne; e++)
        foff[e + 1] = foff[e] + (int64_t)(d->elem_nr[e] / 2 + 1) * 3 * N;
    cplx *fbuf = (cplx *)malloc(sizeof(cplx) * (size_t)foff[ne]);
    /* the transform tables of every Nr before the threads share them */
    for (int e = 0; e < ne; e++)
        if (elem_is3d(d, e)) {
            twiddles(d->elem_nr[e], -1);
            twiddles(d->elem_nr[e], +1);
        }
#pragma omp parallel for schedule(dynamic, 1)
    for (int e = 0; e < ne; e++) element_forces(d, e, displ, fbuf + foff[e]);
    for (int e = 0; e < ne; e++) element_scatter(d, e, fbuf + foff[e], stiff);
    free(fbuf);
    free(foff);
}

/* Threads of the element loop (OpenMP; 1 = serial). */
void oracle_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

/* SolidMoment::apply + SolidElement::addMomentSource (SolidMoment.cpp:29-33,
 * SolidElement.cpp:273-356, no PRT): stress = -pattern * stf, masked above
 * min(nu1_elem, nu1_src); Quad6_NoIntegration; points subtract. */
void oracle_apply_moment_source(const oracle_domain *d, double stf_value,
                                cplx *stiff) {
    int e = d->src_elem;
    if (e < 0) return;
    const int32_t *pts = d->elem_points + (int64_t)e * N;
    int nu1 = d->elem_nr[e] / 2 + 1;
    int nu1c = nu1 < d->src_nu1 ? nu1 : d->src_nu1;
    const double *gq_e = d->elem_gq + (int64_t)e * 6 * N;
    int axial = d->elem_axial[e];
    cplx sig[6 * N], f[3 * N];
    for (int alpha = 0; alpha < nu1; alpha++) {
        if (alpha < nu1c) {
            for (int k = 0; k < 6 * N; k++) {
                cplx pv = d->src_pattern[(int64_t)alpha * 6 * N + k];
                sig[k] = -(pv * stf_value);
            }
            if (d->prt_off && d->prt_off[e] >= 0) {
                /* SolidElement::addMomentSource with a PRT
                 * (SolidElement.cpp:293-331): the stress to RTZ by the
                 * strain rotation with doubled shears, to the spherical
                 * tensor with the unscaled inverse Jacobian, back to SPZ,
                 * computeQuad9_NoIntegration */
                const double *theta = d->elem_theta
                    ? d->elem_theta + (int64_t)e * N : NULL;
                const double *x5 = d->prt_data + d->prt_off[e];
                cplx f9[9 * N];
                for (int k = 3 * N; k < 6 * N; k++) sig[k] *= 2.;
                if (theta) oracle_spz_to_rtz6(theta, sig);
                for (int k = 3 * N; k < 6 * N; k++) sig[k] *= .5;
                PRT_U2S(cplx, x5, sig, f9)
                if (theta) oracle_rtz_to_spz9(theta, f9);
                quad9_noint(gq_e, axial, alpha, f9, f);
            } else {
                oracle_quad6_noint(gq_e, axial, alpha, sig, f);
            }
        } else {
            for (int k = 0; k < 3 * N; k++) f[k] = 0.;
        }
        for (int p = 0; p < N; p++) {
            int ip = pts[p];
            int pnu1 = d->point_nr[ip] / 2 + 1;
            if (alpha < pnu1)
                for (int c = 0; c < 3; c++)
                    stiff[(d->offsets[ip] + alpha) * 3 + c] -= f[c * N + p];
        }
    }
}

/* AxialBoundary::apply, solid part (AxialBoundary.cpp:21-42). */
void oracle_axial_solid(const oracle_domain *d, cplx *stiff) {
    for (int k = 0; k < d->n_axial; k++) {
        int ip = d->axial_points[k];
        int nu = d->point_nr[ip] / 2;
        cplx *s = stiff + d->offsets[ip] * 3;
        s[0] = 0.;
        s[1] = 0.;
        if (nu >= 1) {
            s[3 + 0] = (s[3 + 0] - I * s[3 + 1]) * .5;
            s[3 + 1] = I * s[3 + 0];
            s[3 + 2] = 0.;
            for (int a = 2; a <= nu; a++)
                for (int c = 0; c < 3; c++) s[a * 3 + c] = 0.;
        }
    }
}

/* SolidPoint::computeStiffToAccel (SolidPoint.cpp:28-36) with Mass1D
 * (Mass1D.hpp:25-32: multiply by the stored 1/mass). */
void oracle_stiff_to_accel(const oracle_domain *d, cplx *stiff) {
    for (int ip = 0; ip < d->n_points; ip++) {
        int nr = d->point_nr[ip];
        int nu1 = nr / 2 + 1;
        cplx *s = stiff + d->offsets[ip] * 3;
        if (nr % 2 == 0)
            for (int c = 0; c < 3; c++)
                s[(nu1 - 1) * 3 + c] = creal(s[(nu1 - 1) * 3 + c]);
        if (d->ocean_off && d->ocean_off[ip] >= 0) {
            const double *o = d->ocean_data + d->ocean_off[ip] * 4;
            if (d->ocean_nval[ip] == 1) {
                /* MassOceanLoad1D::computeAccel (MassOceanLoad1D.hpp:
                 * 48-62): (s, z) by the 2x2 factors, phi by 1/m */
                for (int a = 0; a < nu1; a++) {
                    const cplx fs = s[a * 3], fz = s[a * 3 + 2];
                    s[a * 3] = o[0] * fs + o[2] * fz;
                    s[a * 3 + 2] = o[2] * fs + o[1] * fz;
                    s[a * 3 + 1] *= o[3];
                }
                continue;
            }
            /* MassOceanLoad3D::computeAccel (MassOceanLoad3D.cpp:41-56):
             * F = C2R(stiff) per component, a = F / m - (F . k) k per
             * slice, stiff = R2C(a) */
            cplx *col = malloc(sizeof(cplx) * nu1);
            double *f = malloc(sizeof(double) * nr * 3);
            for (int c = 0; c < 3; c++) {
                for (int a = 0; a < nu1; a++) col[a] = s[a * 3 + c];
                oracle_c2r(col, nr, f + (size_t)c * nr);
            }
            for (int k = 0; k < nr; k++) {
                const double *q = o + (size_t)k * 4;
                double fk[3], dot = 0.;
                for (int c = 0; c < 3; c++) {
                    fk[c] = f[(size_t)c * nr + k];
                    dot += fk[c] * q[1 + c];
                }
                for (int c = 0; c < 3; c++)
                    f[(size_t)c * nr + k] = q[0] * fk[c] - dot * q[1 + c];
            }
            for (int c = 0; c < 3; c++) {
                oracle_r2c(f + (size_t)c * nr, nr, col);
                for (int a = 0; a < nu1; a++) s[a * 3 + c] = col[a];
            }
            free(col);
            free(f);
            continue;
        }
        if (d->mass3d_off && d->mass3d_off[ip] >= 0) {
            /* Mass3D::computeAccel (src/core/point/mass/Mass3D.cpp:56-68):
             * per component C2R, times 1 / m slice by slice, R2C */
            const double *m = d->mass3d + d->mass3d_off[ip];
            cplx *col = malloc(sizeof(cplx) * nu1);
            double *r = malloc(sizeof(double) * nr);
            for (int c = 0; c < 3; c++) {
                for (int a = 0; a < nu1; a++) col[a] = s[a * 3 + c];
                oracle_c2r(col, nr, r);
                for (int k = 0; k < nr; k++) r[k] *= 1. / m[k];
                oracle_r2c(r, nr, col);
                for (int a = 0; a < nu1; a++) s[a * 3 + c] = col[a];
            }
            free(col);
            free(r);
            continue;
        }
        double inv = 1. / d->point_mass[ip];
        for (int k = 0; k < nu1 * 3; k++) s[k] *= inv;
    }
}

/* NewmarkTimeScheme::update (NewmarkTimeScheme.hpp:48-59). */
void oracle_newmark(int64_t nmodes, double dt, cplx *stiff, cplx *displ,
                    cplx *veloc, cplx *accel) {
    const double half_dt = .5 * dt;
    const double half_dt_dt = .5 * dt * dt;
    for (int64_t k = 0; k < nmodes * 3; k++) {
        veloc[k] += half_dt * (accel[k] + stiff[k]);
        accel[k] = stiff[k];
        displ[k] += dt * veloc[k] + half_dt_dt * accel[k];
        stiff[k] = 0.;
    }
}

/* ------------------------------------------------------------------ */
/* Absorbing boundaries (src/core/boundary/ABC)                        */
/* ------------------------------------------------------------------ */
/* One Fourier column (nu1 modes, stride `st` complex) of a point field to
 * its Nr slices (SolverFFTW::computeC2R) and back (computeR2C). */
static void col_c2r(const cplx *x, int st, int nr, cplx *tmp, double *r) {
    const int nu1 = nr / 2 + 1;
    for (int a = 0; a < nu1; a++) tmp[a] = x[(int64_t)a * st];
    oracle_c2r(tmp, nr, r);
}

/* AbsorbingBoundary::applyClayton (AbsorbingBoundary.hpp:48-56), solid
 * entries then fluid entries, each in list order:
 *  ClaytonSolid1D::apply (ClaytonSolid1D.cpp:16-29): s -= c0 v_s + c2 v_z,
 *    phi -= c3 v_phi, z -= c2 v_s + c1 v_z;
 *  ClaytonSolid3D::apply (ClaytonSolid3D.cpp:43-67): V = C2R(v), a = rsa V
 *    + (V . k) k per slice, stiff -= R2C(a);
 *  ClaytonFluid1D::apply (ClaytonFluid1D.cpp:15-22): stiff -= v c;
 *  ClaytonFluid3D::apply (ClaytonFluid3D.cpp:33-51): stiff -= R2C(C2R(v)
 *    c). */
void oracle_clayton(const oracle_domain *d, const cplx *veloc, cplx *stiff,
                    const cplx *fveloc, cplx *fstiff) {
    for (int i = 0; i < d->n_clayton_s; i++) {
        const int ip = d->clayton_s_point[i];
        const int nr = d->point_nr[ip], nu1 = nr / 2 + 1;
        const double *c = d->clayton_s_data + 4 * d->clayton_s_off[i];
        const cplx *v = veloc + d->offsets[ip] * 3;
        cplx *s = stiff + d->offsets[ip] * 3;
        if (d->clayton_s_nval[i] == 1) {
            for (int a = 0; a < nu1; a++) {
                const cplx vs = v[a * 3], vp = v[a * 3 + 1], vz = v[a * 3 + 2];
                s[a * 3 + 0] -= c[0] * vs + c[2] * vz;
                s[a * 3 + 2] -= c[2] * vs + c[1] * vz;
                s[a * 3 + 1] -= c[3] * vp;
            }
            continue;
        }
        cplx *tmp = malloc(sizeof(cplx) * nu1);
        double *V = malloc(sizeof(double) * nr * 3);
        double *A = malloc(sizeof(double) * nr * 3);
        for (int q = 0; q < 3; q++) col_c2r(v + q, 3, nr, tmp, V + q * nr);
        for (int k = 0; k < nr; k++) {
            const double *ck = c + 4 * k;
            const double dot = V[k] * ck[1] + V[nr + k] * ck[2] +
                               V[2 * nr + k] * ck[3];
            for (int q = 0; q < 3; q++)
                A[q * nr + k] = ck[0] * V[q * nr + k] + dot * ck[1 + q];
        }
        for (int q = 0; q < 3; q++) {
            oracle_r2c(A + q * nr, nr, tmp);
            for (int a = 0; a < nu1; a++) s[a * 3 + q] -= tmp[a];
        }
        free(tmp);
        free(V);
        free(A);
    }
    for (int i = 0; i < d->n_clayton_f; i++) {
        const int ip = d->clayton_f_point[i];
        const int nr = d->fpoint_nr[ip], nu1 = nr / 2 + 1;
        const double *c = d->clayton_f_data + d->clayton_f_off[i];
        const cplx *v = fveloc + d->foffsets[ip];
        cplx *s = fstiff + d->foffsets[ip];
        if (d->clayton_f_nval[i] == 1) {
            for (int a = 0; a < nu1; a++) s[a] -= v[a] * c[0];
            continue;
        }
        cplx *tmp = malloc(sizeof(cplx) * nu1);
        double *V = malloc(sizeof(double) * nr);
        col_c2r(v, 1, nr, tmp, V);
        for (int k = 0; k < nr; k++) V[k] *= c[k];
        oracle_r2c(V, nr, tmp);
        for (int a = 0; a < nu1; a++) s[a] -= tmp[a];
        free(tmp);
        free(V);
    }
}

/* Sponge::apply (Sponge.hpp:74-96) on one point of nc components, after
 * the acceleration is formed: 1D a -= 2 gamma v + gamma^2 u; 3D a =
 * R2C(C2R(a) - 2 gamma C2R(v) - gamma^2 C2R(u)) slice by slice. */
static void sponge_point(int nc, int nr, int nval, const double *g,
                         const cplx *u, const cplx *v, cplx *acc) {
    const int nu1 = nr / 2 + 1;
    if (nval == 1) {
        const double g2 = 2. * g[0], gg = g[0] * g[0];
        for (int k = 0; k < nu1 * nc; k++) acc[k] -= g2 * v[k] + gg * u[k];
        return;
    }
    cplx *tmp = malloc(sizeof(cplx) * nu1);
    double *A = malloc(sizeof(double) * nr);
    double *U = malloc(sizeof(double) * nr);
    double *V = malloc(sizeof(double) * nr);
    for (int q = 0; q < nc; q++) {
        col_c2r(acc + q, nc, nr, tmp, A);
        col_c2r(u + q, nc, nr, tmp, U);
        col_c2r(v + q, nc, nr, tmp, V);
        for (int k = 0; k < nr; k++)
            A[k] -= 2. * g[k] * V[k] + g[k] * g[k] * U[k];
        oracle_r2c(A, nr, tmp);
        for (int a = 0; a < nu1; a++) acc[(int64_t)a * nc + q] = tmp[a];
    }
    free(tmp);
    free(A);
    free(U);
    free(V);
}

/* AbsorbingBoundary::applySponge (AbsorbingBoundary.hpp:59-66): solid
 * sponges, then fluid sponges. */
void oracle_sponge(const oracle_domain *d, const cplx *displ,
                   const cplx *veloc, cplx *accel, const cplx *fdispl,
                   const cplx *fveloc, cplx *faccel) {
    for (int i = 0; i < d->n_sponge_s; i++) {
        const int ip = d->sponge_s_point[i];
        const int64_t o = d->offsets[ip] * 3;
        sponge_point(3, d->point_nr[ip], d->sponge_s_nval[i],
                     d->sponge_s_data + d->sponge_s_off[i], displ + o,
                     veloc + o, accel + o);
    }
    for (int i = 0; i < d->n_sponge_f; i++) {
        const int ip = d->sponge_f_point[i];
        const int64_t o = d->foffsets[ip];
        sponge_point(1, d->fpoint_nr[ip], d->sponge_f_nval[i],
                     d->sponge_f_data + d->sponge_f_off[i], fdispl + o,
                     fveloc + o, faccel + o);
    }
}

/* Station::interpolate + computeFourierAtPhiExp (Station.hpp:53-73,
 * Station.cpp:13-38, src/shared/eigen/eigen_tools.hpp:79-106):
 * in-plane weighted sum over the non-zero weights, then
 * Re sum_alpha c_alpha u_alpha e^{i alpha phi}, c_0 = 1, c_alpha = 2. */
void oracle_record(const oracle_domain *d, const cplx *displ, double *out) {
    for (int r = 0; r < d->n_rec; r++) {
        int e = d->rec_elem[r];
        const int32_t *pts = d->elem_points + (int64_t)e * N;
        int nu1 = d->elem_nr[e] / 2 + 1;
        const double *w = d->rec_weights + (int64_t)r * N;
        double phi = d->rec_phi[r];
        for (int c = 0; c < 3; c++) {
            double acc = 0.;
            for (int alpha = 0; alpha < nu1; alpha++) {
                cplx xd = 0.;
                for (int p = 0; p < N; p++) {
                    if (fabs(w[p]) < 2.220446049250313e-16) continue;
                    int ip = pts[p];
                    int pnu1 = d->point_nr[ip] / 2 + 1;
                    cplx uv = alpha < pnu1
                        ? displ[(d->offsets[ip] + alpha) * 3 + c] : 0.;
                    xd += uv * w[p];
                }
                cplx two_exp = alpha == 0 ? 1. : 2. * cexp(I * alpha * phi);
                acc += creal(two_exp * xd);
            }
            out[r * 3 + c] = acc;
        }
    }
}

/* NewmarkTimeScheme::solve (NewmarkTimeScheme.cpp:41-108), phase order:
 * sources -> stiffness -> BC before (none here) -> record -> stability ->
 * BC after (axial) -> stiff to accel -> Newmark. */
/* FluidElement::computeStiff, 1D acoustic (Fourier path), no PRT
 * (FluidElement.cpp:91-148): collectDisplFromPoints (zero-padded,
 * FluidPoint.hpp:82-94) -> computeGrad3 -> Acoustic::strainToStress
 * (K p_i, Acoustic.hpp:80-92) -> computeQuad3 -> subtracting scatter
 * (FluidPoint.hpp:97-103). */
/* FluidElement::displToStiff with a 3D medium (FluidElement.cpp:101-148,
 * the cardinal branch): computeGrad3 for every order, gFFT_N3.computeC2R
 * of the 75 series, K per slice (Acoustic::strainToStress _3D_CD,
 * Acoustic.hpp:80-92), computeR2C, computeQuad3, subtracting scatter. */
static void fluid_element_stiff3d(const oracle_domain *d, int e,
                                  const cplx *chi, cplx *fstiff) {
    const int32_t *pts = d->felem_points + (int64_t)e * N;
    const int nr = d->felem_nr[e];
    const int nu1 = nr / 2 + 1;
    const double *gq_e = d->felem_gq + (int64_t)e * 6 * N;
    const double *K3 = d->fk3d + d->fk3d_off[e];
    const int axial = d->felem_axial[e];
    cplx *PI = malloc(sizeof(cplx) * nu1 * 3 * N);
    cplx *col = malloc(sizeof(cplx) * nu1);
    double *r = malloc(sizeof(double) * nr);
    cplx p[N], q[N];
    for (int alpha = 0; alpha < nu1; alpha++) {
        for (int k = 0; k < N; k++) {
            int ip = pts[k];
            int pnu1 = d->fpoint_nr[ip] / 2 + 1;
            p[k] = alpha < pnu1 ? chi[d->foffsets[ip] + alpha] : 0.;
        }
        oracle_grad3(gq_e, axial, alpha, p, PI + (int64_t)alpha * 3 * N);
    }
    for (int c = 0; c < 3; c++)
        for (int k = 0; k < N; k++) {
            for (int a = 0; a < nu1; a++) col[a] = PI[(a * 3 + c) * N + k];
            oracle_c2r(col, nr, r);
            for (int sl = 0; sl < nr; sl++) r[sl] *= K3[sl * N + k];
            oracle_r2c(r, nr, col);
            for (int a = 0; a < nu1; a++) PI[(a * 3 + c) * N + k] = col[a];
        }
    for (int alpha = 0; alpha < nu1; alpha++) {
        oracle_quad3(gq_e, axial, alpha, PI + (int64_t)alpha * 3 * N, q);
        for (int k = 0; k < N; k++) {
            int ip = pts[k];
            int pnu1 = d->fpoint_nr[ip] / 2 + 1;
            if (alpha < pnu1) fstiff[d->foffsets[ip] + alpha] -= q[k];
        }
    }
    free(PI);
    free(col);
    free(r);
}

void oracle_fluid_element_stiff(const oracle_domain *d, int e,
                                const cplx *chi, cplx *fstiff) {
    if (d->felem_kind && d->felem_kind[e] == 1) {
        fluid_element_stiff3d(d, e, chi, fstiff);
        return;
    }
    const int32_t *pts = d->felem_points + (int64_t)e * N;
    int nu1 = d->felem_nr[e] / 2 + 1;
    const double *gq_e = d->felem_gq + (int64_t)e * 6 * N;
    const double *K = d->felem_k + (int64_t)e * N;
    int axial = d->felem_axial[e];
    cplx p[N], pi[3 * N], q[N];
    for (int alpha = 0; alpha < nu1; alpha++) {
        for (int k = 0; k < N; k++) {
            int ip = pts[k];
            int pnu1 = d->fpoint_nr[ip] / 2 + 1;
            p[k] = alpha < pnu1 ? chi[d->foffsets[ip] + alpha] : 0.;
        }
        oracle_grad3(gq_e, axial, alpha, p, pi);
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < N; k++) pi[c * N + k] *= K[k];
        oracle_quad3(gq_e, axial, alpha, pi, q);
        for (int k = 0; k < N; k++) {
            int ip = pts[k];
            int pnu1 = d->fpoint_nr[ip] / 2 + 1;
            if (alpha < pnu1) fstiff[d->foffsets[ip] + alpha] -= q[k];
        }
    }
}

/* SolidFluidBoundary::apply -> SolidFluidCoupling::apply, 1D coupling
 * (SolidFluidCoupling.cpp:30-37, SolidFluidCoupling1D.hpp:26-38), per
 * pair and in this order: fluid stiff += ns u_s + nz u_z; solid stiff
 * (s, z) -= (n / m_fluid) fluid stiff.  One rank: the unassembled and
 * assembled normals coincide. */
/* SolidFluidCoupling3D (SolidFluidCoupling3D.cpp:55-94), per slice:
 * fluid stiff += R2C(n . C2R(u_s)); solid stiff -= R2C(C2R(fluid stiff)
 * n / m_fluid), with m_fluid the fluid point's mass per slice (1D or 3D,
 * GLLPoint.cpp:129-143). */
static void coupling3d(const oracle_domain *d, int k, const cplx *displ,
                       cplx *stiff, cplx *fstiff) {
    int sp = d->sf_solid[k], fp = d->sf_fluid[k];
    int nr = d->point_nr[sp];
    int nu1 = nr / 2 + 1;
    const double *nn = d->sf3d_normal + d->sf3d_off[k] * 3;
    const double *m3 = (d->fmass3d_off && d->fmass3d_off[fp] >= 0)
                           ? d->fmass3d + d->fmass3d_off[fp] : NULL;
    const cplx *u = displ + d->offsets[sp] * 3;
    cplx *s = stiff + d->offsets[sp] * 3;
    cplx *f = fstiff + d->foffsets[fp];
    double *r = malloc(sizeof(double) * nr * 4);
    double *fr = r + 3 * nr;
    cplx *col = malloc(sizeof(cplx) * nu1);
    for (int c = 0; c < 3; c++) {
        for (int a = 0; a < nu1; a++) col[a] = u[a * 3 + c];
        oracle_c2r(col, nr, r + c * nr);
    }
    for (int sl = 0; sl < nr; sl++)
        fr[sl] = nn[sl * 3] * r[sl] + nn[sl * 3 + 1] * r[nr + sl] +
                 nn[sl * 3 + 2] * r[2 * nr + sl];
    oracle_r2c(fr, nr, col);
    for (int a = 0; a < nu1; a++) f[a] += col[a];
    oracle_c2r(f, nr, fr);
    for (int c = 0; c < 3; c++) {
        for (int sl = 0; sl < nr; sl++) {
            double mf = m3 ? m3[sl] : d->fpoint_mass[fp];
            r[sl] = fr[sl] * ((1. / mf) * nn[sl * 3 + c]);
        }
        oracle_r2c(r, nr, col);
        for (int a = 0; a < nu1; a++) s[a * 3 + c] -= col[a];
    }
    free(r);
    free(col);
}

void oracle_coupling(const oracle_domain *d, const cplx *displ, cplx *stiff,
                     cplx *fstiff) {
    for (int k = 0; k < d->n_sf; k++) {
        if (d->sf3d_off && d->sf3d_off[k] >= 0) {
            coupling3d(d, k, displ, stiff, fstiff);
            continue;
        }
        int sp = d->sf_solid[k], fp = d->sf_fluid[k];
        int nu1 = d->point_nr[sp] / 2 + 1;
        double ns = d->sf_normal[2 * k], nz = d->sf_normal[2 * k + 1];
        double mf = d->fpoint_mass[fp];
        double ns_m = ns / mf, nz_m = nz / mf;
        const cplx *u = displ + d->offsets[sp] * 3;
        cplx *s = stiff + d->offsets[sp] * 3;
        cplx *f = fstiff + d->foffsets[fp];
        for (int a = 0; a < nu1; a++) {
            f[a] += ns * u[a * 3 + 0];
            f[a] += nz * u[a * 3 + 2];
        }
        for (int a = 0; a < nu1; a++) {
            s[a * 3 + 0] -= ns_m * f[a];
            s[a * 3 + 2] -= nz_m * f[a];
        }
    }
}

/* AxialBoundary::apply, fluid part (AxialBoundary.cpp:44-47: orders >= 1
 * zero), then FluidSurfaceBoundary::apply (FluidSurfaceBoundary.cpp:
 * 20-25: stiff zero). */
void oracle_fluid_bc(const oracle_domain *d, cplx *fstiff) {
    for (int k = 0; k < d->n_faxial; k++) {
        int ip = d->faxial_points[k];
        int nu1 = d->fpoint_nr[ip] / 2 + 1;
        for (int a = 1; a < nu1; a++) fstiff[d->foffsets[ip] + a] = 0.;
    }
    for (int k = 0; k < d->n_fsurf; k++) {
        int ip = d->fsurf_points[k];
        int nu1 = d->fpoint_nr[ip] / 2 + 1;
        for (int a = 0; a < nu1; a++) fstiff[d->foffsets[ip] + a] = 0.;
    }
}

/* FluidPoint::computeStiffToAccel (FluidPoint.cpp:28-53): Nyquist
 * imaginary part zero (even Nr), divide by the mass: Mass1D, or Mass3D
 * for one component (Mass3D.cpp:41-53: C2R, times 1/m per slice, R2C). */
void oracle_fluid_stiff_to_accel(const oracle_domain *d, cplx *fstiff) {
    for (int ip = 0; ip < d->n_fpoints; ip++) {
        int nr = d->fpoint_nr[ip];
        int nu1 = nr / 2 + 1;
        cplx *s = fstiff + d->foffsets[ip];
        if (nr % 2 == 0) s[nu1 - 1] = creal(s[nu1 - 1]);
        if (d->fmass3d_off && d->fmass3d_off[ip] >= 0) {
            const double *m = d->fmass3d + d->fmass3d_off[ip];
            double *r = malloc(sizeof(double) * nr);
            oracle_c2r(s, nr, r);
            for (int sl = 0; sl < nr; sl++) r[sl] *= 1. / m[sl];
            oracle_r2c(r, nr, s);
            free(r);
            continue;
        }
        double inv = 1. / d->fpoint_mass[ip];
        for (int k = 0; k < nu1; k++) s[k] *= inv;
    }
}

/* NewmarkTimeScheme::solve (NewmarkTimeScheme.cpp:40-104) with solid and
 * fluid points: source -> stiffness (solid, fluid) -> solid-fluid
 * coupling -> record -> axial and fluid-surface boundaries -> stiff to
 * accel -> Newmark on solid and fluid points. */
int oracle_run_coupled(const oracle_domain *d, int step_begin, int nsteps,
                       cplx *displ, cplx *veloc, cplx *accel, cplx *stiff,
                       cplx *fdispl, cplx *fveloc, cplx *faccel,
                       cplx *fstiff, double *seismo) {
    int64_t nmf = d->n_fpoints > 0 ? d->foffsets[d->n_fpoints] : 0;
    for (int ts = step_begin; ts < step_begin + nsteps; ts++) {
        if (d->stf) oracle_apply_moment_source(d, d->stf[ts], stiff);
        solid_stiffness(d, displ, stiff);
        for (int e = 0; e < d->n_felem; e++)
            oracle_fluid_element_stiff(d, e, fdispl, fstiff);
        oracle_clayton(d, veloc, stiff, fveloc, fstiff);
        oracle_coupling(d, displ, stiff, fstiff);
        if (seismo && d->n_rec > 0)
            oracle_record(d, displ, seismo + (int64_t)(ts - step_begin) *
                                                 d->n_rec * 3);
        for (int64_t k = 0; k < d->offsets[d->n_points] * 3; k++)
            if (!isfinite(creal(displ[k])) || !isfinite(cimag(displ[k])))
                return 1;
        oracle_axial_solid(d, stiff);
        oracle_fluid_bc(d, fstiff);
        oracle_stiff_to_accel(d, stiff);
        oracle_fluid_stiff_to_accel(d, fstiff);
        oracle_sponge(d, displ, veloc, stiff, fdispl, fveloc, fstiff);
        oracle_newmark(d->offsets[d->n_points], d->dt, stiff, displ, veloc,
                       accel);
        /* one component per mode: the 3-component routine on nmf / 3
         * would not cover a remainder, so march the fluid directly */
        const double hdt = .5 * d->dt, hdt2 = .5 * d->dt * d->dt;
        for (int64_t k = 0; k < nmf; k++) {
            fveloc[k] += hdt * (faccel[k] + fstiff[k]);
            faccel[k] = fstiff[k];
            fdispl[k] += d->dt * fveloc[k] + hdt2 * faccel[k];
            fstiff[k] = 0.;
        }
    }
    return 0;
}

int oracle_run(const oracle_domain *d, int step_begin, int nsteps,
               cplx *displ, cplx *veloc, cplx *accel, cplx *stiff,
               double *seismo) {
    for (int ts = step_begin; ts < step_begin + nsteps; ts++) {
        if (d->stf) oracle_apply_moment_source(d, d->stf[ts], stiff);
        solid_stiffness(d, displ, stiff);
        oracle_clayton(d, veloc, stiff, NULL, NULL);
        if (seismo && d->n_rec > 0)
            oracle_record(d, displ, seismo + (int64_t)(ts - step_begin) *
                                                 d->n_rec * 3);
        for (int64_t k = 0; k < d->offsets[d->n_points] * 3; k++)
            if (!isfinite(creal(displ[k])) || !isfinite(cimag(displ[k])))
                return 1;
        oracle_axial_solid(d, stiff);
        oracle_stiff_to_accel(d, stiff);
        oracle_sponge(d, displ, veloc, stiff, NULL, NULL, NULL);
        oracle_newmark(d->offsets[d->n_points], d->dt, stiff, displ, veloc,
                       accel);
    }
    return 0;
}

/* SymplecticTimeScheme::solve (SymplecticTimeScheme.cpp:30-103) with
 * SymplecticTimeScheme::launch/update (SymplecticTimeScheme.hpp:37-70):
 * per step the receivers, the source, u += kappa_0 dt v, then four
 * sub-steps of stiffness (solid, fluid), Clayton, coupling, axis and
 * fluid surface, stiff to accel, sponge, v += pi_k dt a, u += kappa_k dt v.
 * The coefficients are the position-extended Forest-Ruth-like ones of
 * Omelyan, Mryglod and Folk (2002), SymplecticTimeScheme.hpp:81-100.
 * fdispl .. fstiff may be NULL without fluid points. */
int oracle_run_symplectic(const oracle_domain *d, int step_begin,
                          int nsteps, cplx *displ, cplx *veloc, cplx *stiff,
                          cplx *fdispl, cplx *fveloc, cplx *fstiff,
                          double *seismo) {
    const double A = 0.1786178958448091, B = -0.2123418310626054,
                 G = -0.06626458266981849;
    const double kap[5] = {A, G, 1. - 2. * (G + A), G, A};
    const double pi_[5] = {0., .5 - B, B, B, .5 - B};
    const int64_t nm = d->offsets[d->n_points] * 3;
    const int64_t nmf = d->n_fpoints > 0 ? d->foffsets[d->n_fpoints] : 0;
    for (int ts = step_begin; ts < step_begin + nsteps; ts++) {
        if (seismo && d->n_rec > 0)
            oracle_record(d, displ, seismo + (int64_t)(ts - step_begin) *
                                                 d->n_rec * 3);
        if (d->stf) oracle_apply_moment_source(d, d->stf[ts], stiff);
        for (int64_t k = 0; k < nm; k++) displ[k] += kap[0] * d->dt * veloc[k];
        for (int64_t k = 0; k < nmf; k++)
            fdispl[k] += kap[0] * d->dt * fveloc[k];
        for (int isub = 1; isub <= 4; isub++) {
            solid_stiffness(d, displ, stiff);
            for (int e = 0; e < d->n_felem; e++)
                oracle_fluid_element_stiff(d, e, fdispl, fstiff);
            oracle_clayton(d, veloc, stiff, fveloc, fstiff);
            if (nmf > 0) oracle_coupling(d, displ, stiff, fstiff);
            if (isub == 1)
                for (int64_t k = 0; k < nm; k++)
                    if (!isfinite(creal(displ[k])) ||
                        !isfinite(cimag(displ[k])))
                        return 1;
            oracle_axial_solid(d, stiff);
            if (nmf > 0) oracle_fluid_bc(d, fstiff);
            oracle_stiff_to_accel(d, stiff);
            if (nmf > 0) oracle_fluid_stiff_to_accel(d, fstiff);
            oracle_sponge(d, displ, veloc, stiff, fdispl, fveloc, fstiff);
            const double pdt = pi_[isub] * d->dt, kdt = kap[isub] * d->dt;
            for (int64_t k = 0; k < nm; k++) {
                veloc[k] += pdt * stiff[k];
                displ[k] += kdt * veloc[k];
                stiff[k] = 0.;
            }
            for (int64_t k = 0; k < nmf; k++) {
                fveloc[k] += pdt * fstiff[k];
                fdispl[k] += kdt * fveloc[k];
                fstiff[k] = 0.;
            }
        }
    }
    return 0;
}

int64_t *oracle_make_offsets(int n, const int32_t *nr) {
    return make_offsets(n, nr);
}

void oracle_free(void *p) { free(p); }
