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
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

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

/* ------------------------------------------------------------------ */
/* Domain / time loop                                                  */
/* ------------------------------------------------------------------ */
static int64_t *make_offsets(int n, const int32_t *nr) {
    int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
    off[0] = 0;
    for (int i = 0; i < n; i++) off[i + 1] = off[i] + nr[i] / 2 + 1;
    return off;
}

/* SolidElement::computeStiff (SolidElement.cpp:213-222) for the Fourier
 * path of a 1D isotropic element, no PRT:
 * collectDisplFromPoints (SolidElement.cpp:102-109, zero-padded,
 * SolidPoint.hpp:82-98) -> computeGrad6 -> strainToStress_FR per mode ->
 * computeQuad6 -> addStiffToPoints (subtract, SolidPoint.hpp:101-109). */
void oracle_solid_element_stiff(const oracle_domain *d, int e,
                                const cplx *displ, cplx *stiff) {
    const int32_t *pts = d->elem_points + (int64_t)e * N;
    int nu1 = d->elem_nr[e] / 2 + 1;
    const double *gq_e = d->elem_gq + (int64_t)e * 6 * N;
    const double *lam = d->elem_iso + (int64_t)e * 2 * N;
    const double *mu = lam + N;
    int axial = d->elem_axial[e];
    cplx u[3 * N], eps[6 * N], sig[6 * N], f[3 * N];
    for (int alpha = 0; alpha < nu1; alpha++) {
        for (int p = 0; p < N; p++) {
            int ip = pts[p];
            int pnu1 = d->point_nr[ip] / 2 + 1;
            for (int c = 0; c < 3; c++)
                u[c * N + p] = alpha < pnu1
                    ? displ[(d->offsets[ip] + alpha) * 3 + c] : 0.;
        }
        oracle_grad6(gq_e, axial, alpha, u, eps);
        oracle_iso_stress(lam, mu, eps, sig);
        oracle_quad6(gq_e, axial, alpha, sig, f);
        for (int p = 0; p < N; p++) {
            int ip = pts[p];
            int pnu1 = d->point_nr[ip] / 2 + 1;
            if (alpha < pnu1)
                for (int c = 0; c < 3; c++)
                    stiff[(d->offsets[ip] + alpha) * 3 + c] -= f[c * N + p];
        }
    }
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
            oracle_quad6_noint(gq_e, axial, alpha, sig, f);
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
int oracle_run(const oracle_domain *d, int step_begin, int nsteps,
               cplx *displ, cplx *veloc, cplx *accel, cplx *stiff,
               double *seismo) {
    for (int ts = step_begin; ts < step_begin + nsteps; ts++) {
        if (d->stf) oracle_apply_moment_source(d, d->stf[ts], stiff);
        for (int e = 0; e < d->n_elem; e++)
            oracle_solid_element_stiff(d, e, displ, stiff);
        if (seismo && d->n_rec > 0)
            oracle_record(d, displ, seismo + (int64_t)(ts - step_begin) *
                                                 d->n_rec * 3);
        for (int64_t k = 0; k < d->offsets[d->n_points] * 3; k++)
            if (!isfinite(creal(displ[k])) || !isfinite(cimag(displ[k])))
                return 1;
        oracle_axial_solid(d, stiff);
        oracle_stiff_to_accel(d, stiff);
        oracle_newmark(d->offsets[d->n_points], d->dt, stiff, displ, veloc,
                       accel);
    }
    return 0;
}

int64_t *oracle_make_offsets(int n, const int32_t *nr) {
    return make_offsets(n, nr);
}

void oracle_free(void *p) { free(p); }
