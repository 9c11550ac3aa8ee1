// axs3d_adjoint.hip -- device code of the adjoint (sensitivity) run.
//
// The reference has no adjoint solver; this is the discrete adjoint of its
// time loop (NewmarkTimeScheme.cpp:41-108, the central-difference form of
// NewmarkTimeScheme.hpp:48-71) for isotropic elastic elements
// (Isotropic.hpp:91-124); this file holds the densities of the Fourier-path
// elements and of the mass, and the adjoint force (the per-slice densities
// of 3D elements are k_cd_sens).  With u_{n+1} = 2 u_n - u_{n-1} +
// dt^2 M^-1 P (f_n - K u_n), receivers r_n = R u_n and dJ/dr_n = G_n, the
// adjoint field psi obeys the same recursion backwards in time with the
// force R^* G_n / dt^2, and
//   dJ/dtheta = -dt^2 sum_n ( <psi_n, dK/dtheta u_n> + <psi_n, dM/dtheta a_n> )
// in the inner product <x, y> = Re sum_p sum_alpha w_alpha conj(x) y, w_0 =
// 1, w_alpha = 2 (the weights of the receiver synthesis, eigen_tools.hpp:
// 79-106), in which K, M and the axis/Nyquist projection P are
// self-adjoint.  For lambda, mu and rho at GLL point k of element e:
//   <psi, dK/dlambda u> = IF w Re conj(tr eps(psi)) tr eps(u)
//   <psi, dK/dmu u>     = IF w Re (2 sum_ii conj eps_ii(psi) eps_ii(u) +
//                                  sum_shear conj gam(psi) gam(u))
//   <psi, dM/drho a>    = IF w Re conj(psi_p) . a_p
// (the element stiffness is Quad(sigma(Grad u)) with Quad the integral-
// factor-weighted transpose of Grad, GradientQuadrature.hpp:207-324; the
// mass is sum rho IF, GLLPoint.cpp:54-77).  With SLS attenuation the
// stress is C eps_n - sum_{j>=1} h_j D eps_{n-j} with h_j fixed by (alpha,
// beta, gamma) and D symmetric (Attenuation.hpp:92-155): the adjoint field
// runs the same recursion with memory variables of its own, and the
// densities above are those of the unrelaxed moduli in C.
#include <hip/hip_runtime.h>

#include "axs3d_common.hpp"
#include "axs3d_element.hpp"
#include "axs3d_launch.hpp"

namespace axs3d {

namespace {

__device__ __forceinline__ double rdot(cpx<double> x, cpx<double> y) {
    return x.re * y.re + x.im * y.im;   // Re conj(x) y
}

// One block per element, 10 Fourier orders x 25 GLL points per pass over
// the element's orders; the densities of each GLL point are summed over
// the orders in a fixed order and added to the kernels: lambda, mu, rho
// to kern[e][k][0..2] (isotropic domains), or the parameters of the
// element's material type (mat_density) to kernm and rho to kernr.
template <int MT>
__device__ __forceinline__ void sens_element(
    const SensArgs &a, int e, cpx<double> (*sU)[3][NPEM],
    cpx<double> (*sP)[3][NPEM], const double (*sG)[NPEM],
    double (*red)[NPEM]) {
    constexpr int NP = MT == 0 ? 2 : MT == 1 ? 5 : 21;
    const int t = threadIdx.x;
    const int slot = t / NPEM;
    const int ipnt = t - slot * NPEM;
    const bool lane_ok = slot < EMB;
    const int i = ipnt / NPED, j = ipnt - (ipnt / NPED) * NPED;
    const bool axial = a.elem_axial[e] != 0;
    const bool axial_row0 = axial && (i == 0);
    const int nu1e = a.elem_nu1[e];
    // 3D elements: only the mass density here (k_cd_sens does the moduli)
    const bool kind3d = a.elem_kind != nullptr && a.elem_kind[e] == 1;
    const int p = lane_ok ? a.elem_points[e * NPEM + ipnt] : 0;
    const int nu1p = lane_ok ? a.point_nu1[p] : 0;
    const size_t poff = lane_ok ? a.point_off[p] : 0;
    PointOps<double> g{};
    double ifac = 0.;
    double rs1 = 0., rc1 = 1., rs2 = 0., rc2 = 1.;
    const bool rtz = MT != 0 && a.rot != nullptr;
    if (lane_ok) {
        const double *gq = a.gq + (size_t)e * 6 * NPEM + ipnt;
        g.dsdxii = gq[0];
        g.dsdeta = gq[NPEM];
        g.dzdxii = gq[2 * NPEM];
        g.dzdeta = gq[3 * NPEM];
        g.invs = gq[4 * NPEM];
        ifac = gq[5 * NPEM];
        if (rtz) {
            const double *r = a.rot + (size_t)e * 4 * NPEM + ipnt;
            rs1 = r[0];
            rc1 = r[NPEM];
            rs2 = r[2 * NPEM];
            rc2 = r[3 * NPEM];
        }
    }
    PointG<double> G;
    load_point_g<double>(sG, axial, i, j, G);
    double acc[NP];
#pragma unroll
    for (int c = 0; c < NP; c++) acc[c] = 0.;
    double kr = 0.;
    for (int a0 = 0; a0 < nu1e; a0 += EMB) {
        const int alpha = a0 + slot;
        const bool act = lane_ok && alpha < nu1e;
        cpx<double> u[3] = {{0, 0}, {0, 0}, {0, 0}};
        cpx<double> ps[3] = {{0, 0}, {0, 0}, {0, 0}};
        cpx<double> ac[3] = {{0, 0}, {0, 0}, {0, 0}};
        if (act && alpha < nu1p) {
            const size_t off = (poff + alpha) * 6;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                u[c] = {a.u[off + 2 * c], a.u[off + 2 * c + 1]};
                ps[c] = {a.psi[off + 2 * c], a.psi[off + 2 * c + 1]};
                ac[c] = {a.a[off + 2 * c], a.a[off + 2 * c + 1]};
            }
        }
        if (lane_ok) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                sU[slot][c][ipnt] = u[c];
                sP[slot][c][ipnt] = ps[c];
            }
        }
        __syncthreads();
        if (act) {
            const double w = alpha == 0 ? 1. : 2.;
            cpx<double> eu[6], ep[6];
            grad6_point<double>(sU[slot], u, (double)alpha, axial_row0, i, j,
                                G, g, eu);
            grad6_point<double>(sP[slot], ps, (double)alpha, axial_row0, i,
                                j, G, g, ep);
            if (!kind3d) {
                if (rtz) {
                    rot_spz_rtz6<cpx<double>, double>(eu, rs1, rc1, rs2, rc2);
                    rot_spz_rtz6<cpx<double>, double>(ep, rs1, rc1, rs2, rc2);
                }
                mat_density<MT>(ep, eu, rdot,
                                [&](int c, double v) { acc[c] += w * v; });
            }
            kr += w * (rdot(ps[0], ac[0]) + rdot(ps[1], ac[1]) +
                       rdot(ps[2], ac[2]));
        }
        __syncthreads();
    }
    // sum over the order slots, one value at a time: the moduli, then rho
#pragma unroll
    for (int c = 0; c <= NP; c++) {
        if (lane_ok) red[slot][ipnt] = (c < NP ? acc[c] : kr) * ifac;
        __syncthreads();
        if (t < NPEM) {
            double sum = 0.;
            for (int s = 0; s < EMB; s++) sum += red[s][t];
            double *dst;
            if (a.kernm == nullptr)
                dst = a.kern + ((size_t)e * NPEM + t) * 3 + (c < NP ? c : 2);
            else if (c < NP)
                dst = a.kernm + a.mat_off[e] + (size_t)c * NPEM + t;
            else
                dst = a.kernr + (size_t)e * NPEM + t;
            *dst += a.scale * sum;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(STIFF_BLOCK) void k_sensitivity(SensArgs a) {
    __shared__ cpx<double> sU[EMB][3][NPEM];
    __shared__ cpx<double> sP[EMB][3][NPEM];
    __shared__ double sG[2][NPEM];
    __shared__ double red[EMB][NPEM];
    const int t = threadIdx.x;
    if (t < 2 * NPEM) sG[t / NPEM][t % NPEM] = a.gtab[t];
    __syncthreads();
    const int e = blockIdx.x;
    const int mt = a.elem_mtype != nullptr ? a.elem_mtype[e] : 0;
    if (mt == 0)
        sens_element<0>(a, e, sU, sP, sG, red);
    else if (mt == 1)
        sens_element<1>(a, e, sU, sP, sG, red);
    else
        sens_element<2>(a, e, sU, sP, sG, red);
}

// The adjoint force of one step, R^* G / dt^2, into the receivers'
// element-local forces (which the point gather subtracts, so the force is
// subtracted here): one thread per (GLL point, order); each thread walks
// the receivers in order, so receivers sharing an element do not race.
template <typename Real>
__global__ __launch_bounds__(256) void k_adjoint_source(
    int nrec, const int32_t *__restrict__ rec_elem,
    const double *__restrict__ rec_w, const double *__restrict__ rec_phi,
    const double *__restrict__ g, double scale, int nu1max,
    const int32_t *__restrict__ elem_points,
    const int32_t *__restrict__ elem_nu1,
    const int32_t *__restrict__ point_nu1,
    const uint32_t *__restrict__ elem_eb_off, Real *__restrict__ eb) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NPEM * nu1max) return;
    const int ipnt = t % NPEM;
    const int alpha = t / NPEM;
    for (int r = 0; r < nrec; r++) {
        const int e = rec_elem[r];
        if (alpha >= elem_nu1[e]) continue;
        const int p = elem_points[e * NPEM + ipnt];
        if (alpha >= point_nu1[p]) continue;
        const double w = rec_w[r * NPEM + ipnt] * scale;
        if (w == 0.) continue;
        double sn, cs;
        sincos((double)alpha * rec_phi[r], &sn, &cs);
        Real *dst = eb + ((size_t)elem_eb_off[e] +
                          (size_t)ipnt * elem_nu1[e] + alpha) * 6;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            // w G e^{-i alpha phi}
            const double gv = g[r * 3 + c] * w;
            dst[2 * c] = (Real)((double)dst[2 * c] - gv * cs);
            dst[2 * c + 1] = (Real)((double)dst[2 * c + 1] + gv * sn);
        }
    }
}

}  // namespace

void launch_sensitivity(hipStream_t st, const SensArgs &a) {
    if (a.n_elem <= 0) return;
    hipLaunchKernelGGL(k_sensitivity, dim3(a.n_elem), dim3(STIFF_BLOCK), 0,
                       st, a);
}

template <typename Real>
void launch_adjoint_source(hipStream_t st, int nrec, const int32_t *rec_elem,
                           const double *rec_w, const double *rec_phi,
                           const double *g, double scale, int nu1max,
                           const int32_t *elem_points,
                           const int32_t *elem_nu1,
                           const int32_t *point_nu1,
                           const uint32_t *elem_eb_off, Real *eb) {
    if (nrec <= 0 || nu1max <= 0) return;
    const int n = NPEM * nu1max;
    hipLaunchKernelGGL((k_adjoint_source<Real>), dim3((n + 255) / 256),
                       dim3(256), 0, st, nrec, rec_elem, rec_w, rec_phi, g,
                       scale, nu1max, elem_points, elem_nu1, point_nu1,
                       elem_eb_off, eb);
}

#define AXS3D_ADJ_INST(Real)                                                  \
    template void launch_adjoint_source<Real>(                                \
        hipStream_t, int, const int32_t *, const double *, const double *,    \
        const double *, double, int, const int32_t *, const int32_t *,        \
        const int32_t *, const uint32_t *, Real *);
AXS3D_ADJ_INST(float)
AXS3D_ADJ_INST(double)

}  // namespace axs3d
