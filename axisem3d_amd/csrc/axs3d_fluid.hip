// axs3d_fluid.hip -- fluid (acoustic) elements and fluid points for gfx950.
//
// The fluid unknown is the displacement potential chi, one complex value
// per (point, Fourier order).  Reference (src = SOLVER/src):
//   FluidElement::displToStiff, 1D acoustic, no PRT
//     (src/core/element/FluidElement.cpp:101-148):
//     computeGrad3 (GradientQuadrature.hpp:163-175) -> K p_i
//     (Acoustic::strainToStress, Acoustic.hpp:80-92) -> computeQuad3
//     (GradientQuadrature.hpp:234-252) -> subtracting scatter
//     (FluidPoint.hpp:97-103)
//   SolidFluidCoupling1D::coupleSolidToFluid (SolidFluidCoupling1D.hpp:
//     26-31), AxialBoundary::apply fluid part (AxialBoundary.cpp:44-47),
//   FluidSurfaceBoundary::apply (FluidSurfaceBoundary.cpp:20-25),
//   FluidPoint::computeStiffToAccel (FluidPoint.cpp:28-53), Newmark.
// Same layout as the Fourier-path solid kernel: one thread per (element,
// order, GLL point), ten element-modes per 256-thread block, LDS tiles
// for the two 5x5 contractions, element-local forces [e][ipnt][alpha]
// gathered by the point kernel in ascending element order.
#include "axs3d_common.hpp"
#include "axs3d_element.hpp"
#include "axs3d_launch.hpp"

namespace axs3d {

template <typename Real>
__global__ __launch_bounds__(STIFF_BLOCK) void k_fluid_stiff_fr(
    FluidArgs<Real> a) {
    __shared__ cpx<Real> sP[EMB][NPEM];
    __shared__ cpx<Real> sX[EMB][NPEM];
    __shared__ cpx<Real> sY[EMB][NPEM];
    __shared__ Real sG[2][NPEM];

    const int t = threadIdx.x;
    if (t < 2 * NPEM) sG[t / NPEM][t % NPEM] = a.gtab[t];
    const int eml = t / NPEM;
    const int ipnt = t - eml * NPEM;
    const int em = xcd_block() * EMB + eml;
    const bool active = (eml < EMB) && (em < a.n_em);
    const int i = ipnt / NPED, j = ipnt - (ipnt / NPED) * NPED;
    const int emx = active ? eml : 0;

    int e = 0, alpha = 0, nu1p = 0;
    uint32_t off = 0;
    bool axial = false;
    cpx<Real> p = {0, 0};
    if (active) {
        const EMItem it = a.em_list[em];
        e = it.elem;
        alpha = it.alpha;
        const int ip = a.elem_points[e * NPEM + ipnt];
        nu1p = a.point_nu1[ip];
        off = a.point_off[ip];
        axial = a.elem_axial[e] != 0;
        if (alpha < nu1p) {
            const double *src = a.chi + (size_t)(off + alpha) * 2;
            p = {(Real)src[0], (Real)src[1]};
        }
        sP[eml][ipnt] = p;
    }
    __syncthreads();

    PointG<Real> G;
    load_point_g<Real>(sG, axial, i, j, G);
    PointOps<Real> g{}, q{};
    const Real ralpha = (Real)alpha;
    const bool axial_row0 = axial && (i == 0);
    cpx<Real> Lt = {0, 0};
    if (active) {
        load_point_ops(a.elem_ops, e, ipnt, g, q);
        // computeGrad3: dp/ds, dp/dz from the two contractions; i alpha p
        // / s, on the axis row its L'Hospital limit i alpha dp/ds
        cpx<Real> x = {0, 0}, y = {0, 0};
#pragma unroll
        for (int k = 0; k < NPED; k++) {
            x = cfma(sP[eml][k * NPED + j], G.gxi_col[k], x);
            y = cfma(sP[eml][i * NPED + k], G.g_col[k], y);
        }
        const cpx<Real> dpds = cadd(cmul(x, g.dzdeta), cmul(y, g.dzdxii));
        const cpx<Real> dpdz = cadd(cmul(x, g.dsdeta), cmul(y, g.dsdxii));
        const cpx<Real> ps = axial_row0 ? ialpha(ralpha, dpds)
                                        : cmul(ialpha(ralpha, p), g.invs);
        // Acoustic: stress = K strain
        const Real K = a.elem_k[(size_t)e * NPEM + ipnt];
        const cpx<Real> q0 = cmul(dpds, K), q1 = cmul(ps, K);
        const cpx<Real> q2 = cmul(dpdz, K);
        // computeQuad3: X, Y from (q0, q2), the 1/s term from i alpha q1,
        // its axial L'Hospital corrections folded into row 0 of X, Y
        cpx<Real> X, Y;
        quad_xy(q0, q2, q, X, Y);
        const cpx<Real> b = ialpha(ralpha, q1);
        Lt = cmul(b, -q.invs);
        if (axial_row0) {
            X = csub(X, cmul(b, q.dzdeta));
            Y = csub(Y, cmul(b, q.dzdxii));
        }
        sX[eml][ipnt] = X;
        sY[eml][ipnt] = Y;
    }
    __syncthreads();
    if (active && alpha < nu1p) {
        cpx<Real> acc = {0, 0}, bsum = {0, 0};
#pragma unroll
        for (int k = 0; k < NPED; k++) {
            acc = cfma(sX[emx][k * NPED + j], G.gxi_row[k], acc);
            bsum = cfma(sY[emx][i * NPED + k], G.g_row[k], bsum);
        }
        const cpx<Real> F = cadd(cadd(acc, bsum), Lt);
        Real *dst = a.eb + ((size_t)a.elem_eb_off[e] +
                            (size_t)ipnt * a.elem_nu1[e] + alpha) * 2;
        dst[0] = F.re;
        dst[1] = F.im;
    }
}

constexpr uint32_t FADJ_NONE = 0xffffffffu;
// slot 3 of a fluid point shared by more than four elements (as ADJ_MORE)
constexpr uint32_t FADJ_MORE = 0x80000000u;

template <typename Real>
__device__ __forceinline__ void gather_fluid(const uint32_t *__restrict__ adj,
                                             const Real *__restrict__ eb,
                                             int p, int alpha, Real (&f)[2]) {
    f[0] = f[1] = 0;
    const uint4 a4 = *reinterpret_cast<const uint4 *>(adj + (size_t)p * 4);
    const uint32_t base[4] = {a4.x, a4.y, a4.z, a4.w};
    auto add = [&](uint32_t b) {
        const Real *src = eb + ((size_t)b + alpha) * 2;
        f[0] -= src[0];
        f[1] -= src[1];
    };
#pragma unroll
    for (int k = 0; k < 3; k++)
        if (base[k] != FADJ_NONE) add(base[k]);
    if (base[3] != FADJ_NONE) {
        if (base[3] & FADJ_MORE) {
            for (uint32_t j = base[3] & ~FADJ_MORE; adj[j] != FADJ_NONE; j++)
                add(adj[j]);
        } else {
            add(base[3]);
        }
    }
}

template <typename Real>
__global__ __launch_bounds__(256) void k_fluid_assemble(
    int64_t n_modes, const int32_t *__restrict__ mode_point,
    const uint32_t *__restrict__ point_off,
    const uint32_t *__restrict__ adj, const Real *__restrict__ eb,
    Real *__restrict__ stiff) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_modes) return;
    const int p = mode_point[k];
    const int alpha = (int)(k - (int64_t)point_off[p]);
    Real f[2];
    gather_fluid(adj, eb, p, alpha, f);
    stiff[(size_t)k * 2] += f[0];
    stiff[(size_t)k * 2 + 1] += f[1];
}

struct FluidPointArgs {
    int64_t n_modes;
    const int32_t *mode_point;
    const uint32_t *point_off;
    const int32_t *point_nr;
    const uint8_t *point_bc;       // bit 0 axis, bit 1 free surface
    const double *inv_mass;
    const uint32_t *adj;
    double *chi, *chi_v, *chi_a;
    double dt, half_dt, half_dt_dt;
    int *flag;
    CouplingGeom cpl;
    const double *solid_displ;
    const uint8_t *defer;          // [point] 1: finished by k_fluid_defer
    AbcGeom abc;                   // Clayton and sponge terms (k_abc)
    int scheme;                    // 0 Newmark, 1 symplectic sub-step
    double sym_pi, sym_kappa;
};

// Fluid point update, one thread per (point, order), phases as for the
// solid points; COUPLE adds the normal solid displacement of the coupled
// solid point (coupleSolidToFluid) and leaves the result in cstiff for
// the solid points' coupleFluidToSolid, after the Clayton boundaries
// (ClaytonFluid1D/3D: s -= the k_abc term); AXIAL applies the axis (orders
// >= 1 zero) and the free surface (all zero); SPONGE alone subtracts the
// sponge term from the acceleration held in stiff (the fused MASS | MARCH
// form does so between the mass and the march).
template <typename Real, bool GATHER, bool AXIAL, bool MASS, bool MARCH,
          bool COUPLE, bool SPONGE>
__global__ __launch_bounds__(256) void k_fluid_point(
    FluidPointArgs a, const Real *__restrict__ eb, Real *__restrict__ stiff) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n_modes) return;
    const int p = a.mode_point[k];
    const int alpha = (int)(k - (int64_t)a.point_off[p]);
    const int ai = a.abc.idx != nullptr ? a.abc.idx[p] : -1;
    double s[2];
    if (GATHER) {
        Real f[2];
        gather_fluid(a.adj, eb, p, alpha, f);
        s[0] = (double)f[0];
        s[1] = (double)f[1];
    } else {
        s[0] = (double)stiff[(size_t)k * 2];
        s[1] = (double)stiff[(size_t)k * 2 + 1];
    }
    if (SPONGE) {
        if (ai >= 0) {
            const int nr = a.point_nr[p];
            sponge_sub<1>(a.abc, ai, alpha, (nr & 1) ? -1 : nr / 2, s);
            stiff[(size_t)k * 2] = (Real)s[0];
            stiff[(size_t)k * 2 + 1] = (Real)s[1];
        }
        return;
    }
    if (COUPLE && ai >= 0) {
        const double *c =
            a.abc.cterm + ((size_t)a.abc.term_off[ai] + alpha) * 2;
        s[0] -= c[0];
        s[1] -= c[1];
    }
    if (COUPLE && a.cpl.n_pairs > 0) {
        const int pi = a.cpl.fluid_pair[p];
        if (pi >= 0) {
            const double *u = a.solid_displ +
                              ((size_t)a.cpl.soff[pi] + alpha) * 6;
            const double ns = a.cpl.n[2 * pi], nz = a.cpl.n[2 * pi + 1];
            s[0] += ns * u[0];
            s[1] += ns * u[1];
            s[0] += nz * u[4];
            s[1] += nz * u[5];
            a.cpl.cstiff[2 * (size_t)k] = s[0];
            a.cpl.cstiff[2 * (size_t)k + 1] = s[1];
        }
    }
    if (a.defer != nullptr && a.defer[p]) {
        // a 3D mass or a 3D coupling: the remaining phases follow in
        // k_fluid_defer, from the stiffness gathered so far
        stiff[(size_t)k * 2] = (Real)s[0];
        stiff[(size_t)k * 2 + 1] = (Real)s[1];
        return;
    }
    if (AXIAL) {
        const uint8_t bc = a.point_bc[p];
        if (((bc & 1) && alpha >= 1) || (bc & 2)) s[0] = s[1] = 0;
    }
    if (MASS) {
        const int nr = a.point_nr[p];
        if ((nr & 1) == 0 && alpha == nr / 2) s[1] = 0;
        const double im = a.inv_mass[p];
        s[0] *= im;
        s[1] *= im;
        if (MARCH && ai >= 0)
            sponge_sub<1>(a.abc, ai, alpha, (nr & 1) ? -1 : nr / 2, s);
    }
    if (!MARCH) {
        stiff[(size_t)k * 2] = (Real)s[0];
        stiff[(size_t)k * 2 + 1] = (Real)s[1];
        return;
    }
    double2 *up = reinterpret_cast<double2 *>(a.chi + (size_t)k * 2);
    double2 *vp = reinterpret_cast<double2 *>(a.chi_v + (size_t)k * 2);
    double2 *ap = reinterpret_cast<double2 *>(a.chi_a + (size_t)k * 2);
    if (a.scheme == 1) {
        // SymplecticTimeScheme::update: v += pi dt a, u += kappa dt v
        const double2 u0 = *up, v0 = *vp;
        const double pdt = a.sym_pi * a.dt, kdt = a.sym_kappa * a.dt;
        const double vr = v0.x + pdt * s[0], vi = v0.y + pdt * s[1];
        const double ur = u0.x + kdt * vr, ui = u0.y + kdt * vi;
        *vp = make_double2(vr, vi);
        *up = make_double2(ur, ui);
        const bool bad = !isfinite(ur) || !isfinite(ui);
        const bool bad0 = !isfinite(u0.x) || !isfinite(u0.y);
        if (!GATHER) stiff[(size_t)k * 2] = stiff[(size_t)k * 2 + 1] = 0;
        if (bad0 || bad) atomicOr(a.flag, (bad0 ? 1 : 0) | (bad ? 2 : 0));
        return;
    }
    // all three loaded (16-byte vectors) before anything is stored
    const double2 u0 = *up, v0 = *vp, a0 = *ap;
    const double vr = v0.x + a.half_dt * (a0.x + s[0]);
    const double vi = v0.y + a.half_dt * (a0.y + s[1]);
    const double ur = u0.x + (a.dt * vr + a.half_dt_dt * s[0]);
    const double ui = u0.y + (a.dt * vi + a.half_dt_dt * s[1]);
    *vp = make_double2(vr, vi);
    *ap = make_double2(s[0], s[1]);
    *up = make_double2(ur, ui);
    const bool bad = !isfinite(ur) || !isfinite(ui);
    const bool bad0 = !isfinite(u0.x) || !isfinite(u0.y);
    if (!GATHER) stiff[(size_t)k * 2] = stiff[(size_t)k * 2 + 1] = 0;
    // bit 0: started from a non-finite potential, bit 1: left one
    if (bad0 || bad) atomicOr(a.flag, (bad0 ? 1 : 0) | (bad ? 2 : 0));
}

template <typename Real>
void launch_fluid_stiff_fr(hipStream_t st, const FluidArgs<Real> &a) {
    if (a.n_em <= 0) return;
    const int grid = (a.n_em + EMB - 1) / EMB;
    hipLaunchKernelGGL(k_fluid_stiff_fr<Real>, dim3(grid), dim3(STIFF_BLOCK),
                       0, st, a);
}

template <typename Real>
void launch_fluid_assemble(hipStream_t st, const PointGeom &g,
                           const Real *eb, Real *stiff) {
    if (g.n_modes <= 0) return;
    const int64_t grid = (g.n_modes + 255) / 256;
    hipLaunchKernelGGL(k_fluid_assemble<Real>, dim3(grid), dim3(256), 0, st,
                       g.n_modes, g.mode_point, g.point_off, g.adj, eb,
                       stiff);
}

template <typename Real>
void launch_fluid_point(hipStream_t st, int phases, const PointGeom &g,
                        const Real *eb, Real *stiff, double *chi,
                        double *chi_v, double *chi_a, double dt, int *flag,
                        const CouplingGeom &cpl, const double *solid_displ) {
    if (g.n_modes <= 0) return;
    const int64_t grid = (g.n_modes + 255) / 256;
    FluidPointArgs a;
    a.n_modes = g.n_modes;
    a.mode_point = g.mode_point;
    a.point_off = g.point_off;
    a.point_nr = g.point_nr;
    a.point_bc = g.point_axial;
    a.inv_mass = g.inv_mass;
    a.adj = g.adj;
    a.chi = chi;
    a.chi_v = chi_v;
    a.chi_a = chi_a;
    a.dt = dt;
    a.half_dt = .5 * dt;
    a.half_dt_dt = .5 * dt * dt;
    a.flag = flag;
    a.cpl = cpl;
    a.solid_displ = solid_displ;
    a.defer = g.mass3d;
    a.abc = g.abc;
    a.scheme = g.scheme;
    a.sym_pi = g.sym_pi;
    a.sym_kappa = g.sym_kappa;
    if (cpl.n_pairs <= 0 && g.abc.idx == nullptr) phases &= ~PHASE_COUPLE;
    if (g.abc.idx == nullptr) phases &= ~PHASE_SPONGE;
    if (phases == 0) return;
#define AXS3D_FPOINT_CASE(G, A, M, T, C, S)                                   \
    case (A) | ((M) << 1) | ((T) << 2) | ((G) << 3) | ((C) << 4) |           \
        ((S) << 5):                                                          \
        hipLaunchKernelGGL((k_fluid_point<Real, G, A, M, T, C, S>),            \
                           dim3(grid), dim3(256), 0, st, a, eb, stiff);       \
        break;
    switch (phases) {
        AXS3D_FPOINT_CASE(0, 1, 0, 0, 0, 0)
        AXS3D_FPOINT_CASE(0, 0, 1, 0, 0, 0)
        AXS3D_FPOINT_CASE(0, 0, 0, 1, 0, 0)
        AXS3D_FPOINT_CASE(0, 0, 0, 0, 1, 0)
        AXS3D_FPOINT_CASE(0, 0, 0, 0, 0, 1)
        AXS3D_FPOINT_CASE(1, 1, 1, 1, 0, 0)
        AXS3D_FPOINT_CASE(1, 1, 1, 1, 1, 0)
        default:
            break;
    }
#undef AXS3D_FPOINT_CASE
}

#define AXS3D_FLUID_INSTANTIATE(Real)                                          \
    template void launch_fluid_stiff_fr<Real>(hipStream_t,                    \
                                              const FluidArgs<Real> &);      \
    template void launch_fluid_assemble<Real>(hipStream_t, const PointGeom &, \
                                              const Real *, Real *);         \
    template void launch_fluid_point<Real>(                                   \
        hipStream_t, int, const PointGeom &, const Real *, Real *, double *,  \
        double *, double *, double, int *, const CouplingGeom &,              \
        const double *);

AXS3D_FLUID_INSTANTIATE(float)
AXS3D_FLUID_INSTANTIATE(double)

}  // namespace axs3d
