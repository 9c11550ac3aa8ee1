// axs3d_kernels.hip -- CDNA4 (gfx950) kernels of the stiffness path.
//
// Fourier-path solid element stiffness (1D material, no PRT): one thread
// per (element, Fourier order alpha, GLL point); ten element-modes per
// 256-thread block (four wave64s).  The element-mode displacement tile and
// the quadrature pre-products are staged in LDS so the two 5x5 GLL
// contractions (gradient and quadrature) read neighbours from LDS instead
// of HBM.  The element kernel writes its element-local forces F (no
// atomics); the point kernel then gathers, for every (point, mode), the
// contributions of the elements sharing the point in ascending element
// order and subtracts them, the order of SolidPoint::gatherStiffFromElement
// over Domain::computeStiffness's element loop.  Element-local force
// buffer layout: [element][ipnt][alpha][3] complex, so a point's modes are
// contiguous for the gather.
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
#include "axs3d_element.hpp"
#include "axs3d_launch.hpp"

namespace axs3d {

// CG4 attenuation points (1,1),(1,3),(3,1),(3,3) (AttBuilder.cpp:141-147).
__device__ __forceinline__ int att_slot_fr(int p, int npa) {
    if (npa == NPEM) return p;
    return p == 6 ? 0 : p == 8 ? 1 : p == 16 ? 2 : p == 18 ? 3 : -1;
}

// MT: material type of the element-modes [em0, em1) of em_list (0
// isotropic, 1 TI, 2 anisotropic); em_list is grouped by type.
// AX: the element-modes [em0, em1) all belong to axial elements (GLJ in
// xi, L'Hospital terms on the axis row) or all to non-axial ones; the
// element-mode list keeps the axial elements of each type at its end, so
// that the axis code and its registers stay out of the launch that does
// nearly all the work.
// PRT: some elements carry a particle relabelling transformation
// (SolidElement::displToStiff, SolidElement.cpp:115-159): theirs take the
// 9-component gradient, rotated to RTZ on spherical meshes whatever the
// material, to the undulated strain (one operand set for all orders: the
// undulation is axisymmetric), and the stress back through the transpose
// and computeQuad9.
template <typename Real, bool ATT, int MT, bool AX, bool PRT>
__global__ __launch_bounds__(STIFF_BLOCK) void k_solid_stiff_fr(
    StiffArgs<Real> a, int em0, int em1) {
    __shared__ cpx<Real> sU[EMB][3][NPEM];
    __shared__ cpx<Real> sX[EMB][3][NPEM];
    __shared__ cpx<Real> sY[EMB][3][NPEM];
    __shared__ Real sG[2][NPEM];

    const int t = threadIdx.x;
    if (t < 2 * NPEM) sG[t / NPEM][t % NPEM] = a.gtab[t];
    const int eml = t / NPEM;
    const int ipnt = t - eml * NPEM;
    const int em = em0 + xcd_block() * EMB + eml;
    const bool active = (eml < EMB) && (em < em1);
    const int i = ipnt / NPED, j = ipnt - (ipnt / NPED) * NPED;
    const int emx = active ? eml : 0;

    int e = 0, alpha = 0, nu1p = 0;
    uint32_t off = 0;
    bool axial = false;
    cpx<Real> u[3] = {{0, 0}, {0, 0}, {0, 0}};
    if (active) {
        const EMItem it = a.em_list[em];
        e = it.elem;
        alpha = it.alpha;
        const int p = a.elem_points[e * NPEM + ipnt];
        nu1p = a.point_nu1[p];
        off = a.point_off[p];
        axial = AX;
        if (alpha < nu1p) {
            const double *src = a.displ + (size_t)(off + alpha) * 6;
#pragma unroll
            for (int c = 0; c < 3; c++)
                u[c] = {(Real)src[2 * c], (Real)src[2 * c + 1]};
        }
#pragma unroll
        for (int c = 0; c < 3; c++) sU[eml][c][ipnt] = u[c];
    }
    __syncthreads();

    PointG<Real> G;
    load_point_g<Real>(sG, axial, i, j, G);
    cpx<Real> sig[6];
    PointOps<Real> g{}, q{};
    const Real ralpha = (Real)alpha;
    const bool axial_row0 = axial && (i == 0);
    const Real *xp = nullptr;
    if (PRT && active && a.prt_idx[e] >= 0)
        xp = a.prt + (size_t)a.prt_idx[e] * 8 * NPEM + ipnt;
    Real rs1 = 0, rc1 = 1, rs2 = 0, rc2 = 1;
    if (active) {
        load_point_ops(a.elem_ops, e, ipnt, g, q);
        cpx<Real> eps[6];
        const Real *mat = a.elem_mat + a.mat_off[e] + ipnt;
        // TI and anisotropic media work in RTZ on spherical meshes
        // (SolidElement.cpp:177-193)
        const bool rtz = MT != 0 && a.rot != nullptr && xp == nullptr;
        if ((rtz || xp != nullptr) && a.rot != nullptr) {
            const Real *r = a.rot + (size_t)e * 4 * NPEM + ipnt;
            rs1 = r[0];
            rc1 = r[NPEM];
            rs2 = r[2 * NPEM];
            rc2 = r[3 * NPEM];
        }
        if (PRT && xp != nullptr) {
            cpx<Real> g9[9];
            grad9_point<Real>(sU[eml], u, ralpha, axial_row0, i, j, G, g, g9);
            if (a.rot != nullptr)
                rot9<cpx<Real>, Real>(g9, rs1, rc1, rs2, rc2);
            prt_sph_to_und6<cpx<Real>, Real>(g9, xp, eps);
        } else {
            grad6_point<Real>(sU[eml], u, ralpha, axial_row0, i, j, G, g, eps);
            if (rtz) rot_spz_rtz6<cpx<Real>, Real>(eps, rs1, rc1, rs2, rc2);
        }
        if constexpr (MT == 0) {
            iso_hooke<Real>(eps, mat[0], mat[NPEM], sig);
        } else if constexpr (MT == 1) {
            ti_hooke<cpx<Real>, Real>(eps, mat[0], mat[NPEM], mat[2 * NPEM],
                                      mat[3 * NPEM], mat[4 * NPEM], sig);
        } else {
            aniso_hooke<cpx<Real>, Real>(
                eps, [&](int k) { return mat[k * NPEM]; }, sig);
        }
        if (ATT) {
            // AttenuationFull/CG4::apply per Fourier order with 1D moduli
            // (AttenuationCG4.hpp:312-330, Attenuation.hpp:92-140):
            // sigma -= sum_l m_l; m_l = alpha_l m_l + beta_l ds_old;
            // ds = iso(dlambda, dmu)(eps); m_l += gamma_l ds
            const int npa = a.npa, nsls = a.nsls;
            const int qa = att_slot_fr(ipnt, npa);
            if (qa >= 0) {
                const Real *dlm = a.att1d + (size_t)e * 2 * npa + qa;
                cpx<Real> dsn[6];
                iso_hooke<Real>(eps, dlm[0], dlm[npa], dsn);
                // the memory variables and their recursion in FP64 (the
                // stress increments in the working precision)
                cpx<double> *mv = reinterpret_cast<cpx<double> *>(a.mv_fr) +
                                  (size_t)em * nsls * 6 * npa + qa;
                cpx<Real> *ds = reinterpret_cast<cpx<Real> *>(a.ds_fr) +
                                (size_t)em * 6 * npa + qa;
                cpx<double> sd[6];
#pragma unroll
                for (int d = 0; d < 6; d++)
                    sd[d] = {(double)sig[d].re, (double)sig[d].im};
                for (int l = 0; l < nsls; l++)
#pragma unroll
                    for (int d = 0; d < 6; d++)
                        sd[d] = csub(sd[d], mv[(l * 6 + d) * npa]);
#pragma unroll
                for (int d = 0; d < 6; d++)
                    sig[d] = {(Real)sd[d].re, (Real)sd[d].im};
                cpx<double> dso[6], dn[6];
#pragma unroll
                for (int d = 0; d < 6; d++) {
                    dso[d] = {(double)ds[d * npa].re, (double)ds[d * npa].im};
                    dn[d] = {(double)dsn[d].re, (double)dsn[d].im};
                }
                for (int l = 0; l < nsls; l++) {
                    const double al = a.abg[l];
                    const double be = a.abg[nsls + l];
                    const double ga = a.abg[2 * nsls + l];
#pragma unroll
                    for (int d = 0; d < 6; d++) {
                        cpx<double> &m = mv[(l * 6 + d) * npa];
                        cpx<double> x = cmul(m, al);
                        x = cfma(dso[d], be, x);
                        m = cfma(dn[d], ga, x);
                    }
                }
#pragma unroll
                for (int d = 0; d < 6; d++) ds[d * npa] = dsn[d];
            }
        }
        if (rtz) rot_rtz_spz6<cpx<Real>, Real>(sig, rs1, rc1, rs2, rc2);
    }
    cpx<Real> F[3];
    if (PRT) {
        cpx<Real> L[3];
        if (active && xp != nullptr) {
            cpx<Real> f9[9];
            prt_und_to_sph9<cpx<Real>, Real>(sig, xp + 4 * NPEM, f9);
            if (a.rot != nullptr)
                rot9<cpx<Real>, Real>(f9, -rs1, rc1, -rs2, rc2);
            quad9_prep<Real>(f9, ralpha, axial_row0, q, sX[emx], sY[emx],
                             ipnt, L);
        } else if (active) {
            quad6_prep<Real>(sig, ralpha, axial_row0, q, sX[emx], sY[emx],
                             ipnt, L);
        }
        __syncthreads();
        if (active) quad6_finish<Real>(sX[emx], sY[emx], i, j, G, L, F);
    } else {
        quad6_point<Real>(sig, ralpha, axial_row0, q, sX[emx], sY[emx], ipnt,
                          i, j, G, active, F);
    }
    if (active && alpha < nu1p) {
        Real *dst = a.eb + ((size_t)a.elem_eb_off[e] +
                            (size_t)ipnt * a.elem_nu1[e] + alpha) * 6;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            dst[2 * c] = F[c].re;
            dst[2 * c + 1] = F[c].im;
        }
    }
}

// Moment source: stress = -pattern * stf (SolidElement.cpp:279-280), then
// computeQuad6_NoIntegration (GradientQuadrature.hpp:363-390) and the
// subtracting scatter.  One thread per (alpha < nu1c, GLL point).  With
// TO_ELEM the force is added to the source element's element-local force
// (gathered later with the element stiffness), otherwise it is subtracted
// from the point stiffness.  The points of one element are distinct, so
// neither form races.
template <typename Real, bool TO_ELEM>
__global__ __launch_bounds__(STIFF_BLOCK) void k_moment_source(
    int elem, int nu1c, const Real *__restrict__ pattern, const double *stf,
    int tstep, int nstf, const int *__restrict__ step_ctr,
    const int32_t *__restrict__ elem_points,
    const uint8_t *__restrict__ elem_axial, const Real *__restrict__ elem_ops,
    const uint32_t *__restrict__ point_off,
    const int32_t *__restrict__ point_nu1, Real *__restrict__ stiff,
    const uint32_t *__restrict__ elem_eb_off,
    const int32_t *__restrict__ elem_nu1, Real *__restrict__ eb,
    const Real *__restrict__ gtab, const Real *__restrict__ prt_e,
    const Real *__restrict__ rot_e) {
    __shared__ cpx<Real> sX[EMB][3][NPEM];
    __shared__ cpx<Real> sY[EMB][3][NPEM];
    __shared__ Real sG[2][NPEM];
    // a replayed step takes its index from the device counter; outside
    // the STF the whole block leaves before the first barrier
    if (step_ctr) tstep = step_ctr[0];
    if (tstep < 0 || tstep >= nstf) return;
    const int t = threadIdx.x;
    if (t < 2 * NPEM) sG[t / NPEM][t % NPEM] = gtab[t];
    __syncthreads();
    const int eml = t / NPEM;
    const int ipnt = t - eml * NPEM;
    const int alpha = blockIdx.x * EMB + eml;
    const bool active = (eml < EMB) && (alpha < nu1c);
    const int i = ipnt / NPED, j = ipnt - (ipnt / NPED) * NPED;
    const bool axial = elem_axial[elem] != 0;
    PointG<Real> G;
    load_point_g<Real>(sG, axial, i, j, G);
    PointOps<Real> q{};
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
    cpx<Real> F[3], L[3];
    const int emx = active ? eml : 0;
    if (active && prt_e != nullptr) {
        // with a PRT (SolidElement.cpp:293-331): the stress to RTZ by the
        // strain rotation with doubled shears, to the spherical tensor with
        // the unscaled inverse Jacobian, back to SPZ, computeQuad9_
        // NoIntegration
        Real rs1 = 0, rc1 = 1, rs2 = 0, rc2 = 1;
        if (rot_e != nullptr) {
            rs1 = rot_e[ipnt];
            rc1 = rot_e[NPEM + ipnt];
            rs2 = rot_e[2 * NPEM + ipnt];
            rc2 = rot_e[3 * NPEM + ipnt];
#pragma unroll
            for (int k = 3; k < 6; k++) sig[k] = sig[k] * (Real)2;
            rot_spz_rtz6<cpx<Real>, Real>(sig, rs1, rc1, rs2, rc2);
#pragma unroll
            for (int k = 3; k < 6; k++) sig[k] = sig[k] * (Real)0.5;
        }
        cpx<Real> f9[9];
        prt_und_to_sph9<cpx<Real>, Real>(sig, prt_e + ipnt, f9);
        if (rot_e != nullptr)
            rot9<cpx<Real>, Real>(f9, -rs1, rc1, -rs2, rc2);
        quad9_prep<Real>(f9, (Real)alpha, axial && i == 0, q, sX[emx],
                         sY[emx], ipnt, L);
    } else if (active) {
        quad6_prep<Real>(sig, (Real)alpha, axial && i == 0, q, sX[emx],
                         sY[emx], ipnt, L);
    }
    __syncthreads();
    if (active) quad6_finish<Real>(sX[emx], sY[emx], i, j, G, L, F);
    if (active) {
        const int p = elem_points[elem * NPEM + ipnt];
        if (alpha < point_nu1[p]) {
            if (TO_ELEM) {
                Real *dst = eb + ((size_t)elem_eb_off[elem] +
                                  (size_t)ipnt * elem_nu1[elem] + alpha) * 6;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    dst[2 * c] += F[c].re;
                    dst[2 * c + 1] += F[c].im;
                }
            } else {
                Real *dst = stiff + (size_t)(point_off[p] + alpha) * 6;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    dst[2 * c] -= F[c].re;
                    dst[2 * c + 1] -= F[c].im;
                }
            }
        }
    }
}

// Sum of the element-local forces of the elements sharing point p at mode
// alpha, in ascending element order.  adj[p][0..MAXADJ) holds the
// element-local offset of (element, ipnt) in [element][ipnt][alpha]
// triples, padded with ADJ_NONE; a point of a conforming quad mesh has at
// most four elements, so MAXADJ = 4 covers the structured meshes and the
// fixed width lets every load of the point issue at once.
constexpr uint32_t ADJ_NONE = 0xffffffffu;
// slot 3 of a point shared by more than four elements: ADJ_MORE | j, the
// rest of its list at adj[j], adj[j + 1], ... up to an ADJ_NONE
constexpr uint32_t ADJ_MORE = 0x80000000u;

template <typename Real>
__device__ __forceinline__ void gather_forces(
    const uint32_t *__restrict__ adj, const Real *__restrict__ eb, int p,
    int alpha, Real (&f)[6]) {
#pragma unroll
    for (int q = 0; q < 6; q++) f[q] = 0;
    const uint4 a4 = *reinterpret_cast<const uint4 *>(adj + (size_t)p * 4);
    const uint32_t base[4] = {a4.x, a4.y, a4.z, a4.w};
    auto add = [&](uint32_t b) {
        const Real *src = eb + ((size_t)b + alpha) * 6;
#pragma unroll
        for (int q = 0; q < 6; q++) f[q] -= src[q];
    };
#pragma unroll
    for (int k = 0; k < 3; k++)
        if (base[k] != ADJ_NONE) add(base[k]);
    if (base[3] != ADJ_NONE) {
        if (base[3] & ADJ_MORE) {
            // element order kept: the overflow list follows slots 0..2
            for (uint32_t j = base[3] & ~ADJ_MORE; adj[j] != ADJ_NONE; j++)
                add(adj[j]);
        } else {
            add(base[3]);
        }
    }
}

// Stage form of the assembly (Domain::computeStiffness followed by the
// points' gatherStiffFromElement): stiff -= sum of element-local forces.
template <typename Real>
__global__ __launch_bounds__(256) void k_assemble(
    int64_t n_modes, const int32_t *__restrict__ mode_point,
    const uint32_t *__restrict__ point_off,
    const uint32_t *__restrict__ adj, const Real *__restrict__ eb,
    Real *__restrict__ stiff) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_modes) return;
    const int p = mode_point[k];
    const int alpha = (int)(k - (int64_t)point_off[p]);
    Real f[6];
    gather_forces(adj, eb, p, alpha, f);
    Real *sp = stiff + (size_t)k * 6;
#pragma unroll
    for (int q = 0; q < 6; q++) sp[q] += f[q];
}

// MassOceanLoad1D::computeAccel (MassOceanLoad1D.hpp:48-62) on one mode's
// (s, phi, z) re/im, FP64: (s, z) by the symmetric 2x2 factors o[0], o[1]
// (diagonal) and o[2], phi by o[3] = 1/m.
template <typename T>
__device__ __forceinline__ void ocean_load_1d(const double *o, const T *s,
                                              double *out) {
#pragma unroll
    for (int ri = 0; ri < 2; ri++) {
        const double fs = (double)s[ri], fz = (double)s[4 + ri];
        out[ri] = o[0] * fs + o[2] * fz;
        out[4 + ri] = o[2] * fs + o[1] * fz;
        out[2 + ri] = (double)s[2 + ri] * o[3];
    }
}

// Point update, one thread per (point, mode).  Flags select the phases:
//   GATHER assemble the element-local forces (bulk step; no point
//          stiffness buffer is read or written)
//   AXIAL  AxialBoundary::apply, solid (AxialBoundary.cpp:21-42)
//   MASS   Nyquist masking + Mass1D (SolidPoint.cpp:28-36, Mass1D.hpp:31)
//   MARCH  NewmarkTimeScheme::update (NewmarkTimeScheme.hpp:48-59)
// With MASS and MARCH fused the acceleration is formed in FP64 from the
// stiffness (FP32 or FP64) and never stored back.
struct PointArgs {
    int64_t n_modes;
    const int32_t *mode_point;
    const uint32_t *point_off;
    const int32_t *point_nr;
    const uint8_t *point_axial;
    const double *inv_mass;
    const uint8_t *mass3d;        // [point] 1: 3D mass (or null: none)
    const int32_t *ocean_idx;     // [point] MassOceanLoad1D entry, or null
    const double *ocean1d;        // [entry][4]
    const uint32_t *adj;
    double *displ, *veloc, *accel;
    double dt, half_dt, half_dt_dt;
    const double *point_dt;       // per-point time steps, or null
    int *flag;
    CouplingGeom cpl;
    AbcGeom abc;
    int scheme;                   // 0 Newmark, 1 symplectic sub-step
    double sym_pi, sym_kappa;
};

//   COUPLE the Clayton-Enquist boundaries (AbsorbingBoundary::
//          applyClayton: s -= the k_abc force term), then
//          SolidFluidCoupling1D::coupleFluidToSolid (SolidFluidCoupling1D
//          .hpp:33-37): s -= (ns, nz) / m_fluid * fluid stiffness, after
//          the fluid points have added the solid displacement
//   SPONGE the sponge on the acceleration held in stiff
//          (AbsorbingBoundary::applySponge), alone; the fused MASS | MARCH
//          form applies it between the mass and the march
template <typename Real, bool GATHER, bool AXIAL, bool MASS, bool MARCH,
          bool COUPLE, bool SPONGE>
__global__ __launch_bounds__(256) void k_point(PointArgs a,
                                               const Real *__restrict__ eb,
                                               Real *__restrict__ stiff) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n_modes) return;
    const int p = a.mode_point[k];
    const int alpha = (int)(k - (int64_t)a.point_off[p]);
    const int ai = a.abc.idx != nullptr ? a.abc.idx[p] : -1;
    Real s[6];
    if (GATHER) {
        gather_forces(a.adj, eb, p, alpha, s);
    } else {
#pragma unroll
        for (int q = 0; q < 6; q++) s[q] = stiff[(size_t)k * 6 + q];
    }
    if (SPONGE) {
        if (ai >= 0) {
            const int nr = a.point_nr[p];
            sponge_sub<3>(a.abc, ai, alpha, (nr & 1) ? -1 : nr / 2, s);
#pragma unroll
            for (int q = 0; q < 6; q++) stiff[(size_t)k * 6 + q] = s[q];
        }
        return;
    }
    if (COUPLE && ai >= 0) {
        const double *c =
            a.abc.cterm + ((size_t)a.abc.term_off[ai] + alpha) * 6;
#pragma unroll
        for (int q = 0; q < 6; q++) s[q] = (Real)((double)s[q] - c[q]);
    }
    if (COUPLE && a.cpl.n_pairs > 0) {
        const int pi = a.cpl.solid_pair[p];
        if (pi >= 0) {
            const double f_re = a.cpl.cstiff[2 * ((size_t)a.cpl.foff[pi] +
                                                   alpha)];
            const double f_im = a.cpl.cstiff[2 * ((size_t)a.cpl.foff[pi] +
                                                   alpha) + 1];
            const double ns = a.cpl.n_m[2 * pi], nz = a.cpl.n_m[2 * pi + 1];
            s[0] = (Real)((double)s[0] - ns * f_re);
            s[1] = (Real)((double)s[1] - ns * f_im);
            s[4] = (Real)((double)s[4] - nz * f_re);
            s[5] = (Real)((double)s[5] - nz * f_im);
        } else if (a.cpl.solid_pair3d != nullptr) {
            // SolidFluidCoupling3D::coupleFluidToSolid, computed with the
            // deferred fluid points
            const int j = a.cpl.solid_pair3d[p];
            if (j >= 0) {
                const double *c = a.cpl.csolid3d +
                                  ((size_t)a.cpl.c3off[j] + alpha) * 6;
#pragma unroll
                for (int q = 0; q < 6; q++)
                    s[q] = (Real)((double)s[q] - c[q]);
            }
        }
    }
    if (AXIAL && a.point_axial[p]) {
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
        const int nr = a.point_nr[p];
        if ((nr & 1) == 0 && alpha == nr / 2) s[1] = s[3] = s[5] = 0;
        if (a.mass3d != nullptr && a.mass3d[p]) {
            // 3D mass: the division slice by slice (and the march) follow
            // in k_mass3d, which reads this assembled stiffness
#pragma unroll
            for (int q = 0; q < 6; q++) stiff[(size_t)k * 6 + q] = s[q];
            return;
        }
    }
    const int oi = MASS && a.ocean_idx != nullptr ? a.ocean_idx[p] : -1;
    if (MASS && !MARCH) {
        if (oi >= 0) {
            double o[6];
            ocean_load_1d(a.ocean1d + 4 * (size_t)oi, s, o);
#pragma unroll
            for (int q = 0; q < 6; q++) s[q] = (Real)o[q];
        } else {
            const Real im = (Real)a.inv_mass[p];
#pragma unroll
            for (int q = 0; q < 6; q++) s[q] *= im;
        }
    }
    if (!MARCH) {
#pragma unroll
        for (int q = 0; q < 6; q++) stiff[(size_t)k * 6 + q] = s[q];
        return;
    }
    double a_new[6];
    if (MASS) {
        if (oi >= 0) {
            ocean_load_1d(a.ocean1d + 4 * (size_t)oi, s, a_new);
        } else {
            const double im = a.inv_mass[p];
#pragma unroll
            for (int q = 0; q < 6; q++) a_new[q] = (double)s[q] * im;
        }
        if (ai >= 0) {
            const int nr = a.point_nr[p];
            sponge_sub<3>(a.abc, ai, alpha, (nr & 1) ? -1 : nr / 2, a_new);
        }
    } else {
#pragma unroll
        for (int q = 0; q < 6; q++) a_new[q] = (double)s[q];
    }
    // all of u, v, a are loaded (16-byte vectors) before anything is
    // stored, so the nine loads are in flight together
    double2 *up = reinterpret_cast<double2 *>(a.displ + (size_t)k * 6);
    double2 *vp = reinterpret_cast<double2 *>(a.veloc + (size_t)k * 6);
    double2 *ap = reinterpret_cast<double2 *>(a.accel + (size_t)k * 6);
    double2 u2[3], v2[3], a2[3];
    bool bad = false, bad0 = false;
    double dt = a.dt, half_dt = a.half_dt, half_dt_dt = a.half_dt_dt;
    if (a.point_dt != nullptr) {
        dt = a.point_dt[p];
        half_dt = .5 * dt;
        half_dt_dt = .5 * dt * dt;
    }
    if (a.scheme == 1) {
        // SymplecticTimeScheme::update (SymplecticTimeScheme.hpp:37-46):
        // v += pi dt a, u += kappa dt v; no acceleration field
        const double pdt = a.sym_pi * dt, kdt = a.sym_kappa * dt;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            u2[c] = up[c];
            v2[c] = vp[c];
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            bad0 |= !isfinite(u2[c].x) || !isfinite(u2[c].y);
            const double vr = v2[c].x + pdt * a_new[2 * c];
            const double vi = v2[c].y + pdt * a_new[2 * c + 1];
            const double ur = u2[c].x + kdt * vr;
            const double ui = u2[c].y + kdt * vi;
            vp[c] = make_double2(vr, vi);
            up[c] = make_double2(ur, ui);
            bad |= !isfinite(ur) || !isfinite(ui);
        }
        if (!GATHER) {
#pragma unroll
            for (int q = 0; q < 6; q++) stiff[(size_t)k * 6 + q] = 0;
        }
        if (bad0 || bad) atomicOr(a.flag, (bad0 ? 1 : 0) | (bad ? 2 : 0));
        return;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        u2[c] = up[c];
        v2[c] = vp[c];
        a2[c] = ap[c];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        bad0 |= !isfinite(u2[c].x) || !isfinite(u2[c].y);
        const double vr = v2[c].x + half_dt * (a2[c].x + a_new[2 * c]);
        const double vi = v2[c].y + half_dt * (a2[c].y + a_new[2 * c + 1]);
        const double ur = u2[c].x + (dt * vr + half_dt_dt * a_new[2 * c]);
        const double ui =
            u2[c].y + (dt * vi + half_dt_dt * a_new[2 * c + 1]);
        v2[c] = make_double2(vr, vi);
        a2[c] = make_double2(a_new[2 * c], a_new[2 * c + 1]);
        u2[c] = make_double2(ur, ui);
        bad |= !isfinite(ur) || !isfinite(ui);
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        vp[c] = v2[c];
        ap[c] = a2[c];
        up[c] = u2[c];
    }
    if (!GATHER) {
#pragma unroll
        for (int q = 0; q < 6; q++) stiff[(size_t)k * 6 + q] = 0;
    }
    // bit 0: the displacement this step started from was non-finite (what
    // Domain::checkStability sees at the start of a step); bit 1: the one
    // it leaves is
    if (bad0 || bad) atomicOr(a.flag, (bad0 ? 1 : 0) | (bad ? 2 : 0));
}


// SymplecticTimeScheme::launch (SymplecticTimeScheme.hpp:51-58): u +=
// kappa_0 dt v over n doubles (the fields' complex components).
__global__ __launch_bounds__(256) void k_sym_launch(double *__restrict__ u,
                                                    const double *__restrict__ v,
                                                    int64_t n, double kdt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] += kdt * v[i];
}

void launch_sym_launch(hipStream_t st, double *u, const double *v, int64_t n,
                       double kdt) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_sym_launch, dim3((unsigned)((n + 255) / 256)),
                       dim3(256), 0, st, u, v, n, kdt);
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
    double *__restrict__ out, const int *__restrict__ step_ctr,
    int max_samples, const double *__restrict__ rec_rot) {
    const int r = blockIdx.x;
    if (r >= nrec) return;
    // a replayed step: sample = step - first recorded step (counter words
    // 0 and 1), nothing outside the seismogram buffer
    if (step_ctr) {
        const int isample = step_ctr[0] - step_ctr[1];
        if (isample < 0 || isample >= max_samples) return;
        out += (size_t)isample * nrec * 3;
    }
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
        // Station::rotateField (Station.hpp:79-118) as Station::rotate
        // (Station.cpp:41-55): about axis k, v_i' = v_i c + v_j s, v_j' =
        // -v_i s + v_j c with i = k + 1, j = k + 2 (mod 3)
        if (rec_rot != nullptr) {
            for (int slot = 0; slot < 2; slot++) {
                const double *q = rec_rot + (size_t)r * 6 + 3 * slot;
                const int k = (int)q[0];
                if (k < 0) continue;
                const int i = (k + 1) % 3, j = (k + 2) % 3;
                const double vi = acc[i], vj = acc[j];
                acc[i] = vi * q[1] + vj * q[2];
                acc[j] = -vi * q[2] + vj * q[1];
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) out[r * 3 + c] = acc[c];
    }
}

// ------------------------------------------------------------ launchers
template <typename Real>
void launch_solid_stiff_fr(hipStream_t st, const StiffArgs<Real> &a) {
    if (a.n_em <= 0) return;
    // one launch per material type present (em_list grouped by type)
    for (int mt = 0; mt < 3; mt++) {
        const int eb = a.em_type_begin[mt], ee = a.em_type_begin[mt + 1];
        const int ea = a.em_ax_begin[mt];
        const bool att = a.nsls > 0;
        for (int ax = 0; ax < 2; ax++) {
            const int em0 = ax ? ea : eb, em1 = ax ? ee : ea;
            if (em1 <= em0) continue;
            const int grid = (em1 - em0 + EMB - 1) / EMB;
#define AXS3D_FR2(ATT_, MT_, PRT_)                                            \
    do {                                                                      \
        if (ax)                                                               \
            hipLaunchKernelGGL((k_solid_stiff_fr<Real, ATT_, MT_, true,       \
                                                 PRT_>),                      \
                               dim3(grid), dim3(STIFF_BLOCK), 0, st, a, em0,  \
                               em1);                                          \
        else                                                                  \
            hipLaunchKernelGGL((k_solid_stiff_fr<Real, ATT_, MT_, false,      \
                                                 PRT_>),                      \
                               dim3(grid), dim3(STIFF_BLOCK), 0, st, a, em0,  \
                               em1);                                          \
    } while (0)
#define AXS3D_FR(ATT_, MT_)                                                   \
    do {                                                                      \
        if (a.prt_idx != nullptr) AXS3D_FR2(ATT_, MT_, true);                 \
        else AXS3D_FR2(ATT_, MT_, false);                                     \
    } while (0)
            if (mt == 0) {
                if (att) AXS3D_FR(true, 0); else AXS3D_FR(false, 0);
            } else if (mt == 1) {
                if (att) AXS3D_FR(true, 1); else AXS3D_FR(false, 1);
            } else {
                if (att) AXS3D_FR(true, 2); else AXS3D_FR(false, 2);
            }
#undef AXS3D_FR
#undef AXS3D_FR2
        }
    }
}

template <typename Real>
void launch_moment_source(hipStream_t st, bool to_elem,
                          const StiffArgs<Real> &a, int elem, int nu1c,
                          const Real *pattern, const double *stf, int tstep,
                          int nstf, const int *step_ctr, Real *stiff,
                          const Real *prt_e) {
    if (nu1c <= 0) return;
    const int grid = (nu1c + EMB - 1) / EMB;
    const Real *rot_e = a.rot ? a.rot + (size_t)elem * 4 * NPEM : nullptr;
    if (to_elem)
        hipLaunchKernelGGL((k_moment_source<Real, true>), dim3(grid),
                           dim3(STIFF_BLOCK), 0, st, elem, nu1c, pattern, stf,
                           tstep, nstf, step_ctr, a.elem_points, a.elem_axial, a.elem_ops,
                           a.point_off, a.point_nu1, stiff, a.elem_eb_off,
                           a.elem_nu1, a.eb, a.gtab, prt_e, rot_e);
    else
        hipLaunchKernelGGL((k_moment_source<Real, false>), dim3(grid),
                           dim3(STIFF_BLOCK), 0, st, elem, nu1c, pattern, stf,
                           tstep, nstf, step_ctr, a.elem_points, a.elem_axial, a.elem_ops,
                           a.point_off, a.point_nu1, stiff, a.elem_eb_off,
                           a.elem_nu1, a.eb, a.gtab, prt_e, rot_e);
}

template <typename Real>
void launch_assemble(hipStream_t st, const PointGeom &g, const Real *eb,
                     Real *stiff) {
    if (g.n_modes <= 0) return;
    const int block = 256;
    const int64_t grid = (g.n_modes + block - 1) / block;
    hipLaunchKernelGGL(k_assemble<Real>, dim3(grid), dim3(block), 0, st,
                       g.n_modes, g.mode_point, g.point_off, g.adj, eb,
                       stiff);
}

template <typename Real>
void launch_point(hipStream_t st, int phases, const PointGeom &g,
                  const Real *eb, Real *stiff, double *displ, double *veloc,
                  double *accel, double dt, int *flag,
                  const CouplingGeom &cpl) {
    if (g.n_modes <= 0) return;
    const int block = 256;
    const int64_t grid = (g.n_modes + block - 1) / block;
    PointArgs a;
    a.n_modes = g.n_modes;
    a.mode_point = g.mode_point;
    a.point_off = g.point_off;
    a.point_nr = g.point_nr;
    a.point_axial = g.point_axial;
    a.inv_mass = g.inv_mass;
    a.mass3d = g.mass3d;
    a.ocean_idx = g.ocean_idx;
    a.ocean1d = g.ocean1d;
    a.adj = g.adj;
    a.displ = displ;
    a.veloc = veloc;
    a.accel = accel;
    a.dt = dt;
    a.half_dt = .5 * dt;
    a.half_dt_dt = .5 * dt * dt;
    a.point_dt = g.point_dt;
    a.flag = flag;
    a.cpl = cpl;
    a.abc = g.abc;
    a.scheme = g.scheme;
    a.sym_pi = g.sym_pi;
    a.sym_kappa = g.sym_kappa;
    if (cpl.n_pairs <= 0 && g.abc.idx == nullptr) phases &= ~PHASE_COUPLE;
    if (g.abc.idx == nullptr) phases &= ~PHASE_SPONGE;
    if (phases == 0) return;
#define AXS3D_POINT_CASE(G, A, M, T, C, S)                                    \
    case (A) | ((M) << 1) | ((T) << 2) | ((G) << 3) | ((C) << 4) |           \
        ((S) << 5):                                                          \
        hipLaunchKernelGGL((k_point<Real, G, A, M, T, C, S>), dim3(grid),      \
                           dim3(block), 0, st, a, eb, stiff);                 \
        break;
    switch (phases) {
        AXS3D_POINT_CASE(0, 1, 0, 0, 0, 0)
        AXS3D_POINT_CASE(0, 0, 1, 0, 0, 0)
        AXS3D_POINT_CASE(0, 0, 0, 1, 0, 0)
        AXS3D_POINT_CASE(0, 0, 0, 0, 1, 0)
        AXS3D_POINT_CASE(0, 0, 0, 0, 0, 1)
        AXS3D_POINT_CASE(0, 1, 1, 1, 0, 0)
        AXS3D_POINT_CASE(1, 1, 1, 1, 0, 0)
        AXS3D_POINT_CASE(1, 1, 1, 1, 1, 0)
        default:
            break;
    }
#undef AXS3D_POINT_CASE
}

void launch_record(hipStream_t st, int nrec, const int32_t *rec_elem,
                   const double *rec_w, const double *rec_phi,
                   const int32_t *elem_points, const int32_t *elem_nu1,
                   const uint32_t *point_off, const int32_t *point_nu1,
                   const double *displ, double *out, const int *step_ctr,
                   int max_samples, const double *rec_rot) {
    if (nrec <= 0) return;
    hipLaunchKernelGGL(k_record, dim3(nrec), dim3(64), 0, st, nrec, rec_elem,
                       rec_w, rec_phi, elem_points, elem_nu1, point_off,
                       point_nu1, displ, out, step_ctr, max_samples,
                       rec_rot);
}

// The step counter of replayed steps: [0] the step, [1] the first
// recorded step.
__global__ void k_step_set(int *ctr, int tstep, int record_step0) {
    if (threadIdx.x == 0) {
        ctr[0] = tstep;
        ctr[1] = record_step0;
    }
}

// flag[0]: raised by the point kernels, bit 0 when a step started from a
// non-finite displacement (what Domain::checkStability tests at the start
// of a step), bit 1 when a step left one; flag[1]: the first step (from
// 0) that started from one (-1: none yet).
__global__ void k_step_advance(int *ctr, int *flag) {
    if (threadIdx.x == 0) {
        if ((flag[0] & 1) != 0 && flag[1] < 0) flag[1] = ctr[0];
        ctr[0] += 1;
    }
}

__global__ void k_flag_note(int *flag, int tstep) {
    if (threadIdx.x == 0 && (flag[0] & 1) != 0 && flag[1] < 0)
        flag[1] = tstep;
}

void launch_step_set(hipStream_t st, int *ctr, int tstep, int record_step0) {
    hipLaunchKernelGGL(k_step_set, dim3(1), dim3(64), 0, st, ctr, tstep,
                       record_step0);
}

void launch_step_advance(hipStream_t st, int *ctr, int *flag) {
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, st, ctr, flag);
}

void launch_flag_note(hipStream_t st, int *flag, int tstep) {
    hipLaunchKernelGGL(k_flag_note, dim3(1), dim3(64), 0, st, flag, tstep);
}

#define AXS3D_INSTANTIATE(Real)                                               \
    template void launch_solid_stiff_fr<Real>(hipStream_t,                   \
                                              const StiffArgs<Real> &);      \
    template void launch_moment_source<Real>(                                 \
        hipStream_t, bool, const StiffArgs<Real> &, int, int, const Real *,   \
        const double *, int, int, const int *, Real *, const Real *);         \
    template void launch_assemble<Real>(hipStream_t, const PointGeom &,       \
                                        const Real *, Real *);                \
    template void launch_point<Real>(hipStream_t, int, const PointGeom &,     \
                                     const Real *, Real *, double *,          \
                                     double *, double *, double, int *,       \
                                     const CouplingGeom &);

AXS3D_INSTANTIATE(float)
AXS3D_INSTANTIATE(double)

}  // namespace axs3d
