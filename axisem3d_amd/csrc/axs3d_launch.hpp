// axs3d_launch.hpp -- host-side launchers of the device kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "axs3d_common.hpp"
#include "axs3d_fft.hpp"

namespace axs3d {

// point-update phase bits
enum PointPhase {
    PHASE_AXIAL = 1,     // axis (and, for fluid points, free surface)
    PHASE_MASS = 2,
    PHASE_MARCH = 4,
    PHASE_GATHER = 8,
    PHASE_COUPLE = 16,   // Clayton boundaries, then solid-fluid coupling
    PHASE_SPONGE = 32    // sponge on the acceleration (per-phase form; the
                         // fused MASS | MARCH form applies it itself)
};

// Solid-fluid coupling pairs (one solid and one fluid point at the same
// GLL position, same Nr).
struct CouplingGeom {
    int n_pairs;
    const int32_t *solid_pair;     // [solid point] pair or -1
    const int32_t *fluid_pair;     // [fluid point] pair or -1
    const uint32_t *soff;          // [pair] solid mode offset
    const uint32_t *foff;          // [pair] fluid mode offset
    const double *n;               // [pair][2] ns, nz (fluid -> solid)
    const double *n_m;             // [pair][2] ns / m_fluid, nz / m_fluid
    double *cstiff;                // [fluid mode] fluid stiffness after
                                   // the solid -> fluid term
    // 3D pairs (SolidFluidCoupling3D): the fluid -> solid term of each,
    // computed with the fluid points that wait for it (k_fluid_defer)
    const int32_t *solid_pair3d = nullptr;  // [solid point] 3D pair or -1
    const uint32_t *c3off = nullptr;        // [3D pair] mode offset
    const double *csolid3d = nullptr;       // [mode][3] complex
};

// Fluid elements (acoustic, 1D), element-modes of the Fourier path.
template <typename Real>
struct FluidArgs {
    int n_em;
    const EMItem *em_list;
    const int32_t *elem_points;
    const uint8_t *elem_axial;
    const Real *elem_ops;          // [e][10][25] as for solids
    const Real *elem_k;            // [e][25] K = 1 / rho
    const uint32_t *point_off;
    const int32_t *point_nu1;
    const double *chi;             // fluid displacement (potential)
    const uint32_t *elem_eb_off;   // element-local force offset, modes
    const int32_t *elem_nu1;
    Real *eb;                      // [e][ipnt][alpha] complex
    const Real *gtab;
};

// Fluid elements with a 3D medium (cardinal path): their element-modes in
// f.em_list, K per slice and the FFT tables.
template <typename Real>
struct FluidCardArgs {
    FluidArgs<Real> f;             // f.em_list: 3D fluid element-modes
    const int32_t *list3d;         // 3D fluid elements, launch order
    const int32_t *elem_nr;        // fluid element Nr
    const Real *k3d;               // per element [25][Nr] at k3d_off[e]
    const int64_t *k3d_off;
    const cpx<Real> *twiddle;
    const int64_t *tw_off;
    const uint32_t *plan;
    const uint16_t *slice_pos;
    const int64_t *pos_off;
    Real *scratch;                 // batch scratch, complex
    const uint32_t *scr_off;       // [fluid element] [3][25][nu1] part
};

// Device operands of the solid element kernels.
template <typename Real>
struct StiffArgs {
    int n_em;
    const EMItem *em_list;
    const int32_t *elem_points;
    const uint8_t *elem_axial;
    const Real *elem_ops;
    const Real *elem_mat;          // 1D material, packed per element
    const uint32_t *mat_off;       // offset of element e in elem_mat
    const Real *rot;               // [e][4][25] sin, cos of theta, 2 theta
                                   // (spherical RTZ), or null
    // particle relabelling of undulated geometry: per element its operand
    // set in prt ([8][25]: X0..X3, X0..X3 times |J|) or -1; null: none
    const int32_t *prt_idx = nullptr;
    const Real *prt = nullptr;
    int em_type_begin[4];          // em_list grouped by material type
    int em_ax_begin[3];            // within type t, the axial elements'
                                   // element-modes from em_ax_begin[t]
    const uint32_t *point_off;
    const int32_t *point_nu1;
    const double *displ;
    const uint32_t *elem_eb_off;   // element-local force offset, triples
    const int32_t *elem_nu1;
    Real *eb;                      // element-local forces
    const Real *gtab;
    // attenuation of the Fourier-path (1D-material) elements, per element
    // mode em of em_list (nsls = 0: none)
    int nsls;
    int npa;                       // 4 (CG4) or 25 (FULL)
    const double *abg;             // [3][nsls] alpha, beta, gamma
    const Real *att1d;             // [e][2][npa] dlambda, dmu
    double *mv_fr;                 // [em][nsls][6][npa] complex, FP64
    Real *ds_fr;                   // [em][6][npa] complex
};

// Largest element Nr of the cardinal kernels (one padded complex series
// of Nr per wave in LDS, sized per launch).
constexpr int CARD_NR_CAP = FFT_NR_CAP;
// Minimum waves per SIMD the stress kernel's register allocation allows:
// five in FP32; in FP64 the LDS of an Nr of several hundred holds at most
// two or three blocks per CU anyway, and three waves leave the radix-16
// passes their registers.
constexpr int CARD_OCC = 5;
template <typename Real>
struct CardOcc {
    static constexpr int value = sizeof(Real) == 4 ? CARD_OCC : 3;
};
// FFT plan words per Nr (see axs3d_fft.hpp).
constexpr int CARD_PLAN_W = FFT_PLAN_W;

// Device operands of the cardinal-path (3D material) element kernel.
template <typename Real>
struct CardArgs {
    StiffArgs<Real> s;
    int n3d;                       // number of 3D elements; s.em_list is
                                   // their element-modes in list3d order
    const EMItem *em_ax;           // the element-modes of the axial ones
    const int32_t *list3d;         // 3D elements grouped by material type,
                                   // each group heaviest first
    int type_begin[4];             // group boundaries in list3d
    const int32_t *elem_nr;
    const int64_t *slice_off;      // first slice of each element
    // Per-slice operands are stored per element as [component][Nr], at
    // element offset slice_off * (components per slice).
    const Real *prop3d;            // [np][25] parameters of the type
    const int64_t *prop_off;       // element offsets in prop3d
    const cpx<Real> *twiddle;      // pass twiddle tables, per Nr
    const int64_t *tw_off;         // offset of each Nr's table
    const uint32_t *plan;          // [Nr][CARD_PLAN_W] FFT passes
    const uint16_t *slice_pos;     // buffer position of each slice after
    const int64_t *pos_off;        // the inverse transform, per Nr
    Real *scratch;                 // batch scratch, complex
    const uint32_t *scr_off;       // [element] its part of its batch's
                                   // scratch ([150][nu1] complex), in
                                   // complex values
    // attenuation (nsls = 0: none)
    int nsls;
    int npa;                       // 4 (CG4) or 25 (FULL)
    const double *abg;             // [3][nsls] alpha, beta, gamma
    const Real *att3d;             // [2][npa] dlambda, dmu
    double *memvar;                // [nsls][6][npa], FP64
    Real *dstress;                 // [6][npa]
};

// What the point kernels read of the absorbing boundaries (AbcArgs): per
// point the index of its boundary record (-1: none), the record's term
// offset, and whether its sponge is 3D (an acceleration that goes through
// C2R and R2C loses the imaginary part of its order 0, Sponge.hpp:84-92).
struct AbcGeom {
    const int32_t *idx = nullptr;        // [point] record or -1; null: none
    const int64_t *term_off = nullptr;
    const int64_t *sp3_off = nullptr;
    const double *cterm = nullptr;       // [term][nc] complex
    const double *sterm = nullptr;
};

// Sponge term of mode alpha of boundary record ai subtracted from an
// acceleration (Sponge::apply, Sponge.hpp:74-96; the term from k_abc); a
// 3D sponge's C2R/R2C round trip drops the imaginary parts of order 0 and
// of the Nyquist order nyq (nr / 2 for an even Nr, else -1).
template <int NC, typename T>
__device__ __forceinline__ void sponge_sub(const AbcGeom &g, int ai,
                                           int alpha, int nyq, T *acc) {
    if ((alpha == 0 || alpha == nyq) && g.sp3_off[ai] >= 0)
#pragma unroll
        for (int c = 0; c < NC; c++) acc[2 * c + 1] = (T)0;
    const double *t =
        g.sterm + ((size_t)g.term_off[ai] + alpha) * (2 * NC);
#pragma unroll
    for (int q = 0; q < 2 * NC; q++) acc[q] = (T)((double)acc[q] - t[q]);
}

// Device point description used by the gather and point kernels.
struct PointGeom {
    // absorbing boundaries of these points (idx null: none)
    AbcGeom abc;
    int64_t n_modes;
    const int32_t *mode_point;
    const uint32_t *point_off;
    const int32_t *point_nr;
    const uint8_t *point_axial;
    const double *inv_mass;
    const uint8_t *mass3d;         // [point] 1: 3D mass, or null
    // MassOceanLoad1D: [point] entry in ocean1d (-1: none), or null;
    // ocean1d [entry][4] FP64 factors (MassOceanLoad1D.hpp:20-30)
    const int32_t *ocean_idx = nullptr;
    const double *ocean1d = nullptr;
    const uint32_t *adj;           // [point][4] element-local offsets,
                                   // ascending element, padded with ~0u
    // independent instances batched in one domain: the time step of each
    // point's instance (null: one time step for all)
    const double *point_dt = nullptr;
    // the march: 0 Newmark; 1 a symplectic sub-step, v += sym_pi dt a, u
    // += sym_kappa dt v (SymplecticTimeScheme::update)
    int scheme = 0;
    double sym_pi = 0., sym_kappa = 0.;
};

// Absorbing boundaries (AbsorbingBoundary, src/core/boundary/ABC) of the
// points of one kind (solid: 3 components per mode, fluid: 1), merged per
// point on the host: the Clayton-Enquist force term and the Kosloff sponge
// acceleration term of a step are linear in the velocity and displacement
// the step starts from, so k_abc forms both before the point kernels,
// which subtract them in the reference's phases (Clayton before the
// solid-fluid coupling, the sponge after the mass).
//   cl1   [i][4] solid 1D Clayton sums (c0 s-s, c1 z-z, c2 s-z, c3 phi) /
//         [i][1] fluid area / (rho vp); zeros without a 1D Clayton
//   cl3   per slice, from cl3_off[i] (-1: none): solid the symmetric
//         operator rsa I + k k^T summed over the point's entries, 6 values
//         (ss, pp, zz, sp, sz, pz); fluid area / (rho vp)
//   sp1   [i][2] 1D sponge (2 sum gamma, sum gamma^2); zeros without one
//   sp3   per slice (2 gamma, gamma^2) from sp3_off[i] (-1: none)
//   cterm, sterm  [term_off[i] + alpha][nc] complex FP64 terms
template <typename Real>
struct AbcArgs {
    int n;                         // boundary points
    int nr_max;                    // their largest Nr with a 3D part
    const int32_t *point;          // [i] point
    const uint32_t *point_off;
    const int32_t *point_nr;
    const double *displ, *veloc;   // the fields the step starts from
    const double *cl1;
    const int64_t *cl3_off;
    const Real *cl3;
    const double *sp1;
    const int64_t *sp3_off;
    const Real *sp3;
    const int64_t *term_off;
    double *cterm, *sterm;
    const cpx<Real> *twiddle;
    const int64_t *tw_off;
    const uint32_t *plan;
    const uint16_t *slice_pos;
    const int64_t *pos_off;
};

template <typename Real, int NC>
void launch_abc(hipStream_t st, const AbcArgs<Real> &a);

// Solid points with a 3D mass (Mass3D): inverse mass per slice in Real,
// the FFT tables of their Nr and the fields the march updates.
template <typename Real>
struct Mass3DArgs {
    int n;                         // points with a 3D mass
    int nr_max;                    // their largest Nr (LDS of a launch)
    const int32_t *point;          // [i] point index
    const int64_t *inv_off;        // [i] offset of its Nr inverse masses
    const Real *inv_mass;          // 1 / m per slice
    // MassOceanLoad3D: [i] offset of the point's k per slice ([slice][3]
    // in ocean_k; -1: plain Mass3D), or null for none
    const int64_t *k_off = nullptr;
    const Real *ocean_k = nullptr;
    const uint32_t *point_off;
    const int32_t *point_nr;
    const cpx<Real> *twiddle;
    const int64_t *tw_off;
    const uint32_t *plan;
    const uint16_t *slice_pos;
    const int64_t *pos_off;
    double *displ, *veloc, *accel;
    double dt;
    int *flag;
    AbcGeom abc;                   // sponge terms of the march
    int scheme = 0;                // the march as PointGeom's
    double sym_pi = 0., sym_kappa = 0.;
};

// Fluid points whose update needs transforms over the azimuth: a 3D mass
// (Mass3D::computeAccel for one component, Mass3D.cpp:41-53) and/or a 3D
// coupling (SolidFluidCoupling3D.cpp:55-94).  The fluid point kernel
// leaves their gathered stiffness in fstiff; k_fluid_defer does their
// phases (coupling, axis/surface, mass, march) and the fluid -> solid term
// of their 3D pairs, which the solid point kernel subtracts.
template <typename Real>
struct FluidDeferArgs {
    int n;                         // deferred fluid points
    int nr_max;                    // their largest Nr (LDS of a launch)
    const int32_t *point;          // [i] fluid point
    const int32_t *pair3d;         // [i] 3D pair or -1
    const int64_t *minv_off;       // [i] offset of Nr inverse masses or -1
    const Real *minv;              // 1 / m per slice (3D mass)
    const double *inv_mass;        // [fluid point] 1 / m (1D mass)
    const uint32_t *point_off;     // fluid
    const int32_t *point_nr;
    const uint8_t *point_bc;       // bit 0 axis, bit 1 free surface
    Real *fstiff;
    double *chi, *chi_v, *chi_a;
    double dt;
    int *flag;
    // 3D pairs
    const int32_t *pair_solid;     // [j] solid point
    const int64_t *n_off;          // [j] offset of its [Nr][3] normals
    const Real *nu;                // unassembled normals (solid -> fluid)
    const Real *nam;               // assembled normals / m_fluid per slice
    const uint32_t *solid_off;     // solid point_off
    const double *solid_displ;
    double *csolid3d;              // [mode][3] complex, at c3off[j]
    const uint32_t *c3off;
    // FFT tables
    const cpx<Real> *twiddle;
    const int64_t *tw_off;
    const uint32_t *plan;
    const uint16_t *slice_pos;
    const int64_t *pos_off;
    AbcGeom abc;                   // sponge terms of the march
    int scheme = 0;                // the march as PointGeom's
    double sym_pi = 0., sym_kappa = 0.;
};
template <typename Real>
void launch_fluid_defer(hipStream_t st, int phases,
                        const FluidDeferArgs<Real> &a);

// Mass3D::computeAccel of the 3D-mass points from the assembled stiffness
// (written there by the point kernel): with march, the Newmark update of
// their modes; without, the acceleration back into stiff (stage form).
template <typename Real>
void launch_mass3d(hipStream_t st, bool march, const Mass3DArgs<Real> &a,
                   Real *stiff);

template <typename Real>
void launch_solid_stiff_fr(hipStream_t st, const StiffArgs<Real> &a);

// A batch of the cardinal path: list3d positions [begin, end), all of
// material type mtype, whose strain and stress fit the scratch together.
struct CardBatch {
    int mtype, begin, end;
    int em_begin, em_end;          // its element-modes in the 3D list
    int ax_begin, ax_end;          // its axial elements' element-modes in
                                   // the axial list
    int nr_max = 0;                // largest Nr of its elements
};
// Default scratch of one batch (bytes).
constexpr size_t CARD_SCRATCH_BYTES = (size_t)4 << 30;

// The cardinal path of one batch: gradient, stress, quadrature, in that
// order on one stream.
template <typename Real>
void launch_cd_grad(hipStream_t st, const CardArgs<Real> &a,
                    const CardBatch &bt);
template <typename Real>
void launch_cd_stress(hipStream_t st, const CardArgs<Real> &a,
                      const CardBatch &bt);
template <typename Real>
void launch_cd_quad(hipStream_t st, const CardArgs<Real> &a,
                    const CardBatch &bt);

// The cardinal path of a batch of 3D fluid elements: gradient, K per
// slice through the FFT, quadrature into the fluid element-local forces.
template <typename Real>
void launch_cdf_batch(hipStream_t st, const FluidCardArgs<Real> &a,
                      const CardBatch &bt);

// resident stress blocks per CU (with or without attenuation) for the
// LDS of an Nr of nr_max
template <typename Real>
int card_blocks_per_cu(bool att, int nr_max);

template <typename Real>
void launch_moment_source(hipStream_t st, bool to_elem,
                          const StiffArgs<Real> &a, int elem, int nu1c,
                          const Real *pattern, const double *stf, int tstep,
                          int nstf, const int *step_ctr, Real *stiff,
                          const Real *prt_e = nullptr);

template <typename Real>
void launch_assemble(hipStream_t st, const PointGeom &g, const Real *eb,
                     Real *stiff);

template <typename Real>
void launch_point(hipStream_t st, int phases, const PointGeom &g,
                  const Real *eb, Real *stiff, double *displ, double *veloc,
                  double *accel, double dt, int *flag,
                  const CouplingGeom &cpl);

template <typename Real>
void launch_fluid_stiff_fr(hipStream_t st, const FluidArgs<Real> &a);

template <typename Real>
void launch_fluid_assemble(hipStream_t st, const PointGeom &g,
                           const Real *eb, Real *stiff);

// point_axial of a fluid PointGeom: bit 0 axis, bit 1 free surface
template <typename Real>
void launch_fluid_point(hipStream_t st, int phases, const PointGeom &g,
                        const Real *eb, Real *stiff, double *chi,
                        double *chi_v, double *chi_a, double dt, int *flag,
                        const CouplingGeom &cpl, const double *solid_displ);

// u += kdt v over n doubles (SymplecticTimeScheme::launch)
void launch_sym_launch(hipStream_t st, double *u, const double *v, int64_t n,
                       double kdt);

void launch_record(hipStream_t st, int nrec, const int32_t *rec_elem,
                   const double *rec_w, const double *rec_phi,
                   const int32_t *elem_points, const int32_t *elem_nu1,
                   const uint32_t *point_off, const int32_t *point_nu1,
                   const double *displ, double *out, const int *step_ctr,
                   int max_samples, const double *rec_rot = nullptr);

// adjoint: sensitivity densities of one step, kern[e][k][3] += scale *
// (lambda, mu, rho) densities of (psi, u, a); FP64 throughout
struct SensArgs {
    int n_elem;
    const int32_t *elem_points;
    const uint8_t *elem_axial;
    const double *gq;              // [e][6][25], row 5 the integral factor
    const int32_t *elem_nu1;
    const uint32_t *point_off;
    const int32_t *point_nu1;
    const double *gtab;            // [2][25] G_GLL, G_GLJ
    const double *psi, *u, *a;     // [mode][3] complex
    double *kern;                  // [e][25][3] lambda, mu, rho (kernm
                                   // null: isotropic domains)
    double scale;
    const uint8_t *elem_kind;      // 1: 3D element (its moduli per slice
                                   // elsewhere), or null
    // any media (kernm non-null): the densities of the parameters of each
    // element's material type (elem_mtype; null: all isotropic) into kernm
    // + mat_off[e] as [param][25], the mass density into kernr[e][25];
    // TI and anisotropic elements on spherical meshes form both strains in
    // RTZ with rot ([e][4][25] sin, cos of theta and of 2 theta; null:
    // Cartesian)
    const uint8_t *elem_mtype = nullptr;
    const int64_t *mat_off = nullptr;
    double *kernm = nullptr;
    double *kernr = nullptr;
    const double *rot = nullptr;
};
void launch_sensitivity(hipStream_t st, const SensArgs &a);
// per-slice lambda, mu densities of the 3D elements of one batch, from the
// strain rows of psi (a.scratch) and of u (scrb); see k_cd_sens
// (the moduli of the batch's material type: kern3d + prop_off[e] as
// [slice][param][25], the layout of axs3d_set_solid_material3d)
template <typename Real>
void launch_cd_sens(hipStream_t st, const CardArgs<Real> &a, const Real *scrb,
                    const CardBatch &bt, const double *ifac, double *kern3d,
                    double scale);

// adjoint force R^* g * scale into the receivers' element-local forces
template <typename Real>
void launch_adjoint_source(hipStream_t st, int nrec, const int32_t *rec_elem,
                           const double *rec_w, const double *rec_phi,
                           const double *g, double scale, int nu1max,
                           const int32_t *elem_points,
                           const int32_t *elem_nu1,
                           const int32_t *point_nu1,
                           const uint32_t *elem_eb_off, Real *eb);

// step counter of a replayed (graph) step; see k_step_set
void launch_step_set(hipStream_t st, int *ctr, int tstep, int record_step0);
void launch_step_advance(hipStream_t st, int *ctr, int *flag);
// eager steps: note tstep as the first blown-up step once flag[0] is up
void launch_flag_note(hipStream_t st, int *flag, int tstep);

}  // namespace axs3d
