/*
 * axs3d.h -- C ABI of the MI355X-native AxiSEM3D stiffness path.
 *
 * The reference has no FFI: its boundary is the C++ virtual interface of
 * Domain / Element / Point / TimeScheme (src = SOLVER/src):
 *   TimeScheme::solve()                     src/core/time/TimeScheme.hpp:66
 *   Domain phase methods                    src/core/domain/Domain.hpp:158-207
 *   SolidElement::computeStiff()            src/core/element/SolidElement.cpp:213-222
 *   Elastic::strainToStress_FR/_CD          src/core/element/material/elastic/Elastic.hpp:63-71
 *   Mass::computeAccel                      src/core/point/mass/Mass.hpp:26-30
 *   NewmarkTimeScheme::update               src/core/time/NewmarkTimeScheme.hpp:48-71
 * Each entry point below names the reference function it replaces.  The
 * operands are the ones the reference preloop hands to those classes
 * (SE_Model::release, src/preloop/se_model/SE_Model.cpp:207-243), passed
 * as plain host arrays; the library copies them to device memory.
 *
 * Conventions
 *  - Every call returns 0 on success, non-zero on error; the message, in
 *    the reference's "Class::method || detail" format, is available from
 *    axs3d_last_error().
 *  - One host thread per context.  A context owns its HIP stream and all
 *    its device memory; host arrays may be freed after each call returns.
 *  - The axs3d_set_* and axs3d_add_* calls describe the domain before
 *    axs3d_finalize (except axs3d_set_graph, axs3d_set_record_origin and
 *    axs3d_set_field, which act on a finalized domain).  Once finalized,
 *    the domain is immutable in shape, as the reference's is after its
 *    preloop: a second axs3d_finalize or a set-up call returns 1.
 *  - A kernel launch the runtime rejects makes the call return 1 with
 *    the HIP error in the message.
 *  - Complex arrays are interleaved (re, im) doubles.
 *  - Point fields are [mode][3] complex per point, points concatenated in
 *    index order, mode count Nr/2+1 per point (SolidPoint::Fields,
 *    src/core/point/SolidPoint.hpp:115-121).
 *  - Element operands are [element][k][25], ipnt = ipol * 5 + jpol.
 *  - Precision: AXS3D_FP64 computes everything in FP64 (the reference
 *    built with -Ddouble=true); AXS3D_FP32_FP64 computes gradient,
 *    stress and quadrature in FP32 and keeps displacement, velocity,
 *    acceleration, mass division, source-time function, receivers and
 *    the SLS memory variables in FP64.
 */
#ifndef AXS3D_H
#define AXS3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct axs3d_ctx axs3d_ctx;

enum axs3d_precision { AXS3D_FP64 = 0, AXS3D_FP32_FP64 = 1 };

enum axs3d_field {
    AXS3D_FIELD_DISPL = 0,
    AXS3D_FIELD_VELOC = 1,
    AXS3D_FIELD_ACCEL = 2,
    AXS3D_FIELD_STIFF = 3,
    /* fluid points: displacement potential chi and its derivatives,
     * complex [fluid point offset + alpha] */
    AXS3D_FIELD_FLUID_DISPL = 4,
    AXS3D_FIELD_FLUID_VELOC = 5,
    AXS3D_FIELD_FLUID_ACCEL = 6,
    AXS3D_FIELD_FLUID_STIFF = 7
};

enum axs3d_kernel_id {
    AXS3D_KERNEL_STIFFNESS = 0,   /* Fourier-path (1D) element stiffness */
    AXS3D_KERNEL_POINT = 1,       /* fused BC/mass/Newmark point update  */
    AXS3D_KERNEL_SOURCE = 2,
    AXS3D_KERNEL_RECORD = 3,
    AXS3D_KERNEL_STIFFNESS_3D = 4,/* cardinal-path (3D) element stiffness */
    AXS3D_KERNEL_FLUID_STIFFNESS = 5, /* fluid element stiffness       */
    AXS3D_KERNEL_FLUID_POINT = 6, /* fluid coupling/BC/mass/Newmark     */
    AXS3D_KERNEL_CD_GRAD = 7,     /* cardinal path: gradient, per batch  */
    AXS3D_KERNEL_CD_STRESS = 8,   /* cardinal path: FFT, stress, FFT     */
    AXS3D_KERNEL_CD_QUAD = 9,     /* cardinal path: quadrature           */
    AXS3D_NUM_KERNELS = 10
};

/* Context.  Replaces the construction of Domain + NewmarkTimeScheme in
 * main() (src/main.cpp:503-548). */
int axs3d_ctx_create(int device, int precision, axs3d_ctx **out);
int axs3d_ctx_destroy(axs3d_ctx *ctx);
const char *axs3d_last_error(const axs3d_ctx *ctx);
int axs3d_precision_of(const axs3d_ctx *ctx);

/* GLL and GLJ derivative matrices, G[i*5+j] = l_i'(x_j).
 * Replaces GradientQuadrature::setGMat (GradientQuadrature.hpp:510-516). */
int axs3d_set_gmat(axs3d_ctx *ctx, const double *g_gll, const double *g_glj);

/* Solid GLL points: Nr per point and assembled 1D mass.  Replaces
 * GLLPoint::release for solid points (src/preloop/se_model/gll/
 * GLLPoint.cpp:40-77) with Mass1D (src/core/point/mass/Mass1D.hpp:19-22). */
int axs3d_set_solid_points(axs3d_ctx *ctx, int n, const int32_t *nr,
                           const double *mass);

/* Solid elements: point indices [n][25], axial flags and the GradQuad
 * operands gq[n][6][25] = dsdxii, dsdeta, dzdxii, dzdeta, inv_s, ifact
 * (inverse-Jacobian entries under the reference's Jacobian-entry names,
 * Quad::createGradient, src/preloop/se_model/quad/Quad.hpp:157-184, and
 * the integral factor, Quad.cpp:404-435).  Replaces the GradientQuadrature
 * constructor (GradientQuadrature.hpp:122-137) and SolidElement's point
 * set (src/core/element/Element.cpp:28-43). */
int axs3d_set_solid_elements(axs3d_ctx *ctx, int n, const int32_t *points,
                             const uint8_t *axial, const double *gq);

/* 1D isotropic solid material, lam_mu[n][2][25] = lambda, mu.  Replaces
 * the Isotropic 1D constructor (src/core/element/material/elastic/
 * anisotropy/Isotropic.hpp:24-34) built by Material::createIsotropic
 * (src/preloop/se_model/quad/physical/Material.cpp:566-582). */
int axs3d_set_solid_iso1d(axs3d_ctx *ctx, const double *lam_mu);

/* Element kinds, one per solid element: 0 = 1D material (Fourier path),
 * 1 = 3D isotropic material (cardinal path: FFT to Nr slices, stress per
 * slice, FFT back).  Replaces the Fourier/cardinal branch of
 * SolidElement::displToStiff (src/core/element/SolidElement.cpp:160-198)
 * selected by Elastic::is1D (Elastic.hpp:37).  Set before 3D properties. */
int axs3d_set_element_kinds(axs3d_ctx *ctx, const uint8_t *kind);

/* 3D isotropic material of the 3D elements in [e_begin, e_end): for each
 * such element in order, Nr slices (Nr = element Nr, max over its points)
 * of [2][25] = lambda, mu at phi_k = 2 pi k / Nr.  May be called in
 * chunks.  Replaces the Isotropic 3D constructor (Isotropic.hpp:36-46,
 * Property<N> 3D data Nr x 25, src/core/element/common/Property.hpp:39-42). */
int axs3d_set_solid_iso3d(axs3d_ctx *ctx, int e_begin, int e_end,
                          const double *lam_mu);

/* Fluid points: Nr and assembled mass (integral factor / kappa) per
 * point, as for solid points (GLLPoint::release / FluidPoint,
 * src/preloop/se_model/gll/GLLPoint.cpp:78-112; Material::getMass,
 * Material.cpp:308-346). */
int axs3d_set_fluid_points(axs3d_ctx *ctx, int n, const int32_t *nr,
                           const double *mass);

/* Fluid elements (FluidElement, src/core/element/FluidElement.cpp):
 * points [n][25] -> fluid point, axial flags, GradQuad operands
 * [n][6][25] as for axs3d_set_solid_elements. */
int axs3d_set_fluid_elements(axs3d_ctx *ctx, int n, const int32_t *points,
                             const uint8_t *axial, const double *gq);

/* 1D acoustic medium of the fluid elements: K = 1 / rho [n][25]
 * (Acoustic 1D constructor, Acoustic.hpp:20-25, from
 * Material::createAcoustic, Material.cpp:350-360). */
int axs3d_set_fluid_k1d(axs3d_ctx *ctx, const double *k);

/* Fluid element kinds, one per fluid element: 0 = 1D medium (Fourier
 * path), 1 = 3D medium (cardinal path: C2R of the three gradient
 * components, K per slice, R2C).  Replaces the branch of
 * FluidElement::displToStiff (src/core/element/FluidElement.cpp:101-148)
 * that Acoustic's 1D/3D property selects.  After the fluid points and
 * elements, before the 3D medium. */
int axs3d_set_fluid_kinds(axs3d_ctx *ctx, const uint8_t *kind);

/* 3D acoustic medium of the 3D fluid elements in [e_begin, e_end): per
 * such element in order, Nr slices (element Nr, max over its points) of
 * [25] K = 1 / rho at phi_k = 2 pi k / Nr.  May be called in chunks.
 * Replaces the Acoustic 3D constructor (src/core/element/material/
 * acoustic/Acoustic.hpp:27-35) built by Material::createAcoustic
 * (src/preloop/se_model/quad/physical/Material.cpp:350-360). */
int axs3d_set_fluid_k3d(axs3d_ctx *ctx, int e_begin, int e_end,
                        const double *k);

/* Fluid points on the axis (AxialBoundary::addPoint, fluid) and on the
 * fluid free surface (FluidSurfaceBoundary::addPoint,
 * src/core/boundary/surface/FluidSurfaceBoundary.hpp). */
int axs3d_set_fluid_axial_points(axs3d_ctx *ctx, int n,
                                 const int32_t *points);
int axs3d_set_fluid_surface_points(axs3d_ctx *ctx, int n,
                                   const int32_t *points);

/* Solid-fluid coupling pairs: solid point, fluid point (same Nr) and
 * normals [n][4] = ns, nz unassembled, ns, nz assembled (pointing from
 * fluid to solid); the fluid mass is the fluid point's.  Replaces the
 * SolidFluidCoupling1D constructor (src/core/boundary/solid_fluid/
 * SolidFluidCoupling1D.hpp:18-24, GLLPoint.cpp:114-150). */
int axs3d_set_solid_fluid_coupling(axs3d_ctx *ctx, int n,
                                   const int32_t *solid_points,
                                   const int32_t *fluid_points,
                                   const double *normals);
/* Fluid points with a 3D mass: points and their Nr masses per slice, in
 * order.  Replaces Mass3D for fluid points (src/core/point/mass/
 * Mass3D.cpp:41-53, built by GLLPoint::release, GLLPoint.cpp:78-112).  A
 * coupled fluid point with a 3D mass needs its pair in
 * axs3d_set_coupling3d. */
int axs3d_set_fluid_mass3d(axs3d_ctx *ctx, int n, const int32_t *points,
                           const double *mass);
/* 3D solid-fluid coupling of the listed pairs (indices into the pairs of
 * axs3d_set_solid_fluid_coupling): normals (s, phi, z) per slice of the
 * solid point's Nr, unassembled and assembled, [Nr][3] per pair in order;
 * the fluid mass per slice is the fluid point's (1D or 3D).  Replaces the
 * SolidFluidCoupling3D constructor (src/core/boundary/solid_fluid/
 * SolidFluidCoupling3D.cpp:15-24, GLLPoint.cpp:129-143). */
int axs3d_set_coupling3d(axs3d_ctx *ctx, int n, const int32_t *pairs,
                         const double *n_unassembled,
                         const double *n_assembled);

/* Number of fluid point-modes (length of a fluid field / 2). */
int axs3d_num_fluid_modes(const axs3d_ctx *ctx, int64_t *nmodes);

/* Material type per element: 0 isotropic, 1 transversely isotropic
 * (radial axis), 2 fully anisotropic; the choice Material::createElastic
 * makes (src/preloop/se_model/quad/physical/Material.cpp:366-379).  Set
 * before any material; default all isotropic. */
int axs3d_set_material_types(axs3d_ctx *ctx, const uint8_t *mtype);

/* 1D material of every element, packed in element order: per element
 * np x 25 values, np = 2 (lambda, mu), 5 (A, C, F, L, N) or 21 (C11 C12 ..
 * C16 C22 .. C26 C33 .. C66) by type (entries of 3D elements ignored).
 * Replaces the Isotropic / TransverselyIsotropic / Anisotropic 1D
 * constructors (Isotropic.hpp:26-34, TransverselyIsotropic.hpp:26-40,
 * Anisotropic.hpp:27-60), after Material::createTISO / createAnisotropic
 * (Material.cpp:385-560). */
int axs3d_set_solid_material1d(axs3d_ctx *ctx, const double *params);

/* 3D material of the 3D elements in [e_begin, e_end): per element in
 * order, Nr slices of [np][25] (np by type as above).  May be called in
 * chunks.  Replaces the 3D constructors of the same classes (Property<N>
 * 3D data, src/core/element/common/Property.hpp:39-42). */
int axs3d_set_solid_material3d(axs3d_ctx *ctx, int e_begin, int e_end,
                               const double *params);

/* Colatitude theta [n_elem][25] of every GLL point on a spherical mesh
 * (geodesy::sz2rtheta); TI and anisotropic elements then form stress in
 * RTZ (Element::createCoordTransform, src/core/element/Element.cpp:
 * 133-151; CoordTransformSpherical::transformSPZ_RTZ6 / RTZ_SPZ6,
 * src/core/element/crd/CoordTransformSpherical.hpp:186-249).  Not set:
 * Cartesian mesh (CoordTransformCartesian, no rotation). */
int axs3d_set_rtz_theta(axs3d_ctx *ctx, const double *theta);

/* Standard linear solids: alpha, beta, gamma per SLS
 * (AttBuilder::setAlphaBetaGamma, src/preloop/se_model/quad/physical/
 * AttBuilder.cpp:183-198 -> Attenuation::setAlphaBetaGamma,
 * src/core/element/material/elastic/attenuation/Attenuation.hpp:144-150);
 * cg4 = 1 selects AttenuationCG4 (4 points), 0 AttenuationFull. */
int axs3d_set_attenuation(axs3d_ctx *ctx, int nsls, int cg4,
                          const double *alpha, const double *beta,
                          const double *gamma);

/* dlambda, dmu of the attenuation of the 3D elements in [e_begin, e_end):
 * for each such element, Nr slices of [2][npa], npa = 4 (CG4, points
 * (1,1),(1,3),(3,1),(3,3), already scaled by the CG4 weights) or 25.
 * Replaces the AttenuationCG4/Full 3D constructors built by
 * AttBuilder::createAttenuation (AttBuilder.cpp:103-178). */
int axs3d_set_attenuation3d(axs3d_ctx *ctx, int e_begin, int e_end,
                            const double *dlam_dmu);

/* dlambda, dmu of the attenuation of the 1D (Fourier-path) elements:
 * [n_elem][2][npa] in element order (entries of 3D elements are ignored),
 * npa as above, CG4 values already scaled by the CG4 weights.  Replaces
 * the AttenuationCG4/Full 1D constructors built by
 * AttBuilder::createAttenuation (AttBuilder.cpp:103-178;
 * AttenuationCG4.hpp:40-47). */
int axs3d_set_attenuation1d(axs3d_ctx *ctx, const double *dlam_dmu);

/* Solid points on the axis.  Replaces AxialBoundary::addPoint
 * (src/core/boundary/axial/AxialBoundary.hpp). */
int axs3d_set_axial_points(axs3d_ctx *ctx, int n, const int32_t *points);

/* 3D mass of solid points: for each listed point in order, its Nr values
 * of the assembled mass at phi_k = 2 pi k / Nr (after
 * axs3d_set_solid_points).  These points divide the stiffness by the mass
 * slice by slice (C2R, times 1 / m, R2C) instead of by their 1D mass.
 * Replaces Mass3D (src/core/point/mass/Mass3D.hpp:17-22, Mass3D.cpp:
 * 56-68), which GLLPoint::release creates for a 3D density (src/preloop/
 * se_model/gll/GLLPoint.cpp:54-77). */
int axs3d_set_solid_mass3d(axs3d_ctx *ctx, int n, const int32_t *points,
                           const double *mass);

/* Ocean load on solid points (after axs3d_set_solid_points; replaces the
 * listed points' mass).  Entry i acts on points[i] with nval[i] values of
 * 4 numbers each: nval 1 gives MassOceanLoad1D, the factors 1/m cos^2 t +
 * 1/(m + m0) sin^2 t, 1/m sin^2 t + 1/(m + m0) cos^2 t, (1/(m + m0) -
 * 1/m) cos t sin t and 1/m (FP64, MassOceanLoad1D.hpp:20-30); nval = the
 * point's Nr gives MassOceanLoad3D, per slice 1/m and k = sqrt(m0 / (m (m
 * + m0))) n (MassOceanLoad3D.cpp:13-20, in the working precision), m the
 * point's mass, m0 its ocean mass, t the colatitude of its surface normal
 * n.  Replaces the ocean-load branches of GLLPoint::release (src/preloop/
 * se_model/gll/GLLPoint.cpp:54-87).  A listed point must not also have a
 * 3D mass from axs3d_set_solid_mass3d. */
int axs3d_set_ocean_load(axs3d_ctx *ctx, int n, const int32_t *points,
                         const int32_t *nval, const double *values);

/* Undulated geometry (topography, ellipticity, undulated interfaces) of
 * solid elements elems[n] by the particle relabelling transformation:
 * x[n][5][25] = X0, X1, X2, X3 (the inverse Jacobian of the relabelling,
 * eq. 13 of Leng et al. 2019) and XJ (its |J|) at each GLL point, one set
 * for all slices (an axisymmetric undulation).  Replaces the element's PRT
 * (src/core/element/PRT/PRT.hpp:57-80, built by Undulation::createPRT,
 * src/preloop/se_model/quad/physical/Undulation.cpp:124-168) and the PRT
 * branch of SolidElement::displToStiff (SolidElement.cpp:115-159); the
 * point masses passed to axs3d_set_solid_points already include |J|
 * (Undulation::getMassJacobian).  A moment source in such an element
 * takes the PRT branch of SolidElement::addMomentSource (:293-331).  Not
 * with the adjoint run. */
int axs3d_set_solid_prt(axs3d_ctx *ctx, int n, const int32_t *elems,
                        const double *x);

/* Moment-tensor source: element, Fourier pattern [nu1][6][25] complex and
 * the source-time function sampled at every step (FP64).  Replaces
 * SolidMoment (src/core/source/elem_src/solid/SolidMoment.cpp:15-33) with
 * its GaussianSTF (src/core/source/elem_src/stf/GaussianSTF.hpp:35-58). */
int axs3d_set_moment_source(axs3d_ctx *ctx, int elem, int nu1,
                            const double *pattern, int nsteps,
                            const double *stf);
/* One more moment source, for batched instances (no reference
 * counterpart: the reference runs one simulation per process).  Each
 * source lies in its own element; every STF has the same nsteps. */
int axs3d_add_moment_source(axs3d_ctx *ctx, int elem, int nu1,
                            const double *pattern, int nsteps,
                            const double *stf);
/* Batched independent instances in one domain (no reference counterpart):
 * point_inst[p] is the instance of solid point p (after
 * axs3d_set_solid_points) and dt[i] the time step of instance i, used by
 * the Newmark update of its points.  The instances' meshes must not share
 * points; their elements, sources and receivers are simply concatenated.
 * Not with fluid or 3D-mass points. */
int axs3d_set_instances(axs3d_ctx *ctx, int n, const int32_t *point_inst,
                        const double *dt);

/* Receivers: element, in-plane weights [n][25], azimuth phi.  Replaces
 * StationSolid::setElement (src/core/output/station-wise/station/
 * StationSolid.cpp:18-27, Station.cpp:13-38).  max_samples bounds the
 * device seismogram buffer (one sample per recorded step). */
int axs3d_set_receivers(axs3d_ctx *ctx, int n, const int32_t *elem,
                        const double *weights, const double *phi,
                        int max_samples);
/* Station frames: per receiver the wavefield coordinate system wcs (0
 * spz, 1 RTZ, 2 ENZ, 3 xyz; channel::WavefieldCS, src/core/output/
 * channel.hpp:24), theta (radians) and back azimuth (radians), the
 * arguments of the Station constructor (Station.hpp:24-29, built at
 * StationOutput.cpp:429-447); cartesian != 0 for a Cartesian mesh.  The
 * receiver kernel then records the displacement rotated as
 * Station::rotateField does (Station.hpp:79-118).  Without this call the
 * seismograms stay in (s, phi, z). */
int axs3d_set_receiver_frames(axs3d_ctx *ctx, int n, const int32_t *wcs,
                              const double *theta,
                              const double *back_azimuth, int cartesian);

/* ---------------- absorbing boundaries (src/core/boundary/ABC) ------- */
/* Entries are added (repeated points allowed, as a point on two boundary
 * element edges gets two); they are merged per point at finalize.
 * Clayton-Enquist on solid points, 1D form: per entry rho vp, rho vs, the
 * area (|n| of the element's integrated edge normal) and theta (its angle
 * to z), the ClaytonSolid1D constructor (ClaytonSolid1D.hpp:19-33, built at
 * GLLPoint.cpp:172-178). */
int axs3d_add_clayton_solid(axs3d_ctx *ctx, int n, const int32_t *points,
                            const double *rho_vp, const double *rho_vs,
                            const double *area, const double *theta);
/* 3D form, ClaytonSolid3D (ClaytonSolid3D.hpp:19-27, GLLPoint.cpp:
 * 180-191): per entry Nr (of its point) values of rho vp, rho vs, area and
 * unit normal [Nr][3] (s, phi, z), concatenated in entry order. */
int axs3d_add_clayton_solid3d(axs3d_ctx *ctx, int n, const int32_t *points,
                              const double *rho_vp, const double *rho_vs,
                              const double *area,
                              const double *unit_normal);
/* Clayton-Enquist on fluid points, ClaytonFluid1D (per_slice = 0: one
 * rho vp and area per entry) or ClaytonFluid3D (Nr of each), GLLPoint.cpp:
 * 156-170. */
int axs3d_add_clayton_fluid(axs3d_ctx *ctx, int n, const int32_t *points,
                            int per_slice, const double *rho_vp,
                            const double *area);
/* Kosloff sponge, Sponge<SolidPoint> (fluid = 0) or Sponge<FluidPoint>
 * (Sponge.hpp:30-45, GLLPoint.cpp:193-215): gamma per entry (per_slice =
 * 0) or Nr values per entry. */
int axs3d_add_sponge(axs3d_ctx *ctx, int fluid, int n, const int32_t *points,
                     int per_slice, const double *gamma);

/* Time step.  Replaces TimeScheme::setTime (src/core/time/TimeScheme.hpp)
 * as used by NewmarkTimeScheme::update's dt constants
 * (NewmarkTimeScheme.hpp:65-67). */
int axs3d_set_dt(axs3d_ctx *ctx, double dt);

/* Time scheme: 0 = NewmarkTimeScheme (default), 1 = SymplecticTimeScheme
 * (src/core/time/SymplecticTimeScheme.cpp:18-142: per step the receivers,
 * the source, u += kappa_0 dt v, then four sub-steps of stiffness,
 * boundaries, mass and v += pi_k dt a, u += kappa_k dt v).  axs3d_run
 * follows the chosen scheme; one time step for the whole domain. */
int axs3d_set_time_scheme(axs3d_ctx *ctx, int scheme);

/* Ends set-up: checks sizes and compatibility (the reference's
 * checkCompatibility calls, e.g. src/core/element/common/Property.hpp:53-80)
 * and builds the device work lists. */
int axs3d_finalize(axs3d_ctx *ctx);

/* ---------------- per-phase entry points (Domain, Domain.cpp:285-345) -- */
/* Domain::applySources (Domain.cpp:285-295). */
int axs3d_apply_sources(axs3d_ctx *ctx, int tstep);
/* Domain::computeStiffness (Domain.cpp:298-308). */
int axs3d_compute_stiffness(axs3d_ctx *ctx);
/* No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops):
 * axs3d_compute_stiffness with the cardinal path's strain and stress rows
 * in a file-scope device array of 64 MiB per device (held until the
 * process ends) instead of the context's scratch, in batches planned for
 * it at axs3d_finalize; bitwise the same stiffness.  Calls no hipMalloc,
 * hipFree or other allocating routine.  Returns 1 (with the message in
 * axs3d_last_error) when one 3D element's rows exceed 64 MiB.  Every
 * context on a device shares that device's array: calls that overlap on
 * one device (contexts on different streams) race.  Nothing in the array
 * needs to start at zero: each batch writes every row it reads. */
int axs3d_compute_stiffness_static(axs3d_ctx *ctx);
/* Domain::applyBC_BeforeAssemblingStiff (Domain.cpp:313-317): the
 * Clayton-Enquist boundaries, then the solid-fluid coupling. */
int axs3d_apply_bc_before(axs3d_ctx *ctx);
/* Domain::recordOutput for station displacement (Domain.cpp:385-409),
 * stored as sample `isample`. */
int axs3d_record(axs3d_ctx *ctx, int isample);
/* Domain::applyBC_AfterAssemblingStiff (Domain.cpp:322-326). */
int axs3d_apply_bc_after(axs3d_ctx *ctx);
/* Domain::applyBC_AfterComputingAccel (Domain.cpp:330-332): the sponge on
 * the acceleration that axs3d_stiff_to_accel left, with the terms of this
 * step's axs3d_apply_bc_before. */
int axs3d_apply_bc_after_accel(axs3d_ctx *ctx);
/* Domain::computeStiffToAccel (Domain.cpp:335-345): Nyquist masking and
 * mass division, in place in the stiffness field. */
int axs3d_stiff_to_accel(axs3d_ctx *ctx);
/* NewmarkTimeScheme::update on all points (NewmarkTimeScheme.hpp:48-71);
 * expects the stiffness field to hold the acceleration. */
int axs3d_newmark(axs3d_ctx *ctx);
/* SymplecticTimeScheme::launch (SymplecticTimeScheme.hpp:51-58): u +=
 * kappa_0 dt v on all points, once per step after the source. */
int axs3d_symplectic_launch(axs3d_ctx *ctx);
/* SymplecticTimeScheme::update of sub-step isub = 1..4 (Symplectic-
 * TimeScheme.hpp:37-46, 61-70): v += pi_isub dt a, u += kappa_isub dt v
 * with a the acceleration in the stiffness field, which is then zeroed. */
int axs3d_symplectic_update(axs3d_ctx *ctx, int isub);
/* Domain::checkStability (Domain.cpp:574-612): raises when a displacement
 * is non-finite.  tstep is the step number the message names, counted
 * from 1 as NewmarkTimeScheme::solve passes it (tstep + 1 of its loop);
 * runs of axs3d_run instead name the first step that started from a
 * non-finite displacement. */
int axs3d_check_stability(axs3d_ctx *ctx, int tstep);

/* ---------------- bulk ---------------------------------------------- */
/* NewmarkTimeScheme::solve (src/core/time/NewmarkTimeScheme.cpp:18-131)
 * over steps [step_begin, step_begin + nsteps), recording every step into
 * the seismogram buffer at sample (tstep - record_step0).  The
 * BC-after/mass/Newmark phases run as one fused kernel. */
int axs3d_run(axs3d_ctx *ctx, int step_begin, int nsteps);
int axs3d_set_record_origin(axs3d_ctx *ctx, int record_step0);
/* Launch form of axs3d_run, no counterpart in the reference (its steps are
 * host loops): enable != 0 (the default) captures one step as a HIP graph
 * on the first axs3d_run and replays it once per step, with the step index
 * in device memory; 0 launches every kernel of every step from the host.
 * Both forms give bitwise the same fields.  Profiling (axs3d_profile_enable)
 * uses the launch-by-launch form. */
int axs3d_set_graph(axs3d_ctx *ctx, int enable);

/* ---------------- adjoint --------------------------------------------- */
/* No counterpart in the reference (it has no adjoint solver): the
 * discrete adjoint of axs3d_run for solid domains of isotropic elastic
 * elements, 1D or 3D, with or without SLS attenuation (other domains:
 * error; with attenuation the derivatives are with respect to the
 * unrelaxed moduli at fixed anelastic coefficients).  Runs
 * the forward problem from rest over steps [0, nsteps), keeping (u, v, a)
 * (and the SLS memory variables) every checkpoint_interval steps on the
 * device, then the adjoint field (with memory variables of its own)
 * backwards in time with the force R^* djdr[n] / dt^2 (djdr:
 * [nsteps][n_receivers][3] FP64, the derivative of the misfit with respect
 * to the seismogram sample of step n, as axs3d_read_seismograms lays it
 * out), recomputing each interval's forward fields from its checkpoint.
 * kernels ([n_elements][25][3] FP64) receives dJ/dlambda, dJ/dmu, dJ/drho
 * at every GLL point of every element (rho through the assembled mass
 * sum rho IF; lambda and mu zero for 3D elements); kernels3d ([slice][2]
 * [25] FP64, the layout of axs3d_set_solid_iso3d; NULL without 3D
 * elements) dJ/dlambda and dJ/dmu per slice of the 3D elements.  Exact for
 * odd element Nr; an even Nr's Nyquist row enters the transforms with
 * weights that make the stiffness only nearly self-adjoint.  Leaves the
 * wavefields undefined (a later axs3d_run must start from fields set to
 * rest) and the seismograms untouched. */
int axs3d_adjoint_run(axs3d_ctx *ctx, int nsteps, int checkpoint_interval,
                      const double *djdr, double *kernels,
                      double *kernels3d);
/* The same adjoint run for any solid media (no reference counterpart):
 * isotropic, transversely isotropic and anisotropic elements, 1D or 3D,
 * with or without attenuation.  kernels_par receives dJ/d(material
 * parameters) of every element packed as axs3d_set_solid_material1d takes
 * them (per element its type's parameters x 25: lambda, mu; A, C, F, L,
 * N; or the 21 Cij), zero for 3D elements; kernels_rho ([n_elements][25])
 * dJ/drho through the assembled mass; kernels_par3d (NULL without 3D
 * elements) the per-slice parameters of the 3D elements packed as
 * axs3d_set_solid_material3d takes them.  TI and anisotropic elements on
 * spherical meshes take the derivatives in the RTZ frame their moduli are
 * given in.  With attenuation the derivatives are those of the unrelaxed
 * moduli at fixed (dlambda, dmu). */
int axs3d_adjoint_run_media(axs3d_ctx *ctx, int nsteps,
                            int checkpoint_interval, const double *djdr,
                            double *kernels_par, double *kernels_rho,
                            double *kernels_par3d);
/* The per-step densities of axs3d_adjoint_run, unscaled, for one adjoint
 * field psi, forward displacement u and forward acceleration a (each laid
 * out as axs3d_get_field): kernels[e][k] = (<psi, dK/dlambda u>,
 * <psi, dK/dmu u>, <psi, dM/drho a>) at GLL point k of element e. */
int axs3d_adjoint_density(axs3d_ctx *ctx, const double *psi, const double *u,
                          const double *a, double *kernels);
/* Copies samples [0, nsamples) of the seismograms, [sample][rec][3]. */
int axs3d_read_seismograms(axs3d_ctx *ctx, double *out, int nsamples);

/* ---------------- debug hooks ---------------------------------------- */
int axs3d_num_modes(const axs3d_ctx *ctx, int64_t *nmodes);
int axs3d_get_field(axs3d_ctx *ctx, int which, double *out);
int axs3d_set_field(axs3d_ctx *ctx, int which, const double *in);
int axs3d_synchronize(axs3d_ctx *ctx);

/* ---------------- measurement --------------------------------------- */
/* When enabled, every launch of the kernels in axs3d_kernel_id is
 * bracketed by HIP events on the context's stream; read returns the
 * summed milliseconds and the launch count since the last reset. */
int axs3d_profile_enable(axs3d_ctx *ctx, int enable);
int axs3d_profile_read(axs3d_ctx *ctx, int kernel, double *total_ms,
                       int64_t *count);
/* Diagnostic, no reference counterpart: runs the library's azimuthal
 * FFT for Nr slices on the host, pass by pass and lane by lane exactly as
 * the device kernels do (same plan, radices, twiddle table, digit
 * reversal), in FP64, so that the plans can be checked against a DFT
 * without a GPU.  The reference's transform is FFTW's (SolverFFTW.hpp:
 * 101-168); this one follows its sign conventions, unscaled: inverse != 0
 * computes X_k = sum_j x_j exp(+2 pi i jk / nr), else exp(-2 pi i jk /
 * nr); data holds nr interleaved complex values, in and out in natural
 * order.  radices (at least 12 ints) and npasses receive the pass radices
 * chosen with the largest radix max_radix (2..16).  Returns 0, 1 for an
 * Nr with a prime factor above 13, 2 for bad arguments. */
int axs3d_fft_check(int nr, int max_radix, int inverse, double *data,
                    int *radices, int *npasses);

int axs3d_profile_reset(axs3d_ctx *ctx);

/* The cardinal path as set up at axs3d_finalize (no reference
 * counterpart; the reference has no GPU): the stress kernel's minimum
 * waves per SIMD, its resident blocks on the device, and in trial_ms3 the
 * time (ms) of one cardinal pass with the chosen batches, the cardinal
 * scratch size chosen by the autotuner (MiB; it times one pass per
 * candidate size) and the number of sizes timed. */
int axs3d_autotune_info(const axs3d_ctx *ctx, int *cardinal_occupancy,
                        int *cardinal_blocks, double *trial_ms3);
/* Every choice the autotuner made at axs3d_finalize (no reference
 * counterpart): out[0] the chosen cardinal pass (ms), [1] the scratch
 * size (MiB), [2] scratch sizes timed, [3] the FFT plans' largest pass
 * radix, [4] radices timed, then (radix, pass ms) for each radix timed;
 * at most n values, their count in *written. */
int axs3d_autotune_report(const axs3d_ctx *ctx, double *out, int n,
                          int *written);

#ifdef __cplusplus
}
#endif

#endif /* AXS3D_H */
