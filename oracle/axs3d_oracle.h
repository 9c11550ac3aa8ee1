/*
 * axs3d_oracle.h -- FP64 CPU restatement of the AxiSEM3D stiffness path.
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline); the product
 * path never includes, links or loads it.
 */
#ifndef AXS3D_ORACLE_H
#define AXS3D_ORACLE_H

#include <complex.h>
#include <stdint.h>

typedef struct {
    /* solid points */
    int n_points;
    const int32_t *point_nr;      /* Nr per point                        */
    const double *point_mass;     /* assembled 1D mass per point         */
    const int64_t *offsets;       /* prefix sum of Nr/2+1, n_points + 1  */
    /* solid elements */
    int n_elem;
    const int32_t *elem_points;   /* [e][25]                             */
    const int32_t *elem_nr;       /* element Nr = max of its points      */
    const uint8_t *elem_axial;
    const double *elem_gq;        /* [e][6][25]                          */
    const double *elem_iso;       /* [e][2][25] lambda, mu (1D elements) */
    const uint8_t *elem_kind;     /* NULL or per element: 0 1D (Fourier
                                     path), 1 3D isotropic (cardinal)    */
    const int64_t *elem_slice_off;/* 3D: first slice of element e        */
    const double *prop3d;         /* 3D: [slice][2][25] lambda, mu       */
    /* attenuation (standard linear solids) */
    int nsls;                     /* 0: none                             */
    int att_cg4;                  /* 1: CG4 (4 points), 0: FULL (25)     */
    const double *att_abg;        /* [3][nsls] alpha, beta, gamma        */
    const double *att1d;          /* [e][2][npa] dlambda, dmu (1D)       */
    const double *att3d;          /* [slice][2][npa] dlambda, dmu (3D)   */
    void *att_state;              /* owned, see oracle_att_alloc          */
    /* axis */
    int n_axial;
    const int32_t *axial_points;
    /* moment source */
    int src_elem;                 /* -1: none                            */
    int src_nu1;
    const double complex *src_pattern;  /* [nu1][6][25]                  */
    const double *stf;                  /* per step                      */
    /* receivers */
    int n_rec;
    const int32_t *rec_elem;
    const double *rec_weights;    /* [r][25]                             */
    const double *rec_phi;
    /* time */
    double dt;
    /* material types (NULL: all isotropic from elem_iso / prop3d by
     * elem_slice_off); with types, parameters are packed per element:
     * 1D mat1d + mat1d_off[e] = [np][25], 3D prop3d + prop_off[e] =
     * [Nr][np][25], np = 2, 5, 21 for isotropic, TI, anisotropic */
    const uint8_t *elem_mtype;
    const double *mat1d;
    const int64_t *mat1d_off;
    const int64_t *prop_off;
    const double *elem_theta;     /* [e][25] colatitude (spherical), or
                                     NULL (Cartesian)                    */
    /* fluid (displacement potential, 1 component per mode) */
    int n_fpoints;
    const int32_t *fpoint_nr;
    const double *fpoint_mass;    /* assembled IF / kappa                */
    const int64_t *foffsets;      /* prefix sum of Nr/2+1                */
    int n_felem;
    const int32_t *felem_points;  /* [e][25] -> fluid point              */
    const int32_t *felem_nr;
    const uint8_t *felem_axial;
    const double *felem_gq;       /* [e][6][25]                          */
    const double *felem_k;        /* [e][25] K = 1 / rho                 */
    int n_faxial;
    const int32_t *faxial_points;
    int n_fsurf;
    const int32_t *fsurf_points;
    /* solid-fluid coupling */
    int n_sf;
    const int32_t *sf_solid;
    const int32_t *sf_fluid;
    const double *sf_normal;      /* [pair][2] (ns, nz), fluid -> solid   */
    /* 3D mass (NULL: all 1D): per solid point the offset of its Nr masses
     * in mass3d, or -1 for a 1D mass */
    const int64_t *mass3d_off;
    const double *mass3d;
    /* 3D acoustic media (NULL: all fluid elements 1D): per fluid element
     * 0 = 1D (felem_k), 1 = 3D with K per slice at fk3d + fk3d_off[e],
     * [Nr][25] */
    const uint8_t *felem_kind;
    const double *fk3d;
    const int64_t *fk3d_off;
    /* 3D fluid mass (NULL: all 1D): per fluid point the offset of its Nr
     * masses in fmass3d, or -1 */
    const int64_t *fmass3d_off;
    const double *fmass3d;
    /* 3D solid-fluid coupling (NULL: all 1D): per pair the offset of its
     * Nr normals (s, phi, z) in sf3d_normal, or -1 for the 1D normal */
    const int64_t *sf3d_off;
    const double *sf3d_normal;
    /* absorbing boundaries, one list per kind: entry i acts on point
     * *_point[i] with nval[i] values (1: the 1D form, Nr: one per slice,
     * the 3D form) starting at value *_off[i] of *_data.
     *  clayton_s (ClaytonSolid1D/3D): 4 numbers per value, 1D the
     *    constructor's RSA cos^2 + RPA sin^2, RSA sin^2 + RPA cos^2,
     *    (RPA - RSA) cos sin, RSA (ClaytonSolid1D.hpp:22-33); 3D per slice
     *    rsa and k = sqrt((rpa - rsa) area) n (ClaytonSolid3D.hpp:20-27);
     *  clayton_f (ClaytonFluid1D/3D): area / (rho vp);
     *  sponge_s, sponge_f (Sponge): gamma. */
    int n_clayton_s;
    const int32_t *clayton_s_point, *clayton_s_nval;
    const int64_t *clayton_s_off;
    const double *clayton_s_data;
    int n_clayton_f;
    const int32_t *clayton_f_point, *clayton_f_nval;
    const int64_t *clayton_f_off;
    const double *clayton_f_data;
    int n_sponge_s;
    const int32_t *sponge_s_point, *sponge_s_nval;
    const int64_t *sponge_s_off;
    const double *sponge_s_data;
    int n_sponge_f;
    const int32_t *sponge_f_point, *sponge_f_nval;
    const int64_t *sponge_f_off;
    const double *sponge_f_data;
    /* ocean load on solid points (NULL: none): per solid point the offset
     * of its values in ocean_data, or -1, and their count in ocean_nval,
     * 4 numbers per value.  1D (1 value): the MassOceanLoad1D constructor's
     * 1/m cos^2 t + 1/(m + m0) sin^2 t, 1/m sin^2 t + 1/(m + m0) cos^2 t,
     * (1/(m + m0) - 1/m) cos t sin t and 1/m (MassOceanLoad1D.hpp:20-30);
     * 3D (Nr values, one per slice): 1/m and k = sqrt(m0 / (m (m + m0))) n
     * (MassOceanLoad3D.cpp:13-20).  The point's own mass is not used. */
    const int64_t *ocean_off;
    const int32_t *ocean_nval;
    const double *ocean_data;
    /* particle relabelling transformation of undulated geometry (NULL:
     * none): per solid element the offset of its operands in prt_data, or
     * -1; one set for all slices (1D undulation), [5][25] X0, X1, X2, X3,
     * XJ (PRT.hpp:57-80, Undulation::createPRT). */
    const int64_t *prt_off;
    const double *prt_data;
} oracle_domain;

void oracle_set_gmat(const double *g_gll, const double *g_glj);
void oracle_grad6(const double *gq_e, int axial, int alpha,
                  const double complex *u, double complex *e);
void oracle_quad6(const double *gq_e, int axial, int alpha,
                  const double complex *sij, double complex *fi);
void oracle_quad6_noint(const double *gq_e, int axial, int alpha,
                        const double complex *sij, double complex *fi);
void oracle_iso_stress(const double *lam, const double *mu,
                       const double complex *e, double complex *s);
int oracle_mat_nparams(int mt);
void oracle_material_stress(int mt, const double *par,
                            const double complex *e, double complex *s);
void oracle_spz_to_rtz6(const double *theta, double complex *e);
void oracle_grad9(const double *gq_e, int axial, int alpha,
                  const double complex *u, double complex *g);
void oracle_quad9(const double *gq_e, int axial, int alpha,
                  const double complex *f, double complex *fi);
void oracle_spz_to_rtz9(const double *theta, double complex *g);
void oracle_rtz_to_spz9(const double *theta, double complex *g);
void oracle_prt_sph_to_und6(const double *x, const double complex *g9,
                            double complex *e6);
void oracle_prt_und_to_sph9(const double *x, const double complex *s6,
                            double complex *f9);
void oracle_rtz_to_spz6(const double *theta, double complex *s);
void oracle_grad3(const double *gq_e, int axial, int alpha,
                  const double complex *p, double complex *pi);
void oracle_quad3(const double *gq_e, int axial, int alpha,
                  const double complex *qi, double complex *q);
void oracle_fluid_element_stiff(const oracle_domain *d, int e,
                                const double complex *chi,
                                double complex *fstiff);
void oracle_coupling(const oracle_domain *d, const double complex *displ,
                     double complex *stiff, double complex *fstiff);
void oracle_fluid_bc(const oracle_domain *d, double complex *fstiff);
void oracle_fluid_stiff_to_accel(const oracle_domain *d,
                                 double complex *fstiff);
int oracle_run_coupled(const oracle_domain *d, int step_begin, int nsteps,
                       double complex *displ, double complex *veloc,
                       double complex *accel, double complex *stiff,
                       double complex *fdispl, double complex *fveloc,
                       double complex *faccel, double complex *fstiff,
                       double *seismo);
void oracle_solid_element_stiff(const oracle_domain *d, int e,
                                const double complex *displ,
                                double complex *stiff);
void oracle_apply_moment_source(const oracle_domain *d, double stf_value,
                                double complex *stiff);
void oracle_axial_solid(const oracle_domain *d, double complex *stiff);
void oracle_stiff_to_accel(const oracle_domain *d, double complex *stiff);
void oracle_newmark(int64_t nmodes, double dt, double complex *stiff,
                    double complex *displ, double complex *veloc,
                    double complex *accel);
void oracle_record(const oracle_domain *d, const double complex *displ,
                   double *out);
/* Clayton-Enquist boundaries on the solid and fluid stiffness
 * (AbsorbingBoundary::applyClayton), and the Kosloff sponge on the solid
 * and fluid accelerations (AbsorbingBoundary::applySponge). */
void oracle_clayton(const oracle_domain *d, const double complex *veloc,
                    double complex *stiff, const double complex *fveloc,
                    double complex *fstiff);
void oracle_sponge(const oracle_domain *d, const double complex *displ,
                   const double complex *veloc, double complex *accel,
                   const double complex *fdispl,
                   const double complex *fveloc, double complex *faccel);
int oracle_run(const oracle_domain *d, int step_begin, int nsteps,
               double complex *displ, double complex *veloc,
               double complex *accel, double complex *stiff, double *seismo);
int oracle_run_symplectic(const oracle_domain *d, int step_begin,
                          int nsteps, double complex *displ,
                          double complex *veloc, double complex *stiff,
                          double complex *fdispl, double complex *fveloc,
                          double complex *fstiff, double *seismo);
int64_t *oracle_make_offsets(int n, const int32_t *nr);
void oracle_r2c(const double *r, int nr, double complex *c);
void oracle_c2r(const double complex *c, int nr, double *r);
void *oracle_att_alloc(const oracle_domain *d);
void oracle_att_free(void *state, int n_elem);
void oracle_att_apply_fr(const oracle_domain *d, int e, int alpha,
                         const double complex *strain,
                         double complex *stress);
void oracle_att_apply_cd(const oracle_domain *d, int e, int slice,
                         const double *strain, double *stress);
void oracle_free(void *p);

/* Threads of the element loop of oracle_run / oracle_run_coupled
 * (OpenMP; the result does not depend on the count). */
void oracle_set_threads(int n);

#endif
