// axs3d_launch.hpp -- host-side launchers of the device kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "axs3d_common.hpp"

namespace axs3d {

// point-update phase bits
enum PointPhase { PHASE_AXIAL = 1, PHASE_MASS = 2, PHASE_MARCH = 4 };

template <typename Real>
void launch_solid_stiff_fr(hipStream_t st, int n_em, const EMItem *em_list,
                           const int32_t *elem_points,
                           const uint8_t *elem_axial, const Real *elem_ops,
                           const Real *elem_mat, const uint32_t *point_off,
                           const int32_t *point_nu1, const double *displ,
                           Real *stiff, const Real *gtab);

template <typename Real>
void launch_moment_source(hipStream_t st, int elem, int nu1c,
                          const Real *pattern, const double *stf, int tstep,
                          const int32_t *elem_points,
                          const uint8_t *elem_axial, const Real *elem_ops,
                          const uint32_t *point_off, const int32_t *point_nu1,
                          Real *stiff, const Real *gtab);

template <typename Real>
void launch_point(hipStream_t st, int phases, int64_t n_modes,
                  const int32_t *mode_point, const uint32_t *point_off,
                  const int32_t *point_nr, const uint8_t *point_axial,
                  const double *inv_mass, Real *stiff, double *displ,
                  double *veloc, double *accel, double dt, int *flag);

void launch_record(hipStream_t st, int nrec, const int32_t *rec_elem,
                   const double *rec_w, const double *rec_phi,
                   const int32_t *elem_points, const int32_t *elem_nu1,
                   const uint32_t *point_off, const int32_t *point_nu1,
                   const double *displ, double *out);

}  // namespace axs3d
