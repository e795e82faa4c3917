// Launchers for the wrfdyn kernels (one translation unit per phase).
#pragma once

#include "wrfdyn_common.h"

namespace wd {

// RK-stage setup: transport velocities, tendencies, small_step_prep,
// calc_p_rho step 0, rw_tend/ph_tend and calc_coef_w.
hipError_t launch_stage_setup(const DGrid& g, const DNamelist& nl, const DState& s, const DWork& w,
                              int rk_step, double dts, hipStream_t st);

// One acoustic substep: advance_uv, then the fused column kernel.
hipError_t launch_acoustic_substep(const DGrid& g, const DNamelist& nl, const DState& s, const DWork& w,
                                   double dts, hipStream_t st);

// small_step_finish + grid%p refresh.
hipError_t launch_stage_finish(const DGrid& g, const DState& s, const DWork& w, hipStream_t st);

}  // namespace wd
