// gm_theta_f32.hip - the polar-quadrature apply with the FP32 Toeplitz phase
// (mixed-precision Krylov applies; gm_theta_apply.hpp).
#include "gm_theta_apply.hpp"

namespace gmk {

template hipError_t launch_tt<float>(int, const ApplyArgs *, const ApplyArgs *, const int *, int, int, hipStream_t);

}  // namespace gmk
