// Interior evaluation launch wrapper (rs_eval.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "rs_device.hpp"

namespace iegpu {

// Number of source chunks the partial sums are split into.
int eval_chunks(int64_t n_points);

// out (sol_dim per target) = K_domain mu + U_forward alpha; partial holds
// eval_chunks(n_points) * sol_dim * nt doubles.
void launch_eval(const DevGeom& g, int nt, const double* targets,
                 const int* in_domain, const double* mu, int n_holes,
                 const double* hole_centers, const double* alpha,
                 double* partial, double* out, hipStream_t st);

}  // namespace iegpu
