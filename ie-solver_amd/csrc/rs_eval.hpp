// Interior evaluation launch wrapper (rs_eval.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "rs_device.hpp"

namespace iegpu {

// Number of source chunks the partial sums are split into.
int eval_chunks(int64_t n_points);

// The allocation-free evaluation keeps its partial sums in a static device
// array of kEvalScratchElems doubles (4 MiB) per GPU.
constexpr int64_t kEvalScratchElems = int64_t(1) << 19;
constexpr int kEvalMaxDevices = 64;
// Partial-sum doubles an evaluation of n_targets needs.
int64_t eval_scratch_need(int64_t n_points, int sol_dim, int64_t n_targets);
// Device address of that array on `device`, which must be the current
// device; looked up once per device (hipGetSymbolAddress) and kept.
// Returns 0 or a hipError_t.
int eval_scratch_address(int device, double** ptr);

// out (sol_dim per target) = K_domain mu + U_forward alpha; partial holds
// eval_chunks(n_points) * sol_dim * nt doubles.
void launch_eval(const DevGeom& g, int nt, const double* targets,
                 const int* in_domain, const double* mu, int n_holes,
                 const double* hole_centers, const double* alpha,
                 double* partial, double* out, hipStream_t st);

}  // namespace iegpu
