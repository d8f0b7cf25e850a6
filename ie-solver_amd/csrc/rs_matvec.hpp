// FP64 operator apply used by iterative refinement (rs_matvec.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "rs_device.hpp"

namespace iegpu {

// Size of the partial-sum buffer launch_matvec needs.
int matvec_partials(const DevGeom& g);

// partial[p] = problem p's rank-one sum: sum_j w_j (n_j . x_j) for Stokes,
// sum_j w_j x_j for Laplace Neumann (one workgroup per problem).
void launch_wn_dot(const DevGeom& g, const double* x, double* partial, hipStream_t st);

// y = K x (ndof), FP64, matrix-free.
void launch_matvec(const DevGeom& g, const double* x, double* y, double* partial,
                   hipStream_t st);

// y <- b - y (b may be NULL: zero); partial (2 * residual_blocks()) gets
// per-block sums of r^2 and b^2.
int residual_blocks();
void launch_residual(const double* b, double* y, int64_t n, double* partial,
                     hipStream_t st);
// y += x.
void launch_axpy(double* y, const double* x, int64_t n, hipStream_t st);

// Border of the multiply connected system: y += U alpha; z = Psi x - alpha.
void launch_border_apply(const double* U, const double* PsiT, const double* x,
                         const double* alpha, int64_t ndof, int h, double* y,
                         double* z, hipStream_t st);

}  // namespace iegpu
