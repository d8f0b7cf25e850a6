// FP64 fast multipole apply of the second-kind boundary operator
// (rs_fmm.hip), the O(N) counterpart of launch_matvec for one boundary.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "fmm_tree.hpp"
#include "rs_device.hpp"

namespace iegpu {

// Device view of an FmmHostTree plus the expansion storage.
struct FmmDev {
  int order;       // coefficients per expansion (32, 48 or 64)
  int depth;       // leaf level
  int n_leaves, first_leaf;
  double side;     // root square side
  const double* center;  // 2 per box
  const int* parent;
  const int* child_off;
  const int* child_cnt;
  const int* vl_off;
  const int* vl;
  const int* pt_off;
  const int* nb_off;
  const int* nb;
  const int* perm;
  const double* sgeo;  // sorted nodes, 6 each: x, y, n0, n1, weight, curvature
  double* xs;          // sorted density, sol_dim per node
  double* mexp;        // multipole coefficients, (box * n_exp + e) * order complex
  double* lexp;        // local coefficients, same layout
  const double* t_m2m;
  const double* t_m2l;
  const double* t_l2l;
};

bool fmm_order_supported(int order);
// Expansions per box: 3 for Stokes (charges c1, c2, c4), else 1.
int fmm_n_exp(int pde);

// sgeo <- the geometry in the tree's sorted order.
void launch_fmm_gather(const DevGeom& g, const FmmDev& f, double* sgeo, hipStream_t st);

// y = K x (ndof), FP64, for a single boundary (g.n_prob == 1); `partial`
// holds matvec_partials(g) doubles; level_off as in FmmHostTree.
void launch_fmm_matvec(const DevGeom& g, const FmmDev& f,
                       const std::vector<int>& level_off, const double* x,
                       double* y, double* partial, hipStream_t st);

}  // namespace iegpu
