// Uniform-depth quadtree of one boundary's nodes for the FP64 fast
// multipole operator apply (rs_fmm.hip).  Host side: built once per
// geometry, uploaded as flat arrays.
#pragma once

#include <cstdint>
#include <vector>

namespace iegpu {

// Highest coefficient count the kernels take (one wave lane per coefficient).
constexpr int kFmmMaxOrder = 64;
// Deepest leaf level: 15 bits per coordinate in a 30-bit Morton key.
constexpr int kFmmMaxDepth = 15;

struct FmmHostTree {
  int n_points = 0;
  int depth = 0;  // leaf level L >= 2; levels 2..L hold boxes
  double lo_x = 0, lo_y = 0, side = 0;  // root square [lo, lo + side)^2
  // Boxes of level l (2 <= l <= L) are [level_off[l - 2], level_off[l - 1]),
  // each level in Morton order, so a box's children are contiguous.
  std::vector<int> level_off;
  std::vector<double> center;      // 2 per box
  std::vector<int> parent;         // box of level l - 1; -1 on level 2
  std::vector<int> child_off;      // children on level l + 1: [child_off[b],
  std::vector<int> child_cnt;      //   child_off[b] + child_cnt[b])
  std::vector<int> vl_off, vl;     // interaction lists: children of the
                                   // parent's neighbours not adjacent to b
  std::vector<int> pt_off;         // leaf i holds sorted nodes [pt_off[i], pt_off[i + 1])
  std::vector<int> nb_off, nb;     // adjacent leaves of leaf i (itself included)
  std::vector<int> perm;           // sorted position -> node index

  int n_boxes() const { return level_off.empty() ? 0 : level_off.back(); }
  int first_leaf() const { return level_off[depth - 2]; }
  int n_leaves() const { return level_off[depth - 1] - level_off[depth - 2]; }

  // Nodes (2 doubles each); the leaf level is the first at which the mean
  // count of the non-empty leaves is at most `leaf_points` (or level 15).
  void build(const double* pts, int n, double leaf_points);
};

// Translation tables, 64 x 64 doubles each (row-major), from exact integer
// binomials:  m2m[m][k] = C(k, m) 2^-m,  m2l[k][l] = C(k + l, l),
// l2l[l][m] = C(l, m) 2^-(m + 1).
void fmm_tables(std::vector<double>& m2m, std::vector<double>& m2l,
                std::vector<double>& l2l);

}  // namespace iegpu
