// Host-side quadtree for the RS factorization, flattened into level-major
// arrays for the device.
//
// Tree rules restated from ie_solver/quadtree/quadtree.cpp:
//   root square [min-0.01, max+0.01]^2 over all coordinates   (:92-117)
//   points inserted one at a time; a leaf splits when its DOF count
//   exceeds the leaf cap                                       (:230-269)
//   children {bl, tl, tr, br}; appended to their level as
//   bl, br, tr, tl                                             (:275-359)
//   neighbours: siblings, then parent's neighbours' children whose
//   bottom-left corners lie within side*sqrt(2)+1e-5, then for leaves the
//   deeper boxes touching them (wall test as written in :185-227)
//                                                              (:131-175)
#pragma once

#include <vector>

namespace iegpu {

struct HostTree {
  int n_points = 0;
  int sol_dim = 1;
  int leaf_cap = 64;
  int n_boxes = 0;
  int n_levels = 0;
  // Independent problems in this tree (a forest when > 1): problem p owns
  // points [point_off[p], point_off[p+1]) and its root is slot p.
  int n_roots = 1;
  std::vector<int> point_off;

  // All per-box arrays are indexed by "slot": boxes sorted level-major, in
  // the reference's per-level order.  box_id[slot] is the creation index
  // (the reference's QuadTreeNode::id order, counted from the root).
  std::vector<int> level_start;  // n_levels + 1
  std::vector<int> box_id;
  std::vector<int> slot_of_id;
  std::vector<int> level, parent, leaf;
  std::vector<int> child;  // 4 per slot: bl, tl, tr, br (-1 if none)
  std::vector<double> bl_x, bl_y, side, cx, cy;
  // The split point of every box, computed as the build does (a point goes
  // to the left / bottom child when below it).
  std::vector<double> mid_x, mid_y;

  // Original DOFs of every box (CSR; only leaves' lists are read on device).
  std::vector<int> orig_off, orig_cnt, orig_dofs;
  // Neighbour slots per box (reference order, all levels).
  std::vector<int> nbr_off, nbr_cnt, nbr;
  // ID row-source boxes per box:
  //   level >= 3: neighbours at the same or a coarser level (near field,
  //               filtered by distance, plus proxy rows);
  //   level 2:    the other level-2 boxes, then the level-1 leaves;
  //   level 1:    the other level-1 boxes.
  std::vector<int> src_off, src_cnt, src;

  void build(const double* points, int n_points, int sol_dim, int leaf_cap);
  // One tree per problem, merged level by level: level l of the forest is
  // level l of problem 0, then of problem 1, ...  DOF indices are global
  // (offset by sol_dim * point_off[p]); creation ids are numbered problem
  // after problem.
  void build_forest(const double* points, const int* point_off, int n_problems,
                    int sol_dim, int leaf_cap);
  // Puts moved points into the kept cells (QuadTree::perturb's descent of
  // added points, ie_solver/quadtree/quadtree.cpp:680-697): every point
  // descends from its problem's root to a leaf, and the original-DOF lists
  // are rebuilt.  old_leaf / new_leaf get each point's leaf slot before and
  // after.  A leaf may grow to 2 leaf_cap DOFs (the reference splits a
  // leaf past leaf_cap).  Returns false, leaving the tree unchanged, if a
  // point lies outside its root cell or a leaf would exceed that; the
  // caller then rebuilds the tree.
  bool reinsert(const double* points, std::vector<int>* old_leaf,
                std::vector<int>* new_leaf);
};

}  // namespace iegpu
