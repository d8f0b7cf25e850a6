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

// What HostTree::update changed.
struct TreeUpdate {
  bool restructured = false;   // boxes were split or merged: slots renumbered
  bool lists_changed = false;  // some leaf's DOF list changed
  int n_split = 0;             // leaves subdivided
  int n_merged = 0;            // boxes that became leaves, their subtrees removed
  // Restructured: per slot after the update, the box's slot and creation
  // id before (-1: a new box).
  std::vector<int> old_slot, old_id;
  // Slots (after the update) from which the re-factorization marks start:
  // the old and new leaf of every changed point, every new box, every box
  // whose leaf flag or row-source list changed.
  std::vector<int> seeds;
};

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

  // Original DOFs of every leaf (CSR, in DOF order; an internal box's list
  // is empty here, the DOFs below it are its leaves' lists: dofs_below).
  // Only leaves' lists are read on the device.
  std::vector<int> orig_off, orig_cnt, orig_dofs;
  // Right and top edges of every box (the corners as the build computes
  // them: a child's edges are its parent's edges or the parent's split
  // point), and the leaf slot of every point.
  std::vector<double> tr_x, tr_y;
  std::vector<int> point_leaf;
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
  // Neighbour and row-source lists formed again over the whole tree from
  // the boxes (quadtree.cpp:752-801); update() forms them only where they
  // can change and must give the same lists.
  void relink_all();
  // Every DOF below box `slot`, in DOF order.
  std::vector<int> dofs_below(int slot) const;
  // New node positions for the same points (QuadTree::perturb,
  // ie_solver/quadtree/quadtree.cpp:520-801), `changed` listing every point
  // whose data changed.  A point whose coordinates changed leaves its leaf
  // and descends from its problem's root (:648-668); then every leaf that
  // received a point and holds more than leaf_cap DOFs is subdivided
  // (:726-733, node_subdivide :275-432), and on the old path of every moved
  // point the first box from the root that is not a leaf and holds fewer
  // than leaf_cap DOFs becomes a leaf, its descendants removed (:689-712,
  // :735-741, consolidate_node :464-517).  Neighbour and row-source lists
  // are recomputed over the whole tree when boxes were split or merged
  // (:743-801).  Surviving boxes keep their order within a level; new
  // boxes follow them in creation order (level l of a forest is still
  // level l of problem 0, then of problem 1, ...).  Leaf DOF lists stay in
  // DOF order (the reference appends moved DOFs at the end).  Returns
  // false, leaving the tree unchanged, if a point lies outside its root
  // cell; the caller then rebuilds the tree.
  bool update(const double* old_points, const double* points, const std::vector<int>& changed,
              TreeUpdate* out);
};

}  // namespace iegpu
