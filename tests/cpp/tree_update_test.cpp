// CPU checks of HostTree::update (tree.cpp), run by tests/test_host.py.
// Frames of local deformations of closed curves (one boundary, Laplace and
// Stokes DOFs; a forest of three boundaries) drive leaves over and under the
// leaf capacity.  After every update:
//   * the neighbour and row-source lists, formed only where they can change,
//     equal those of relink_all (every list formed again over the tree);
//   * every point lies in the cell of its leaf, every DOF is in exactly one
//     leaf list, lists are in DOF order;
//   * leaves hold at most leaf_cap DOFs, boxes that are not leaves at least
//     leaf_cap; collapsing the boxes that are not leaves but hold exactly
//     leaf_cap DOFs gives the cells of a fresh build on the new points (the
//     reference merges only below the capacity);
//   * parents, children, levels, slots and ids agree; a surviving box has
//     the cell it had, at the id update() reports.
// Exit status 0 on success; prints the failing check otherwise.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <set>
#include <tuple>
#include <vector>

#include "../../ie-solver_amd/csrc/tree.hpp"

using iegpu::HostTree;
using iegpu::TreeUpdate;

namespace {

int fails = 0;
void check(bool ok, const char* what, int a = 0, int b = 0) {
  if (!ok) {
    if (fails < 10) std::printf("FAIL %s (%d, %d)\n", what, a, b);
    fails++;
  }
}

// Wavy closed curve around (cx, cy) with a bump of height amp over an arc
// of half-width w at angle t0.
std::vector<double> curve(int n, int seed, double cx, double cy, double amp, double t0,
                          double w) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> U(-0.015, 0.015);
  double a[12], b[12];
  for (int k = 0; k < 12; k++) {
    a[k] = U(rng);
    b[k] = U(rng);
  }
  std::vector<double> p(2 * (size_t)n);
  for (int i = 0; i < n; i++) {
    const double t = 2 * M_PI * i / n;
    double r = 0.3;
    for (int k = 2; k < 12; k++) r += a[k] * std::cos(k * t) + b[k] * std::sin(k * t);
    const double d = std::remainder(t - t0, 2 * M_PI);
    if (std::fabs(d) < w) r += amp * std::pow(std::cos(M_PI * d / (2 * w)), 2);
    p[2 * i] = cx + r * std::cos(t);
    p[2 * i + 1] = cy + r * std::sin(t);
  }
  return p;
}

std::vector<int> list_of(const std::vector<int>& off, const std::vector<int>& cnt,
                         const std::vector<int>& v, int s) {
  return std::vector<int>(v.begin() + off[s], v.begin() + off[s] + cnt[s]);
}

void check_tree(const HostTree& T, const std::vector<double>& pts, const char* tag) {
  const int ns = T.n_boxes, sd = T.sol_dim, cap = T.leaf_cap;
  // relink_all gives the same lists
  HostTree G = T;
  G.relink_all();
  for (int s = 0; s < ns; s++) {
    check(list_of(T.nbr_off, T.nbr_cnt, T.nbr, s) == list_of(G.nbr_off, G.nbr_cnt, G.nbr, s),
          "neighbour list differs from the tree-wide pass", s, T.level[s]);
    check(list_of(T.src_off, T.src_cnt, T.src, s) == list_of(G.src_off, G.src_cnt, G.src, s),
          "row sources differ from the tree-wide pass", s, T.level[s]);
  }
  // structure
  check((int)T.level_start.size() == T.n_levels + 1 && T.level_start[T.n_levels] == ns,
        "level_start");
  for (int l = 0; l < T.n_levels; l++)
    for (int s = T.level_start[l]; s < T.level_start[l + 1]; s++)
      check(T.level[s] == l, "slot outside its level", s, l);
  std::vector<int> seen(ns, 0);
  for (int s = 0; s < ns; s++) {
    check(T.slot_of_id[T.box_id[s]] == s, "box id map", s);
    seen[T.box_id[s]]++;
    const int p = T.parent[s];
    check((p < 0) == (s < T.n_roots), "roots", s);
    if (p >= 0) {
      int hits = 0;
      for (int q = 0; q < 4; q++) hits += T.child[4 * (size_t)p + q] == s;
      check(hits == 1 && T.level[p] + 1 == T.level[s], "parent and child disagree", s, p);
    }
    int nch = 0;
    for (int q = 0; q < 4; q++) nch += T.child[4 * (size_t)s + q] >= 0;
    check(nch == (T.leaf[s] ? 0 : 4), "children of a box", s, nch);
  }
  for (int id = 0; id < ns; id++) check(seen[id] == 1, "ids are a permutation", id);
  // points and leaf lists
  std::vector<int> below(ns, 0), owner(T.n_points * sd, -1);
  for (int s = 0; s < ns; s++) {
    if (!T.leaf[s]) {
      check(T.orig_cnt[s] == 0, "a box that is not a leaf holds a list", s);
      continue;
    }
    const std::vector<int> l = list_of(T.orig_off, T.orig_cnt, T.orig_dofs, s);
    for (size_t i = 0; i < l.size(); i++) {
      check(i == 0 || l[i] > l[i - 1], "leaf list out of DOF order", s);
      check(owner[l[i]] < 0, "DOF in two leaves", l[i], s);
      owner[l[i]] = s;
    }
    below[s] = (int)l.size();
    check(below[s] <= cap, "leaf over the capacity", s, below[s]);
  }
  for (int d = 0; d < T.n_points * sd; d++) check(owner[d] >= 0, "DOF in no leaf", d);
  for (int p = 0; p < T.n_points; p++) {
    const int s = T.point_leaf[p];
    check(owner[sd * p] == s, "point_leaf disagrees with the lists", p, s);
    const double x = pts[2 * p], y = pts[2 * p + 1];
    check(x >= T.bl_x[s] && x <= T.tr_x[s] && y >= T.bl_y[s] && y <= T.tr_y[s],
          "point outside its leaf's cell", p, s);
  }
  for (int s = ns - 1; s >= 0; s--)
    if (!T.leaf[s]) {
      below[s] = 0;
      for (int q = 0; q < 4; q++) below[s] += below[T.child[4 * (size_t)s + q]];
      check(below[s] >= cap, "box that is not a leaf under the capacity", s, below[s]);
    }
  (void)tag;
}

// Cells (problem, level, bl_x, bl_y, leaf) of T after collapsing every box
// that is not a leaf but holds at most leaf_cap DOFs.
std::set<std::tuple<int, int, double, double, int>> collapsed_cells(const HostTree& T) {
  const int ns = T.n_boxes;
  std::vector<int> below(ns, 0), prob(ns, 0);
  for (int s = 0; s < ns; s++) {
    below[s] = T.orig_cnt[s];
    prob[s] = s < T.n_roots ? s : prob[T.parent[s]];
  }
  for (int s = ns - 1; s >= 0; s--)
    if (!T.leaf[s])
      for (int q = 0; q < 4; q++) below[s] += below[T.child[4 * (size_t)s + q]];
  std::vector<char> gone(ns, 0), leafc(ns, 0);
  std::set<std::tuple<int, int, double, double, int>> out;
  for (int s = 0; s < ns; s++) {
    const int p = T.parent[s];
    if (p >= 0 && (gone[p] || leafc[p])) {
      gone[s] = 1;
      continue;
    }
    leafc[s] = T.leaf[s] || below[s] <= T.leaf_cap;
    out.insert({prob[s], T.level[s], T.bl_x[s], T.bl_y[s], (int)leafc[s]});
  }
  return out;
}

std::vector<int> changed_points(const std::vector<double>& a, const std::vector<double>& b) {
  std::vector<int> ch;
  for (size_t k = 0; k < a.size() / 2; k++)
    if (a[2 * k] != b[2 * k] || a[2 * k + 1] != b[2 * k + 1]) ch.push_back((int)k);
  return ch;
}

struct Stats {
  int frames = 0, restructured = 0, splits = 0, merges = 0, rebuilt = 0, compared = 0;
};

void run_frame(HostTree& T, std::vector<double>& p, const std::vector<double>& q,
               const std::vector<int>& poff, Stats* st, const char* tag) {
  HostTree before = T;
  const std::vector<int> ch = changed_points(p, q);
  TreeUpdate u;
  const bool kept = T.update(p.data(), q.data(), ch, &u);
  if (!kept) {
    st->rebuilt++;
    T = HostTree();
    T.build_forest(q.data(), poff.data(), (int)poff.size() - 1, before.sol_dim, before.leaf_cap);
    p = q;
    return;
  }
  st->frames++;
  st->restructured += u.restructured;
  st->splits += u.n_split;
  st->merges += u.n_merged;
  check_tree(T, q, tag);
  // a fresh build's cells, where its root cells are the kept ones (a
  // build sizes them to the new points)
  HostTree F;
  F.build_forest(q.data(), poff.data(), (int)poff.size() - 1, T.sol_dim, T.leaf_cap);
  bool same_roots = true;
  for (int r = 0; r < T.n_roots; r++)
    same_roots = same_roots && F.bl_x[r] == T.bl_x[r] && F.bl_y[r] == T.bl_y[r] &&
                 F.tr_x[r] == T.tr_x[r] && F.tr_y[r] == T.tr_y[r];
  if (same_roots) {
    st->compared++;
    check(collapsed_cells(T) == collapsed_cells(F), "cells differ from a fresh build", st->frames);
  }
  // surviving boxes keep their cells; the ids map
  if (u.restructured) {
    check((int)u.old_slot.size() == T.n_boxes && (int)u.old_id.size() == T.n_boxes,
          "slot map size");
    for (int s = 0; s < T.n_boxes; s++) {
      const int o = u.old_slot[s];
      if (o < 0) continue;
      check(before.box_id[o] == u.old_id[s], "old id", s, o);
      check(before.level[o] == T.level[s] && before.bl_x[o] == T.bl_x[s] &&
                before.bl_y[o] == T.bl_y[s] && before.side[o] == T.side[s],
            "surviving box moved", s, o);
    }
  } else {
    check(T.n_boxes == before.n_boxes && T.nbr == before.nbr && T.child == before.child,
          "structure changed without a restructure");
  }
  p = q;
}

}  // namespace

int main() {
  Stats st;
  for (int sd : {1, 2}) {
    for (int n : {8192, 32768}) {
      for (int seed = 0; seed < 2; seed++) {
        const int cap = 64;
        std::vector<double> p = curve(n, seed, 0.5, 0.5, 0, 0, 0.1);
        HostTree T;
        T.build(p.data(), n, sd, cap);
        const std::vector<int> poff = {0, n};
        for (int f = 0; f < 10; f++) {
          const double amp = (f % 3 == 0 ? 2e-3 : f % 3 == 1 ? -1.5e-2 : 3e-2) * (1 + seed);
          std::vector<double> q = curve(n, seed, 0.5, 0.5, amp, 0.6 * f + seed, 0.08 + 0.04 * (f % 4));
          run_frame(T, p, q, poff, &st, "curve");
        }
        // crowd 300 nodes into a small square, then release them
        std::vector<double> q = p;
        for (int k = 0; k < 300; k++) {
          q[2 * k] = 0.52 + 2e-4 * (k % 17);
          q[2 * k + 1] = 0.47 + 2e-4 * (k / 17);
        }
        run_frame(T, p, q, poff, &st, "crowd");
        run_frame(T, p, curve(n, seed, 0.5, 0.5, 0, 0, 0.1), poff, &st, "release");
      }
    }
  }
  // a forest of three boundaries: one of them moves, the others' boxes keep
  // their lists (mapped)
  {
    const int n0 = 8192, n1 = 16384, n2 = 4096;
    std::vector<double> a = curve(n0, 3, 0.5, 0.5, 0, 0, 0.1), b = curve(n1, 4, 0.2, 0.7, 0, 0, 0.1),
                        c = curve(n2, 5, 0.8, 0.3, 0, 0, 0.1);
    std::vector<double> p(a);
    p.insert(p.end(), b.begin(), b.end());
    p.insert(p.end(), c.begin(), c.end());
    const std::vector<int> poff = {0, n0, n0 + n1, n0 + n1 + n2};
    HostTree T;
    T.build_forest(p.data(), poff.data(), 3, 1, 64);
    for (int f = 0; f < 6; f++) {
      std::vector<double> q = p;
      const std::vector<double> bb =
          curve(n1, 4, 0.2, 0.7, (f % 2 ? -2e-2 : 3e-2), 0.9 * f, 0.1 + 0.03 * f);
      std::copy(bb.begin(), bb.end(), q.begin() + 2 * n0);
      HostTree before = T;
      run_frame(T, p, q, poff, &st, "forest");
      // boxes of problems 0 and 2 keep their lists
      std::vector<int> prob(T.n_boxes, 0), oprob(before.n_boxes, 0);
      for (int s = 0; s < T.n_boxes; s++) prob[s] = s < T.n_roots ? s : prob[T.parent[s]];
      for (int s = 0; s < before.n_boxes; s++)
        oprob[s] = s < before.n_roots ? s : oprob[before.parent[s]];
      int others_new = 0, others_old = 0;
      for (int s = 0; s < T.n_boxes; s++) others_new += prob[s] != 1;
      for (int s = 0; s < before.n_boxes; s++) others_old += oprob[s] != 1;
      check(others_new == others_old, "another problem's boxes changed", f);
    }
  }
  std::printf("%d frames kept (%d restructured: %d leaves split, %d merged; %d compared with a "
              "fresh build), %d rebuilt; %d failures\n",
              st.frames, st.restructured, st.splits, st.merges, st.compared, st.rebuilt, fails);
  if (st.splits == 0 || st.merges == 0 || st.compared < st.frames / 2) {
    std::printf("FAIL the frames exercised no split, no merge or few fresh builds\n");
    return 1;
  }
  return fails ? 1 : 0;
}
