// Quadtree construction (host).  Rules and citations in tree.hpp.
#include "tree.hpp"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <stdexcept>

namespace iegpu {

namespace {

struct Box {
  int level = 0;
  int parent = -1;
  int child[4] = {-1, -1, -1, -1};  // bl, tl, tr, br
  bool leaf = true;
  double c[8] = {0};  // corners BL, TL, TR, BR as (x, y)
  double side = 0;
  std::vector<int> orig;
  std::vector<int> nbrs;
};

class Builder {
 public:
  Builder(const double* pts, int sol_dim, int cap)
      : pts_(pts), sd_(sol_dim), cap_(cap) {}

  std::vector<Box> boxes;
  std::vector<std::vector<int>> levels;

  void run(int n_points) {
    double mn = pts_[0], mx = pts_[0];
    for (int i = 0; i < 2 * n_points; i++) {
      mn = std::fmin(mn, pts_[i]);
      mx = std::fmax(mx, pts_[i]);
    }
    const double lo = mn - 0.01, hi = mx + 0.01;
    boxes.emplace_back();
    Box& r = boxes[0];
    const double rc[8] = {lo, lo, lo, hi, hi, hi, hi, lo};
    std::memcpy(r.c, rc, sizeof(rc));
    r.side = hi - lo;
    levels.push_back({0});
    for (int i = 0; i < n_points; i++) insert(0, pts_[2 * i], pts_[2 * i + 1], i);
    link_neighbours();
  }

 private:
  const double* pts_;
  int sd_, cap_;

  static void mids(const Box& b, double* mx, double* my) {
    *mx = ((b.c[6] - b.c[0]) / 2.0) + b.c[0];
    *my = ((b.c[3] - b.c[1]) / 2.0) + b.c[1];
  }
  static int quadrant(double x, double y, double mx, double my) {
    if (x < mx) return (y < my) ? 0 : 1;  // bl : tl
    return (y < my) ? 3 : 2;              // br : tr
  }

  void insert(int b, double x, double y, int pi) {
    // Iterative descent: every box on the path gets the DOFs; the leaf at
    // the bottom splits once it exceeds the cap.
    for (;;) {
      for (int i = 0; i < sd_; i++) boxes[b].orig.push_back(sd_ * pi + i);
      double mx, my;
      mids(boxes[b], &mx, &my);
      const int q = quadrant(x, y, mx, my);
      const int ch = boxes[b].child[q];
      if (ch >= 0) {
        b = ch;
        continue;
      }
      if (boxes[b].leaf && (int)boxes[b].orig.size() > cap_) split(b);
      return;
    }
  }

  void split(int b) {
    boxes[b].leaf = false;
    double mx, my;
    mids(boxes[b], &mx, &my);
    const double* c = boxes[b].c;
    const double geo[4][8] = {
        {c[0], c[1], c[2], my, mx, my, mx, c[7]},  // bl
        {c[0], my, c[2], c[3], mx, c[5], mx, my},  // tl
        {mx, my, mx, c[3], c[4], c[5], c[6], my},  // tr
        {mx, c[1], mx, my, c[4], my, c[6], c[7]},  // br
    };
    const int lv = boxes[b].level + 1;
    const double s = boxes[b].side / 2.0;
    int ids[4];
    for (int q = 0; q < 4; q++) {
      ids[q] = (int)boxes.size();
      boxes.emplace_back();
      Box& k = boxes.back();
      std::memcpy(k.c, geo[q], sizeof(geo[q]));
      k.level = lv;
      k.side = s;
      k.parent = b;
    }
    for (int q = 0; q < 4; q++) boxes[b].child[q] = ids[q];
    if ((int)levels.size() < lv + 1) levels.emplace_back();
    // level order bl, br, tr, tl
    levels[lv].push_back(ids[0]);
    levels[lv].push_back(ids[3]);
    levels[lv].push_back(ids[2]);
    levels[lv].push_back(ids[1]);
    const std::vector<int> dofs = boxes[b].orig;
    for (size_t i = 0; i < dofs.size(); i += sd_) {
      const int pi = dofs[i] / sd_;
      const int q = quadrant(pts_[2 * pi], pts_[2 * pi + 1], mx, my);
      for (int d = 0; d < sd_; d++) boxes[ids[q]].orig.push_back(dofs[i] + d);
    }
    for (int q = 0; q < 4; q++)
      if ((int)boxes[ids[q]].orig.size() > cap_) split(ids[q]);
  }

  void touch_descendants(int big, int small) {
    const Box& B = boxes[big];
    const double top = B.c[3], bottom = B.c[1], left = B.c[0], right = B.c[4];
    const Box& S = boxes[small];
    for (int i = 0; i < 8; i += 2) {
      const double x = S.c[i], y = S.c[i + 1];
      const bool on_tb = (x > left && x < right) &&
                         (std::fabs(y - top) < 1e-14 || std::fabs(y - bottom) < 1e-14);
      const bool on_lr = (y < top && y < bottom) &&
                         (std::fabs(x - left) < 1e-14 || std::fabs(x - right) < 1e-14);
      if (on_tb || on_lr) {
        boxes[big].nbrs.push_back(small);
        boxes[small].nbrs.push_back(big);
        break;
      }
    }
    for (int q = 0; q < 4; q++) {
      const int ch = boxes[small].child[q];
      if (ch >= 0) touch_descendants(big, ch);
    }
  }

  void link_neighbours() {
    for (size_t lv = 0; lv < levels.size(); lv++) {
      for (int a : levels[lv]) {
        const int p = boxes[a].parent;
        if (p >= 0) {
          for (int q = 0; q < 4; q++)
            if (boxes[p].child[q] != a) boxes[a].nbrs.push_back(boxes[p].child[q]);
          const std::vector<int> pn = boxes[p].nbrs;
          for (int u : pn) {
            for (int q = 0; q < 4; q++) {
              const int cz = boxes[u].child[q];
              if (cz < 0 || boxes[cz].level != boxes[a].level) continue;
              const double dx = boxes[a].c[0] - boxes[cz].c[0];
              const double dy = boxes[a].c[1] - boxes[cz].c[1];
              if (std::sqrt(dx * dx + dy * dy) < boxes[a].side * std::sqrt(2.0) + 1e-5)
                boxes[a].nbrs.push_back(cz);
            }
          }
        }
        if (boxes[a].leaf) {
          for (size_t t = 0; t < boxes[a].nbrs.size(); t++) {
            const int u = boxes[a].nbrs[t];
            if (boxes[u].level != boxes[a].level) continue;
            for (int q = 0; q < 4; q++) {
              const int ch = boxes[u].child[q];
              if (ch >= 0) touch_descendants(a, ch);
            }
          }
        }
      }
    }
  }
};

}  // namespace

void HostTree::build(const double* points, int np, int sd, int cap) {
  if (np <= 0) throw std::invalid_argument("empty boundary");
  n_points = np;
  sol_dim = sd;
  leaf_cap = cap;
  n_roots = 1;
  point_off = {0, np};
  Builder B(points, sd, cap);
  B.run(np);
  n_boxes = (int)B.boxes.size();
  n_levels = (int)B.levels.size();

  level_start.assign(n_levels + 1, 0);
  box_id.clear();
  for (int l = 0; l < n_levels; l++) {
    level_start[l] = (int)box_id.size();
    for (int id : B.levels[l]) box_id.push_back(id);
  }
  level_start[n_levels] = (int)box_id.size();
  slot_of_id.assign(n_boxes, -1);
  for (int s = 0; s < n_boxes; s++) slot_of_id[box_id[s]] = s;

  level.resize(n_boxes);
  parent.resize(n_boxes);
  leaf.resize(n_boxes);
  child.assign(4 * n_boxes, -1);
  bl_x.resize(n_boxes);
  bl_y.resize(n_boxes);
  side.resize(n_boxes);
  cx.resize(n_boxes);
  cy.resize(n_boxes);
  mid_x.resize(n_boxes);
  mid_y.resize(n_boxes);
  orig_off.resize(n_boxes);
  orig_cnt.resize(n_boxes);
  orig_dofs.clear();
  nbr_off.resize(n_boxes);
  nbr_cnt.resize(n_boxes);
  nbr.clear();
  for (int s = 0; s < n_boxes; s++) {
    const Box& b = B.boxes[box_id[s]];
    level[s] = b.level;
    parent[s] = b.parent >= 0 ? slot_of_id[b.parent] : -1;
    leaf[s] = b.leaf ? 1 : 0;
    for (int q = 0; q < 4; q++) child[4 * s + q] = b.child[q] >= 0 ? slot_of_id[b.child[q]] : -1;
    bl_x[s] = b.c[0];
    bl_y[s] = b.c[1];
    side[s] = b.side;
    // box centre as in ie_solver/kernel/kernel.cpp:296-297
    cx[s] = b.c[0] + b.side / 2.0;
    cy[s] = b.c[1] + b.side / 2.0;
    mid_x[s] = ((b.c[6] - b.c[0]) / 2.0) + b.c[0];
    mid_y[s] = ((b.c[3] - b.c[1]) / 2.0) + b.c[1];
    orig_off[s] = (int)orig_dofs.size();
    orig_cnt[s] = (int)b.orig.size();
    orig_dofs.insert(orig_dofs.end(), b.orig.begin(), b.orig.end());
    nbr_off[s] = (int)nbr.size();
    nbr_cnt[s] = (int)b.nbrs.size();
    for (int u : b.nbrs) nbr.push_back(slot_of_id[u]);
  }

  src_off.assign(n_boxes, 0);
  src_cnt.assign(n_boxes, 0);
  src.clear();
  for (int s = 0; s < n_boxes; s++) {
    src_off[s] = (int)src.size();
    const int lv = level[s];
    if (lv >= 3) {
      for (int t = 0; t < nbr_cnt[s]; t++) {
        const int u = nbr[nbr_off[s] + t];
        if (level[u] <= lv) src.push_back(u);
      }
    } else if (lv >= 1) {
      for (int u = level_start[lv]; u < level_start[lv + 1]; u++)
        if (u != s) src.push_back(u);
      if (lv == 2)
        for (int u = level_start[1]; u < level_start[2]; u++)
          if (leaf[u]) src.push_back(u);
    }
    src_cnt[s] = (int)src.size() - src_off[s];
  }
}

void HostTree::build_forest(const double* points, const int* poff, int np_, int sd,
                            int cap) {
  if (np_ < 1) throw std::invalid_argument("no problems");
  if (np_ == 1) {
    build(points, poff[1] - poff[0], sd, cap);
    return;
  }
  std::vector<HostTree> sub(np_);
  int nl = 0;
  for (int p = 0; p < np_; p++) {
    sub[p].build(points + 2 * (int64_t)poff[p], poff[p + 1] - poff[p], sd, cap);
    nl = std::max(nl, sub[p].n_levels);
  }
  n_points = poff[np_];
  sol_dim = sd;
  leaf_cap = cap;
  n_roots = np_;
  point_off.assign(poff, poff + np_ + 1);
  n_levels = nl;
  // forest slot of every (problem, local slot); level-major, problem order
  std::vector<std::vector<int>> fslot(np_);
  for (int p = 0; p < np_; p++) fslot[p].assign(sub[p].n_boxes, -1);
  level_start.assign(nl + 1, 0);
  int ns = 0;
  for (int l = 0; l < nl; l++) {
    level_start[l] = ns;
    for (int p = 0; p < np_; p++) {
      const HostTree& S = sub[p];
      if (l >= S.n_levels) continue;
      for (int t = S.level_start[l]; t < S.level_start[l + 1]; t++) fslot[p][t] = ns++;
    }
  }
  level_start[nl] = ns;
  n_boxes = ns;
  box_id.assign(ns, 0);
  slot_of_id.assign(ns, 0);
  level.assign(ns, 0);
  parent.assign(ns, -1);
  leaf.assign(ns, 0);
  child.assign(4 * (size_t)ns, -1);
  bl_x.assign(ns, 0);
  bl_y.assign(ns, 0);
  side.assign(ns, 0);
  cx.assign(ns, 0);
  cy.assign(ns, 0);
  mid_x.assign(ns, 0);
  mid_y.assign(ns, 0);
  orig_off.assign(ns, 0);
  orig_cnt.assign(ns, 0);
  nbr_off.assign(ns, 0);
  nbr_cnt.assign(ns, 0);
  src_off.assign(ns, 0);
  src_cnt.assign(ns, 0);
  orig_dofs.clear();
  nbr.clear();
  src.clear();
  int id0 = 0;
  for (int p = 0; p < np_; p++) {
    const HostTree& S = sub[p];
    const std::vector<int>& f = fslot[p];
    auto remap = [&f](int t) { return t >= 0 ? f[t] : -1; };
    for (int t = 0; t < S.n_boxes; t++) {
      const int s = f[t];
      box_id[s] = id0 + S.box_id[t];
      level[s] = S.level[t];
      parent[s] = remap(S.parent[t]);
      leaf[s] = S.leaf[t];
      for (int q = 0; q < 4; q++) child[4 * (size_t)s + q] = remap(S.child[4 * (size_t)t + q]);
      bl_x[s] = S.bl_x[t];
      bl_y[s] = S.bl_y[t];
      side[s] = S.side[t];
      cx[s] = S.cx[t];
      cy[s] = S.cy[t];
      mid_x[s] = S.mid_x[t];
      mid_y[s] = S.mid_y[t];
    }
    id0 += S.n_boxes;
  }
  for (int s = 0; s < ns; s++) slot_of_id[box_id[s]] = s;
  // CSR lists in forest slot order
  std::vector<int> prob_of(ns, 0), local(ns, 0);
  for (int p = 0; p < np_; p++)
    for (int t = 0; t < sub[p].n_boxes; t++) {
      prob_of[fslot[p][t]] = p;
      local[fslot[p][t]] = t;
    }
  for (int s = 0; s < ns; s++) {
    const int p = prob_of[s], t = local[s];
    const HostTree& S = sub[p];
    const std::vector<int>& f = fslot[p];
    const int dof0 = sd * poff[p];
    orig_off[s] = (int)orig_dofs.size();
    orig_cnt[s] = S.orig_cnt[t];
    for (int u = 0; u < S.orig_cnt[t]; u++) orig_dofs.push_back(dof0 + S.orig_dofs[S.orig_off[t] + u]);
    nbr_off[s] = (int)nbr.size();
    nbr_cnt[s] = S.nbr_cnt[t];
    for (int u = 0; u < S.nbr_cnt[t]; u++) nbr.push_back(f[S.nbr[S.nbr_off[t] + u]]);
    src_off[s] = (int)src.size();
    src_cnt[s] = S.src_cnt[t];
    for (int u = 0; u < S.src_cnt[t]; u++) src.push_back(f[S.src[S.src_off[t] + u]]);
  }
}

bool HostTree::reinsert(const double* points, std::vector<int>* old_leaf,
                        std::vector<int>* new_leaf) {
  std::vector<int> was(n_points, -1), now(n_points, -1);
  for (int s = 0; s < n_boxes; s++)
    if (leaf[s])
      for (int i = 0; i < orig_cnt[s]; i += sol_dim)
        was[orig_dofs[orig_off[s] + i] / sol_dim] = s;
  std::vector<int> fill(n_boxes, 0);
  for (int r = 0; r < n_roots; r++)
    for (int p = point_off[r]; p < point_off[r + 1]; p++) {
      const double x = points[2 * p], y = points[2 * p + 1];
      if (!(x >= bl_x[r] && x <= bl_x[r] + side[r] && y >= bl_y[r] && y <= bl_y[r] + side[r]))
        return false;
      int s = r;
      while (!leaf[s]) {
        const int q = (x < mid_x[s]) ? ((y < mid_y[s]) ? 0 : 1) : ((y < mid_y[s]) ? 3 : 2);
        s = child[4 * (size_t)s + q];
      }
      now[p] = s;
      fill[s] += sol_dim;
      if (fill[s] > 2 * leaf_cap) return false;
    }
  // Every box lists the DOFs of the points below it, in point order (the
  // order the build's insertion gives).
  std::vector<int> cnt(n_boxes, 0);
  for (int p = 0; p < n_points; p++)
    for (int s = now[p]; s >= 0; s = parent[s]) cnt[s] += sol_dim;
  std::vector<int> off(n_boxes, 0);
  int tot = 0;
  for (int s = 0; s < n_boxes; s++) {
    off[s] = tot;
    tot += cnt[s];
  }
  std::vector<int> dofs(tot), pos(off);
  for (int p = 0; p < n_points; p++)
    for (int s = now[p]; s >= 0; s = parent[s])
      for (int d = 0; d < sol_dim; d++) dofs[pos[s]++] = sol_dim * p + d;
  orig_off = off;
  orig_cnt = cnt;
  orig_dofs.swap(dofs);
  if (old_leaf) old_leaf->swap(was);
  if (new_leaf) new_leaf->swap(now);
  return true;
}

}  // namespace iegpu
