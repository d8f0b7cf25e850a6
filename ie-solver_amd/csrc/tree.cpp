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

void mids(const Box& b, double* mx, double* my) {
  *mx = ((b.c[6] - b.c[0]) / 2.0) + b.c[0];
  *my = ((b.c[3] - b.c[1]) / 2.0) + b.c[1];
}
int quadrant(double x, double y, double mx, double my) {
  if (x < mx) return (y < my) ? 0 : 1;  // bl : tl
  return (y < my) ? 3 : 2;              // br : tr
}
// Corners of child q (bl, tl, tr, br) of a box with corners c split at
// (mx, my).
void child_corners(const double* c, double mx, double my, int q, double* out) {
  const double geo[4][8] = {
      {c[0], c[1], c[2], my, mx, my, mx, c[7]},  // bl
      {c[0], my, c[2], c[3], mx, c[5], mx, my},  // tl
      {mx, my, mx, c[3], c[4], c[5], c[6], my},  // tr
      {mx, c[1], mx, my, c[4], my, c[6], c[7]},  // br
  };
  std::memcpy(out, geo[q], sizeof(geo[q]));
}

void touch_descendants(std::vector<Box>& boxes, int big, int small) {
  const Box& B = boxes[big];
  const double top = B.c[3], bottom = B.c[1], left = B.c[0], right = B.c[4];
  const Box& S = boxes[small];
  // Every corner below `small` lies in its cell [c0, c4] x [c1, c3]: when
  // no point of the cell passes either test below, no box of the subtree
  // does (the tolerance is doubled to stay on the safe side).
  {
    constexpr double e = 2e-14;
    const double x0 = S.c[0], x1 = S.c[4], y0 = S.c[1], y1 = S.c[3];
    const bool tb = x0 < right && x1 > left &&
                    ((y0 <= top + e && y1 >= top - e) || (y0 <= bottom + e && y1 >= bottom - e));
    const bool lr = y0 < top && y0 < bottom &&
                    ((x0 <= left + e && x1 >= left - e) || (x0 <= right + e && x1 >= right - e));
    if (!tb && !lr) return;
  }
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
    if (ch >= 0) touch_descendants(boxes, big, ch);
  }
}

// Neighbour lists of one tree, levels visited in order (quadtree.cpp:131-175).
void link_neighbours(std::vector<Box>& boxes, const std::vector<std::vector<int>>& levels) {
  for (size_t lv = 0; lv < levels.size(); lv++) {
    for (int a : levels[lv]) {
      const int p = boxes[a].parent;
      if (p >= 0) {
        for (int q = 0; q < 4; q++)
          if (boxes[p].child[q] != a) boxes[a].nbrs.push_back(boxes[p].child[q]);
        const std::vector<int>& pn = boxes[p].nbrs;  // p != a: not appended to here
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
            if (ch >= 0) touch_descendants(boxes, a, ch);
          }
        }
      }
    }
  }
}

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
    link_neighbours(boxes, levels);
  }

 private:
  const double* pts_;
  int sd_, cap_;

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
    double geo[4][8];
    for (int q = 0; q < 4; q++) child_corners(boxes[b].c, mx, my, q, geo[q]);
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
  tr_x.resize(n_boxes);
  tr_y.resize(n_boxes);
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
    tr_x[s] = b.c[4];
    tr_y[s] = b.c[3];
    side[s] = b.side;
    // box centre as in ie_solver/kernel/kernel.cpp:296-297
    cx[s] = b.c[0] + b.side / 2.0;
    cy[s] = b.c[1] + b.side / 2.0;
    mid_x[s] = ((b.c[6] - b.c[0]) / 2.0) + b.c[0];
    mid_y[s] = ((b.c[3] - b.c[1]) / 2.0) + b.c[1];
    orig_off[s] = (int)orig_dofs.size();
    orig_cnt[s] = b.leaf ? (int)b.orig.size() : 0;
    if (b.leaf) orig_dofs.insert(orig_dofs.end(), b.orig.begin(), b.orig.end());
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
  point_leaf.assign(np, -1);
  for (int s = 0; s < n_boxes; s++)
    for (int i = 0; i < orig_cnt[s]; i += sd) point_leaf[orig_dofs[orig_off[s] + i] / sd] = s;
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
  tr_x.assign(ns, 0);
  tr_y.assign(ns, 0);
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
      tr_x[s] = S.tr_x[t];
      tr_y[s] = S.tr_y[t];
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
  point_leaf.assign(n_points, -1);
  for (int p = 0; p < np_; p++)
    for (int i = 0; i < sub[p].n_points; i++) point_leaf[poff[p] + i] = fslot[p][sub[p].point_leaf[i]];
}

namespace {

// The descendant test of the neighbour rule (touch_descendants above) on
// slot arrays: does box `small` pass it against leaf `big`?
bool touch_hit(const HostTree& T, int big, int small) {
  const double top = T.tr_y[big], bottom = T.bl_y[big], left = T.bl_x[big], right = T.tr_x[big];
  const double xs[4] = {T.bl_x[small], T.bl_x[small], T.tr_x[small], T.tr_x[small]};
  const double ys[4] = {T.bl_y[small], T.tr_y[small], T.tr_y[small], T.bl_y[small]};
  for (int i = 0; i < 4; i++) {
    const double x = xs[i], y = ys[i];
    const bool on_tb = (x > left && x < right) &&
                       (std::fabs(y - top) < 1e-14 || std::fabs(y - bottom) < 1e-14);
    const bool on_lr = (y < top && y < bottom) &&
                       (std::fabs(x - left) < 1e-14 || std::fabs(x - right) < 1e-14);
    if (on_tb || on_lr) return true;
  }
  return false;
}

// The boxes of the subtree at `small` that pass the test against `big`, in
// the order the tree-wide pass appends them (pre-order, children bl, tl,
// tr, br), skipping subtrees whose cell no box of can pass.
void touch_walk(const HostTree& T, int big, int small, std::vector<int>* out) {
  constexpr double e = 2e-14;
  const double top = T.tr_y[big], bottom = T.bl_y[big], left = T.bl_x[big], right = T.tr_x[big];
  const double x0 = T.bl_x[small], x1 = T.tr_x[small], y0 = T.bl_y[small], y1 = T.tr_y[small];
  const bool tb = x0 < right && x1 > left &&
                  ((y0 <= top + e && y1 >= top - e) || (y0 <= bottom + e && y1 >= bottom - e));
  const bool lr = y0 < top && y0 < bottom &&
                  ((x0 <= left + e && x1 >= left - e) || (x0 <= right + e && x1 >= right - e));
  if (!tb && !lr) return;
  if (touch_hit(T, big, small)) out->push_back(small);
  for (int q = 0; q < 4; q++) {
    const int ch = T.child[4 * (size_t)small + q];
    if (ch >= 0) touch_walk(T, big, ch, out);
  }
}

// Row-source lists of every box from the neighbour lists (the build's
// rule, within each problem; sprob = problem of every slot).
void fill_src(HostTree& T, const std::vector<int>& sprob) {
  const int ns = T.n_boxes;
  T.src_off.assign(ns, 0);
  T.src_cnt.assign(ns, 0);
  T.src.clear();
  for (int s = 0; s < ns; s++) {
    T.src_off[s] = (int)T.src.size();
    const int lv = T.level[s];
    if (lv >= 3) {
      for (int t = 0; t < T.nbr_cnt[s]; t++) {
        const int w = T.nbr[T.nbr_off[s] + t];
        if (T.level[w] <= lv) T.src.push_back(w);
      }
    } else if (lv >= 1) {
      for (int w = T.level_start[lv]; w < T.level_start[lv + 1]; w++)
        if (w != s && sprob[w] == sprob[s]) T.src.push_back(w);
      if (lv == 2)
        for (int w = T.level_start[1]; w < T.level_start[2]; w++)
          if (T.leaf[w] && sprob[w] == sprob[s]) T.src.push_back(w);
    }
    T.src_cnt[s] = (int)T.src.size() - T.src_off[s];
  }
}

std::vector<int> slot_problems(const HostTree& T) {
  std::vector<int> pr(T.n_boxes, 0);
  for (int s = 0; s < T.n_boxes; s++) pr[s] = s < T.n_roots ? s : pr[T.parent[s]];
  return pr;
}

}  // namespace

void HostTree::relink_all() {
  std::vector<Box> bx(n_boxes);
  const std::vector<int> pr = slot_problems(*this);
  std::vector<std::vector<std::vector<int>>> plev(n_roots,
                                                  std::vector<std::vector<int>>(n_levels));
  for (int s = 0; s < n_boxes; s++) {
    Box& b = bx[s];
    b.level = level[s];
    b.parent = parent[s];
    for (int q = 0; q < 4; q++) b.child[q] = child[4 * (size_t)s + q];
    b.leaf = leaf[s] != 0;
    const double c[8] = {bl_x[s], bl_y[s], bl_x[s], tr_y[s], tr_x[s], tr_y[s], tr_x[s], bl_y[s]};
    std::memcpy(b.c, c, sizeof(c));
    b.side = side[s];
    plev[pr[s]][level[s]].push_back(s);
  }
  for (int r = 0; r < n_roots; r++) link_neighbours(bx, plev[r]);
  nbr_off.assign(n_boxes, 0);
  nbr_cnt.assign(n_boxes, 0);
  nbr.clear();
  for (int s = 0; s < n_boxes; s++) {
    nbr_off[s] = (int)nbr.size();
    nbr_cnt[s] = (int)bx[s].nbrs.size();
    nbr.insert(nbr.end(), bx[s].nbrs.begin(), bx[s].nbrs.end());
  }
  fill_src(*this, pr);
}

std::vector<int> HostTree::dofs_below(int slot) const {
  std::vector<int> v, stack{slot};
  while (!stack.empty()) {
    const int s = stack.back();
    stack.pop_back();
    if (leaf[s]) {
      v.insert(v.end(), orig_dofs.begin() + orig_off[s],
               orig_dofs.begin() + orig_off[s] + orig_cnt[s]);
    } else {
      for (int q = 0; q < 4; q++)
        if (child[4 * (size_t)s + q] >= 0) stack.push_back(child[4 * (size_t)s + q]);
    }
  }
  std::sort(v.begin(), v.end());
  return v;
}

bool HostTree::update(const double* old_pts, const double* pts, const std::vector<int>& changed,
                      TreeUpdate* out) {
  TreeUpdate u;
  const int sd = sol_dim;
  auto descend = [&](int s, double x, double y) {
    while (!leaf[s]) s = child[4 * (size_t)s + quadrant(x, y, mid_x[s], mid_y[s])];
    return s;
  };
  // Points whose coordinates changed: old leaf, new leaf in the kept cells.
  std::vector<int> mv, from, to;
  for (int p : changed) {
    const double x = pts[2 * (size_t)p], y = pts[2 * (size_t)p + 1];
    if (x == old_pts[2 * (size_t)p] && y == old_pts[2 * (size_t)p + 1]) continue;
    const int r =
        (int)(std::upper_bound(point_off.begin(), point_off.end(), p) - point_off.begin()) - 1;
    if (!(x >= bl_x[r] && x <= tr_x[r] && y >= bl_y[r] && y <= tr_y[r])) return false;
    mv.push_back(p);
    const int l0 = point_leaf[p];
    from.push_back(l0);
    // strictly inside its old leaf's cell: the descent ends there again
    const bool inside = x > bl_x[l0] && x < tr_x[l0] && y > bl_y[l0] && y < tr_y[l0];
    to.push_back(inside ? l0 : descend(r, x, y));
  }
  bool moves = false;
  for (size_t k = 0; k < mv.size(); k++) moves = moves || to[k] != from[k];
  if (!moves) {
    for (int p : changed) u.seeds.push_back(point_leaf[p]);
    *out = std::move(u);
    return true;
  }
  // DOFs below every box after the moves (children lie in later slots).
  std::vector<int> below(n_boxes, 0);
  for (int s = 0; s < n_boxes; s++)
    if (leaf[s]) below[s] = orig_cnt[s];
  for (size_t k = 0; k < mv.size(); k++) {
    below[from[k]] -= sd;
    below[to[k]] += sd;
  }
  for (int s = n_boxes - 1; s >= 0; s--)
    if (!leaf[s]) {
      int b = 0;
      for (int q = 0; q < 4; q++) b += below[child[4 * (size_t)s + q]];
      below[s] = b;
    }
  // Leaves to subdivide (a leaf that received a point and holds more than
  // leaf_cap DOFs) and boxes to merge (on a moved point's old path, the
  // first box from the root that is not a leaf and holds fewer than
  // leaf_cap DOFs).  A box to subdivide holds more than leaf_cap DOFs, so
  // none lies below a box to merge.
  std::vector<char> flag(n_boxes, 0);
  std::vector<int> splits, merges, path;
  for (size_t k = 0; k < mv.size(); k++) {
    const int s = to[k];
    if (below[s] > leaf_cap && !flag[s]) {
      flag[s] = 1;
      splits.push_back(s);
    }
  }
  std::vector<char> walked(n_boxes, 0);
  for (size_t k = 0; k < mv.size(); k++) {
    if (walked[from[k]]) continue;  // the same path as before
    walked[from[k]] = 1;
    path.clear();
    for (int s = from[k]; s >= 0; s = parent[s]) path.push_back(s);
    for (int i = (int)path.size() - 1; i >= 0; i--) {
      const int s = path[i];
      if (!leaf[s] && below[s] < leaf_cap) {
        if (!flag[s]) {
          flag[s] = 2;
          merges.push_back(s);
        }
        break;
      }
    }
  }
  // The leaf of every changed point before and after; point_leaf takes the
  // moves now (put back if the restructuring below throws).
  for (int p : changed) u.seeds.push_back(point_leaf[p]);
  for (size_t k = 0; k < mv.size(); k++) point_leaf[mv[k]] = to[k];
  for (int p : changed) u.seeds.push_back(point_leaf[p]);
  if (splits.empty() && merges.empty()) {
    // new lists of the leaves that lost or gained points; the others copied
    std::vector<int> lidx(n_boxes, -1);
    std::vector<std::vector<int>> lists;
    for (size_t k = 0; k < mv.size(); k++) {
      if (to[k] == from[k]) continue;
      for (int l : {from[k], to[k]})
        if (lidx[l] < 0) {
          lidx[l] = (int)lists.size();
          lists.emplace_back();
          for (int i = orig_off[l]; i < orig_off[l] + orig_cnt[l]; i += sd) {
            const int q = orig_dofs[i] / sd;
            if (point_leaf[q] == l) lists.back().push_back(q);
          }
        }
    }
    for (size_t k = 0; k < mv.size(); k++)
      if (to[k] != from[k]) lists[lidx[to[k]]].push_back(mv[k]);
    for (auto& v : lists) std::sort(v.begin(), v.end());
    std::vector<int> off(n_boxes), cnt(n_boxes), dofs;
    dofs.reserve(orig_dofs.size());
    for (int s = 0; s < n_boxes; s++) {
      off[s] = (int)dofs.size();
      if (lidx[s] >= 0) {
        for (int q : lists[lidx[s]])
          for (int d = 0; d < sd; d++) dofs.push_back(sd * q + d);
      } else {
        dofs.insert(dofs.end(), orig_dofs.begin() + orig_off[s],
                    orig_dofs.begin() + orig_off[s] + orig_cnt[s]);
      }
      cnt[s] = (int)dofs.size() - off[s];
    }
    orig_off.swap(off);
    orig_cnt.swap(cnt);
    orig_dofs.swap(dofs);
    u.lists_changed = true;
    *out = std::move(u);
    return true;
  }
  struct Restore {
    std::vector<int>& pl;
    const std::vector<int>& mv;
    const std::vector<int>& from;
    bool armed;
    ~Restore() {
      if (armed)
        for (size_t k = 0; k < mv.size(); k++) pl[mv[k]] = from[k];
    }
  } restore{point_leaf, mv, from, true};
  // old leaf of every changed point (the first half of the seeds so far)
  const std::vector<int> old_leaf(u.seeds.begin(), u.seeds.begin() + changed.size());
  u.seeds.clear();

  // Boxes split or merged.  Extended box table (structure of arrays): the
  // current slots, then the new boxes in creation order.
  std::vector<int> xlevel(level), xparent(parent), xchild(child), xleaf(leaf);
  std::vector<double> xbl_x(bl_x), xbl_y(bl_y), xtr_x(tr_x), xtr_y(tr_y), xside(side),
      xcx(cx), xcy(cy), xmid_x(mid_x), xmid_y(mid_y);
  std::vector<char> alive(n_boxes, 1);
  std::vector<int> prob(n_boxes, 0);
  for (int s = 0; s < n_boxes; s++) prob[s] = s < n_roots ? s : prob[parent[s]];
  std::vector<int> merged_into(n_boxes, -1);
  std::vector<int> stack;
  for (int a : merges) {
    stack.assign(&xchild[4 * (size_t)a], &xchild[4 * (size_t)a] + 4);
    while (!stack.empty()) {
      const int d = stack.back();
      stack.pop_back();
      if (d < 0) continue;
      alive[d] = 0;
      merged_into[d] = a;
      for (int q = 0; q < 4; q++) stack.push_back(xchild[4 * (size_t)d + q]);
    }
    xleaf[a] = 1;
    for (int q = 0; q < 4; q++) xchild[4 * (size_t)a + q] = -1;
  }
  // Points of the leaves to subdivide, in point order.
  std::vector<int> split_idx(n_boxes, -1);
  for (size_t i = 0; i < splits.size(); i++) split_idx[splits[i]] = (int)i;
  std::vector<std::vector<int>> split_pts(splits.size());
  // Leaves whose point lists change (merged boxes, leaves that lost or
  // gained points) get their lists here; new leaves get theirs from the
  // subdivision; every other leaf keeps its list.
  std::vector<int> lidx(n_boxes, -1);
  std::vector<std::vector<int>> lists;
  auto take = [&](int l) {
    if (l >= 0 && lidx[l] < 0 && alive[l] && xleaf[l] && split_idx[l] < 0) {
      lidx[l] = (int)lists.size();
      lists.emplace_back();
    }
  };
  for (int a : merges) take(a);
  for (size_t k = 0; k < mv.size(); k++)
    if (to[k] != from[k]) {
      take(from[k]);
      take(to[k]);
    }
  // Points whose leaf changes come from the old lists of the removed
  // leaves, the leaves to subdivide and the leaves that lost or gained
  // points, and from the moves; `relocated` holds (point, leaf) where the
  // leaf is not point_leaf's.
  std::vector<std::pair<int, int>> relocated;
  auto visit = [&](int q) {
    int l = point_leaf[q];
    if (!alive[l]) {
      l = merged_into[l];
      relocated.push_back({q, l});
    }
    if (split_idx[l] >= 0) split_pts[split_idx[l]].push_back(q);
    else if (lidx[l] >= 0) lists[lidx[l]].push_back(q);
  };
  for (int e = 0; e < n_boxes; e++)
    if (leaf[e] && (!alive[e] || split_idx[e] >= 0 || lidx[e] >= 0))
      for (int i = orig_off[e]; i < orig_off[e] + orig_cnt[e]; i += sd) {
        const int q = orig_dofs[i] / sd;
        if (point_leaf[q] == e) visit(q);
      }
  for (size_t k = 0; k < mv.size(); k++)
    if (to[k] != from[k]) visit(mv[k]);
  for (auto& v : split_pts) std::sort(v.begin(), v.end());
  for (auto& v : lists) std::sort(v.begin(), v.end());
  std::vector<std::vector<int>> lvl_new(n_levels + 1);
  // node_subdivide: children appended to their level as bl, br, tr, tl;
  // DOFs handed down in order; children over the cap subdivided in the
  // order bl, tl, tr, br.
  struct Frame {
    int box;
    std::vector<int> pts;
  };
  std::vector<Frame> work;
  for (size_t i = 0; i < splits.size(); i++) {
    work.push_back({splits[i], std::move(split_pts[i])});
    while (!work.empty()) {
      Frame f = std::move(work.back());
      work.pop_back();
      const int b = f.box;
      const int lv = xlevel[b] + 1;
      if (lv >= 64) throw std::runtime_error("quadtree deeper than 64 levels");
      xleaf[b] = 0;
      const double mx = xmid_x[b], my = xmid_y[b];
      const double c[8] = {xbl_x[b], xbl_y[b], xbl_x[b], xtr_y[b],
                           xtr_x[b], xtr_y[b], xtr_x[b], xbl_y[b]};
      const double hs = xside[b] / 2.0;
      int ids[4];
      for (int q = 0; q < 4; q++) {
        double k[8];
        child_corners(c, mx, my, q, k);
        ids[q] = (int)xlevel.size();
        xlevel.push_back(lv);
        xparent.push_back(b);
        for (int z = 0; z < 4; z++) xchild.push_back(-1);
        xleaf.push_back(1);
        xbl_x.push_back(k[0]);
        xbl_y.push_back(k[1]);
        xtr_x.push_back(k[4]);
        xtr_y.push_back(k[3]);
        xside.push_back(hs);
        xcx.push_back(k[0] + hs / 2.0);
        xcy.push_back(k[1] + hs / 2.0);
        xmid_x.push_back(((k[6] - k[0]) / 2.0) + k[0]);
        xmid_y.push_back(((k[3] - k[1]) / 2.0) + k[1]);
        alive.push_back(1);
        prob.push_back(prob[b]);
      }
      for (int q = 0; q < 4; q++) xchild[4 * (size_t)b + q] = ids[q];
      if ((int)lvl_new.size() < lv + 1) lvl_new.resize(lv + 1);
      for (int q : {0, 3, 2, 1}) lvl_new[lv].push_back(ids[q]);
      std::vector<int> part[4];
      for (int p : f.pts) part[quadrant(pts[2 * (size_t)p], pts[2 * (size_t)p + 1], mx, my)].push_back(p);
      // depth first in the order bl, tl, tr, br: pushed in reverse
      lidx.resize(xlevel.size(), -1);
      for (int q = 3; q >= 0; q--) {
        if ((int)part[q].size() * sd > leaf_cap) {
          work.push_back({ids[q], std::move(part[q])});
        } else {
          for (int p : part[q]) relocated.push_back({p, ids[q]});
          lidx[ids[q]] = (int)lists.size();
          lists.push_back(std::move(part[q]));
        }
      }
    }
  }
  const int n_ext = (int)xlevel.size();

  // Slot order: per level, per problem, the surviving boxes in their
  // order, then the new ones in creation order.
  int nl = 0;
  for (int e = 0; e < n_ext; e++)
    if (alive[e]) nl = std::max(nl, xlevel[e] + 1);
  std::vector<int> order, lstart(nl + 1, 0);
  order.reserve(n_ext);
  for (int l = 0; l < nl; l++) {
    lstart[l] = (int)order.size();
    int i = l < n_levels ? level_start[l] : 0;
    const int i_end = l < n_levels ? level_start[l + 1] : 0;
    std::vector<int> fresh;
    if (l < (int)lvl_new.size()) fresh = lvl_new[l];
    std::stable_sort(fresh.begin(), fresh.end(),
                     [&prob](int a, int b) { return prob[a] < prob[b]; });
    size_t j = 0;
    for (int r = 0; r < n_roots; r++) {
      for (; i < i_end && prob[i] == r; i++)
        if (alive[i]) order.push_back(i);
      for (; j < fresh.size() && prob[fresh[j]] == r; j++) order.push_back(fresh[j]);
    }
  }
  lstart[nl] = (int)order.size();
  const int ns = (int)order.size();
  std::vector<int> slot_of(n_ext, -1);
  for (int s = 0; s < ns; s++) slot_of[order[s]] = s;

  HostTree nt;
  nt.n_points = n_points;
  nt.sol_dim = sd;
  nt.leaf_cap = leaf_cap;
  nt.n_roots = n_roots;
  nt.point_off = point_off;
  nt.n_boxes = ns;
  nt.n_levels = nl;
  nt.level_start = lstart;
  // Creation ids, problem after problem: surviving boxes in their id order,
  // then the new ones in creation order.
  {
    std::vector<std::vector<int>> fresh_of(n_roots);
    for (int e = n_boxes; e < n_ext; e++)
      if (alive[e]) fresh_of[prob[e]].push_back(e);
    nt.box_id.assign(ns, 0);
    nt.slot_of_id.assign(ns, 0);
    int id = 0, r = 0;
    auto emit = [&](int e) {
      nt.box_id[slot_of[e]] = id;
      nt.slot_of_id[id++] = slot_of[e];
    };
    for (int old_id = 0; old_id < n_boxes; old_id++) {
      const int e = slot_of_id[old_id];
      for (; r < prob[e]; r++)
        for (int f : fresh_of[r]) emit(f);
      if (alive[e]) emit(e);
    }
    for (; r < n_roots; r++)
      for (int f : fresh_of[r]) emit(f);
  }
  auto gather = [&order, ns](const auto& src, auto& dst) {
    dst.resize(ns);
    for (int s = 0; s < ns; s++) dst[s] = src[order[s]];
  };
  gather(xlevel, nt.level);
  gather(xleaf, nt.leaf);
  gather(xbl_x, nt.bl_x);
  gather(xbl_y, nt.bl_y);
  gather(xtr_x, nt.tr_x);
  gather(xtr_y, nt.tr_y);
  gather(xside, nt.side);
  gather(xcx, nt.cx);
  gather(xcy, nt.cy);
  gather(xmid_x, nt.mid_x);
  gather(xmid_y, nt.mid_y);
  nt.parent.resize(ns);
  nt.child.resize(4 * (size_t)ns);
  for (int s = 0; s < ns; s++) {
    const int e = order[s];
    nt.parent[s] = xparent[e] >= 0 ? slot_of[xparent[e]] : -1;
    for (int q = 0; q < 4; q++) {
      const int c = xchild[4 * (size_t)e + q];
      nt.child[4 * (size_t)s + q] = c >= 0 ? slot_of[c] : -1;
    }
  }

  // Neighbour lists (quadtree.cpp:752-801 recomputes them all), here only
  // where they can change.  A list is [coarser leaves that touch the box, by
  // slot] ++ [same-level neighbours: siblings, then the children of the
  // parent's same-level neighbours within reach] ++ [for a leaf: the boxes
  // below its same-level neighbours that it touches], the order of the
  // tree-wide pass in level order.  Levels are visited in order and each
  // part is formed again only where it can change: the same-level part of
  // new boxes and of the children of a box whose same-level part changed
  // or that neighbours a box whose children changed; the touches of new,
  // split or merged leaves, of leaves whose same-level part changed and of
  // leaves next to a box above a change; the coarser part of every box a
  // removed, split, merged or re-formed leaf touched (before or after).
  std::vector<char> inC(ns, 0), ancC(ns, 0), nbrC(ns, 0), slprop(ns, 0), rcchg(ns, 0), sm(ns, 0);
  for (int e : splits) sm[slot_of[e]] = inC[slot_of[e]] = 1;
  for (int e : merges) sm[slot_of[e]] = inC[slot_of[e]] = 1;
  for (int s = 0; s < ns; s++)
    if (order[s] >= n_boxes && !nt.leaf[s]) inC[s] = 1;
  for (int s = 0; s < ns; s++)
    if (inC[s])
      for (int w = s; w >= 0 && !ancC[w]; w = nt.parent[w]) ancC[w] = 1;
  // Removed boxes: their same-level neighbours' children lose cousins; the
  // boxes a removed leaf touched lose it.
  for (int e = 0; e < n_boxes; e++) {
    if (alive[e]) continue;
    for (int t = nbr_off[e]; t < nbr_off[e] + nbr_cnt[e]; t++) {
      const int w = nbr[t];
      if (!alive[w]) continue;
      if (level[w] == level[e] && !leaf[e]) nbrC[slot_of[w]] = 1;
      if (level[w] > level[e] && leaf[e]) rcchg[slot_of[w]] = 1;
    }
  }
  nt.nbr_off.assign(ns, 0);
  nt.nbr_cnt.assign(ns, 0);
  nt.nbr.reserve(nbr.size() + 1024);
  std::vector<char> relisted(ns, 0), mark(ns, 0);
  std::vector<int> o_rc, o_sl, o_td, rc, sl, td;
  auto has_gap = [](const std::vector<int>& v) {
    return std::find(v.begin(), v.end(), -1) != v.end();
  };
  for (int a = 0; a < ns; a++) {
    const int e = order[a] < n_boxes ? order[a] : -1;
    const int p = nt.parent[a];
    const int lv = nt.level[a];
    if (e >= 0 && !(p >= 0 && (slprop[p] || nbrC[p] || inC[p])) && !rcchg[a] && !sm[a] &&
        !inC[a]) {
      // nothing around the box changed: its list as it was, when every
      // entry survives and no same-level neighbour lies above a change
      const size_t at = nt.nbr.size();
      const int n0 = nbr_cnt[e];
      nt.nbr.resize(at + n0);
      const int* src_l = nbr.data() + nbr_off[e];
      int* dst_l = nt.nbr.data() + at;
      bool ok = true;
      for (int t = 0; t < n0; t++) {
        const int w = src_l[t];
        const int m = slot_of[w];
        dst_l[t] = m;
        ok = ok && alive[w];
      }
      if (ok && nt.leaf[a])
        for (int t = 0; ok && t < n0; t++)
          ok = !(level[src_l[t]] == level[e] && ancC[dst_l[t]]);
      if (ok) {
        nt.nbr_off[a] = (int)at;
        nt.nbr_cnt[a] = n0;
        continue;
      }
      nt.nbr.resize(at);
    }
    relisted[a] = 1;
    o_rc.clear();
    o_sl.clear();
    o_td.clear();
    if (e >= 0)
      for (int t = nbr_off[e]; t < nbr_off[e] + nbr_cnt[e]; t++) {
        const int w = nbr[t];
        const int m = alive[w] ? slot_of[w] : -1;
        (level[w] < level[e] ? o_rc : level[w] == level[e] ? o_sl : o_td).push_back(m);
      }
    // same-level part
    const bool sl_re = e < 0 || has_gap(o_sl) || (p >= 0 && (slprop[p] || nbrC[p] || inC[p]));
    if (sl_re) {
      sl.clear();
      if (p >= 0) {
        for (int q = 0; q < 4; q++) {
          const int c = nt.child[4 * (size_t)p + q];
          if (c != a) sl.push_back(c);
        }
        const double lim = nt.side[a] * std::sqrt(2.0) + 1e-5;
        for (int t = nt.nbr_off[p]; t < nt.nbr_off[p] + nt.nbr_cnt[p]; t++) {
          const int u = nt.nbr[t];
          if (nt.level[u] != nt.level[p]) continue;
          for (int q = 0; q < 4; q++) {
            const int cz = nt.child[4 * (size_t)u + q];
            if (cz < 0 || nt.level[cz] != lv) continue;
            const double dx = nt.bl_x[a] - nt.bl_x[cz], dy = nt.bl_y[a] - nt.bl_y[cz];
            if (std::sqrt(dx * dx + dy * dy) < lim) sl.push_back(cz);
          }
        }
      }
    } else {
      sl = o_sl;
    }
    const bool sl_chg = e < 0 || sl != o_sl;
    slprop[a] = sl_chg ? 1 : 0;
    // touches (leaves)
    bool td_re = false;
    if (nt.leaf[a]) {
      td_re = e < 0 || sm[a] || sl_chg || has_gap(o_td);
      for (size_t i = 0; !td_re && i < sl.size(); i++) td_re = ancC[sl[i]] != 0;
    }
    if (td_re) {
      td.clear();
      for (int u : sl)
        for (int q = 0; q < 4; q++) {
          const int ch = nt.child[4 * (size_t)u + q];
          if (ch >= 0) touch_walk(nt, a, ch, &td);
        }
    } else {
      td.clear();
      if (nt.leaf[a]) td = o_td;
    }
    // boxes that a touches now and did not before, or the other way round
    // (all of them when a leaf was split or is new), re-form their coarser
    // part; a keeps its place among the others' coarser leaves
    if (td_re || (e >= 0 && leaf[e] && !nt.leaf[a])) {
      for (int w : td) mark[w] = 1;
      if (e >= 0 && leaf[e])
        for (int w : o_td)
          if (w >= 0) {
            if (mark[w]) mark[w] = 2;
            else rcchg[w] = 1;
          }
      for (int w : td) {
        if (mark[w] == 1) rcchg[w] = 1;
        mark[w] = 0;
      }
    }
    // coarser part: on each coarser level, the leaves among the same-level
    // neighbours of a's ancestor there that touch a
    const bool rc_re = e < 0 || rcchg[a] || has_gap(o_rc);
    if (rc_re) {
      rc.clear();
      for (int w = p; w >= 0; w = nt.parent[w])
        for (int t = nt.nbr_off[w]; t < nt.nbr_off[w] + nt.nbr_cnt[w]; t++) {
          const int u = nt.nbr[t];
          if (nt.level[u] == nt.level[w] && nt.leaf[u] && touch_hit(nt, u, a)) rc.push_back(u);
        }
      std::sort(rc.begin(), rc.end());
    } else {
      rc = o_rc;
    }
    nt.nbr_off[a] = (int)nt.nbr.size();
    nt.nbr.insert(nt.nbr.end(), rc.begin(), rc.end());
    nt.nbr.insert(nt.nbr.end(), sl.begin(), sl.end());
    nt.nbr.insert(nt.nbr.end(), td.begin(), td.end());
    nt.nbr_cnt[a] = (int)nt.nbr.size() - nt.nbr_off[a];
    // a box whose children changed: its same-level neighbours' children
    // look again
    if (inC[a]) {
      for (int w : sl) nbrC[w] = 1;
      for (int w : o_sl)
        if (w >= 0) nbrC[w] = 1;
    }
  }
  // Row sources as the build forms them, within each problem.
  {
    std::vector<int> sprob(ns);
    for (int t = 0; t < ns; t++) sprob[t] = prob[order[t]];
    fill_src(nt, sprob);
  }
  nt.point_leaf.resize(n_points);
  for (int p = 0; p < n_points; p++) nt.point_leaf[p] = slot_of[point_leaf[p]];
  for (const auto& r : relocated) nt.point_leaf[r.first] = slot_of[r.second];
  nt.orig_off.assign(ns, 0);
  nt.orig_cnt.assign(ns, 0);
  nt.orig_dofs.reserve(orig_dofs.size());
  for (int t = 0; t < ns; t++) {
    const int e = order[t];
    nt.orig_off[t] = (int)nt.orig_dofs.size();
    if (!nt.leaf[t]) continue;
    if (lidx[e] >= 0) {
      for (int q : lists[lidx[e]])
        for (int d = 0; d < sd; d++) nt.orig_dofs.push_back(sd * q + d);
    } else {
      nt.orig_dofs.insert(nt.orig_dofs.end(), orig_dofs.begin() + orig_off[e],
                          orig_dofs.begin() + orig_off[e] + orig_cnt[e]);
    }
    nt.orig_cnt[t] = (int)nt.orig_dofs.size() - nt.orig_off[t];
  }

  // Marking seeds (new slots) and the slot map.
  auto now_slot = [&](int old) { return alive[old] ? slot_of[old] : slot_of[merged_into[old]]; };
  for (size_t i = 0; i < changed.size(); i++) {
    u.seeds.push_back(now_slot(old_leaf[i]));
    u.seeds.push_back(nt.point_leaf[changed[i]]);
  }
  u.old_slot.assign(ns, -1);
  u.old_id.assign(ns, -1);
  for (int s = 0; s < ns; s++) {
    const int e = order[s];
    if (e >= n_boxes) {
      u.seeds.push_back(s);
      continue;
    }
    u.old_slot[s] = e;
    u.old_id[s] = box_id[e];
    if (!relisted[s] && nt.level[s] >= 3) continue;  // list kept: the same row sources
    bool same = nt.leaf[s] == leaf[e] && nt.src_cnt[s] == src_cnt[e];
    for (int t = 0; same && t < src_cnt[e]; t++) {
      const int w = src[src_off[e] + t];
      same = alive[w] && slot_of[w] == nt.src[nt.src_off[s] + t];
    }
    if (!same) u.seeds.push_back(s);
  }
  u.restructured = true;
  u.lists_changed = true;
  u.n_split = (int)splits.size();
  u.n_merged = (int)merges.size();
  restore.armed = false;
  *this = std::move(nt);
  *out = std::move(u);
  return true;
}

}  // namespace iegpu
