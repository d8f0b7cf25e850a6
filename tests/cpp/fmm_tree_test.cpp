// CPU checks of the FMM's quadtree and translation algebra (fmm_tree.hpp,
// rs_fmm.hip), run by tests/test_host.py:
//   * the sorted order is a permutation and the leaves partition it;
//   * parents, children and levels agree, children are contiguous;
//   * interaction lists hold same-level boxes at least one box apart;
//   * every (target leaf, source leaf) pair is counted exactly once: either
//     adjacent at the leaf level, or through exactly one interaction list
//     of the target's ancestors (itself included);
//   * the expansions and translations, evaluated in the order rs_fmm.hip
//     uses them (scalar complex arithmetic, same tables, same box-centred
//     Stokes origin), reproduce the direct sum of the Laplace, Laplace
//     Neumann and Stokes kernel entries to 1e-13 relative at order 48.
// Exit status 0 on success; prints the failing check otherwise.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "../../ie-solver_amd/csrc/fmm_tree.hpp"

using iegpu::FmmHostTree;
using C = std::complex<double>;

namespace {

int fails = 0;
void check(bool ok, const char* what, int a = 0, int b = 0) {
  if (!ok) {
    if (fails < 10) std::printf("FAIL %s (%d, %d)\n", what, a, b);
    fails++;
  }
}

struct Curve {
  std::vector<double> pts, nrm, w;
};

// A seven-armed star r(t) = 0.3 (1 + 0.2 cos 7t) around (0.5, 0.5).
Curve star(int n) {
  Curve c;
  c.pts.resize(2 * n);
  c.nrm.resize(2 * n);
  c.w.resize(n);
  for (int i = 0; i < n; i++) {
    const double t = 2 * M_PI * i / n;
    const double r = 0.3 * (1 + 0.2 * std::cos(7 * t)), dr = -0.3 * 0.2 * 7 * std::sin(7 * t);
    c.pts[2 * i] = 0.5 + r * std::cos(t);
    c.pts[2 * i + 1] = 0.5 + r * std::sin(t);
    const double tx = dr * std::cos(t) - r * std::sin(t), ty = dr * std::sin(t) + r * std::cos(t);
    const double sp = std::hypot(tx, ty);
    c.nrm[2 * i] = ty / sp;
    c.nrm[2 * i + 1] = -tx / sp;
    c.w[i] = sp * 2 * M_PI / n;
  }
  return c;
}

int level_of(const FmmHostTree& T, int b) {
  for (int l = 2; l <= T.depth; l++)
    if (b < T.level_off[l - 1]) return l;
  return -1;
}

void check_structure(const FmmHostTree& T) {
  const int n = T.n_points, L = T.depth, nb = T.n_boxes();
  std::vector<int> seen(n, 0);
  for (int p : T.perm) {
    check(p >= 0 && p < n, "perm range", p);
    if (p >= 0 && p < n) seen[p]++;
  }
  for (int i = 0; i < n; i++) check(seen[i] == 1, "perm is a permutation", i);
  check(T.pt_off.front() == 0 && T.pt_off.back() == n, "leaf ranges cover the nodes");
  for (int i = 0; i < T.n_leaves(); i++) check(T.pt_off[i] < T.pt_off[i + 1], "leaf non-empty", i);
  for (int b = 0; b < nb; b++) {
    const int l = level_of(T, b);
    if (l > 2) {
      const int p = T.parent[b];
      check(level_of(T, p) == l - 1, "parent level", b, p);
      check(b >= T.child_off[p] && b < T.child_off[p] + T.child_cnt[p], "child of parent", b, p);
    } else {
      check(T.parent[b] == -1, "level 2 has no parent", b);
    }
    if (l < L) {
      check(T.child_cnt[b] >= 1 && T.child_cnt[b] <= 4, "child count", b);
      for (int c = T.child_off[b]; c < T.child_off[b] + T.child_cnt[b]; c++)
        check(T.parent[c] == b, "children contiguous", b, c);
    }
    const double s = std::ldexp(T.side, -l);
    for (int v = T.vl_off[b]; v < T.vl_off[b + 1]; v++) {
      const int src = T.vl[v];
      check(level_of(T, src) == l, "interaction list level", b, src);
      const double dx = std::fabs(T.center[2 * b] - T.center[2 * src]) / s;
      const double dy = std::fabs(T.center[2 * b + 1] - T.center[2 * src + 1]) / s;
      check(std::max(dx, dy) > 1.5 && std::max(dx, dy) < 3.5, "interaction list separation", b, src);
    }
  }
  // pair coverage
  const int nl = T.n_leaves(), f = T.first_leaf();
  std::vector<std::vector<int>> anc(nl);  // ancestors from the leaf up to level 2
  for (int i = 0; i < nl; i++)
    for (int b = f + i; b >= 0; b = T.parent[b]) anc[i].push_back(b);
  std::vector<std::map<int, int>> vset(nb);
  for (int b = 0; b < nb; b++)
    for (int v = T.vl_off[b]; v < T.vl_off[b + 1]; v++) vset[b][T.vl[v]]++;
  for (int t = 0; t < nl; t++) {
    std::vector<int> cnt(nl, 0);
    for (int q = T.nb_off[t]; q < T.nb_off[t + 1]; q++) cnt[T.nb[q]]++;
    for (int s = 0; s < nl; s++) {
      for (size_t a = 0; a < anc[t].size() && a < anc[s].size(); a++) {
        auto it = vset[anc[t][a]].find(anc[s][a]);
        if (it != vset[anc[t][a]].end()) cnt[s] += it->second;
      }
      check(cnt[s] == 1, "pair counted once", t, s);
    }
  }
}

// The FMM in scalar complex arithmetic, in rs_fmm.hip's order.
double fmm_vs_direct(const Curve& g, int pde, int P, double leaf_points) {
  const int n = int(g.w.size()), NE = pde == 2 ? 3 : 1, SD = pde == 2 ? 2 : 1;
  FmmHostTree T;
  T.build(g.pts.data(), n, leaf_points);
  std::vector<double> m2m, m2l, l2l;
  iegpu::fmm_tables(m2m, m2l, l2l);
  std::mt19937_64 rng(11 + pde);
  std::normal_distribution<double> nd;
  std::vector<double> x(SD * n);
  for (auto& v : x) v = nd(rng);
  const int nb = T.n_boxes(), L = T.depth;
  std::vector<C> M(size_t(nb) * NE * P), Lc(size_t(nb) * NE * P);
  auto ctr = [&](int b) { return C(T.center[2 * b], T.center[2 * b + 1]); };
  auto at = [&](std::vector<C>& v, int b, int e, int k) -> C& { return v[(size_t(b) * NE + e) * P + k]; };
  for (int i = 0; i < T.n_leaves(); i++) {
    const int b = T.first_leaf() + i;
    const double s = std::ldexp(T.side, -L);
    for (int j = T.pt_off[i]; j < T.pt_off[i + 1]; j++) {
      const int o = T.perm[j];
      const C y(g.pts[2 * o], g.pts[2 * o + 1]), nn(g.nrm[2 * o], g.nrm[2 * o + 1]);
      C q[3];
      if (pde == 2) {
        const C mu(x[2 * o], x[2 * o + 1]);
        q[0] = g.w[o] / (4 * M_PI) * nn * mu;
        q[1] = g.w[o] / (2 * M_PI) * (std::conj(nn) * mu).real();
        q[2] = q[0] * std::conj(y - ctr(b));
      } else if (pde == 1) {
        q[0] = g.w[o] * x[o] / (2 * M_PI);
      } else {
        q[0] = -g.w[o] * x[o] / (2 * M_PI) * nn;
      }
      const C u = (y - ctr(b)) / s;
      C pk = 1;
      for (int k = 0; k < P; k++, pk *= u)
        for (int e = 0; e < NE; e++) at(M, b, e, k) += q[e] * pk;
    }
  }
  for (int l = L - 1; l >= 2; l--)
    for (int b = T.level_off[l - 2]; b < T.level_off[l - 1]; b++) {
      const double s = std::ldexp(T.side, -l);
      for (int c = T.child_off[b]; c < T.child_off[b] + T.child_cnt[b]; c++) {
        const C d = (ctr(c) - ctr(b)) / s, shift = std::conj(ctr(c) - ctr(b));
        for (int k = 0; k < P; k++) {
          C pw = 1;
          for (int m = k; m >= 0; m--, pw *= d)
            for (int e = 0; e < NE; e++)
              at(M, b, e, k) += m2m[m * 64 + k] * pw *
                                (at(M, c, e, m) + (e == 2 ? shift * at(M, c, 0, m) : C(0)));
        }
      }
    }
  for (int l = 2; l <= L; l++)
    for (int b = T.level_off[l - 2]; b < T.level_off[l - 1]; b++) {
      const double s = std::ldexp(T.side, -l);
      if (l > 2) {
        const int p = T.parent[b];
        const C d = (ctr(b) - ctr(p)) / (2 * s), shift = std::conj(ctr(p) - ctr(b));
        for (int j = 0; j < P; j++) {
          C pw = 1;
          for (int ll = j; ll < P; ll++, pw *= d)
            for (int e = 0; e < NE; e++)
              at(Lc, b, e, j) += l2l[ll * 64 + j] * pw *
                                 (at(Lc, p, e, ll) + (e == 2 ? shift * at(Lc, p, 0, ll) : C(0)));
        }
      }
      for (int v = T.vl_off[b]; v < T.vl_off[b + 1]; v++) {
        const int src = T.vl[v];
        const C w = s / (ctr(b) - ctr(src)), shift = std::conj(ctr(src) - ctr(b));
        for (int e = 0; e < NE; e++)
          for (int j = 0; j < P; j++) {
            C S = 0;
            for (int k = 0; k < P; k++)
              S += m2l[k * 64 + j] * std::pow(w, k) *
                   (at(M, src, e, k) + (e == 2 ? shift * at(M, src, 0, k) : C(0)));
            at(Lc, b, e, j) += ((j & 1) ? -1.0 : 1.0) * std::pow(w, j + 1) * S;
          }
      }
    }
  auto entry = [&](int o, int so, double& a0, double& a1) {
    const double r0 = g.pts[2 * o] - g.pts[2 * so], r1 = g.pts[2 * o + 1] - g.pts[2 * so + 1];
    const double d2 = r0 * r0 + r1 * r1;
    if (pde == 2) {
      const double pot = g.w[so] / M_PI * (r0 * g.nrm[2 * so] + r1 * g.nrm[2 * so + 1]) / (d2 * d2);
      const double proj = (r0 * x[2 * so] + r1 * x[2 * so + 1]) * pot;
      a0 += r0 * proj;
      a1 += r1 * proj;
    } else if (pde == 1) {
      a0 += g.w[so] / (2 * M_PI) * x[so] * (r0 * g.nrm[2 * o] + r1 * g.nrm[2 * o + 1]) / d2;
    } else {
      a0 += -g.w[so] / (2 * M_PI) * x[so] * (r0 * g.nrm[2 * so] + r1 * g.nrm[2 * so + 1]) / d2;
    }
  };
  double num = 0, den = 0;
  for (int i = 0; i < T.n_leaves(); i++) {
    const int b = T.first_leaf() + i;
    const double s = std::ldexp(T.side, -L);
    for (int it = T.pt_off[i]; it < T.pt_off[i + 1]; it++) {
      const int o = T.perm[it];
      const C t(g.pts[2 * o], g.pts[2 * o + 1]), tn(g.nrm[2 * o], g.nrm[2 * o + 1]);
      const C u = (t - ctr(b)) / s;
      C v[3], dv[3];
      for (int e = 0; e < NE; e++)
        for (int ll = P - 1; ll >= 0; ll--) {
          if (ll > 0) dv[e] = dv[e] * u + double(ll) * at(Lc, b, e, ll);
          v[e] = v[e] * u + at(Lc, b, e, ll);
        }
      double a0, a1 = 0;
      if (pde == 2) {
        const C D3 = -dv[0] / (s * s), D4 = -dv[2] / (s * s);
        const C uu = v[0] / s + std::conj(v[1] / s) + (t - ctr(b)) * std::conj(D3) - std::conj(D4);
        a0 = uu.real();
        a1 = uu.imag();
      } else if (pde == 1) {
        a0 = (tn * v[0] / s).real();
      } else {
        a0 = (v[0] / s).real();
      }
      for (int q = T.nb_off[i]; q < T.nb_off[i + 1]; q++)
        for (int jj = T.pt_off[T.nb[q]]; jj < T.pt_off[T.nb[q] + 1]; jj++)
          if (T.perm[jj] != o) entry(o, T.perm[jj], a0, a1);
      double e0 = 0, e1 = 0;
      for (int so = 0; so < n; so++)
        if (so != o) entry(o, so, e0, e1);
      num += (a0 - e0) * (a0 - e0) + (a1 - e1) * (a1 - e1);
      den += e0 * e0 + e1 * e1;
    }
  }
  return std::sqrt(num / den);
}

}  // namespace

int main() {
  for (int n : {100, 3000, 20000}) {
    const Curve g = star(n);
    FmmHostTree T;
    T.build(g.pts.data(), n, 32);
    std::printf("star %d: depth %d, %d boxes, %d leaves\n", n, T.depth, T.n_boxes(), T.n_leaves());
    check_structure(T);
  }
  {  // every node on one point: the deepest level, one leaf
    std::vector<double> same(2 * 50, 0.25);
    FmmHostTree T;
    T.build(same.data(), 50, 32);
    check(T.depth == iegpu::kFmmMaxDepth && T.n_leaves() == 1, "coincident nodes");
    check_structure(T);
  }
  const Curve g = star(4000);
  for (int pde = 0; pde < 3; pde++) {
    const double err = fmm_vs_direct(g, pde, 48, 32);
    std::printf("pde %d order 48: FMM vs direct %.2e\n", pde, err);
    check(err <= 1e-13, "translations reproduce the direct sum", pde);
  }
  if (fails) std::printf("%d failures\n", fails);
  return fails ? 1 : 0;
}
