// Host builder of the FMM's uniform-depth quadtree (fmm_tree.hpp).
//
// Nodes are sorted by 30-bit Morton key (three 10-bit counting-sort
// passes); the boxes of every level are the distinct key prefixes of the
// sorted keys, so each level comes out in Morton order and a box's children
// are one contiguous run of the next level.  Neighbour and interaction
// lists are found by binary search of a level's keys.

#include "fmm_tree.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace iegpu {

namespace {

uint32_t spread_bits(uint32_t v) {  // 15 bits -> even positions of 30
  v &= 0x7fffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

uint32_t compact_bits(uint32_t v) {  // inverse of spread_bits
  v &= 0x55555555u;
  v = (v | (v >> 1)) & 0x33333333u;
  v = (v | (v >> 2)) & 0x0f0f0f0fu;
  v = (v | (v >> 4)) & 0x00ff00ffu;
  v = (v | (v >> 8)) & 0x0000ffffu;
  return v;
}

uint32_t morton(uint32_t ix, uint32_t iy) { return spread_bits(ix) | (spread_bits(iy) << 1); }

// Index of `key` among the sorted keys of a level, or -1.
int find_key(const std::vector<uint32_t>& keys, uint32_t key) {
  auto it = std::lower_bound(keys.begin(), keys.end(), key);
  return (it != keys.end() && *it == key) ? int(it - keys.begin()) : -1;
}

}  // namespace

void FmmHostTree::build(const double* pts, int n, double leaf_points) {
  if (n <= 0) throw std::invalid_argument("FMM tree of no nodes");
  n_points = n;
  double x0 = pts[0], x1 = pts[0], y0 = pts[1], y1 = pts[1];
  for (int i = 1; i < n; i++) {
    x0 = std::min(x0, pts[2 * i]);
    x1 = std::max(x1, pts[2 * i]);
    y0 = std::min(y0, pts[2 * i + 1]);
    y1 = std::max(y1, pts[2 * i + 1]);
  }
  double ext = std::max(x1 - x0, y1 - y0);
  if (!(ext > 0)) ext = 1.0;
  side = ext * (1 + 1e-9);
  lo_x = 0.5 * (x0 + x1) - 0.5 * side;
  lo_y = 0.5 * (y0 + y1) - 0.5 * side;
  const double cells = double(1 << kFmmMaxDepth);
  std::vector<uint32_t> key(n), key2(n);
  std::vector<int> idx(n), idx2(n);
  for (int i = 0; i < n; i++) {
    auto cell = [&](double v, double lo) {
      const double c = std::floor((v - lo) / side * cells);
      return uint32_t(std::min(std::max(c, 0.0), cells - 1));
    };
    key[i] = morton(cell(pts[2 * i], lo_x), cell(pts[2 * i + 1], lo_y));
    idx[i] = i;
  }
  for (int shift = 0; shift < 30; shift += 10) {  // stable LSD radix sort
    std::vector<int> cnt(1025, 0);
    for (int i = 0; i < n; i++) cnt[((key[i] >> shift) & 1023) + 1]++;
    for (int b = 0; b < 1024; b++) cnt[b + 1] += cnt[b];
    for (int i = 0; i < n; i++) {
      const int d = cnt[(key[i] >> shift) & 1023]++;
      key2[d] = key[i];
      idx2[d] = idx[i];
    }
    key.swap(key2);
    idx.swap(idx2);
  }
  perm = idx;
  // leaf level: first at which the non-empty leaves hold <= leaf_points on average
  depth = kFmmMaxDepth;
  for (int l = 2; l <= kFmmMaxDepth; l++) {
    const int sh = 2 * (kFmmMaxDepth - l);
    int distinct = 1;
    for (int i = 1; i < n; i++) distinct += (key[i] >> sh) != (key[i - 1] >> sh);
    if (double(n) / distinct <= leaf_points) {
      depth = l;
      break;
    }
  }
  const int L = depth;
  // distinct keys per level, leaf level first
  std::vector<std::vector<uint32_t>> lk(L + 1);
  {
    const int sh = 2 * (kFmmMaxDepth - L);
    pt_off.clear();
    for (int i = 0; i < n; i++) {
      const uint32_t k = key[i] >> sh;
      if (i == 0 || k != lk[L].back()) {
        lk[L].push_back(k);
        pt_off.push_back(i);
      }
    }
    pt_off.push_back(n);
  }
  for (int l = L - 1; l >= 2; l--)
    for (uint32_t k : lk[l + 1])
      if (lk[l].empty() || (k >> 2) != lk[l].back()) lk[l].push_back(k >> 2);
  level_off.assign(L, 0);
  level_off[0] = 0;
  for (int l = 2; l <= L; l++) level_off[l - 1] = level_off[l - 2] + int(lk[l].size());
  const int nb_total = level_off[L - 1];
  center.assign(2 * size_t(nb_total), 0.0);
  parent.assign(nb_total, -1);
  child_off.assign(nb_total, 0);
  child_cnt.assign(nb_total, 0);
  vl_off.assign(nb_total + 1, 0);
  vl.clear();
  for (int l = 2; l <= L; l++) {
    const double s = std::ldexp(side, -l);
    const int base = level_off[l - 2];
    const int dim = 1 << l;
    for (int i = 0; i < int(lk[l].size()); i++) {
      const int b = base + i;
      const int ix = int(compact_bits(lk[l][i])), iy = int(compact_bits(lk[l][i] >> 1));
      center[2 * b] = lo_x + (ix + 0.5) * s;
      center[2 * b + 1] = lo_y + (iy + 0.5) * s;
      if (l > 2) parent[b] = level_off[l - 3] + find_key(lk[l - 1], lk[l][i] >> 2);
      if (l < L) {
        const auto first = std::lower_bound(lk[l + 1].begin(), lk[l + 1].end(), lk[l][i] << 2);
        const auto last = std::lower_bound(first, lk[l + 1].end(), (lk[l][i] << 2) + 4);
        child_off[b] = level_off[l - 1] + int(first - lk[l + 1].begin());
        child_cnt[b] = int(last - first);
      }
      const int px = ix >> 1, py = iy >> 1;
      for (int a = -1; a <= 1; a++)
        for (int c = -1; c <= 1; c++) {
          const int qx = px + a, qy = py + c;
          if (qx < 0 || qy < 0 || 2 * qx >= dim || 2 * qy >= dim) continue;
          for (int ca = 0; ca < 2; ca++)
            for (int cb = 0; cb < 2; cb++) {
              const int sx = 2 * qx + ca, sy = 2 * qy + cb;
              if (std::abs(sx - ix) <= 1 && std::abs(sy - iy) <= 1) continue;
              const int j = find_key(lk[l], morton(sx, sy));
              if (j >= 0) vl.push_back(base + j);
            }
        }
      vl_off[b + 1] = int(vl.size());
    }
  }
  // adjacent leaves
  const int nl = int(lk[L].size());
  nb_off.assign(nl + 1, 0);
  nb.clear();
  for (int i = 0; i < nl; i++) {
    const int ix = int(compact_bits(lk[L][i])), iy = int(compact_bits(lk[L][i] >> 1));
    for (int a = -1; a <= 1; a++)
      for (int c = -1; c <= 1; c++) {
        const int sx = ix + a, sy = iy + c;
        if (sx < 0 || sy < 0 || sx >= (1 << L) || sy >= (1 << L)) continue;
        const int j = find_key(lk[L], morton(sx, sy));
        if (j >= 0) nb.push_back(j);
      }
    nb_off[i + 1] = int(nb.size());
  }
}

void fmm_tables(std::vector<double>& m2m, std::vector<double>& m2l,
                std::vector<double>& l2l) {
  const int K = kFmmMaxOrder;
  // Pascal's triangle in 128-bit integers: C(127, 63) < 2^127
  std::vector<unsigned __int128> C(size_t(2 * K) * (2 * K), 0);
  for (int nn = 0; nn < 2 * K; nn++) {
    C[size_t(nn) * 2 * K] = 1;
    for (int k = 1; k <= nn; k++)
      C[size_t(nn) * 2 * K + k] = C[size_t(nn - 1) * 2 * K + k - 1] +
                                  (k < nn ? C[size_t(nn - 1) * 2 * K + k] : 0);
  }
  auto binom = [&](int nn, int k) { return double(C[size_t(nn) * 2 * K + k]); };
  m2m.assign(size_t(K) * K, 0.0);
  m2l.assign(size_t(K) * K, 0.0);
  l2l.assign(size_t(K) * K, 0.0);
  for (int a = 0; a < K; a++)
    for (int b = 0; b < K; b++) {
      if (a <= b) m2m[size_t(a) * K + b] = std::ldexp(binom(b, a), -a);  // [m][k]
      m2l[size_t(a) * K + b] = binom(a + b, b);                            // [k][l]
      if (b <= a) l2l[size_t(a) * K + b] = std::ldexp(binom(a, b), -(b + 1));  // [l][m]
    }
}

}  // namespace iegpu
