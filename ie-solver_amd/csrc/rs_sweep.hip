// Register-sweep decouple of the compressed boxes with |active| <= 64 (one
// wave per box): Schur complement, X_rr^{-1}, L, U and the carried block,
// skel_factorization.cpp:77-128.
#include <hip/hip_runtime.h>

#include "rs_device.hpp"
#include "rs_kernels.hpp"

namespace iegpu {

namespace {

// ---------------------------------------------------------------------------
// Register-sweep decouple: one wave per compressed box with |active| <= 64.
//
// Lane i holds row i of the box's block in pivot order (skeleton positions
// first) as NB vectors of 16 registers; columns are register indices.
// With G = [I -T; 0 I] the wave forms, without a workgroup barrier,
//   Z = G^T K_BN G = [[A11, K_sr], [K_rs, X_rr]],  A11 := -(children's
//   carried blocks on the skeleton),
// in two steps: W = K_BN G (each lane its own row, T rows broadcast from
// LDS) and Z = G^T W (the skeleton rows of W broadcast from LDS).  It then
// exchanges each redundant column c against a pivot row p(c) (Gauss-Jordan
// with partial pivoting among the redundant rows, rows not moved; the pivot
// row of a step reaches every lane through LDS).  Afterwards
//   row p(c),      skeleton column a   = -U(c, a)
//   row p(c),      redundant column c2 =  X_rr^{-1}(c, p(c2))
//   skeleton row i, redundant column c =  L(i, p(c))
//   skeleton row i, skeleton column a  = -(carried block)(i, a)
// where U = X_rr^{-1} K_rs, L = K_sr X_rr^{-1} and the carried block is
// the children's update plus K_sr X_rr^{-1} K_rs: one sweep of r n^2
// multiply-adds yields the inverse and the three products, the flops of
// the four separate products.  A pivot row is not rescaled when it is
// exchanged; its scale -1 / pivot stays in a per-lane register and is
// applied when the factors are written.
// ---------------------------------------------------------------------------

template <typename R>
struct Vec16;
template <>
struct Vec16<double> {
  typedef double T __attribute__((ext_vector_type(16)));
  typedef double A __attribute__((ext_vector_type(16), aligned(16)));  // in LDS
};
template <>
struct Vec16<float> {
  typedef float T __attribute__((ext_vector_type(16)));
  typedef float A __attribute__((ext_vector_type(16), aligned(16)));
};

// The top 32 bits of |v| (exponent and leading mantissa bits, sign clear).
__device__ inline unsigned mag_bits(double v) {
  return (unsigned)((unsigned long long)__double_as_longlong(v) >> 32) & 0x7fffffffu;
}
__device__ inline unsigned mag_bits(float v) { return __float_as_uint(v) & 0x7fffffffu; }

// Maximum over the 64 lanes, in every lane (lanes outside a row_bcast mask
// read 0, the neutral element).
__device__ inline unsigned wave_max_u32(unsigned v) {
  v = max(v, (unsigned)dpp_i32<dpp::kQuadSwap1, 0xF>((int)v));
  v = max(v, (unsigned)dpp_i32<dpp::kQuadSwap2, 0xF>((int)v));
  v = max(v, (unsigned)dpp_i32<dpp::kRowHalfMirror, 0xF>((int)v));
  v = max(v, (unsigned)dpp_i32<dpp::kRowMirror, 0xF>((int)v));
  v = max(v, (unsigned)dpp_i32<dpp::kRowBcast15, 0xA>((int)v));
  v = max(v, (unsigned)dpp_i32<dpp::kRowBcast31, 0xC>((int)v));
  return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

// LDS of one box of the sweep decouple (NMAX = 16 NB positions).
template <typename R, int NB>
struct SweepLds {
  static constexpr int NMAX = 16 * NB;
  // T rows, row a at a (r + 16) + 16 after 16 zeros, so that the columns of
  // a 16-register block below the first redundant column, and those past
  // the last, read zeros: k (r + 16) + 16 <= (n + 16)^2 / 4 + 16.
  static constexpr int kTr = (NMAX + 16) * (NMAX + 16) / 4 + 16;
  static constexpr int kRows = 8;  // skeleton rows of W per broadcast chunk
  alignas(16) R tr[kTr];
  union alignas(16) {
    struct {
      double2 xy[NMAX], nrm[NMAX], wk[NMAX];
      int comp[NMAX];
    } geo;
    R w[kRows * NMAX];
  } u;
  alignas(16) R piv_row[NMAX];
  int pinfo[NMAX];  // child (q << 16 | local position), -1: none
  int pr[NMAX];     // pivot row of each redundant column
  const R* chS[4];
  int chN[4];
};

// Kernel entry without branches (the sweep evaluates whole rows at once):
// the coincident-point value is selected, not branched to.
template <int PDE>
__device__ inline double sweep_entry(const DofGeo& t, const DofGeo& s) {
  const double r0 = t.x - s.x, r1 = t.y - s.y;
  const bool same = t.x == s.x && t.y == s.y;
  const double d2 = same ? 1.0 : r0 * r0 + r1 * r1;
  const double i2 = recip(d2);
  if constexpr (PDE == kPdeStokes) {
    const double scale = 1.0 / M_PI;
    const int a = t.c, bb = s.c;
    const double tna = a ? t.n2 : t.n1;
    const double snb = bb ? s.n2 : s.n1;
    const double pot = s.w * scale * (r0 * s.n1 + r1 * s.n2) * (i2 * i2);
    const double ra = a ? r1 : r0, rb = bb ? r1 : r0;
    const double off = pot * ra * rb + s.w * tna * snb;
    const double dp = -0.5 * s.k * s.w * scale;
    const double d00 = -0.5 + dp * s.n2 * s.n2 + s.w * tna * s.n1;
    const double d11 = -0.5 + dp * s.n1 * s.n1 + s.w * tna * s.n2;
    const double d01 = -dp * s.n1 * s.n2 + s.w * t.n2 * s.n1;
    const double dg = (a == 0 && bb == 0) ? d00 : ((a == 1 && bb == 1) ? d11 : d01);
    return same ? dg : off;
  } else if constexpr (PDE == kPdeLaplaceNeumann) {
    const double scale = 1.0 / (2 * M_PI);
    const double off = s.w + s.w * scale * (r0 * t.n1 + r1 * t.n2) * i2;
    return same ? s.w - 0.5 + 0.5 * s.k * s.w * scale : off;
  } else {
    const double scale = 1.0 / (2 * M_PI);
    const double off = -s.w * scale * (r0 * s.n1 + r1 * s.n2) * i2;
    return same ? 0.5 + 0.5 * s.k * s.w * scale : off;
  }
}

// Boxes with nmin < |active| <= 16 NB.
template <typename R, int PDE, int NB>
__global__ __launch_bounds__(64, (NB <= 2 ? 4 : 2)) void k_decouple_sweep(
    DevGeom g, DevTree t, DevState s, int start, int nmin) {
  using V = typename Vec16<R>::T;
  using VA = typename Vec16<R>::A;
  using L = SweepLds<R, NB>;
  constexpr int NMAX = L::NMAX;
  __shared__ L sh;

  const int b = start + blockIdx.x;
  const int lane = threadIdx.x;
  if (!pools_ok(s) || box_clean(s, b)) return;
  const int n = s.act_cnt[b], k = s.rank[b];
  if (n <= nmin || n > NMAX || k == 0) return;  // another launch's box
  const int r = n - k;
  const int* act = s.act_list + (int64_t)b * s.ncap;
  R* F = reinterpret_cast<R*>(s.fac_pool) + s.fac_off[b];
  const FacLayout fl = FacLayout::make(k, r);

  // Children's carried blocks, the geometry of every position and each
  // position's child (position i = perm[i] in the active list).
  if (lane < 4) {
    const int c = t.leaf[b] ? -1 : t.child[4 * b + lane];
    sh.chS[lane] = (c >= 0) ? carried_block<R>(s, c) : nullptr;
    sh.chN[lane] = (c >= 0) ? s.out_cnt[c] : 0;
  }
  if (lane < NMAX) {
    DofGeo d;
    int info = -1;
    if (lane < n) {
      const int ai = s.perm[(int64_t)b * s.ncap + lane];
      d = dof_geo(g, act[ai]);
      if (!t.leaf[b]) {
        for (int q = 0; q < 4; q++) {
          const int c = t.child[4 * b + q];
          const int p0 = s.chpos[4 * b + q];
          if (c >= 0 && ai >= p0 && ai < p0 + s.out_cnt[c]) info = (q << 16) | (ai - p0);
        }
      }
    } else {  // far away with zero weight: every entry with it is 0
      d.x = d.y = 1e10;
      d.n1 = d.n2 = d.w = d.k = 0;
      d.c = 0;
    }
    sh.u.geo.xy[lane] = make_double2(d.x, d.y);
    sh.u.geo.nrm[lane] = make_double2(d.n1, d.n2);
    sh.u.geo.wk[lane] = make_double2(d.w, d.k);
    sh.u.geo.comp[lane] = d.c;
    sh.pinfo[lane] = info;
  }
  // T rows (F.T is k x r column-major) with their zero padding.
  for (int idx = lane; idx < k * (r + 16) + 16; idx += 64) sh.tr[idx] = R(0);
  __syncthreads();
  for (int idx = lane; idx < k * r; idx += 64) {
    const int a = idx % k, jp = idx / k;
    sh.tr[a * (r + 16) + 16 + jp] = F[fl.T + idx];
  }

  // Children's update at (own position, j): read from the child's carried
  // block when both positions belong to the same child (else from a valid
  // dummy address, and dropped).
  const int my_info = lane < NMAX ? sh.pinfo[lane] : -1;
  const int my_q = my_info >> 16, my_loc = my_info & 0xffff;
  const R* my_S = my_info >= 0 ? sh.chS[my_q] : nullptr;
  const int my_N = my_info >= 0 ? sh.chN[my_q] : 0;
  auto upd = [&](int j) -> R {
    const int inf = sh.pinfo[j];
    const bool same = my_S != nullptr && inf >= 0 && (inf >> 16) == my_q;
    const R* src = same ? my_S + my_loc + (int64_t)my_N * (inf & 0xffff) : F;
    const R v = *src;
    return same ? v : R(0);
  };
  DofGeo tg;
  {
    const int i = lane < NMAX ? lane : 0;
    const double2 xy = sh.u.geo.xy[i], nr = sh.u.geo.nrm[i], wk = sh.u.geo.wk[i];
    tg.x = xy.x;
    tg.y = xy.y;
    tg.n1 = nr.x;
    tg.n2 = nr.y;
    tg.w = wk.x;
    tg.k = wk.y;
    tg.c = sh.u.geo.comp[i];
  }
  const bool live = lane < n;
  const bool leaf = t.leaf[b] != 0;
  // Row of K_BN: minus the children's carried blocks (all loads in flight
  // together; a leaf has none), then plus the kernel entries, column by
  // column (the register of column j is picked with a wave-uniform index).
  V z[NB];
#pragma unroll
  for (int blk = 0; blk < NB; blk++) {
#pragma unroll
    for (int jj = 0; jj < 16; jj++) z[blk][jj] = R(0);
    if (!leaf && 16 * blk < n) {
#pragma unroll
      for (int jj = 0; jj < 16; jj++) z[blk][jj] = -upd(16 * blk + jj);
    }
  }
#pragma unroll
  for (int blk = 0; blk < NB; blk++) {
    const int j1 = min(n, 16 * blk + 16);
    for (int j = 16 * blk; j < j1; j++) {
      DofGeo sg;
      const double2 xy = sh.u.geo.xy[j], nr = sh.u.geo.nrm[j], wk = sh.u.geo.wk[j];
      sg.x = xy.x;
      sg.y = xy.y;
      sg.n1 = nr.x;
      sg.n2 = nr.y;
      sg.w = wk.x;
      sg.k = wk.y;
      sg.c = sh.u.geo.comp[j];
      const R v = static_cast<R>(sweep_entry<PDE>(tg, sg));
      const int jj = j & 15;
      z[blk][jj] = live ? z[blk][jj] + v : R(0);
    }
  }
  __syncthreads();  // the geometry's LDS is reused below

  // W = K_BN G: z(j) -= z(a) T(a, j - k) for every skeleton a and redundant
  // j.  Blocks below the one holding column k are skeleton columns; in that
  // block, and past column n, the padded T rows read zeros.
  const int bk = k >> 4;
#pragma unroll
  for (int ab = 0; ab < NB; ab++) {
    if (16 * ab < k) {
      const int a_end = min(16, k - 16 * ab);
      for (int aa = 0; aa < a_end; aa++) {
        const int a = 16 * ab + aa;
        const R pa = z[ab][aa];
        const R* trow = sh.tr + a * (r + 16) + 16 - k;
#pragma unroll
        for (int blk = 0; blk < NB; blk++) {
          if (blk >= bk && 16 * blk < n) {
#pragma unroll
            for (int jj = 0; jj < 16; jj++)
              z[blk][jj] = fma(-pa, trow[16 * blk + jj], z[blk][jj]);
          }
        }
      }
    }
  }
  // Z = G^T W: z_i(j) -= T(a, i - k) W(a, j) for every redundant row i,
  // the skeleton rows of W broadcast through LDS, kRows at a time.
  const bool red = lane >= k && lane < n;
  for (int a0 = 0; a0 < k; a0 += L::kRows) {
    const int a1 = min(a0 + L::kRows, k);
    if (lane >= a0 && lane < a1) {
      VA* dst = reinterpret_cast<VA*>(sh.u.w + (lane - a0) * NMAX);
#pragma unroll
      for (int blk = 0; blk < NB; blk++) dst[blk] = z[blk];
    }
    __syncthreads();
    for (int a = a0; a < a1; a++) {
      const R coef = sh.tr[a * (r + 16) + 16 + (red ? lane - k : -16)];
      const R* wrow = sh.u.w + (a - a0) * NMAX;
#pragma unroll
      for (int blk = 0; blk < NB; blk++) {
        if (16 * blk < n) {
#pragma unroll
          for (int jj = 0; jj < 16; jj++)
            z[blk][jj] = fma(-coef, wrow[16 * blk + jj], z[blk][jj]);
        }
      }
    }
    __syncthreads();
  }
  // Skeleton block: minus the children's update.
#pragma unroll
  for (int blk = 0; blk < NB; blk++) {
    const int j1 = min(k, 16 * blk + 16);
    for (int j = 16 * blk; j < j1; j++) {
      const R u = leaf ? R(0) : upd(j);
      const int jj = j & 15;
      z[blk][jj] = lane < k ? -u : z[blk][jj];
    }
  }
  if (s.fwd_pool) {  // X_rr before the sweep, for the forward apply
    R* Xf = reinterpret_cast<R*>(s.fwd_pool) + s.fac_off[b];
#pragma unroll
    for (int blk = 0; blk < NB; blk++) {
      const int j0 = max(k, 16 * blk), j1 = min(n, 16 * blk + 16);
      for (int j = j0; j < j1; j++) {
        const R v = z[blk][j & 15];
        if (red) Xf[(lane - k) + (int64_t)r * (j - k)] = v;
      }
    }
  }

  // Sweep over the redundant columns.
  bool used = !red;
  R sig = R(1);
  int myc = 0, sing = 0;
  VA* prow = reinterpret_cast<VA*>(sh.piv_row);
#pragma unroll
  for (int cb = 0; cb < NB; cb++) {
    const int c0 = max(k, 16 * cb), c1 = min(n, 16 * cb + 16);
    for (int c = c0; c < c1; c++) {
      const int cc = c & 15;
      const R pc = z[cb][cc];
      // Largest |pivot column| among unused redundant rows, first row on
      // ties (keys keep the leading exponent and mantissa bits).
      const unsigned key = used ? 0u : ((mag_bits(pc) & ~63u) | (unsigned)(63 - lane));
      const int p = 63 - (int)(wave_max_u32(key) & 63u);
      const R piv = lane_value(pc, p);
      sing |= (piv == R(0)) ? 1 : 0;
      const R inv = R(1) / piv;
      const bool me = lane == p;
      if (me) {
#pragma unroll
        for (int blk = 0; blk < NB; blk++) prow[blk] = z[blk];
      }
      const R m = me ? R(0) : pc * inv;
#pragma unroll
      for (int blk = 0; blk < NB; blk++) {
        if (16 * blk < n) {
          const V pv = prow[blk];
#pragma unroll
          for (int jj = 0; jj < 16; jj++) z[blk][jj] = fma(-m, pv[jj], z[blk][jj]);
        }
      }
      z[cb][cc] = me ? R(-1) : m;
      sig = me ? -inv : sig;
      myc = me ? c - k : myc;
      used = used || me;
      if (lane == 0) sh.pr[c - k] = p;
    }
  }
  __syncthreads();

  // Factors into the slot (T stays as the ID wrote it).
#pragma unroll
  for (int blk = 0; blk < NB; blk++) {
    const int j1 = min(n, 16 * blk + 16);
    for (int j = 16 * blk; j < j1; j++) {
      const R v = z[blk][j & 15];
      int64_t off;
      R out;
      if (j < k) {  // carried block / U
        off = lane < k ? fl.S + lane + (int64_t)k * j : fl.U + myc + (int64_t)r * j;
        out = lane < k ? -v : -sig * v;
      } else {  // L / X_rr^{-1}
        const int c2 = sh.pr[j - k] - k;
        off = lane < k ? fl.L + lane + (int64_t)k * c2 : fl.X + myc + (int64_t)r * c2;
        out = lane < k ? v : sig * v;
      }
      if (live) F[off] = out;
    }
  }
  if (lane == 0) {
    if (sing) atomicOr(s.err, kErrSingular);
    s.upd_kind[b] = 1;
  }
}

}  // namespace

template <typename R>
void launch_decouple_sweep(const DevGeom& g, const DevTree& t, const DevState& s, int start,
                           int count, hipStream_t st) {
  if (count <= 0) return;
  // Two register classes: |active| <= 32 (two 16-entry blocks per row) and
  // 32 < |active| <= 64 (four blocks).
  if (g.pde == kPdeStokes) {
    k_decouple_sweep<R, kPdeStokes, 2><<<count, 64, 0, st>>>(g, t, s, start, 0);
    k_decouple_sweep<R, kPdeStokes, 4><<<count, 64, 0, st>>>(g, t, s, start, 32);
  } else if (g.pde == kPdeLaplaceNeumann) {
    k_decouple_sweep<R, kPdeLaplaceNeumann, 2><<<count, 64, 0, st>>>(g, t, s, start, 0);
    k_decouple_sweep<R, kPdeLaplaceNeumann, 4><<<count, 64, 0, st>>>(g, t, s, start, 32);
  } else {
    k_decouple_sweep<R, kPdeLaplace, 2><<<count, 64, 0, st>>>(g, t, s, start, 0);
    k_decouple_sweep<R, kPdeLaplace, 4><<<count, 64, 0, st>>>(g, t, s, start, 32);
  }
}

template void launch_decouple_sweep<double>(const DevGeom&, const DevTree&, const DevState&, int,
                                            int, hipStream_t);
template void launch_decouple_sweep<float>(const DevGeom&, const DevTree&, const DevState&, int,
                                           int, hipStream_t);

}  // namespace iegpu
