// RS factorization and solve kernels for gfx950 (CDNA4, wave64).
//
// One workgroup per box of a tree level; levels run bottom-up (factor,
// upward sweep) or top-down (downward sweep) as stream-ordered launches,
// so the whole factorization is a fixed launch sequence with no host
// synchronisation (capturable in a HIP graph).
//
// Reference stages (paths relative to the reference root):
//   k_active    quadtree.cpp:1028-1074 (active = children's skeletons)
//   k_rows      kernel.cpp:293-423     (ID row count: near field + proxy)
//   k_assemble  kernel.cpp:123-591     (ID matrix assembly)
//   k_qrcp      ie_mat.cpp:548-588     (dgeqp3 rank cut + dtrtrs)
//   k_decouple  skel_factorization.cpp:77-128, 323-381
//   k_root_*    skel_factorization.cpp:172-186
//   k_sweep_*   skel_factorization.cpp:388-426, 507-592

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "rs_device.hpp"
#include "rs_kernels.hpp"

namespace iegpu {

namespace {

template <typename R>
struct Eps;
template <>
struct Eps<double> {
  // LAPACK dlamch('Epsilon') = 2^-53; dlaqp2 recomputes a column norm when
  // the downdated estimate falls below sqrt of it.
  static constexpr double tol3z = 1.0536712127723509e-08;
};
template <>
struct Eps<float> {
  static constexpr float tol3z = 2.44140625e-04f;
};

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// ---------------------------------------------------------------------------
// Level bookkeeping
// ---------------------------------------------------------------------------

__global__ void k_reset(DevState s) {
  if (threadIdx.x == 0) {
    s.err[0] = 0;
    s.err[1] = 0;
    *s.fac_top = 0;
  }
  if (threadIdx.x < kMaxLevels) s.njobs[threadIdx.x] = 0;
}

// active(box) = concatenation of the children's outgoing lists (skeleton if
// compressed, active otherwise) in child order bl, tl, tr, br; a leaf's
// active list is its original DOF list.
__global__ __launch_bounds__(64) void k_active(DevTree t, DevState s,
                                               int start) {
  const int b = start + blockIdx.x;
  const int lane = threadIdx.x;
  int* dst = s.act_list + (int64_t)b * s.ncap;
  int n = 0;
  if (t.leaf[b]) {
    n = t.orig_cnt[b];
    if (n > s.ncap) {
      if (lane == 0) atomicOr(s.err, kErrBoxCap);
      n = 0;
    }
    const int* src = t.orig_dofs + t.orig_off[b];
    for (int i = lane; i < n; i += 64) dst[i] = src[i];
  } else {
    int pos = 0;
    for (int q = 0; q < 4; q++) {
      const int c = t.child[4 * b + q];
      const int cnt = s.out_cnt[c];
      if (lane == 0) s.chpos[4 * b + q] = pos;
      if (pos + cnt <= s.ncap) {
        const int* src = s.out_list + (int64_t)c * s.ncap;
        for (int i = lane; i < cnt; i += 64) dst[pos + i] = src[i];
      }
      pos += cnt;
    }
    n = pos;
    if (n > s.ncap) {
      if (lane == 0) atomicOr(s.err, kErrBoxCap);
      n = 0;
    }
  }
  if (lane == 0) {
    s.act_cnt[b] = n;
    atomicMax(s.err + 1, n);
  }
}

// Visits the ID row DOFs of box b in reference order (source boxes in
// order, each list in order, filtered by the proxy circle at level >= 3).
// Called by one full wave; `emit(pos, dof)` receives the compacted index.
template <typename F>
__device__ int for_each_row_dof(const DevGeom& g, const DevTree& t,
                                const DevState& s, int b, bool filter,
                                F emit) {
  const int lane = threadIdx.x & 63;
  const double cx = t.cx[b], cy = t.cy[b], side = t.side[b];
  int cnt = 0;
  const int so = t.src_off[b], sc = t.src_cnt[b];
  for (int u = 0; u < sc; u++) {
    const DofList L = source_dofs(t, s, t.src[so + u]);
    for (int base = 0; base < L.n; base += 64) {
      const int i = base + lane;
      int dof = -1;
      bool keep = false;
      if (i < L.n) {
        dof = L.p[i];
        keep = !filter || in_proxy_circle(g, dof, cx, cy, side);
      }
      const uint64_t mask = __ballot(keep);
      const uint64_t below = (lane == 0) ? 0ull : (mask & ((~0ull) >> (64 - lane)));
      if (keep) emit(cnt + __popcll(below), dof);
      cnt += __popcll(mask);
    }
  }
  return cnt;
}

__device__ inline int proxy_rows(const DevGeom& g, int level) {
  if (level < 3) return 0;
  return (g.pde == kPdeStokes) ? 4 * kNumProxy : 2 * kNumProxy;
}

__global__ __launch_bounds__(64) void k_rows(DevGeom g, DevTree t, DevState s,
                                             int start, int level) {
  const int b = start + blockIdx.x;
  const int n = s.act_cnt[b];
  int m = 0;
  if (n >= kMinCompress) {
    const int cnt =
        for_each_row_dof(g, t, s, b, level >= 3, [](int, int) {});
    m = 2 * cnt + proxy_rows(g, level);
  }
  int kind = kQrcpNone;
  if (m > 0) {
    const int nch = id_chunks(m);
    if (n <= kRegMaxCols && m <= kRegMaxRows) kind = kQrcpReg;
    else if (n <= kRegMaxCols && nch <= kMaxChunks && nch * n <= kRegMaxRows) kind = kQrcpStack;
    else kind = kQrcpLds;
  }
  if (threadIdx.x == 0) {
    s.m_rows[b] = m;
    s.qkind[b] = kind;
    const int64_t ld = m > 2 * n ? m : 2 * n;
    int64_t stack = 0;
    if (kind == kQrcpStack) {
      // Row-compression jobs, one per chunk, and room for the stacked
      // triangles after the ID matrix.
      const int nch = id_chunks(m);
      stack = (int64_t)nch * n * n;
      const int base = atomicAdd(s.njobs + level, nch);
      int* J = s.jobs + (int64_t)2 * kMaxChunks * start + 2 * base;
      for (int q = 0; q < nch; q++) {
        J[2 * q] = b;
        J[2 * q + 1] = q;
      }
    }
    s.id_sz[b] = (m > 0) ? ld * n + stack : 0;
    s.row_sz[b] = (m > 0) ? (m - proxy_rows(g, level)) / 2 : 0;
    s.fac_sz[b] = (n > 0) ? fac_slot_size(n) : 0;
  }
}

// Exclusive scans of the per-box sizes of one level (single workgroup).
// ID scratch and row lists are level-local; factor slots append to the
// running pool offset.
__global__ __launch_bounds__(1024) void k_scan(DevState s, int start,
                                               int count) {
  __shared__ int64_t sh[3][1024];
  const int tid = threadIdx.x;
  const int chunk = (count + 1023) / 1024;
  const int lo = min(count, tid * chunk), hi = min(count, lo + chunk);
  int64_t a = 0, b = 0, c = 0;
  for (int i = lo; i < hi; i++) {
    a += s.id_sz[start + i];
    b += s.row_sz[start + i];
    c += s.fac_sz[start + i];
  }
  sh[0][tid] = a;
  sh[1][tid] = b;
  sh[2][tid] = c;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    int64_t x0 = 0, x1 = 0, x2 = 0;
    if (tid >= o) {
      x0 = sh[0][tid - o];
      x1 = sh[1][tid - o];
      x2 = sh[2][tid - o];
    }
    __syncthreads();
    sh[0][tid] += x0;
    sh[1][tid] += x1;
    sh[2][tid] += x2;
    __syncthreads();
  }
  const int64_t fbase = *s.fac_top;
  int64_t oa = sh[0][tid] - a, ob = sh[1][tid] - b, oc = sh[2][tid] - c;
  for (int i = lo; i < hi; i++) {
    s.id_off[start + i] = oa;
    s.row_off[start + i] = ob;
    s.fac_off[start + i] = fbase + oc;
    oa += s.id_sz[start + i];
    ob += s.row_sz[start + i];
    oc += s.fac_sz[start + i];
  }
  __syncthreads();
  if (tid == 0) {
    const int64_t ta = sh[0][1023], tb = sh[1][1023], tc = sh[2][1023];
    int e = 0;
    if (ta > s.id_cap) e |= kErrIdPool;
    if (tb > s.row_cap) e |= kErrRowPool;
    if (fbase + tc > s.fac_cap) e |= kErrFacPool;
    if (e) atomicOr(s.err, e);
    *s.fac_top = fbase + tc;
  }
}

__device__ inline bool pools_ok(const DevState& s) {
  return (s.err[0] & (kErrIdPool | kErrRowPool | kErrFacPool | kErrBoxCap)) == 0;
}

// ---------------------------------------------------------------------------
// ID matrix assembly: [K(ic, box); K(box, ic)^T; proxy] (m x n,
// column-major, ld m), kernel.cpp:293-423.  Consecutive threads write
// consecutive rows of a column.
// ---------------------------------------------------------------------------

template <typename R>
__global__ __launch_bounds__(256) void k_assemble(DevGeom g, DevTree t,
                                                  DevState s, int start,
                                                  int level) {
  const int b = start + blockIdx.x;
  const int m = s.m_rows[b];
  if (m == 0 || !pools_ok(s)) return;
  const int n = s.act_cnt[b];
  const int tid = threadIdx.x;
  const int* act = s.act_list + (int64_t)b * s.ncap;
  int* rows = s.row_pool + s.row_off[b];
  R* A = reinterpret_cast<R*>(s.id_pool) + s.id_off[b];
  const int P = proxy_rows(g, level);
  const int cnt = (m - P) / 2;
  if (tid < 64)
    for_each_row_dof(g, t, s, b, level >= 3,
                     [&](int pos, int dof) { rows[pos] = dof; });
  __shared__ double pxy[4][kNumProxy];
  const double cx = t.cx[b], cy = t.cy[b];
  const double rad = t.side[b] * kRadiusRatio;
  if (P > 0 && tid < kNumProxy) {
    double px, py, n1, n2;
    proxy_point(cx, cy, rad, tid, &px, &py, &n1, &n2);
    pxy[0][tid] = px;
    pxy[1][tid] = py;
    pxy[2][tid] = n1;
    pxy[3][tid] = n2;
  }
  __syncthreads();
  const double pw = 2.0 * M_PI * rad / kNumProxy;
  for (int j = 0; j < n; j++) {
    const int sd = act[j];
    R* col = A + (int64_t)m * j;
    for (int i = tid; i < m; i += 256) {
      double v;
      if (i < cnt) {
        v = kernel_entry(g, rows[i], sd);
      } else if (i < 2 * cnt) {
        v = kernel_entry(g, sd, rows[i - cnt]);
      } else {
        const int row = i - 2 * cnt;
        if (g.pde != kPdeStokes) {
          const double scale = 1.0 / (2 * M_PI);
          const double bx = g.pts[2 * sd], by = g.pts[2 * sd + 1];
          if (row < kNumProxy) {
            const double r0 = pxy[0][row] - bx, r1 = pxy[1][row] - by;
            v = -g.wts[sd] * scale * (r0 * g.nrm[2 * sd] + r1 * g.nrm[2 * sd + 1]) /
                (r0 * r0 + r1 * r1);
          } else {
            const int p = row - kNumProxy;
            const double r0 = bx - pxy[0][p], r1 = by - pxy[1][p];
            v = -pw * scale * (r0 * pxy[2][p] + r1 * pxy[3][p]) / (r0 * r0 + r1 * r1);
          }
        } else {
          const double scale = 1.0 / M_PI;
          const int PP = 2 * kNumProxy;
          const bool to_proxy = row < PP;
          const int rr = to_proxy ? row : row - PP;
          const int p = rr >> 1, c = rr & 1;
          const double px = pxy[0][p], py = pxy[1][p], n1 = pxy[2][p], n2 = pxy[3][p];
          const int bp = sd >> 1, bc = sd & 1;
          const double bx = g.pts[2 * bp], by = g.pts[2 * bp + 1];
          const double bn1 = g.nrm[2 * bp], bn2 = g.nrm[2 * bp + 1];
          if (to_proxy) {
            const double sw = g.wts[bp];
            const double r0 = px - bx, r1 = py - by;
            const double d2 = r0 * r0 + r1 * r1;
            const double pot = sw * scale * (r0 * bn1 + r1 * bn2) / (d2 * d2);
            v = pot * (c ? r1 : r0) * (bc ? r1 : r0) + sw * (c ? n2 : n1) * (bc ? bn2 : bn1);
          } else {
            const double r0 = bx - px, r1 = by - py;
            const double d2 = r0 * r0 + r1 * r1;
            const double pot = pw * scale * (r0 * n1 + r1 * n2) / (d2 * d2);
            v = pot * (bc ? r1 : r0) * (c ? r1 : r0) + pw * (bc ? bn2 : bn1) * (c ? n2 : n1);
          }
        }
      }
      col[i] = static_cast<R>(v);
    }
  }
}

// ---------------------------------------------------------------------------
// Register-resident QR kernels (1024 threads, 16 waves).  Wave w holds
// columns 4w .. 4w+3 of an m x n matrix (m <= 64 RPL, n <= 64) in the
// four-column interleave of rs_device.hpp: row i on lane i % 64, register
// i / 64, column 4w + (sl ^ quad_class(lane)) in slot sl.  A Householder
// step: the wave owning the pivot column forms the reflector and publishes
// it through LDS; every wave then updates its columns in registers with
// one four-column reduction.  Waves past the last column only take part
// in the barriers.
// ---------------------------------------------------------------------------

constexpr int kXThreads = 1024;
constexpr int kXRows = kRegMaxRows / 64;  // registers per lane per column

// Entries of the column in slot psl (per lane) of the interleave.
template <typename R, int RPL>
__device__ inline void pick_slot(const R (&a)[4][RPL], int psl, R (&x)[RPL]) {
#pragma unroll
  for (int r = 0; r < RPL; r++) x[r] = a[0][r];
#pragma unroll
  for (int sl = 1; sl < 4; sl++) {
    const bool me = psl == sl;
#pragma unroll
    for (int r = 0; r < RPL; r++) x[r] = me ? a[sl][r] : x[r];
  }
}

// Householder reflector of column x at step j (j < 64), LAPACK dlarfg:
// beta = -sign(alpha) ||x(j:)||, tau = (beta - alpha) / beta,
// v = x(j+1:) / (alpha - beta), v_j = 1, v(0:j-1) = 0.  Writes v to vb
// (all 64 RPL rows; rows past the matrix hold zeros in x).
template <typename R, int RPL>
__device__ inline void householder(const R (&x)[RPL], int j, R* vb, R* beta_out,
                                   R* tau_out) {
  const int lane = threadIdx.x & 63;
  const R alpha = lane_value(x[0], j);
  R ss = lane > j ? x[0] * x[0] : R(0);
#pragma unroll
  for (int r = 1; r < RPL; r++) ss += x[r] * x[r];
  ss = wave_total(ss);
  const R xnorm = sqrt(ss);
  R beta, tau, sc;
  if (xnorm == 0) {
    tau = 0;
    beta = alpha;
    sc = 0;
  } else {
    beta = -copysign(hypot(alpha, xnorm), alpha);
    tau = (beta - alpha) / beta;
    sc = 1 / (alpha - beta);
  }
  vb[lane] = lane > j ? x[0] * sc : (lane == j ? R(1) : R(0));
#pragma unroll
  for (int r = 1; r < RPL; r++) vb[lane + 64 * r] = x[r] * sc;
  *beta_out = beta;
  *tau_out = tau;
}

// a <- (I - tau v v^T) a on this wave's columns; `live` says whether the
// class quad_class(lane) is still to be updated (pivoted columns keep
// their R entries).
template <typename R, int RPL>
__device__ inline void apply_reflector(R (&a)[4][RPL], const R* vb, R tau, bool live) {
  const int lane = threadIdx.x & 63;
  R vr[RPL];
#pragma unroll
  for (int r = 0; r < RPL; r++) vr[r] = vb[lane + 64 * r];
  R p[4];
#pragma unroll
  for (int sl = 0; sl < 4; sl++) {
    p[sl] = 0;
#pragma unroll
    for (int r = 0; r < RPL; r++) p[sl] += vr[r] * a[sl][r];
  }
  R t = wave_total4(p);
  t = live ? tau * t : R(0);
  R tw[4];
  spread4(t, tw);
#pragma unroll
  for (int sl = 0; sl < 4; sl++)
#pragma unroll
    for (int r = 0; r < RPL; r++) a[sl][r] -= tw[sl] * vr[r];
}

template <typename R, int RPL>
__device__ inline void load_columns(const R* A, int64_t ld, int row0, int rows, int n,
                                    R (&a)[4][RPL]) {
  const int lane = threadIdx.x & 63, c0 = 4 * (threadIdx.x >> 6);
  const int hq = quad_class(lane);
#pragma unroll
  for (int sl = 0; sl < 4; sl++) {
    const int c = c0 + (sl ^ hq);
#pragma unroll
    for (int r = 0; r < RPL; r++) {
      const int i = lane + 64 * r;
      a[sl][r] = (c < n && i < rows) ? A[row0 + i + ld * c] : R(0);
    }
  }
}

// Row compression of a tall ID matrix: chunk q of box b (rows
// [q ch, q ch + rows), ch = ceil(m / chunks)) is reduced by unpivoted
// Householder QR to its n x n triangle R_q, written to rows [q n, q n + n)
// of the box's stack (ld chunks * n, after the ID matrix in its scratch
// slot).  The stack is Q^T A for an orthogonal Q, so its pivoted QR has
// the same R and column order as that of A.  Grid-strided over the
// level's jobs; one barrier per step (double-buffered reflector).
template <typename R, int RPL>
__global__ __launch_bounds__(kXThreads) void k_rowcomp(DevState s, int start, int level) {
  constexpr int MAXM = 64 * RPL;
  __shared__ R vbuf[2][MAXM];
  __shared__ R sh_tau[2];
  if (!pools_ok(s)) return;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int hq = quad_class(lane);
  const int c0 = 4 * wv;
  const int* jobs = s.jobs + (int64_t)2 * kMaxChunks * start;
  const int njobs = s.njobs[level];
  for (int job = blockIdx.x; job < njobs; job += gridDim.x) {
    const int b = jobs[2 * job], q = jobs[2 * job + 1];
    const int n = s.act_cnt[b], m = s.m_rows[b];
    const int nch = id_chunks(m);
    const int ch = (m + nch - 1) / nch;
    const int r0 = q * ch, rows = min(ch, m - r0);
    R* A = reinterpret_cast<R*>(s.id_pool) + s.id_off[b];
    R* S = A + (int64_t)max(m, 2 * n) * n;
    const int64_t lds = (int64_t)nch * n;
    const bool active = c0 < n;
    R a[4][RPL];
    load_columns(A, m, r0, rows, n, a);
    const int steps = min(rows, n);
    int dmask = 0;
    for (int j = 0; j < steps; j++) {
      if (wv == (j >> 2)) {
        const int psl = (j & 3) ^ hq;
        R x[RPL];
        pick_slot(a, psl, x);
        R beta, tau;
        householder(x, j, vbuf[j & 1], &beta, &tau);
        dmask |= 1 << (j & 3);
        if (lane == j) {
#pragma unroll
          for (int sl = 0; sl < 4; sl++) a[sl][0] = (sl == psl) ? beta : a[sl][0];
        }
        if (lane == 0) sh_tau[j & 1] = tau;
      }
      __syncthreads();
      if (active) apply_reflector(a, vbuf[j & 1], sh_tau[j & 1], ((dmask >> hq) & 1) == 0);
    }
    // Column c was reduced at step c: rows below its diagonal are zero.
    if (active) {
#pragma unroll
      for (int sl = 0; sl < 4; sl++) {
        const int c = c0 + (sl ^ hq);
        if (c < n && lane < n)
          S[(int64_t)q * n + lane + lds * c] = lane > c ? R(0) : a[sl][0];
      }
    }
    __syncthreads();
  }
}

// Interpolative decomposition by column-pivoted Householder QR, stopped at
// the first |R_jj| < tol |R_00| (ie_mat.cpp:548-588).  The pivot rule,
// Householder form and column-norm downdating follow LAPACK dgeqp3/dlaqp2.
// Lanes 0..3 of each wave keep the norms, the logical position and the
// pivot key of the wave's four columns; the pivot is the largest key
// (pivot_key), found by one wave maximum.
template <typename R, int RPL>
__global__ __launch_bounds__(kXThreads) void k_qrcp_id(DevState s, int start, double tol_d) {
  constexpr int MAXM = 64 * RPL;
  __shared__ R vbuf[MAXM];
  __shared__ R rowj[kRegMaxCols];
  __shared__ double keys[kRegMaxCols];
  __shared__ R rbuf[kRegMaxCols * kRegMaxCols];
  __shared__ int colmap[kRegMaxCols];
  __shared__ R sh_tau, sh_thresh;
  __shared__ int sh_stop;

  const int b = start + blockIdx.x;
  const int kind = s.qkind[b];
  if ((kind != kQrcpReg && kind != kQrcpStack) || !pools_ok(s)) return;
  const int n = s.act_cnt[b];
  const R* A = reinterpret_cast<const R*>(s.id_pool) + s.id_off[b];
  int m = s.m_rows[b];
  if (kind == kQrcpStack) {
    A += (int64_t)max(m, 2 * n) * n;
    m = id_chunks(m) * n;
  }
  const R tol = static_cast<R>(tol_d);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int hq = quad_class(lane);
  const int c0 = 4 * wv;
  const bool active = c0 < n;

  R a[4][RPL];
  load_columns(A, m, 0, m, n, a);
  // Lanes 0..3: column c0 + lane.
  const int myc = c0 + (lane & 3);
  R vn1 = 0, vn2 = 0;
  int pos = myc;
  bool avail = lane < 4 && myc < n;
  if (active) {
    R p[4];
#pragma unroll
    for (int sl = 0; sl < 4; sl++) {
      p[sl] = 0;
#pragma unroll
      for (int r = 0; r < RPL; r++) p[sl] += a[sl][r] * a[sl][r];
    }
    vn1 = sqrt(class_by_lane(wave_total4(p)));
    vn2 = vn1;
  }
  if (lane < 4) keys[myc] = avail ? pivot_key(vn1, pos) : -1.0;
  __syncthreads();

  const int kmax = min(m, n);
  int k = kmax;
  R thresh = 0;
  int dmask = 0;  // classes of this wave already pivoted
  for (int j = 0; j < kmax; j++) {
    const double kv = keys[lane];
    const double best = wave_max(kv);
    if (!(best >= 0)) {  // every wave sees the same keys
      k = j;
      break;
    }
    const int p = __ffsll((unsigned long long)__ballot(kv == best)) - 1;
    const int q = pivot_key_pos(best);
    if (lane < 4) {
      if (myc == p) {
        pos = j;
        avail = false;
      } else if (pos == j) {
        pos = q;
      }
    }
    if (wv == (p >> 2)) {
      const int psl = (p & 3) ^ hq;
      R x[RPL];
      pick_slot(a, psl, x);
      R beta, tau;
      householder(x, j, vbuf, &beta, &tau);
      const R th = (j == 0) ? fabs(tol * beta) : thresh;
      const int stop = (j >= 1 && fabs(beta) < th) ? 1 : 0;
      if (!stop) {
        dmask |= 1 << (p & 3);
        if (lane == j) {
#pragma unroll
          for (int sl = 0; sl < 4; sl++) a[sl][0] = (sl == psl) ? beta : a[sl][0];
        }
      }
      if (lane == 0) {
        sh_tau = tau;
        sh_stop = stop;
        if (j == 0) sh_thresh = th;
      }
    }
    __syncthreads();
    if (sh_stop) {
      k = j;
      break;
    }
    if (j == 0) thresh = sh_thresh;
    if (active) {
      apply_reflector(a, vbuf, sh_tau, ((dmask >> hq) & 1) == 0);
      if (lane == j) {
#pragma unroll
        for (int sl = 0; sl < 4; sl++) rowj[c0 + (sl ^ hq)] = a[sl][0];
      }
      wave_sync();
      bool need = false;
      if (avail && vn1 != 0) {
        R temp = fabs(rowj[myc]) / vn1;
        temp = 1 - temp * temp;
        temp = temp > 0 ? temp : R(0);
        const R rat = vn1 / vn2;
        if (temp * rat * rat <= Eps<R>::tol3z) need = true;
        else vn1 = vn1 * sqrt(temp);
      }
      if (__ballot(need)) {
        R p2[4];
#pragma unroll
        for (int sl = 0; sl < 4; sl++) {
          p2[sl] = lane > j ? a[sl][0] * a[sl][0] : R(0);
#pragma unroll
          for (int r = 1; r < RPL; r++) p2[sl] += a[sl][r] * a[sl][r];
        }
        const R t = class_by_lane(wave_total4(p2));
        if (need) {
          const R nv = (j < m - 1) ? sqrt(t) : R(0);
          vn1 = nv;
          vn2 = nv;
        }
      }
      if (lane < 4) keys[myc] = avail ? pivot_key(vn1, pos) : -1.0;
    }
    __syncthreads();
  }

  // Logical order and rows 0..k-1 of R (physical column order, ld k).
  if (lane < 4 && myc < n) colmap[pos] = myc;
  if (active) {
#pragma unroll
    for (int sl = 0; sl < 4; sl++) {
      const int c = c0 + (sl ^ hq);
      if (c < n && lane < k) rbuf[lane + k * c] = a[sl][0];
    }
  }
  __syncthreads();

  const int* act = s.act_list + (int64_t)b * s.ncap;
  int* out = s.out_list + (int64_t)b * s.ncap;
  if (k == 0) {
    for (int i = tid; i < n; i += kXThreads) out[i] = act[i];
    if (tid == 0) {
      s.rank[b] = 0;
      s.out_cnt[b] = n;
    }
    return;
  }
  // T = R11^{-1} R12: one wave per column, lane l holds t_l (k <= 64).
  const int r = n - k;
  R* F = reinterpret_cast<R*>(s.fac_pool) + s.fac_off[b];
  for (int cp = wv; cp < r; cp += kXThreads / 64) {
    const int pcol = colmap[k + cp];
    R t = (lane < k) ? rbuf[lane + k * pcol] : R(0);
    for (int i = k - 1; i >= 0; i--) {
      const int ci = colmap[i];
      const R ti = lane_value(t, i) / rbuf[i + k * ci];
      if (lane == i) t = ti;
      if (lane < i) t -= rbuf[lane + k * ci] * ti;
    }
    if (lane < k) F[lane + (int64_t)k * cp] = t;
  }
  int* red = s.red_list + (int64_t)b * s.ncap;
  int* pm = s.perm + (int64_t)b * s.ncap;
  for (int i = tid; i < n; i += kXThreads) {
    const int c = colmap[i];
    pm[i] = c;
    if (i < k) out[i] = act[c];
    else red[i - k] = act[c];
  }
  if (tid == 0) {
    s.rank[b] = k;
    s.out_cnt[b] = k;
  }
}

// General form (any box the register form does not take): the matrix lives
// in LDS when it fits, else in the box's scratch slot; each wave owns a
// strided set of columns.
template <typename R, int NT>
__global__ __launch_bounds__(NT) void k_qrcp(DevState s, int start, double tol_d) {
  constexpr int NW = NT / 64;
  constexpr int kLdsElems = 131072 / sizeof(R);
  __shared__ R smat[kLdsElems];
  __shared__ R vn1[kMaxBoxDofs];
  __shared__ R vn2[kMaxBoxDofs];
  __shared__ int col[kMaxBoxDofs];
  __shared__ R sh_tau;
  __shared__ int sh_stop;

  const int b = start + blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int n = s.act_cnt[b], m = s.m_rows[b];
  const int kind = s.qkind[b];
  if (pools_ok(s) && (kind == kQrcpReg || kind == kQrcpStack)) return;  // k_qrcp_id's box
  const int* act = s.act_list + (int64_t)b * s.ncap;
  int* out = s.out_list + (int64_t)b * s.ncap;
  if (m == 0 || n < kMinCompress || !pools_ok(s)) {
    for (int i = tid; i < n; i += NT) out[i] = act[i];
    if (tid == 0) {
      s.rank[b] = 0;
      s.out_cnt[b] = n;
    }
    return;
  }
  const R tol = static_cast<R>(tol_d);
  R* A = reinterpret_cast<R*>(s.id_pool) + s.id_off[b];
  const int64_t mn = (int64_t)m * n;
  if (mn <= kLdsElems) {
    for (int64_t i = tid; i < mn; i += NT) smat[i] = A[i];
    A = smat;
  }
  for (int c = tid; c < n; c += NT) col[c] = c;
  __syncthreads();
  for (int c = wv; c < n; c += NW) {
    const R* a = A + (int64_t)m * c;
    R ss = 0;
    for (int i = lane; i < m; i += 64) ss += a[i] * a[i];
    ss = wave_sum(ss);
    if (lane == 0) {
      vn1[c] = sqrt(ss);
      vn2[c] = vn1[c];
    }
  }
  __syncthreads();

  const int kmax = min(m, n);
  int k = 0;
  R thresh = 0;
  for (int j = 0; j < kmax; j++) {
    if (wv == 0) {
      R best = -1;
      int bi = n;
      for (int q = j + lane; q < n; q += 64) {
        const R v = vn1[q];
        if (v > best) {
          best = v;
          bi = q;
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const R ob = __shfl_xor(best, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (bi >= n) bi = j;  // all-NaN column norms: keep the current column
      const int cj = col[j], cb = col[bi];
      const R v1 = vn1[j], v2 = vn2[j];
      wave_sync();
      if (lane == 0 && bi != j) {
        col[j] = cb;
        col[bi] = cj;
        vn1[bi] = v1;
        vn2[bi] = v2;
      }
      const int pc = (bi != j) ? cb : cj;
      R* v = A + (int64_t)m * pc;
      const R alpha = v[j];
      R ss = 0;
      for (int i = j + 1 + lane; i < m; i += 64) ss += v[i] * v[i];
      ss = wave_sum(ss);
      const R xnorm = sqrt(ss);
      R beta, tau;
      if (xnorm == 0) {
        tau = 0;
        beta = alpha;
      } else {
        beta = -copysign(hypot(alpha, xnorm), alpha);
        tau = (beta - alpha) / beta;
        const R sc = 1 / (alpha - beta);
        for (int i = j + 1 + lane; i < m; i += 64) v[i] *= sc;
      }
      if (j == 0) thresh = fabs(tol * beta);
      const int stop = (j >= 1 && fabs(beta) < thresh) ? 1 : 0;
      if (lane == 0) {
        if (!stop) v[j] = beta;
        sh_tau = tau;
        sh_stop = stop;
      }
    }
    __syncthreads();
    if (sh_stop) {
      k = j;
      break;
    }
    const R tau = sh_tau;
    const R* v = A + (int64_t)m * col[j];
    for (int q = j + 1 + wv; q < n; q += NW) {
      R* cq = A + (int64_t)m * col[q];
      R part = 0;
      for (int i = j + 1 + lane; i < m; i += 64) part += v[i] * cq[i];
      part = wave_sum(part);
      const R w = cq[j] + part;
      const R tw = tau * w;
      for (int i = j + 1 + lane; i < m; i += 64) cq[i] -= tw * v[i];
      const R rjc = cq[j] - tw;
      wave_sync();
      if (lane == 0) cq[j] = rjc;
      const R v1 = vn1[q];
      if (v1 != 0) {
        R temp = fabs(rjc) / v1;
        temp = 1 - temp * temp;
        temp = temp > 0 ? temp : R(0);
        const R rat = v1 / vn2[q];
        const R temp2 = temp * rat * rat;
        if (temp2 <= Eps<R>::tol3z) {
          R ss2 = 0;
          for (int i = j + 1 + lane; i < m; i += 64) ss2 += cq[i] * cq[i];
          ss2 = wave_sum(ss2);
          const R nv = (j < m - 1) ? sqrt(ss2) : R(0);
          if (lane == 0) {
            vn1[q] = nv;
            vn2[q] = nv;
          }
        } else if (lane == 0) {
          vn1[q] = v1 * sqrt(temp);
        }
      }
    }
    __syncthreads();
  }
  if (k == 0) {
    for (int i = tid; i < n; i += NT) out[i] = act[i];
    if (tid == 0) {
      s.rank[b] = 0;
      s.out_cnt[b] = n;
    }
    return;
  }
  // T = R11^{-1} R12 (k x r), one thread per column, back substitution.
  const int r = n - k;
  R* F = reinterpret_cast<R*>(s.fac_pool) + s.fac_off[b];
  for (int c = tid; c < r; c += NT) {
    const R* y = A + (int64_t)m * col[k + c];
    R* tc = F + (int64_t)k * c;
    for (int i = k - 1; i >= 0; i--) {
      R acc = y[i];
      for (int l = i + 1; l < k; l++) acc -= A[i + (int64_t)m * col[l]] * tc[l];
      tc[i] = acc / A[i + (int64_t)m * col[i]];
    }
  }
  int* red = s.red_list + (int64_t)b * s.ncap;
  int* pm = s.perm + (int64_t)b * s.ncap;
  for (int i = tid; i < n; i += NT) {
    const int c = col[i];
    pm[i] = c;
    if (i < k) out[i] = act[c];
    else red[i - k] = act[c];
  }
  if (tid == 0) {
    s.rank[b] = k;
    s.out_cnt[b] = k;
  }
}

// ---------------------------------------------------------------------------
// Block LU with partial pivoting (LAPACK dgetrf pivot rule: first index of
// the largest |entry|), in place, column-major with leading dimension ld.
// ---------------------------------------------------------------------------

template <typename R, int NT>
__device__ int block_lu(R* X, int n, int ld, int* ipiv) {
  __shared__ int sh_p;
  const int tid = threadIdx.x, lane = tid & 63;
  int info = 0;
  for (int j = 0; j < n; j++) {
    if (tid < 64) {
      R best = -1;
      int bi = n;
      for (int i = j + lane; i < n; i += 64) {
        const R v = fabs(X[i + (int64_t)ld * j]);
        if (v > best) {
          best = v;
          bi = i;
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const R ob = __shfl_xor(best, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (bi >= n) bi = j;  // NaN column: no interchange
      if (lane == 0) {
        sh_p = bi;
        ipiv[j] = bi;
      }
    }
    __syncthreads();
    const int p = sh_p;
    if (p != j) {
      for (int c = tid; c < n; c += NT) {
        const R a = X[j + (int64_t)ld * c];
        X[j + (int64_t)ld * c] = X[p + (int64_t)ld * c];
        X[p + (int64_t)ld * c] = a;
      }
    }
    __syncthreads();
    const R piv = X[j + (int64_t)ld * j];
    if (piv != 0) {
      const R inv = 1 / piv;
      for (int i = j + 1 + tid; i < n; i += NT) X[i + (int64_t)ld * j] *= inv;
    } else if (info == 0) {
      info = j + 1;
    }
    __syncthreads();
    const int len = n - j - 1;
    const int64_t tot = (int64_t)len * len;
    for (int64_t idx = tid; idx < tot; idx += NT) {
      const int i = j + 1 + (int)(idx % len);
      const int c = j + 1 + (int)(idx / len);
      X[i + (int64_t)ld * c] -= X[i + (int64_t)ld * j] * X[j + (int64_t)ld * c];
    }
    __syncthreads();
  }
  return info;
}

// Carried Schur block of child q at (local i, local j); 0 outside the
// child's diagonal block.
template <typename R>
struct ChildUpdates {
  // The four child blocks and sizes stay in LDS (indexing a private copy
  // by the runtime child id would place it in scratch memory).
  const R* const* S;
  const int* N;
  const signed char* pch;
  const short* ploc;
  __device__ R operator()(int i, int j) const {
    const int q = pch[i];
    if (q < 0 || q != pch[j] || S[q] == nullptr) return R(0);
    return S[q][ploc[i] + (int64_t)N[q] * ploc[j]];
  }
};

template <typename R>
__device__ const R* carried_block(const DevState& s, int c) {
  if (!s.upd_kind[c]) return nullptr;
  const R* F = reinterpret_cast<const R*>(s.fac_pool) + s.fac_off[c];
  const int k = s.rank[c];
  if (k == 0) return F;
  return F + FacLayout::make(k, s.act_cnt[c] - k).S;
}

constexpr int kSmallBox = 64;  // decouple in LDS with lane-per-row solves

// Children's carried blocks and the position map of a box's active list.
template <typename R>
__device__ void load_child_blocks(const DevTree& t, const DevState& s, int b,
                                  int n, const R** chS, int* chN,
                                  signed char* pch, short* ploc, int NT) {
  const int tid = threadIdx.x;
  const bool leaf = t.leaf[b] != 0;
  if (tid < 4) {
    const int c = leaf ? -1 : t.child[4 * b + tid];
    chS[tid] = (c >= 0) ? carried_block<R>(s, c) : nullptr;
    chN[tid] = (c >= 0) ? s.out_cnt[c] : 0;
  }
  __syncthreads();
  for (int i = tid; i < n; i += NT) {
    int q = -1, loc = 0;
    if (!leaf) {
      for (int u = 0; u < 4; u++) {
        const int p0 = s.chpos[4 * b + u];
        if (i >= p0 && i < p0 + chN[u]) {
          q = u;
          loc = i - p0;
        }
      }
    }
    pch[i] = (signed char)q;
    ploc[i] = (short)loc;
  }
}

// Decouple of a compressed box with |active| <= 64: every operand in LDS;
// LU by one wave, triangular solves with one matrix row (or column) per
// lane and the pivots broadcast by shuffles.  Same operations as
// k_decouple below.
template <typename R>
__global__ __launch_bounds__(256) void k_decouple_small(DevGeom g, DevTree t,
                                                        DevState s, int start) {
  constexpr int NT = 256, NW = 4;
  constexpr int kBuf = 8704;  // >= n^2 + r^2 + 3 k r for n <= 64
  __shared__ R buf[kBuf];
  __shared__ signed char pch[kSmallBox];
  __shared__ short ploc[kSmallBox];
  __shared__ short sperm[kSmallBox];
  __shared__ int spiv[kSmallBox], idxf[kSmallBox], idxb[kSmallBox];
  __shared__ const R* chS[4];
  __shared__ int chN[4];
  __shared__ int sh_info;

  const int b = start + blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int n = s.act_cnt[b], k = s.rank[b];
  if (!pools_ok(s) || n == 0 || n > kSmallBox || k == 0) return;
  load_child_blocks<R>(t, s, b, n, chS, chN, pch, ploc, NT);
  for (int i = tid; i < n; i += NT) sperm[i] = (short)s.perm[(int64_t)b * s.ncap + i];
  __syncthreads();
  ChildUpdates<R> upd;
  upd.S = chS;
  upd.N = chN;
  upd.pch = pch;
  upd.ploc = ploc;
  const int r = n - k;
  const FacLayout fl = FacLayout::make(k, r);
  R* F = reinterpret_cast<R*>(s.fac_pool) + s.fac_off[b];
  const int* act = s.act_list + (int64_t)b * s.ncap;
  R* P = buf;
  R* X = P + n * n;
  R* Krs = X + r * r;
  R* Ksr = Krs + r * k;
  R* T = Ksr + k * r;
  R* Lm = P;  // P is dead once X, K_rs and K_sr are formed
  for (int i = tid; i < k * r; i += NT) T[i] = F[fl.T + i];
  for (int idx = tid; idx < n * n; idx += NT) {
    const int ia = sperm[idx % n], ic = sperm[idx / n];
    P[idx] = static_cast<R>(kernel_entry(g, act[ia], act[ic])) - upd(ia, ic);
  }
  __syncthreads();
  for (int idx = tid; idx < r * k; idx += NT) {
    const int i = idx % r, a = idx / r;
    R acc = P[(k + i) + n * a];
    for (int l = 0; l < k; l++) acc -= T[l + k * i] * P[l + n * a];
    Krs[idx] = acc;
  }
  for (int idx = tid; idx < k * r; idx += NT) {
    const int a = idx % k, j = idx / k;
    R acc = P[a + n * (k + j)];
    for (int l = 0; l < k; l++) acc -= P[a + n * l] * T[l + k * j];
    Ksr[idx] = acc;
  }
  __syncthreads();
  for (int idx = tid; idx < r * r; idx += NT) {
    const int i = idx % r, j = idx / r;
    R acc = P[(k + i) + n * (k + j)];
    for (int l = 0; l < k; l++) acc -= T[l + k * i] * P[l + n * (k + j)];
    for (int l = 0; l < k; l++) acc -= Krs[i + r * l] * T[l + k * j];
    X[idx] = acc;
  }
  __syncthreads();
  if (wv == 0) {
    // LU with partial pivoting, lane i = row i.
    int info = 0;
    for (int j = 0; j < r; j++) {
      const bool live = lane >= j && lane < r;
      const int bi = wave_argmax<R>(live ? fabs(X[lane + r * j]) : R(-1),
                                    live ? lane : 0x7fffffff);
      const int p = (bi >= r) ? j : bi;
      if (lane == 0) spiv[j] = p;
      if (p != j && lane < r) {
        const R tmp = X[j + r * lane];
        X[j + r * lane] = X[p + r * lane];
        X[p + r * lane] = tmp;
      }
      wave_sync();
      const R piv = X[j + r * j];
      if (piv == 0) {
        if (info == 0) info = j + 1;
      } else if (lane > j && lane < r) {
        X[lane + r * j] *= 1 / piv;
      }
      wave_sync();
      if (lane > j && lane < r) {
        const R lij = X[lane + r * j];
        for (int c = j + 1; c < r; c++) X[lane + r * c] -= lij * X[j + r * c];
      }
      wave_sync();
    }
    if (lane == 0) {
      sh_info = info;
      for (int i = 0; i < r; i++) {
        idxf[i] = i;
        idxb[i] = i;
      }
      for (int i = 0; i < r; i++) {
        const int p = spiv[i], q = idxf[i];
        idxf[i] = idxf[p];
        idxf[p] = q;
      }
      for (int i = r - 1; i >= 0; i--) {
        const int p = spiv[i], q = idxb[i];
        idxb[i] = idxb[p];
        idxb[p] = q;
      }
    }
  }
  __syncthreads();
  if (sh_info != 0 && tid == 0) atomicOr(s.err, kErrSingular);
  // U = X^{-1} K_rs: wave per column, lane i = row i.
  for (int a = wv; a < k; a += NW) {
    R u = (lane < r) ? Krs[idxf[lane] + r * a] : R(0);
    for (int l = 0; l < r; l++) {
      const R ul = lane_value(u, l);
      if (lane > l && lane < r) u -= X[lane + r * l] * ul;
    }
    for (int i = r - 1; i >= 0; i--) {
      const R ui = lane_value(u, i) / X[i + r * i];
      if (lane == i) u = ui;
      if (lane < i) u -= X[lane + r * i] * ui;
    }
    if (lane < r) F[fl.U + lane + (int64_t)r * a] = u;
  }
  // L = K_sr X^{-1}: wave per row, lane j = column j (X^T z^T = row^T).
  for (int a = wv; a < k; a += NW) {
    R acc = (lane < r) ? Ksr[a + k * lane] : R(0);
    for (int i = 0; i < r; i++) {
      const R yi = lane_value(acc, i) / X[i + r * i];
      if (lane == i) acc = yi;
      if (lane > i && lane < r) acc -= X[i + r * lane] * yi;
    }
    for (int i = r - 1; i >= 0; i--) {
      const R zi = lane_value(acc, i);
      if (lane < i) acc -= X[i + r * lane] * zi;
    }
    const R z = __shfl(acc, lane < r ? idxb[lane] : 0, 64);
    if (lane < r) {
      Lm[a + k * lane] = z;
      F[fl.L + a + (int64_t)k * lane] = z;
    }
  }
  __syncthreads();
  for (int idx = tid; idx < k * k; idx += NT) {
    const int a = idx % k, c = idx / k;
    R acc = 0;
    for (int j = 0; j < r; j++) acc += Lm[a + k * j] * Krs[j + r * c];
    F[fl.S + idx] = upd(sperm[a], sperm[c]) + acc;
  }
  for (int idx = tid; idx < r * r; idx += NT) F[fl.X + idx] = X[idx];
  for (int i = tid; i < r; i += NT) F[fl.piv + i] = static_cast<R>(spiv[i]);
  if (tid == 0) s.upd_kind[b] = 1;
}

// ---------------------------------------------------------------------------
// Decouple one box (skel_factorization.cpp:77-128): with s/r the skeleton
// and redundant positions and K_BN = K(act, act) - (children's carried
// Schur blocks),
//   K_rs = K(r,s) - T^T K(s,s)
//   X_rr = K(r,r) - T^T K(s,r) - K_rs T          (LU, partial pivoting)
//   K_sr = K(s,r) - K(s,s) T
//   L = K_sr X_rr^{-1},  U = X_rr^{-1} K_rs,  schur = L K_rs
// and the box's carried block = (children's update on s,s) + schur.
// Uncompressed boxes carry their children's blocks unchanged.
// ---------------------------------------------------------------------------

template <typename R, int NT>
__global__ __launch_bounds__(NT) void k_decouple(DevGeom g, DevTree t,
                                                 DevState s, int start) {
  constexpr int kLdsElems = 65536 / sizeof(R);
  __shared__ R sbuf[kLdsElems];
  __shared__ signed char pch[kMaxBoxDofs];
  __shared__ short ploc[kMaxBoxDofs];
  __shared__ short sperm[kMaxBoxDofs];
  __shared__ int spiv[kMaxBoxDofs];
  __shared__ const R* chS[4];
  __shared__ int chN[4];

  const int b = start + blockIdx.x;
  const int tid = threadIdx.x;
  const int n = s.act_cnt[b];
  if (!pools_ok(s)) return;
  if (n == 0) {
    if (tid == 0) s.upd_kind[b] = 0;
    return;
  }
  if (n <= kSmallBox && s.rank[b] > 0) return;  // k_decouple_small's box
  const bool leaf = t.leaf[b] != 0;
  if (tid < 4) {
    const int c = leaf ? -1 : t.child[4 * b + tid];
    chS[tid] = (c >= 0) ? carried_block<R>(s, c) : nullptr;
    chN[tid] = (c >= 0) ? s.out_cnt[c] : 0;
  }
  __syncthreads();
  for (int i = tid; i < n; i += NT) {
    int q = -1, loc = 0;
    if (!leaf) {
      for (int u = 0; u < 4; u++) {
        const int p0 = s.chpos[4 * b + u];
        if (i >= p0 && i < p0 + chN[u]) {
          q = u;
          loc = i - p0;
        }
      }
    }
    pch[i] = (signed char)q;
    ploc[i] = (short)loc;
  }
  __syncthreads();
  ChildUpdates<R> upd;
  upd.S = chS;
  upd.N = chN;
  bool any = false;
  for (int u = 0; u < 4; u++) any = any || chS[u] != nullptr;
  upd.pch = pch;
  upd.ploc = ploc;
  const int k = s.rank[b];
  R* F = reinterpret_cast<R*>(s.fac_pool) + s.fac_off[b];
  if (k == 0) {
    if (any) {
      for (int64_t idx = tid; idx < (int64_t)n * n; idx += NT)
        F[idx] = upd((int)(idx % n), (int)(idx / n));
    }
    if (tid == 0) s.upd_kind[b] = any ? 1 : 0;
    return;
  }
  const int r = n - k;
  const FacLayout fl = FacLayout::make(k, r);
  const int* act = s.act_list + (int64_t)b * s.ncap;
  for (int i = tid; i < n; i += NT) sperm[i] = (short)s.perm[(int64_t)b * s.ncap + i];
  __syncthreads();
  R* P = (2 * (int64_t)n * n <= kLdsElems)
             ? sbuf
             : reinterpret_cast<R*>(s.id_pool) + s.id_off[b];
  R* X = P + (int64_t)n * n;
  R* Krs = X + (int64_t)r * r;
  R* Ksr = Krs + (int64_t)r * k;
  for (int64_t idx = tid; idx < (int64_t)n * n; idx += NT) {
    const int ia = sperm[idx % n], ic = sperm[idx / n];
    P[idx] = static_cast<R>(kernel_entry(g, act[ia], act[ic])) - upd(ia, ic);
  }
  __syncthreads();
  const R* T = F + fl.T;
  for (int idx = tid; idx < r * k; idx += NT) {
    const int i = idx % r, a = idx / r;
    R acc = P[(k + i) + (int64_t)n * a];
    for (int l = 0; l < k; l++) acc -= T[l + k * i] * P[l + (int64_t)n * a];
    Krs[idx] = acc;
  }
  for (int idx = tid; idx < k * r; idx += NT) {
    const int a = idx % k, j = idx / k;
    R acc = P[a + (int64_t)n * (k + j)];
    for (int l = 0; l < k; l++) acc -= P[a + (int64_t)n * l] * T[l + k * j];
    Ksr[idx] = acc;
  }
  __syncthreads();
  for (int idx = tid; idx < r * r; idx += NT) {
    const int i = idx % r, j = idx / r;
    R acc = P[(k + i) + (int64_t)n * (k + j)];
    for (int l = 0; l < k; l++) acc -= T[l + k * i] * P[l + (int64_t)n * (k + j)];
    for (int l = 0; l < k; l++) acc -= Krs[i + r * l] * T[l + k * j];
    X[idx] = acc;
  }
  __syncthreads();
  const int info = block_lu<R, NT>(X, r, r, spiv);
  if (info != 0 && tid == 0) atomicOr(s.err, kErrSingular);
  R* FX = F + fl.X;
  R* Fp = F + fl.piv;
  for (int idx = tid; idx < r * r; idx += NT) FX[idx] = X[idx];
  for (int i = tid; i < r; i += NT) Fp[i] = static_cast<R>(spiv[i]);
  // U = X^{-1} K_rs, one thread per column (dgetrs 'N').
  R* U = F + fl.U;
  for (int a = tid; a < k; a += NT) {
    R* u = U + (int64_t)r * a;
    for (int i = 0; i < r; i++) u[i] = Krs[i + r * a];
    for (int i = 0; i < r; i++) {
      const int p = spiv[i];
      if (p != i) {
        const R tmp = u[i];
        u[i] = u[p];
        u[p] = tmp;
      }
    }
    for (int l = 0; l < r; l++) {
      const R ul = u[l];
      for (int i = l + 1; i < r; i++) u[i] -= X[i + r * l] * ul;
    }
    for (int i = r - 1; i >= 0; i--) {
      u[i] /= X[i + r * i];
      const R ui = u[i];
      for (int l = 0; l < i; l++) u[l] -= X[l + r * i] * ui;
    }
  }
  // L = K_sr X^{-1}, one thread per row (dgetrs 'T' on the transpose).
  R* Lm = F + fl.L;
  for (int a = tid; a < k; a += NT) {
    for (int i = 0; i < r; i++) {
      R acc = Ksr[a + k * i];
      for (int l = 0; l < i; l++) acc -= X[l + r * i] * Lm[a + k * l];
      Lm[a + k * i] = acc / X[i + r * i];
    }
    for (int i = r - 1; i >= 0; i--) {
      R acc = Lm[a + k * i];
      for (int l = i + 1; l < r; l++) acc -= X[l + r * i] * Lm[a + k * l];
      Lm[a + k * i] = acc;
    }
    for (int i = r - 1; i >= 0; i--) {
      const int p = spiv[i];
      if (p != i) {
        const R tmp = Lm[a + k * i];
        Lm[a + k * i] = Lm[a + k * p];
        Lm[a + k * p] = tmp;
      }
    }
  }
  __syncthreads();
  R* S = F + fl.S;
  for (int idx = tid; idx < k * k; idx += NT) {
    const int a = idx % k, c = idx / k;
    R acc = 0;
    for (int j = 0; j < r; j++) acc += Lm[a + k * j] * Krs[j + r * c];
    S[idx] = upd(sperm[a], sperm[c]) + acc;
  }
  if (tid == 0) s.upd_kind[b] = 1;
}

// ---------------------------------------------------------------------------
// Root: K(allskel, allskel) - carried blocks of the level-1 boxes, LU.
// ---------------------------------------------------------------------------

template <typename R, int NT>
__global__ __launch_bounds__(NT) void k_root(DevGeom g, DevTree t, DevState s,
                                             int factor_lu) {
  __shared__ signed char pch[kMaxBoxDofs];
  __shared__ short ploc[kMaxBoxDofs];
  __shared__ const R* chS[4];
  __shared__ int chN[4];
  const int b = 0;
  const int tid = threadIdx.x;
  if (tid == 0) {  // the root is never compressed
    s.rank[b] = 0;
    s.m_rows[b] = 0;
    s.out_cnt[b] = s.act_cnt[b];
  }
  if (!pools_ok(s)) return;
  const int n = s.act_cnt[b];
  const bool leaf = t.leaf[b] != 0;
  if (tid < 4) {
    const int c = leaf ? -1 : t.child[4 * b + tid];
    chS[tid] = (c >= 0) ? carried_block<R>(s, c) : nullptr;
    chN[tid] = (c >= 0) ? s.out_cnt[c] : 0;
  }
  __syncthreads();
  for (int i = tid; i < n; i += NT) {
    int q = -1, loc = 0;
    if (!leaf) {
      for (int u = 0; u < 4; u++) {
        const int p0 = s.chpos[4 * b + u];
        if (i >= p0 && i < p0 + chN[u]) {
          q = u;
          loc = i - p0;
        }
      }
    }
    pch[i] = (signed char)q;
    ploc[i] = (short)loc;
  }
  __syncthreads();
  ChildUpdates<R> upd;
  upd.S = chS;
  upd.N = chN;
  upd.pch = pch;
  upd.ploc = ploc;
  const int* act = s.act_list;
  R* A = reinterpret_cast<R*>(s.root_mat);
  for (int64_t idx = tid; idx < (int64_t)n * n; idx += NT) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    A[idx] = static_cast<R>(kernel_entry(g, act[i], act[j])) - upd(i, j);
  }
  __syncthreads();
  if (factor_lu) {
    const int info = block_lu<R, NT>(A, n, n, s.root_piv);
    if (info != 0 && tid == 0) atomicOr(s.err, kErrSingular);
  }
}

// ---------------------------------------------------------------------------
// Solve sweeps (skel_factorization.cpp:507-592).  One wave per box; the
// box's skeleton and redundant entries are gathered to LDS, updated and
// scattered back (boxes of one level touch disjoint DOFs).
// ---------------------------------------------------------------------------

template <typename R>
__device__ void wave_lu_solve(const R* X, const R* piv, int r, double* x) {
  const int lane = threadIdx.x & 63;
  if (lane == 0) {
    for (int i = 0; i < r; i++) {
      const int p = (int)piv[i];
      if (p != i) {
        const double tmp = x[i];
        x[i] = x[p];
        x[p] = tmp;
      }
    }
  }
  wave_sync();
  for (int l = 0; l < r; l++) {
    const double xl = x[l];
    for (int i = l + 1 + lane; i < r; i += 64) x[i] -= (double)X[i + r * l] * xl;
    wave_sync();
  }
  for (int i = r - 1; i >= 0; i--) {
    if (lane == 0) x[i] /= (double)X[i + r * i];
    wave_sync();
    const double xi = x[i];
    for (int l = lane; l < i; l += 64) x[l] -= (double)X[l + r * i] * xi;
    wave_sync();
  }
}

template <typename R>
__global__ __launch_bounds__(64) void k_sweep_up(DevState s, int start,
                                                 SolveVec v, int use_Ut,
                                                 int diag) {
  __shared__ double xs[kMaxBoxDofs];
  __shared__ double xr[kMaxBoxDofs];
  const int b = start + blockIdx.x;
  const int k = s.rank[b];
  if (k == 0) return;
  const int lane = threadIdx.x;
  const int r = s.act_cnt[b] - k;
  const FacLayout fl = FacLayout::make(k, r);
  const R* F = reinterpret_cast<const R*>(s.fac_pool) + s.fac_off[b];
  const int* sk = s.out_list + (int64_t)b * s.ncap;
  const int* rd = s.red_list + (int64_t)b * s.ncap;
  for (int c = 0; c < v.nrhs; c++) {
    double* x = v.x + v.ldx * c;
    for (int a = lane; a < k; a += 64) xs[a] = x[sk[a]];
    for (int i = lane; i < r; i += 64) xr[i] = x[rd[i]];
    wave_sync();
    for (int i = lane; i < r; i += 64) {
      double acc = xr[i];
      const R* tc = F + fl.T + (int64_t)k * i;
      for (int a = 0; a < k; a++) acc -= (double)tc[a] * xs[a];
      xr[i] = acc;
    }
    wave_sync();
    for (int a = lane; a < k; a += 64) {
      double acc = xs[a];
      if (use_Ut) {
        const R* uc = F + fl.U + (int64_t)r * a;
        for (int i = 0; i < r; i++) acc -= (double)uc[i] * xr[i];
      } else {
        const R* L = F + fl.L;
        for (int i = 0; i < r; i++) acc -= (double)L[a + (int64_t)k * i] * xr[i];
      }
      xs[a] = acc;
    }
    wave_sync();
    if (diag) wave_lu_solve(F + fl.X, F + fl.piv, r, xr);
    for (int a = lane; a < k; a += 64) x[sk[a]] = xs[a];
    for (int i = lane; i < r; i += 64) x[rd[i]] = xr[i];
    wave_sync();
  }
}

template <typename R>
__global__ __launch_bounds__(64) void k_diag(DevState s, int start, SolveVec v) {
  __shared__ double xr[kMaxBoxDofs];
  const int b = start + blockIdx.x;
  const int k = s.rank[b];
  if (k == 0) return;
  const int lane = threadIdx.x;
  const int r = s.act_cnt[b] - k;
  const FacLayout fl = FacLayout::make(k, r);
  const R* F = reinterpret_cast<const R*>(s.fac_pool) + s.fac_off[b];
  const int* rd = s.red_list + (int64_t)b * s.ncap;
  for (int c = 0; c < v.nrhs; c++) {
    double* x = v.x + v.ldx * c;
    for (int i = lane; i < r; i += 64) xr[i] = x[rd[i]];
    wave_sync();
    wave_lu_solve(F + fl.X, F + fl.piv, r, xr);
    for (int i = lane; i < r; i += 64) x[rd[i]] = xr[i];
    wave_sync();
  }
}

template <typename R>
__global__ __launch_bounds__(64) void k_sweep_down(DevState s, int start,
                                                   SolveVec v) {
  __shared__ double xs[kMaxBoxDofs];
  __shared__ double xr[kMaxBoxDofs];
  const int b = start + blockIdx.x;
  const int k = s.rank[b];
  if (k == 0) return;
  const int lane = threadIdx.x;
  const int r = s.act_cnt[b] - k;
  const FacLayout fl = FacLayout::make(k, r);
  const R* F = reinterpret_cast<const R*>(s.fac_pool) + s.fac_off[b];
  const int* sk = s.out_list + (int64_t)b * s.ncap;
  const int* rd = s.red_list + (int64_t)b * s.ncap;
  for (int c = 0; c < v.nrhs; c++) {
    double* x = v.x + v.ldx * c;
    for (int a = lane; a < k; a += 64) xs[a] = x[sk[a]];
    for (int i = lane; i < r; i += 64) xr[i] = x[rd[i]];
    wave_sync();
    for (int i = lane; i < r; i += 64) {
      double acc = xr[i];
      const R* U = F + fl.U;
      for (int a = 0; a < k; a++) acc -= (double)U[i + (int64_t)r * a] * xs[a];
      xr[i] = acc;
    }
    wave_sync();
    for (int a = lane; a < k; a += 64) {
      double acc = xs[a];
      const R* T = F + fl.T;
      for (int i = 0; i < r; i++) acc -= (double)T[a + (int64_t)k * i] * xr[i];
      xs[a] = acc;
    }
    wave_sync();
    for (int a = lane; a < k; a += 64) x[sk[a]] = xs[a];
    for (int i = lane; i < r; i += 64) x[rd[i]] = xr[i];
    wave_sync();
  }
}

template <typename R>
__global__ __launch_bounds__(64) void k_root_solve(DevState s, SolveVec v) {
  __shared__ double xa[kMaxBoxDofs];
  __shared__ R piv[kMaxBoxDofs];
  const int lane = threadIdx.x;
  const int n = s.act_cnt[0];
  const int* act = s.act_list;
  const R* A = reinterpret_cast<const R*>(s.root_mat);
  for (int i = lane; i < n; i += 64) piv[i] = static_cast<R>(s.root_piv[i]);
  for (int c = 0; c < v.nrhs; c++) {
    double* x = v.x + v.ldx * c;
    for (int i = lane; i < n; i += 64) xa[i] = x[act[i]];
    wave_sync();
    wave_lu_solve(A, piv, n, xa);
    for (int i = lane; i < n; i += 64) x[act[i]] = xa[i];
    wave_sync();
  }
}

}  // namespace

void launch_reset(const DevState& s, hipStream_t st) {
  k_reset<<<1, 64, 0, st>>>(s);
}

template <typename R>
void launch_factor_stage(int stage, const DevGeom& g, const DevTree& t,
                         const DevState& s, int level, int start, int count,
                         double tol, hipStream_t st) {
  if (count <= 0) return;
  switch (stage) {
    case kStageActive:
      k_active<<<count, 64, 0, st>>>(t, s, start);
      break;
    case kStageRows:
      k_rows<<<count, 64, 0, st>>>(g, t, s, start, level);
      break;
    case kStageScan:
      k_scan<<<1, 1024, 0, st>>>(s, start, count);
      break;
    case kStageAssemble:
      k_assemble<R><<<count, 256, 0, st>>>(g, t, s, start, level);
      break;
    case kStageQrcp:
      k_rowcomp<R, kXRows><<<min(2 * count, 512), kXThreads, 0, st>>>(s, start, level);
      k_qrcp_id<R, kXRows><<<count, kXThreads, 0, st>>>(s, start, tol);
      k_qrcp<R, 512><<<count, 512, 0, st>>>(s, start, tol);
      break;
    case kStageDecouple:
      k_decouple_small<R><<<count, 256, 0, st>>>(g, t, s, start);
      k_decouple<R, 256><<<count, 256, 0, st>>>(g, t, s, start);
      break;
    default:
      break;
  }
}

template <typename R>
void launch_root(const DevGeom& g, const DevTree& t, const DevState& s,
                 bool factor_lu, hipStream_t st) {
  k_active<<<1, 64, 0, st>>>(t, s, 0);
  k_root<R, 1024><<<1, 1024, 0, st>>>(g, t, s, factor_lu ? 1 : 0);
}

template <typename R>
void launch_sweep_up(const DevState& s, int start, int count, SolveVec v,
                     bool use_Ut, bool diag, hipStream_t st) {
  if (count <= 0) return;
  k_sweep_up<R><<<count, 64, 0, st>>>(s, start, v, use_Ut ? 1 : 0, diag ? 1 : 0);
}

template <typename R>
void launch_diag(const DevState& s, int start, int count, SolveVec v,
                 hipStream_t st) {
  if (count <= 0) return;
  k_diag<R><<<count, 64, 0, st>>>(s, start, v);
}

template <typename R>
void launch_sweep_down(const DevState& s, int start, int count, SolveVec v,
                       hipStream_t st) {
  if (count <= 0) return;
  k_sweep_down<R><<<count, 64, 0, st>>>(s, start, v);
}

template <typename R>
void launch_root_solve(const DevState& s, SolveVec v, hipStream_t st) {
  k_root_solve<R><<<1, 64, 0, st>>>(s, v);
}

#define IEGPU_INSTANTIATE(R)                                                   \
  template void launch_factor_stage<R>(int, const DevGeom&, const DevTree&,   \
                                       const DevState&, int, int, int, double, \
                                       hipStream_t);                           \
  template void launch_root<R>(const DevGeom&, const DevTree&,                \
                               const DevState&, bool, hipStream_t);            \
  template void launch_sweep_up<R>(const DevState&, int, int, SolveVec, bool, \
                                   bool, hipStream_t);                         \
  template void launch_diag<R>(const DevState&, int, int, SolveVec,           \
                               hipStream_t);                                   \
  template void launch_sweep_down<R>(const DevState&, int, int, SolveVec,     \
                                     hipStream_t);                             \
  template void launch_root_solve<R>(const DevState&, SolveVec, hipStream_t);

IEGPU_INSTANTIATE(double)
IEGPU_INSTANTIATE(float)

}  // namespace iegpu
