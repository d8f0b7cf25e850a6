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
    s.err[2] = 0;
    s.err[3] = 0;
    // an incremental run appends new factor slots past the kept ones
    if (!s.dirty) *s.fac_top = 0;
  }
  if (threadIdx.x < kMaxLevels) s.njobs[threadIdx.x] = 0;
  for (int i = threadIdx.x; i < 2 * kMaxLevels * kNumLists; i += blockDim.x) s.lcnt[i] = 0;
}

// active(box) = concatenation of the children's outgoing lists (skeleton if
// compressed, active otherwise) in child order bl, tl, tr, br; a leaf's
// active list is its original DOF list.
__global__ __launch_bounds__(64) void k_active(DevTree t, DevState s,
                                               int start) {
  const int b = start + blockIdx.x;
  if (box_clean(s, b)) return;
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
      const int cnt = min(max(s.out_cnt[c], 0), s.ncap);
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
  if (lane == 0) s.act_cnt[b] = n;
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

// Rank-model features of box b (every lane returns them): log2 |active|,
// log2 ID rows, level, log2(near rows + 1), side x mean and max |curvature|
// of the active DOFs, 1 - |mean normal| (how far the curve turns inside
// the box), -log10 tol, Stokes flag.
__device__ inline void box_features(const DevGeom& g, const DevTree& t, const DevState& s, int b,
                                    int level, int n, int m, int cnt, double tol, float* f) {
  const int lane = threadIdx.x & 63;
  const int* act = s.act_list + (int64_t)b * s.ncap;
  double ks = 0, km = 0, nx = 0, ny = 0;
  for (int i = lane; i < n; i += 64) {
    const DofGeo d = dof_geo(g, act[i]);
    const double k = fabs(d.k);
    ks += k;
    km = fmax(km, k);
    nx += d.n1;
    ny += d.n2;
  }
  ks = wave_total(ks);
  km = wave_max(km);
  nx = wave_total(nx);
  ny = wave_total(ny);
  const double side = t.side[b], nn = n > 0 ? n : 1;
  f[0] = (float)log2((double)nn);
  f[1] = (float)log2((double)(m > 0 ? m : 1));
  f[2] = (float)level;
  f[3] = (float)log2((double)cnt + 1.0);
  f[4] = (float)(side * ks / nn);
  f[5] = (float)(side * km);
  f[6] = (float)(1.0 - sqrt(nx * nx + ny * ny) / nn);
  f[7] = (float)(-log10(tol));
  f[8] = g.pde == kPdeStokes ? 1.0f : 0.0f;
}

// Rank predicted by the model for features f (lanes 0..kHid-1 own the
// hidden units; every lane returns the prediction).
__device__ inline float mlp_rank(const float* M, const float* f, float* hb) {
  const int lane = threadIdx.x & 63;
  const int j = lane & (kHid - 1);
  float z = M[kMlpB1 + j];
#pragma unroll
  for (int q = 0; q < kFeat; q++)
    z += M[kMlpW1 + q * kHid + j] * ((f[q] - M[kMlpMu + q]) / M[kMlpSd + q]);
  if (lane < kHid) hb[j] = tanhf(z);
  __syncthreads();
  float z2 = M[kMlpB2 + j];
  for (int i = 0; i < kHid; i++) z2 += M[kMlpW2 + i * kHid + j] * hb[i];
  double o = lane < kHid ? (double)(M[kMlpW3 + j] * tanhf(z2)) : 0.0;
  o = wave_total(o) + M[kMlpB3];
  return (float)o * M[kMlpYSd] + M[kMlpYMu];
}

__global__ __launch_bounds__(64) void k_rows(DevGeom g, DevTree t, DevState s,
                                             int start, int level, double tol) {
  __shared__ float hb[kHid];
  const int b = start + blockIdx.x;
  if (box_clean(s, b)) return;
  const int n = s.act_cnt[b];
  int m = 0;
  if (n >= kMinCompress) {
    const int cnt =
        for_each_row_dof(g, t, s, b, level >= 3, [](int, int) {});
    m = 2 * cnt + proxy_rows(g, level);
  }
  int kind = kQrcpNone;
  if (m > 0) {
    const int nch4 = id_chunks(m, kRegMaxRows), nch8 = id_chunks(m, kRegMaxRows8);
    if (n <= kRegMaxCols && m <= kRegMaxRows) kind = kQrcpReg;
    else if (n <= kRegMaxCols && nch4 <= kMaxChunks && nch4 * n <= kRegMaxRows)
      kind = kQrcpStack;
    else if (n <= kRegMaxCols8 && m <= kRegMaxRows8) kind = kQrcpReg8;
    else if (n <= kRegMaxCols8 && nch8 <= kMaxChunks && nch8 * n <= kRegMaxRows8)
      kind = kQrcpStack8;
    else kind = kQrcpLds;
  }
  if (threadIdx.x == 0) {
    s.m_rows[b] = m;
    s.qkind[b] = kind;
    // Scratch: the assembled matrix (and the general decouple's operands)
    // for the general QRCP, the stacked triangles for row compression;
    // register-QRCP boxes evaluate their entries in place.
    const int64_t ld = m > 2 * n ? m : 2 * n;
    int64_t scratch = kind == kQrcpLds ? ld * n : 0;
    if (kind == kQrcpStack || kind == kQrcpStack8)
      scratch = (int64_t)id_chunks(m, kind == kQrcpStack ? kRegMaxRows : kRegMaxRows8) * n * n;
    s.id_sz[b] = scratch;
    // k_decouple keeps the operands of a box with more than kSmallBox
    // active DOFs (n^2 + r^2 + 2 r k <= 2 n^2 entries) in the decouple pool
    // when they do not fit its LDS buffer.
    s.dec_sz[b] = n > kSmallBox ? 2 * (int64_t)n * n : 0;
    s.row_sz[b] = kind == kQrcpLds ? (m - proxy_rows(g, level)) / 2 : 0;
    // fac_sz holds a box's slot size; in an incremental run a re-factored
    // box whose kept slot fits its new factors reuses it (marked by the
    // negated size, which k_scan restores) instead of taking a new one.
    const int64_t need = (n > 0) ? fac_slot_size(n) : 0;
    const int64_t old = s.fac_sz[b];
    const bool reuse = s.dirty != nullptr && need <= old && s.fac_off[b] + old <= s.kept_top;
    s.fac_sz[b] = reuse ? -old : need;
  }
  // Sketch width: fixed, or from the rank model (0: straight to the full
  // QRCP when the predicted rank does not fit the widest sketch).  Boxes
  // too large for the narrow sketch on the top levels take the wide one.
  const int near = (m - proxy_rows(g, level)) / 2;
  const bool elig = s.rid_mode != 0 && (kind == kQrcpReg || kind == kQrcpStack) &&
                    near <= kRidMaxNear;
  const bool eligw = s.rid_mode != 0 && m > 0 && level <= kRidWideTop && n <= kRidWideCols &&
                     near <= kRidMaxNearW;
  int w = elig ? s.rid_width : (eligw ? kRidRowsW : 0);
  if (m > 0 && (s.feat != nullptr || ((elig || eligw) && s.rid_mode == 2))) {
    float f[kFeat];
    box_features(g, t, s, b, level, n, m, near, tol, f);
    if (s.feat != nullptr && threadIdx.x == 0)
      for (int q = 0; q < kFeat; q++) s.feat[(int64_t)b * kFeat + q] = f[q];
    if ((elig || eligw) && s.rid_mode == 2) {
      // The model sizes the narrow sketch; a box it does not fit tries the
      // wide one when eligible (few boxes, whose ranks the model has seen
      // least), else goes straight to the full QRCP.
      const float kp = mlp_rank(s.rank_model, f, hb) + (float)s.rid_margin;
      w = (elig && kp <= 16 - kRidOver) ? 16
          : (elig && kp <= kRidRows - kRidOver) ? kRidRows
          : eligw ? kRidRowsW
                  : 0;
    }
  }
  if (threadIdx.x == 0) {
    s.sk_w[b] = w;
    s.rid_done[b] = 0;
  }
  // Until its ID succeeds a box passes its active DOFs on uncompressed:
  // boxes a failed (retried) attempt never reaches still hold a consistent
  // outgoing list for the next level's k_active.
  const int* act = s.act_list + (int64_t)b * s.ncap;
  int* out = s.out_list + (int64_t)b * s.ncap;
  for (int i = threadIdx.x; i < n; i += 64) out[i] = act[i];
  if (threadIdx.x == 0) {
    s.rank[b] = 0;
    s.out_cnt[b] = n;
  }
}

// A register-QRCP box the randomized ID does not sketch goes to the full
// QRCP list at the scan.
__device__ inline bool rid_full_qrcp(const DevState& s, int b, int kind) {
  return s.rid_mode != 0 && (kind == kQrcpReg || kind == kQrcpStack) && s.sk_w[b] == 0;
}

// Scanned per box: ID scratch, row-list and factor sizes, register-QRCP
// and general-QRCP list entries, row-compression jobs, sketches, decouple
// operands.
constexpr int kScanQ = 12;
constexpr int kScanLd = 16;      // words per block in scan_blk
constexpr int kScanMulti = 2048; // levels of at least this many boxes: three kernels

__device__ inline void scan_counts(const DevState& s, int b, int64_t (&own)[kScanQ]) {
  if (box_clean(s, b)) return;
  const int kind = s.qkind[b];
  own[0] += s.id_sz[b];
  own[1] += s.row_sz[b];
  own[2] += max(s.fac_sz[b], int64_t(0));
  own[3] += ((kind == kQrcpReg || kind == kQrcpStack) && s.act_cnt[b] > kHalfCols) ? 1 : 0;
  own[7] += ((kind == kQrcpReg || kind == kQrcpStack) && s.act_cnt[b] <= kHalfCols) ? 1 : 0;
  own[4] += (kind == kQrcpLds) ? 1 : 0;
  own[5] += row_jobs(s, b);
  own[6] += (kind == kQrcpReg8 || kind == kQrcpStack8) ? 1 : 0;
  own[8] += (s.sk_w[b] > 0) ? 1 : 0;
  own[9] += rid_full_qrcp(s, b, kind) ? 1 : 0;
  own[10] += s.dec_sz[b];
  own[11] += (s.sk_w[b] == kRidRowsW) ? 1 : 0;
}

// Box b's slots and list entries from the exclusive offsets off (advanced
// past the box).
__device__ inline void scan_place(const DevState& s, int b, int start, int level,
                                  int64_t fbase, int64_t (&off)[kScanQ]) {
  if (box_clean(s, b)) return;
  int* lreg = s.lists + (int64_t)kLstReg * s.nbox + start;
  int* llds = s.lists + (int64_t)kLstLds * s.nbox + start;
  int* lreg8 = s.lists + (int64_t)kLstReg8 * s.nbox + start;
  int* lhalf = s.lists + (int64_t)kLstRegHalf * s.nbox + start;
  int* lrid = s.lists + (int64_t)kLstRid * s.nbox + start;
  int* lfb = s.lists + (int64_t)kLstFallback * s.nbox + start;
  int* jobs = s.jobs + (int64_t)2 * kMaxChunks * start;
  const int kind = s.qkind[b];
  // Scratch of odd and even levels in separate halves: a level's ID runs
  // while the level below it is still being decoupled.
  s.id_off[b] = (level & 1) * s.id_cap + off[0];
  s.row_off[b] = off[1];
  if (s.fac_sz[b] < 0) {  // keeps its slot
    s.fac_sz[b] = -s.fac_sz[b];
  } else {
    s.fac_off[b] = fbase + off[2];
    off[2] += s.fac_sz[b];
  }
  s.dec_off[b] = off[10];
  off[10] += s.dec_sz[b];
  off[0] += s.id_sz[b];
  off[1] += s.row_sz[b];
  if (kind == kQrcpReg || kind == kQrcpStack) {
    if (s.act_cnt[b] > kHalfCols) lreg[off[3]++] = b;
    else lhalf[off[7]++] = b;
  }
  if (kind == kQrcpLds) llds[off[4]++] = b;
  if (kind == kQrcpReg8 || kind == kQrcpStack8) lreg8[off[6]++] = b;
  if (s.sk_w[b] > 0) lrid[off[8]++] = b;
  if (rid_full_qrcp(s, b, kind)) lfb[off[9]++] = b;
  const int nch = row_jobs(s, b);
  for (int q = 0; q < nch; q++) {
    jobs[2 * off[5]] = b;
    jobs[2 * off[5] + 1] = q;
    off[5]++;
  }
}

// The level's list lengths, pool checks and the factor-pool top.
__device__ inline void scan_finish(const DevState& s, int level, const int64_t* tot,
                                   int64_t fbase) {
  const int64_t ta = tot[0], tb = tot[1], tc = tot[2];
  list_slot(s, level, kLstReg)[0] = (int)tot[3];
  list_slot(s, level, kLstLds)[0] = (int)tot[4];
  list_slot(s, level, kLstReg8)[0] = (int)tot[6];
  list_slot(s, level, kLstRegHalf)[0] = (int)tot[7];
  list_slot(s, level, kLstRid)[0] = (int)tot[8];
  list_slot(s, level, kLstRid)[1] = (int)tot[11];  // of which 64-row sketches
  list_slot(s, level, kLstFallback)[0] = (int)tot[9];  // k_rid appends
  s.njobs[level] = (int)tot[5];
  int e = 0;
  if (ta > s.id_cap) {
    e |= kErrIdPool;
    atomicMax(s.err + 2, (int)((ta + 1023) >> 10));  // size needed, in 1024 elements
  }
  if (tb > s.row_cap) {
    e |= kErrRowPool;
    atomicMax(s.err + 3, (int)((tb + 1023) >> 10));
  }
  if (fbase + tc > s.fac_cap) e |= kErrFacPool;
  if (tot[10] > s.dec_cap) {
    e |= kErrDecPool;
    atomicMax(s.err + 1, (int)((tot[10] + 1023) >> 10));
  }
  if (e) atomicOr(s.err, e);
  *s.fac_top = fbase + tc;
}

// Exclusive scan of the kScanQ values of the block's threads (NT a
// multiple of 64): off = sum over lower threads; tot (LDS, kScanQ) = the
// block's totals.  wtot: kScanQ x NT/64 LDS words.
template <int NT>
__device__ inline void block_scan(const int64_t (&own)[kScanQ], int64_t (&off)[kScanQ],
                                  int64_t (*wtot)[NT / 64], int64_t* tot) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  int64_t inc[kScanQ];
#pragma unroll
  for (int q = 0; q < kScanQ; q++) {
    int64_t v = own[q];
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int64_t u = __shfl_up(v, o, 64);
      if (lane >= o) v += u;
    }
    inc[q] = v;
    if (lane == 63) wtot[q][wv] = v;
  }
  __syncthreads();
  if (tid < kScanQ) {
    int64_t run = 0;
    for (int w = 0; w < NT / 64; w++) {
      const int64_t t = wtot[tid][w];
      wtot[tid][w] = run;
      run += t;
    }
    tot[tid] = run;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kScanQ; q++) off[q] = wtot[q][wv] + inc[q] - own[q];
}

// Exclusive scans of the per-box sizes of one level (single workgroup;
// each thread a contiguous chunk of boxes).  ID scratch and row lists are
// level-local; factor slots append to the running pool offset.
// With the randomized ID, a register-QRCP box that gets no sketch (not
// eligible, or sent straight to the full QRCP by the rank model) starts the
// level's full-QRCP list; k_rid appends the boxes whose sketch fails.
__global__ __launch_bounds__(1024) void k_scan(DevState s, int start,
                                               int count, int level) {
  __shared__ int64_t wtot[kScanQ][16];
  __shared__ int64_t tot[kScanQ];
  const int tid = threadIdx.x;
  const int chunk = (count + 1023) / 1024;
  const int lo = min(count, tid * chunk), hi = min(count, lo + chunk);
  int64_t own[kScanQ] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = lo; i < hi; i++) scan_counts(s, start + i, own);
  int64_t off[kScanQ];
  block_scan<1024>(own, off, wtot, tot);
  const int64_t fbase = *s.fac_top;
  for (int i = lo; i < hi; i++) scan_place(s, start + i, start, level, fbase, off);
  __syncthreads();
  if (tid == 0) scan_finish(s, level, tot, fbase);
}

// The same for levels of kScanMulti boxes or more over many workgroups of
// 256 boxes (one workgroup walked them at one chunk of boxes per thread
// at 15-70 us per level): each block's totals (k_scan_blocks), their
// exclusive scan and the level's totals (k_scan_totals; the factor-pool
// base goes after the block offsets), then each box's offsets
// (k_scan_place).  Same offsets and list order as k_scan.
__global__ __launch_bounds__(256) void k_scan_blocks(DevState s, int start, int count) {
  __shared__ int64_t wtot[kScanQ][4];
  __shared__ int64_t tot[kScanQ];
  const int i = blockIdx.x * 256 + threadIdx.x;
  int64_t own[kScanQ] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (i < count) scan_counts(s, start + i, own);
  int64_t off[kScanQ];
  block_scan<256>(own, off, wtot, tot);
  if (threadIdx.x < kScanQ) s.scan_blk[(int64_t)blockIdx.x * kScanLd + threadIdx.x] = tot[threadIdx.x];
}

__global__ __launch_bounds__(1024) void k_scan_totals(DevState s, int nblk, int level) {
  __shared__ int64_t wtot[kScanQ][16];
  __shared__ int64_t tot[kScanQ];
  const int tid = threadIdx.x;
  const int chunk = (nblk + 1023) / 1024;
  const int lo = min(nblk, tid * chunk), hi = min(nblk, lo + chunk);
  int64_t own[kScanQ] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int j = lo; j < hi; j++)
#pragma unroll
    for (int q = 0; q < kScanQ; q++) own[q] += s.scan_blk[(int64_t)j * kScanLd + q];
  int64_t off[kScanQ];
  block_scan<1024>(own, off, wtot, tot);
  for (int j = lo; j < hi; j++)
#pragma unroll
    for (int q = 0; q < kScanQ; q++) {
      int64_t* w = s.scan_blk + (int64_t)j * kScanLd + q;
      const int64_t t = *w;
      *w = off[q];
      off[q] += t;
    }
  __syncthreads();
  if (tid == 0) {
    const int64_t fbase = *s.fac_top;
    s.scan_blk[(int64_t)nblk * kScanLd] = fbase;
    scan_finish(s, level, tot, fbase);
  }
}

__global__ __launch_bounds__(256) void k_scan_place(DevState s, int start, int count,
                                                    int level, int nblk) {
  __shared__ int64_t wtot[kScanQ][4];
  __shared__ int64_t tot[kScanQ];
  const int i = blockIdx.x * 256 + threadIdx.x;
  int64_t own[kScanQ] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (i < count) scan_counts(s, start + i, own);
  int64_t off[kScanQ];
  block_scan<256>(own, off, wtot, tot);
  if (i >= count) return;
  const int64_t* base = s.scan_blk + (int64_t)blockIdx.x * kScanLd;
#pragma unroll
  for (int q = 0; q < kScanQ; q++) off[q] += base[q];
  scan_place(s, start + i, start, level, s.scan_blk[(int64_t)nblk * kScanLd], off);
}


// ---------------------------------------------------------------------------
// ID matrix assembly: [K(ic, box); K(box, ic)^T; proxy] (m x n,
// column-major, ld m), kernel.cpp:293-423.  The register QR kernels
// evaluate their boxes' entries in place; this kernel writes the matrix of
// the boxes the general QRCP takes.
// ---------------------------------------------------------------------------

// Row DOFs and proxy points of box b into LDS (all threads; the caller
// synchronises).  Returns the number of row DOFs (cnt).
__device__ inline int id_setup(const DevGeom& g, const DevTree& t, const DevState& s,
                               int b, int level, int m, int* rows,
                               double (*pxy)[kNumProxy]) {
  const int tid = threadIdx.x;
  if (tid < 64)
    for_each_row_dof(g, t, s, b, level >= 3, [&](int pos, int dof) { rows[pos] = dof; });
  if (level >= 3 && tid < kNumProxy) {
    double px, py, n1, n2;
    proxy_point(t.cx[b], t.cy[b], t.side[b] * kRadiusRatio, tid, &px, &py, &n1, &n2);
    pxy[0][tid] = px;
    pxy[1][tid] = py;
    pxy[2][tid] = n1;
    pxy[3][tid] = n2;
  }
  return (m - proxy_rows(g, level)) / 2;
}

__device__ inline double proxy_weight(const DevTree& t, int b) {
  return 2.0 * M_PI * (t.side[b] * kRadiusRatio) / kNumProxy;
}

template <typename R>
__global__ __launch_bounds__(256) void k_assemble(DevGeom g, DevTree t,
                                                  DevState s, int start,
                                                  int level) {
  __shared__ double pxy[4][kNumProxy];
  if (!pools_ok(s)) return;
  const int nl = list_count(s, level, kLstLds);
  for (int li = blockIdx.x; li < nl; li += gridDim.x) {
    const int b = list_at(s, start, kLstLds, li);
    const int m = s.m_rows[b];
    const int n = s.act_cnt[b];
    const int tid = threadIdx.x;
    const int* act = s.act_list + (int64_t)b * s.ncap;
    int* rows = s.row_pool + s.row_off[b];
    R* A = reinterpret_cast<R*>(s.id_pool) + s.id_off[b];
    const int cnt = id_setup(g, t, s, b, level, m, rows, pxy);
    __syncthreads();
    const double pw = proxy_weight(t, b);
    for (int i = tid; i < m; i += 256) {
      const IdRow rw = id_row(g, rows, cnt, pxy, pw, i);
      for (int j = 0; j < n; j++)
        A[i + (int64_t)m * j] = static_cast<R>(id_entry(g.pde, rw, dof_geo(g, act[j])));
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Register-resident QR kernels (1024 threads, 16 waves).  Wave w holds
// columns CPW w .. CPW w + CPW - 1 of an m x n matrix (m <= 64 RPL,
// n <= 16 CPW) in the column interleave of rs_device.hpp: row i on lane
// i % 64, register i / 64, column CPW w + (sl ^ slot_class(lane)) in slot
// sl.  Two shapes: 4 columns x 8 row registers (n <= 64, m <= 512) and
// 8 columns x 4 row registers (n <= 128, m <= 256).  A Householder step:
// the wave owning the pivot column publishes the reflector through LDS;
// every wave then updates its columns in registers with one interleaved
// column reduction.  Waves past the last column only take part in the
// barriers.
// ---------------------------------------------------------------------------

constexpr int kXThreads = 1024;

template <int CPW>
struct RegShape {
  static constexpr int RPL = CPW == 4 ? 8 : 4;  // row registers per lane
  static constexpr int MAXM = 64 * RPL;
  static constexpr int MAXN = 16 * CPW;
};

// Entries of the column in slot psl (per lane) of the interleave.
template <typename R, int RPL, int CPW>
__device__ inline void pick_slot(const R (&a)[CPW][RPL], int psl, R (&x)[RPL]) {
#pragma unroll
  for (int r = 0; r < RPL; r++) x[r] = a[0][r];
#pragma unroll
  for (int sl = 1; sl < CPW; sl++) {
    const bool me = psl == sl;
#pragma unroll
    for (int r = 0; r < RPL; r++) x[r] = me ? a[sl][r] : x[r];
  }
}

// Row i (wave-uniform) of a column held as x[RPL], in every lane.  Each
// register is read out before the select: selecting first turns into a
// dynamically indexed extract, which the compiler lowers through scratch.
template <typename R, int RPL>
__device__ inline R row_value(const R (&x)[RPL], int i) {
  const int ri = i >> 6, li = i & 63;
  R v = lane_value(x[0], li);
#pragma unroll
  for (int r = 1; r < RPL; r++) {
    const R t = lane_value(x[r], li);
    v = (r == ri) ? t : v;
  }
  return v;
}

// LAPACK dlarfg on (alpha, x) with ||x||^2 = xn2: beta = -sign(alpha)
// ||(alpha, x)||, tau = (beta - alpha) / beta, scale 1 / (alpha - beta).
// (The ID matrices hold moderate entries, so ||(alpha, x)|| is formed
// directly rather than through dlapy2's scaling, and the two quotients use
// independent reciprocals.)
template <typename R>
__device__ inline R recip_any(R d) {
  if constexpr (sizeof(R) == 8) return recip(d);
  else return R(1) / d;
}
template <typename R>
__device__ inline void householder_scalars(R alpha, R xn2, R* beta, R* tau, R* sc) {
  if (xn2 == 0) {
    *tau = 0;
    *beta = alpha;
    *sc = 0;
  } else {
    *beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
    *tau = (*beta - alpha) * recip_any(*beta);
    *sc = recip_any(alpha - *beta);
  }
}

// Reflector v of column x at step j: v(0:j-1) = 0, v_j = 1,
// v(j+1:) = x(j+1:) sc; all 64 RPL rows (rows past the matrix hold zeros).
template <typename R, int RPL>
__device__ inline void write_reflector(const R (&x)[RPL], int j, R sc, R* vb) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int r = 0; r < RPL; r++) {
    const int i = lane + 64 * r;
    vb[i] = i > j ? x[r] * sc : (i == j ? R(1) : R(0));
  }
}

// a <- (I - tau v v^T) a on this wave's columns; `live` says whether the
// class slot_class(lane) is still to be updated (pivoted columns keep
// their R entries).
template <typename R, int RPL, int CPW>
__device__ inline void apply_reflector(R (&a)[CPW][RPL], const R* vb, R tau, bool live) {
  const int lane = threadIdx.x & 63;
  R vr[RPL];
#pragma unroll
  for (int r = 0; r < RPL; r++) vr[r] = vb[lane + 64 * r];
  R p[CPW];
#pragma unroll
  for (int sl = 0; sl < CPW; sl++) {
    R p0 = 0, p1 = 0;  // two chains halve the dependent FMA latency
#pragma unroll
    for (int r = 0; r < RPL; r += 2) {
      p0 += vr[r] * a[sl][r];
      if (r + 1 < RPL) p1 += vr[r + 1] * a[sl][r + 1];
    }
    p[sl] = p0 + p1;
  }
  R t = wave_total_cols<R, CPW>(p);
  t = live ? tau * t : R(0);
  R tw[CPW];
  spread_cols<R, CPW>(t, tw);
#pragma unroll
  for (int sl = 0; sl < CPW; sl++)
#pragma unroll
    for (int r = 0; r < RPL; r++) a[sl][r] -= tw[sl] * vr[r];
}

// Sum of squares of rows > i of each of this wave's columns (i = -1: all
// rows); lanes 0..CPW-1 return column CPW w + lane.
template <typename R, int RPL, int CPW, int MAXN>
__device__ inline R tail_sq(const R (&a)[CPW][RPL], int i) {
  const int lane = threadIdx.x & 63;
  R p[CPW];
#pragma unroll
  for (int sl = 0; sl < CPW; sl++) {
    p[sl] = 0;
#pragma unroll
    for (int r = 0; r < RPL; r++) {
      // With at most 64 columns, i < 64: registers past the first are
      // below row i.
      if (MAXN <= 64 && r > 0) p[sl] += a[sl][r] * a[sl][r];
      else p[sl] += (lane + 64 * r > i) ? a[sl][r] * a[sl][r] : R(0);
    }
  }
  return class_by_lane_cols<R, CPW>(wave_total_cols<R, CPW>(p));
}

// Row i of this wave's columns into dst[column].  With at most 64 columns
// row i < 64 sits in register 0 of lane i; the wide shape
// reads the row out of every register that can hold it (rows < 128) and
// selects, so that no register is indexed at run time (that would move the
// whole matrix to scratch memory).
template <typename R, int RPL, int CPW, int MAXN>
__device__ inline void publish_row(const R (&a)[CPW][RPL], int i, R* dst) {
  const int lane = threadIdx.x & 63, c0 = CPW * wave_id();
  if constexpr (MAXN <= 64) {
    const int cls = slot_class<CPW>(i);  // lane i's class: uniform addresses
    if (lane == i) {
#pragma unroll
      for (int sl = 0; sl < CPW; sl++) dst[c0 + (sl ^ cls)] = a[sl][0];
    }
  } else {
    constexpr int NR = (MAXN + 63) / 64 < RPL ? (MAXN + 63) / 64 : RPL;
    const int li = i & 63, ri = i >> 6;
    const int cls = slot_class<CPW>(li);
#pragma unroll
    for (int sl = 0; sl < CPW; sl++) {
      R v = lane_value(a[sl][0], li);
#pragma unroll
      for (int r = 1; r < NR; r++) {
        const R t = lane_value(a[sl][r], li);
        v = (r == ri) ? t : v;
      }
      if (lane == 0) dst[c0 + (sl ^ cls)] = v;
    }
  }
}

template <typename R, int RPL, int CPW>
__device__ inline void load_columns(const R* A, int64_t ld, int row0, int rows, int n,
                                    R (&a)[CPW][RPL]) {
  const int lane = threadIdx.x & 63, c0 = CPW * wave_id();
  const int cls = slot_class<CPW>(lane);
#pragma unroll
  for (int sl = 0; sl < CPW; sl++) {
    const int c = c0 + (sl ^ cls);
#pragma unroll
    for (int r = 0; r < RPL; r++) {
      const int i = lane + 64 * r;
      a[sl][r] = (c < n && i < rows) ? A[row0 + i + ld * c] : R(0);
    }
  }
}

// Entries of rows [row0, row0 + rows) of a box's ID matrix, evaluated into
// the interleave (range row i on lane i % 64, register i / 64).  The wave's
// columns are read once as wave-uniform values (scalar loads); each row's
// entries are formed in natural column order and then placed in the
// interleaved slots.
template <typename R, int RPL, int CPW>
__device__ inline void eval_columns(const DevGeom& g, const int* act, const int* rowdofs,
                                    int cnt, const double (*pxy)[kNumProxy], double pw,
                                    int row0, int rows, int n, R (&a)[CPW][RPL]) {
  const int lane = threadIdx.x & 63, c0 = CPW * wave_id();
  const int cls = slot_class<CPW>(lane);
  DofGeo cg[CPW];
#pragma unroll
  for (int q = 0; q < CPW; q++) cg[q] = dof_geo(g, act[min(c0 + q, n - 1)]);
#pragma unroll
  for (int r = 0; r < RPL; r++) {
    const int i = lane + 64 * r;
    R nat[CPW];
#pragma unroll
    for (int q = 0; q < CPW; q++) nat[q] = R(0);
    if (i < rows) {
      const IdRow rw = id_row(g, rowdofs, cnt, pxy, pw, row0 + i);
#pragma unroll
      for (int q = 0; q < CPW; q++)
        if (c0 + q < n) nat[q] = static_cast<R>(id_entry(g.pde, rw, cg[q]));
    }
#pragma unroll
    for (int sl = 0; sl < CPW; sl++) {
      const int want = sl ^ cls;
      R v = nat[0];
#pragma unroll
      for (int q = 1; q < CPW; q++) v = (want == q) ? nat[q] : v;
      a[sl][r] = v;
    }
  }
}

// Row compression of a tall ID matrix: chunk q of box b (rows
// [q ch, q ch + rows), ch = ceil(m / chunks)) is reduced by unpivoted
// Householder QR to its n x n triangle R_q, written to rows [q n, q n + n)
// of the box's stack (ld chunks * n, in its scratch slot).  The stack is
// Q^T A for an orthogonal Q, so its pivoted QR has the same R and column
// order as that of A.  Grid-strided over the level's jobs of this shape;
// one barrier per step (double-buffered reflector).
template <typename R, int CPW>
__global__ __launch_bounds__(kXThreads) void k_rowcomp(DevGeom g, DevTree t, DevState s,
                                                       int start, int level) {
  constexpr int RPL = RegShape<CPW>::RPL, MAXM = RegShape<CPW>::MAXM;
  constexpr int MAXN = RegShape<CPW>::MAXN;
  constexpr int kind = CPW == 4 ? kQrcpStack : kQrcpStack8;
  __shared__ R vbuf[2][MAXM];
  __shared__ R sh_tau[2];
  __shared__ int rowdofs[kMaxChunks * kRegMaxRows / 2];
  __shared__ double pxy[4][kNumProxy];
  __shared__ R rdiag[MAXN];  // R_jj of step j (the register keeps alpha)
  if (!pools_ok(s)) return;
  const int tid = threadIdx.x, lane = tid & 63, wv = wave_id();
  const int cls = slot_class<CPW>(lane);
  const int c0 = CPW * wv;
  const int* jobs = s.jobs + (int64_t)2 * kMaxChunks * start;
  // With the randomized ID only its fallbacks need row compression: their
  // (box, chunk) pairs come from the fallback list.
  const bool fb = kind == kQrcpStack && s.rid_mode != 0;
  const int total = fb ? list_count(s, level, kLstFallback) * kMaxChunks : s.njobs[level];
  for (int job = blockIdx.x; job < total; job += gridDim.x) {
    int b, q;
    if (fb) {
      b = list_at(s, start, kLstFallback, job / kMaxChunks);
      q = job % kMaxChunks;
      if (s.qkind[b] != kind || q >= id_chunks(s.m_rows[b], MAXM)) continue;
    } else {
      b = jobs[2 * job];
      q = jobs[2 * job + 1];
      if (s.qkind[b] != kind || s.rid_done[b]) continue;  // rid_done: sketched
    }
    const int n = s.act_cnt[b], m = s.m_rows[b];
    const int nch = id_chunks(m, MAXM);
    const int ch = (m + nch - 1) / nch;
    const int r0 = q * ch, rows = min(ch, m - r0);
    R* S = reinterpret_cast<R*>(s.id_pool) + s.id_off[b];
    const int64_t lds = (int64_t)nch * n;
    const bool active = c0 < n;
    const int* act = s.act_list + (int64_t)b * s.ncap;
    const int cnt = id_setup(g, t, s, b, level, m, rowdofs, pxy);
    __syncthreads();
    R a[CPW][RPL];
    eval_columns<R, RPL, CPW>(g, act, rowdofs, cnt, pxy, proxy_weight(t, b), r0, rows, n, a);
    const int steps = min(rows, n);
    int dmask = 0;
    for (int j = 0; j < steps; j++) {
      if (wv == j / CPW) {
        const int pcls = j % CPW, psl = pcls ^ cls;
        R x[RPL];
        pick_slot<R, RPL, CPW>(a, psl, x);
        const R alpha = row_value<R, RPL>(x, j);
        R ss = 0;
#pragma unroll
        for (int r = 0; r < RPL; r++) ss += (lane + 64 * r > j) ? x[r] * x[r] : R(0);
        R beta, tau, sc;
        householder_scalars(alpha, wave_total(ss), &beta, &tau, &sc);
        write_reflector<R, RPL>(x, j, sc, vbuf[j & 1]);
        dmask |= 1 << pcls;
        if (lane == 0) {
          sh_tau[j & 1] = tau;
          rdiag[j] = beta;
        }
      }
      __syncthreads();
      if (active)
        apply_reflector<R, RPL, CPW>(a, vbuf[j & 1], sh_tau[j & 1], ((dmask >> cls) & 1) == 0);
    }
    // Column c was reduced at step c: rows below its diagonal are zero.
    if (active) {
#pragma unroll
      for (int sl = 0; sl < CPW; sl++) {
        const int c = c0 + (sl ^ cls);
#pragma unroll
        for (int r = 0; r < RPL; r++) {
          const int i = lane + 64 * r;
          if (c < n && i < n)
            S[(int64_t)q * n + i + lds * c] =
                i > c ? R(0) : (i == c && c < steps ? rdiag[c] : a[sl][r]);
        }
      }
    }
    __syncthreads();
  }
}

// Interpolative decomposition by column-pivoted Householder QR, stopped at
// the first |R_jj| < tol |R_00| (ie_mat.cpp:548-588).  The pivot rule,
// Householder form and column-norm downdating follow LAPACK dgeqp3/dlaqp2.
// Lanes 0..CPW-1 of each wave keep the downdated norms, logical position
// and pivot key of the wave's columns; the pivot is the largest key
// (pivot_key), found by one wave maximum; the wave owning the pivot column
// forms the reflector and the stop decision.
template <typename R, int CPW, int RPL, int NW>
__device__ __forceinline__ void qrcp_id_box(const DevGeom& g, const DevTree& t, const DevState& s,
                            int start, int level, int b, double tol_d) {
  constexpr int MAXM = 64 * RPL;
  constexpr int NT = 64 * NW, MAXN = CPW * NW, KH = (MAXN + 63) / 64;
  __shared__ R vbuf[MAXM];
  __shared__ R rowj[MAXN];    // row j after step j (downdating)
  __shared__ R rdiag[MAXN];   // R_jj of step j (the register keeps alpha)
  __shared__ R sh_tau, sh_thresh;
  __shared__ int sh_stop;
  __shared__ double keys[MAXN];
  __shared__ R rbuf[MAXN * MAXN];
  __shared__ int colmap[MAXN];
  __shared__ int rowdofs[MAXM / 2];
  __shared__ double pxy[4][kNumProxy];
  if (s.rid_done[b]) return;  // the wide sketch took it

  const int kind = s.qkind[b];
  const int n = s.act_cnt[b];
  int m = s.m_rows[b];
  const R tol = static_cast<R>(tol_d);
  const int tid = threadIdx.x, lane = tid & 63, wv = wave_id();
  const int cls = slot_class<CPW>(lane);
  const int c0 = CPW * wv;
  const bool active = c0 < n;
  const int* act = s.act_list + (int64_t)b * s.ncap;

  R a[CPW][RPL];
  if (kind == kQrcpReg || kind == kQrcpReg8) {
    const int cnt = id_setup(g, t, s, b, level, m, rowdofs, pxy);
    __syncthreads();
    eval_columns<R, RPL, CPW>(g, act, rowdofs, cnt, pxy, proxy_weight(t, b), 0, m, n, a);
  } else {
    // The stacked triangles of the row compression.
    m = id_chunks(m, MAXM) * n;
    load_columns<R, RPL, CPW>(reinterpret_cast<const R*>(s.id_pool) + s.id_off[b], m, 0, m,
                              n, a);
  }
  // Lanes 0..CPW-1: column c0 + lane.
  const int myc = c0 + (lane % CPW);
  R vn1 = 0, vn2 = 0;
  int pos = myc;
  bool avail = lane < CPW && myc < n;
  if (active) {
    vn1 = sqrt(tail_sq<R, RPL, CPW, MAXN>(a, -1));
    vn2 = vn1;
  }
  if (lane < CPW) keys[myc] = avail ? pivot_key(vn1, pos) : -1.0;
  __syncthreads();

  const int kmax = min(m, n);
  int k = kmax;
  R thresh = 0;
  int dmask = 0;  // classes of this wave already pivoted
  for (int j = 0; j < kmax; j++) {
    double kv[KH];
    double best = -1.0;
#pragma unroll
    for (int h = 0; h < KH; h++) {
      kv[h] = (lane + 64 * h < MAXN) ? keys[lane + 64 * h] : -1.0;
      best = fmax(best, kv[h]);
    }
    best = wave_max(best);
    if (!(best >= 0)) {  // every wave sees the same keys
      k = j;
      break;
    }
    int p = -1;
#pragma unroll
    for (int h = KH - 1; h >= 0; h--) {
      const uint64_t hit = __ballot(kv[h] == best);
      if (hit) p = 64 * h + __ffsll((unsigned long long)hit) - 1;
    }
    const int q = pivot_key_pos(best);
    if (lane < CPW) {
      if (myc == p) {
        pos = j;
        avail = false;
      } else if (pos == j) {
        pos = q;
      }
    }
    if (wv == p / CPW) {
      const int psl = (p % CPW) ^ cls;
      R x[RPL];
      pick_slot<R, RPL, CPW>(a, psl, x);
      // With four columns per wave j < 64: row j sits in register 0.
      const R alpha = MAXN <= 64 ? lane_value(x[0], j) : row_value<R, RPL>(x, j);
      R ss = 0;
#pragma unroll
      for (int r = 0; r < RPL; r++) {
        if (MAXN <= 64 && r > 0) ss += x[r] * x[r];
        else ss += (lane + 64 * r > j) ? x[r] * x[r] : R(0);
      }
      R beta, tau, sc;
      householder_scalars(alpha, wave_total(ss), &beta, &tau, &sc);
      write_reflector<R, RPL>(x, j, sc, vbuf);
      const R th = (j == 0) ? fabs(tol * beta) : thresh;
      const int stop = (j >= 1 && fabs(beta) < th) ? 1 : 0;
      if (!stop) dmask |= 1 << (p % CPW);
      if (lane == 0) {
        sh_tau = tau;
        sh_stop = stop;
        rdiag[j] = beta;
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
      apply_reflector<R, RPL, CPW>(a, vbuf, sh_tau, ((dmask >> cls) & 1) == 0);
      publish_row<R, RPL, CPW, MAXN>(a, j, rowj);
      wave_sync();
      bool need = false;
      if (avail && vn1 != 0) {
        R temp = fabs(rowj[myc]) * recip_any(vn1);
        temp = 1 - temp * temp;
        temp = temp > 0 ? temp : R(0);
        const R rat = vn1 * recip_any(vn2);
        if (temp * rat * rat <= Eps<R>::tol3z) need = true;
        else vn1 = vn1 * sqrt(temp);
      }
      if (__ballot(need)) {  // dlaqp2 recomputes the norm of rows j+1..
        const R t2 = tail_sq<R, RPL, CPW, MAXN>(a, j);
        if (need) {
          const R nv = (j < m - 1) ? sqrt(t2) : R(0);
          vn1 = nv;
          vn2 = nv;
        }
      }
      if (lane < CPW) keys[myc] = avail ? pivot_key(vn1, pos) : -1.0;
    }
    __syncthreads();
  }

  // Logical order and rows 0..k-1 of R (physical column order, ld k).
  if (lane < CPW && myc < n) colmap[pos] = myc;
  if (active) {
#pragma unroll
    for (int sl = 0; sl < CPW; sl++) {
      const int c = c0 + (sl ^ cls);
#pragma unroll
      for (int r = 0; r < RPL; r++) {
        const int i = lane + 64 * r;
        if (c < n && i < k) rbuf[i + k * c] = a[sl][r];
      }
    }
  }
  __syncthreads();
  if (tid < k) {
    rbuf[tid + k * colmap[tid]] = rdiag[tid];
    rdiag[tid] = R(1) / rdiag[tid];  // reciprocal diagonal for the solve below
  }
  __syncthreads();

  int* out = s.out_list + (int64_t)b * s.ncap;
  if (k == 0) {
    for (int i = tid; i < n; i += NT) out[i] = act[i];
    if (tid == 0) {
      s.rank[b] = 0;
      s.out_cnt[b] = n;
    }
    return;
  }
  // T = R11^{-1} R12: one wave per column, lane l holds rows l, l + 64, ...
  const int r = n - k;
  R* F = reinterpret_cast<R*>(s.fac_pool) + s.fac_off[b];
  for (int cp = wv; cp < r; cp += NW) {
    const int pcol = colmap[k + cp];
    R tv[KH];
#pragma unroll
    for (int h = 0; h < KH; h++) {
      const int i = lane + 64 * h;
      tv[h] = (i < k) ? rbuf[i + k * pcol] : R(0);
    }
    for (int i = k - 1; i >= 0; i--) {
      const int ci = colmap[i];
      const R ti = row_value<R, KH>(tv, i) * rdiag[i];
#pragma unroll
      for (int h = 0; h < KH; h++) {
        const int row = lane + 64 * h;
        if (row == i) tv[h] = ti;
        if (row < i) tv[h] -= rbuf[row + k * ci] * ti;
      }
    }
#pragma unroll
    for (int h = 0; h < KH; h++) {
      const int row = lane + 64 * h;
      if (row < k) F[row + (int64_t)k * cp] = tv[h];
    }
  }
  int* red = s.red_list + (int64_t)b * s.ncap;
  int* pm = s.perm + (int64_t)b * s.ncap;
  for (int i = tid; i < n; i += NT) {
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

// Register shapes (columns per wave x row registers per lane x waves):
// 4 x 8 x 8 for n <= 32 (two boxes share a CU), 4 x 8 x 16 for n <= 64,
// 8 x 4 x 16 for n <= 128 with m <= 256.  Four waves per SIMD are needed to
// hide the FP64 latency of the step loop: an 8 x 8 x 8 shape (256
// registers, two waves per SIMD) measured 22% slower than 4 x 8 x 16.
template <int CPW, int RPL, int NW>
struct QrcpShape {
  static constexpr int kMaxN = CPW * NW;
  static constexpr int kList = kMaxN > 64 ? kLstReg8 : (kMaxN <= 32 ? kLstRegHalf : kLstReg);
  static constexpr int kWavesPerSimd = CPW * RPL > 32 ? 2 : 4;  // register budget
};

// One workgroup per list entry for the n <= 64 shapes (a loop over boxes
// would keep more values live across the step loop than the registers
// hold); the few n <= 128 boxes share a small strided grid, and so do the
// randomized ID's few fallbacks (STRIDE: the loop's spills stay there).
template <typename R, int CPW, int RPL, int NW, bool STRIDE, bool LOOP = true>
__global__ __launch_bounds__(64 * NW, (QrcpShape<CPW, RPL, NW>::kWavesPerSimd)) void k_qrcp_id(
    DevGeom g, DevTree t, DevState s, int start, int level, double tol_d) {
  constexpr int L = QrcpShape<CPW, RPL, NW>::kList;
  if (!pools_ok(s)) return;
  const int nl = list_count(s, level, L);
  if constexpr (!STRIDE) {
    if ((int)blockIdx.x < nl)
      qrcp_id_box<R, CPW, RPL, NW>(g, t, s, start, level, list_at(s, start, L, blockIdx.x), tol_d);
  } else {
    // With the randomized ID the n <= 64 shapes take only its fallbacks.
    const bool fb = L != kLstReg8 && s.rid_mode != 0;
    const int LL = fb ? kLstFallback : L;
    const int cnt = list_count(s, level, LL);
    if constexpr (LOOP) {
      for (int li = blockIdx.x; li < cnt; li += gridDim.x) {
        const int b = list_at(s, start, LL, li);
        if (fb && (L == kLstRegHalf) != (s.act_cnt[b] <= kHalfCols)) continue;
        qrcp_id_box<R, CPW, RPL, NW>(g, t, s, start, level, b, tol_d);
        __syncthreads();
      }
    } else {  // the grid covers the list: one box per workgroup
      if ((int)blockIdx.x >= cnt) return;
      const int b = list_at(s, start, LL, blockIdx.x);
      if (fb && (L == kLstRegHalf) != (s.act_cnt[b] <= kHalfCols)) return;
      qrcp_id_box<R, CPW, RPL, NW>(g, t, s, start, level, b, tol_d);
    }
  }
}

// ---------------------------------------------------------------------------
// Randomized interpolative decomposition.  The sketch Y = Omega A (l x n,
// l = sk_w[b] in {16, 32}) is formed on the matrix cores: A's entries are
// evaluated straight into the B operand of v_mfma_{f64,f32}_16x16x4 (lane:
// row k0 + lane / 16, column lane % 16 of a 16-column tile) and Omega, a
// fixed Gaussian table indexed by the ID row, is the A operand.  Y moves
// through LDS to one column per lane and is factored by column-pivoted
// Householder QR with the reference's rank rule (ie_mat.cpp:561-568)
// applied to the sketch, |R_jj| corrected for the dimensions the sketch
// has left; T = R11^{-1} R12 of the sketch is the ID of A (Halko,
// Martinsson & Tropp 2011, sec. 5.2).  A box whose sketch shows no cut
// within l - kRidOver pivots leaves rid_done = 0 and the full QRCP kernels
// take it.
// ---------------------------------------------------------------------------

// Output layout of the 16x16x4 MFMAs: lane l, element v holds D(i, l % 16)
// with i = l / 16 + 4 v (f64) or 4 (l / 16) + v (f32).
template <typename R>
struct Mfma16 {
  typedef R V4 __attribute__((ext_vector_type(4)));
  __device__ static int row(int q, int v) { return sizeof(R) == 8 ? q + 4 * v : 4 * q + v; }
};
__device__ inline Mfma16<double>::V4 mfma16x16x4(double a, double b, Mfma16<double>::V4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline Mfma16<float>::V4 mfma16x16x4(float a, float b, Mfma16<float>::V4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ inline DofGeo pick(bool first, const DofGeo& a, const DofGeo& b) {
  DofGeo d;
  d.x = first ? a.x : b.x;
  d.y = first ? a.y : b.y;
  d.n1 = first ? a.n1 : b.n1;
  d.n2 = first ? a.n2 : b.n2;
  d.w = first ? a.w : b.w;
  d.k = first ? a.k : b.k;
  d.c = first ? a.c : b.c;
  return d;
}

// Rows [s0, s1) of one segment of a box's ID matrix (all rows of a
// segment share their role: TG the row point is the target, NEAR a
// boundary DOF) times Omega into the sketch tiles acc0 (rows 0..15) and
// acc1 (rows 16..31, when two).  Row geometry comes from the LDS table at
// index idx0 + (r - s0) (>> 1 for proxy rows of Stokes); the next 16
// rows' Omega operands are loaded before this iteration's entries.
template <typename R, int PDE, bool TG, bool NEAR>
__device__ __forceinline__ void rid_segment(const double (*tab)[kRidMaxNear + kNumProxy],
                                            const int* rowdofs, const double* omega, int s0,
                                            int s1, int idx0, const DofGeo& cg, bool cv,
                                            bool two, typename Mfma16<R>::V4& acc0,
                                            typename Mfma16<R>::V4& acc1) {
  constexpr int LM = kRidRows;
  constexpr bool stokes = PDE == kPdeStokes;
  const int lane = threadIdx.x & 63, cl = lane & 15, kq = lane >> 4;
  auto load_omega = [&](int k0, R* w0, R* w1) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int r = k0 + 4 * u + kq;
      const double* om = omega + (int64_t)min(r, s1 - 1) * LM + cl;
      w0[u] = r < s1 ? static_cast<R>(om[0]) : R(0);
      w1[u] = (two && r < s1) ? static_cast<R>(om[16]) : R(0);
    }
  };
  R nw0[4], nw1[4];
  load_omega(s0, nw0, nw1);
  for (int k0 = s0; k0 < s1; k0 += 16) {
    R w0[4], w1[4], a[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      w0[u] = nw0[u];
      w1[u] = nw1[u];
    }
    if (k0 + 16 < s1) load_omega(k0 + 16, nw0, nw1);
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int r = min(k0 + 4 * u + kq, s1 - 1);
      const int rel = r - s0;
      const int idx = idx0 + ((stokes && !NEAR) ? rel >> 1 : rel);
      DofGeo rg;
      rg.x = tab[0][idx];
      rg.y = tab[1][idx];
      rg.n1 = tab[2][idx];
      rg.n2 = tab[3][idx];
      rg.w = tab[4][idx];
      rg.k = tab[5][idx];
      rg.c = stokes ? (NEAR ? rowdofs[rel] & 1 : rel & 1) : 0;
      const double e = TG ? kernel_entry(PDE, rg, cg, NEAR) : kernel_entry(PDE, cg, rg, NEAR);
      a[u] = (cv && k0 + 4 * u + kq < s1) ? static_cast<R>(e) : R(0);
    }
    if (two) {
#pragma unroll
      for (int u = 0; u < 4; u++) {
        acc0 = mfma16x16x4(w0[u], a[u], acc0);
        acc1 = mfma16x16x4(w1[u], a[u], acc1);
      }
    } else {
#pragma unroll
      for (int u = 0; u < 4; u++) acc0 = mfma16x16x4(w0[u], a[u], acc0);
    }
  }
}

// One workgroup (4 waves) per box.  The box's row geometry (near DOFs,
// then the proxy circle) is staged in LDS; wave w sketches the 16-column
// tile w (n <= 64) segment by segment (near target rows, near source rows,
// proxy target rows, proxy source rows).  The tiles of Y go to LDS, and
// wave 0 alone runs the pivoted QR of Y, the other waves ending there.
// QR layout: lane c holds column c's active rows, shifted so that the
// step's row is always register 0; row j of R leaves the registers for
// LDS (rr[j][c]) at step j, where the T solve finds R11 and R12.  The
// pivot column reaches every lane through readlane (uniform registers),
// not LDS: 93 registers, five workgroups per CU.
template <typename R, int PDE>
__global__ __launch_bounds__(256, 5) void k_rid(DevGeom g, DevTree t, DevState s, int start,
                                             int level, double tol_d) {
  constexpr int LM = kRidRows, YLD = 65, TR = kRidMaxNear + kNumProxy;
  using V4 = typename Mfma16<R>::V4;
  // Row geometry x, y, n1, n2, w, k during the sketch; Y (row-major, ld
  // YLD), then the rows of R, afterwards.
  __shared__ double tab[6][TR];
  __shared__ int rowdofs[kRidMaxNear];
  __shared__ int spiv[64];  // pivot column (lane) of each step
  static_assert(LM * YLD * sizeof(R) <= sizeof(tab), "Y fits the row table");
  R* ybuf = reinterpret_cast<R*>(&tab[0][0]);
  if (!pools_ok(s)) return;
  if ((int)blockIdx.x >= list_count(s, level, kLstRid)) return;
  const int b = list_at(s, start, kLstRid, blockIdx.x);
  const int tid = threadIdx.x, lane = tid & 63, wv = wave_id();
  const int n = s.act_cnt[b], m = s.m_rows[b], l = s.sk_w[b];
  if (l > kRidRows) return;  // the wide form's box
  const int P = proxy_rows(g, level) / 2;  // rows per proxy block
  const int cnt = (m - 2 * P) / 2;
  const int* act = s.act_list + (int64_t)b * s.ncap;
  if (wv == 0)
    for_each_row_dof(g, t, s, b, level >= 3, [&](int pos, int dof) { rowdofs[pos] = dof; });
  if (level >= 3 && tid < kNumProxy) {
    double px, py, n1, n2;
    proxy_point(t.cx[b], t.cy[b], t.side[b] * kRadiusRatio, tid, &px, &py, &n1, &n2);
    tab[0][cnt + tid] = px;
    tab[1][cnt + tid] = py;
    tab[2][cnt + tid] = n1;
    tab[3][cnt + tid] = n2;
    tab[4][cnt + tid] = proxy_weight(t, b);
    tab[5][cnt + tid] = 0;
  }
  __syncthreads();
  for (int i = tid; i < cnt; i += 256) {
    const DofGeo d = dof_geo(g, rowdofs[i]);
    tab[0][i] = d.x;
    tab[1][i] = d.y;
    tab[2][i] = d.n1;
    tab[3][i] = d.n2;
    tab[4][i] = d.w;
    tab[5][i] = d.k;
  }
  __syncthreads();

  // Sketch: wave wv takes columns 16 wv .. 16 wv + 15.
  const int kq = lane >> 4;
  const int c = 16 * wv + (lane & 15);
  V4 acc0 = V4{0, 0, 0, 0}, acc1 = V4{0, 0, 0, 0};
  if (16 * wv < n) {
    const bool cv = c < n;
    const DofGeo cg = dof_geo(g, act[cv ? c : n - 1]);
    const bool two = l > 16;
    rid_segment<R, PDE, true, true>(tab, rowdofs, s.omega, 0, cnt, 0, cg, cv, two, acc0, acc1);
    rid_segment<R, PDE, false, true>(tab, rowdofs, s.omega, cnt, 2 * cnt, 0, cg, cv, two, acc0,
                                     acc1);
    if (P > 0) {
      rid_segment<R, PDE, true, false>(tab, rowdofs, s.omega, 2 * cnt, 2 * cnt + P, cnt, cg, cv,
                                       two, acc0, acc1);
      rid_segment<R, PDE, false, false>(tab, rowdofs, s.omega, 2 * cnt + P, m, cnt, cg, cv, two,
                                        acc0, acc1);
    }
  }
  __syncthreads();  // the row table becomes Y
  // acc[v] = Y[16 it + row(kq, v)][c]; tiles past n and rows past l are 0.
#pragma unroll
  for (int v = 0; v < 4; v++) {
    ybuf[Mfma16<R>::row(kq, v) * YLD + c] = acc0[v];
    ybuf[(16 + Mfma16<R>::row(kq, v)) * YLD + c] = acc1[v];
  }
  __syncthreads();
  if (wv != 0) return;

  // Pivoted QR of the sketch (wave 0).
  R y[LM];
#pragma unroll
  for (int r = 0; r < LM; r++) y[r] = ybuf[r * YLD + lane];
  R* rr = ybuf;  // rows of R, rr[j * YLD + c]
  const R tol = static_cast<R>(tol_d);
  bool avail = lane < n;
  int k = -1;
  R thresh = 0;
  R tl = 0;  // squared norm of the active rows past the first
  {
    R p4[4] = {0, 0, 0, 0};
#pragma unroll
    for (int r = 1; r < LM; r++) p4[r & 3] += y[r] * y[r];
    tl = (p4[0] + p4[1]) + (p4[2] + p4[3]);
  }
  const int jlim = min(n, l - kRidOver);
  int j = 0;
  for (; j < jlim; j++) {
    const double key = avail ? pivot_key<R, 6>(y[0] * y[0] + tl, lane) : -1.0;
    const double best = wave_max(key);
    const int p = __ffsll((unsigned long long)__ballot(key == best)) - 1;
    const R alpha = lane_value(y[0], p), xt = lane_value(tl, p);
    R beta, tau, sc;
    householder_scalars(alpha, xt, &beta, &tau, &sc);
    // A sketch of l rows keeps l - j dimensions at step j, which shrinks
    // |R_jj| by about sqrt((l - j) / l) against |R_00|: the cut compares
    // |R_jj| sqrt(l / (l - j)), with a factor 2 margin, to tol |R_00|.
    if (j == 0) {
      thresh = fabs(tol * beta);
    } else if (2 * fabs(beta) * sqrt(static_cast<R>(l) / static_cast<R>(l - j)) < thresh) {
      k = j;
      break;
    }
    if (lane == p) {
      spiv[j] = p;
      avail = false;
    }
    // The pivot column's entries below row j, uniform across the wave:
    // the Householder vector is (1, sc yp[1], ..., sc yp[LM-1]).
    R yp[LM];
#pragma unroll
    for (int r = 1; r < LM; r++) yp[r] = lane_value(y[r], p);
    R d4[4] = {0, 0, 0, 0};
#pragma unroll
    for (int r = 1; r < LM; r++) d4[r & 3] += yp[r] * y[r];
    const R w = tau * (y[0] + sc * ((d4[0] + d4[1]) + (d4[2] + d4[3])));
    const R ws = w * sc;
    // Row j of R leaves; the rest shift up one register.
    rr[j * YLD + lane] = (lane == p) ? beta : y[0] - w;
    R t4[4] = {0, 0, 0, 0};
#pragma unroll
    for (int r = 0; r < LM - 1; r++) {
      y[r] = y[r + 1] - ws * yp[r + 1];
      if (r > 0) t4[r & 3] += y[r] * y[r];
    }
    y[LM - 1] = 0;
    tl = (t4[0] + t4[1]) + (t4[2] + t4[3]);
  }
  wave_sync();
  if (k < 0) {
    if (j < n) {  // the sketch ran out before a cut: the full QRCP takes it
      if (lane == 0) {
        const int pos = atomicAdd(list_slot(s, level, kLstFallback), 1);
        s.lists[(int64_t)kLstFallback * s.nbox + start + pos] = b;
      }
      return;
    }
    k = n;  // every column taken: nothing redundant
  }

  // T = R11^{-1} R12 by a column sweep in registers: lane c holds column c
  // of R (rows 0..k-1), R11(i, q) = rr[i][spiv[q]] is read from LDS; the
  // redundant lanes' columns become T.
  constexpr int KMAX = LM - kRidOver;
  const bool red = avail;
  R x[KMAX];
#pragma unroll
  for (int i = 0; i < KMAX; i++) x[i] = i < k ? rr[i * YLD + lane] : R(0);
#pragma unroll
  for (int q = KMAX - 1; q >= 0; q--) {
    if (q < k) {
      const int pq = spiv[q];
      const R tq = x[q] * recip_any(rr[q * YLD + pq]);
      x[q] = tq;
#pragma unroll
      for (int i = 0; i < q; i++) x[i] -= rr[i * YLD + pq] * tq;
    }
  }
  const uint64_t redmask = __ballot(red);
  const int cp = __popcll(redmask & ((1ull << lane) - 1));  // redundant index
  R* F = reinterpret_cast<R*>(s.fac_pool) + s.fac_off[b];
  if (red) {
#pragma unroll
    for (int i = 0; i < KMAX; i++)
      if (i < k) F[i + (int64_t)k * cp] = x[i];
  }
  int* out = s.out_list + (int64_t)b * s.ncap;
  int* rl = s.red_list + (int64_t)b * s.ncap;
  int* pm = s.perm + (int64_t)b * s.ncap;
  if (lane < k) {
    const int cc = spiv[lane];
    pm[lane] = cc;
    out[lane] = act[cc];
  }
  if (red) {
    pm[k + cp] = lane;
    rl[cp] = act[lane];
  }
  if (lane == 0) {
    s.rank[b] = k;
    s.out_cnt[b] = k;
    s.rid_done[b] = 1;
  }
}

// Wide randomized ID: the boxes of the top levels too large for k_rid (up
// to kRidMaxNearW row DOFs and kRidWideCols columns, ranks up to kRidRowsW
// - kRidOver).  Same sketch, rank rule and outputs as k_rid with a sketch
// of kRidRowsW rows from its own table, on 8 waves: wave w sketches the
// 16-column tile w % (4 QW) over every (8 / 4 QW)-th 16-row block, the
// row parts added in a fixed order.  QW waves then run the pivoted QR,
// wave q holding columns 64 q + lane: QW = 1 for n <= 64 (one wave, no
// barrier), QW = 2 for n <= 128 (each step's pivot found from both waves'
// maxima through LDS, the pivot column's Householder vector published by
// its wave; two barriers per step).  A box whose sketch shows no cut keeps
// rid_done = 0: a register-QRCP box joins the fallback list, any other is
// still on its full-QRCP list (whose kernels skip the boxes the sketch
// took).
template <typename R, int PDE, bool TG, bool NEAR>
__device__ __forceinline__ void rid_segment_w(const double (*tab)[kRidMaxNearW + kNumProxy],
                                              const int* rowdofs, const double* omega, int s0,
                                              int s1, int idx0, int part, int parts,
                                              const DofGeo& cg, bool cv,
                                              typename Mfma16<R>::V4 (&acc)[4]) {
  constexpr int LM = kRidRowsW;
  constexpr bool stokes = PDE == kPdeStokes;
  const int lane = threadIdx.x & 63, cl = lane & 15, kq = lane >> 4;
  // the next block's Omega operands are loaded before this block's entries
  auto load_omega = [&](int k0, R (&w)[4][4]) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int r = k0 + 4 * u + kq;
      const double* om = omega + (int64_t)min(r, s1 - 1) * LM + cl;
#pragma unroll
      for (int q = 0; q < 4; q++) w[q][u] = r < s1 ? static_cast<R>(om[16 * q]) : R(0);
    }
  };
  const int step = 16 * parts;
  R nw[4][4];
  load_omega(s0 + 16 * part, nw);
  for (int k0 = s0 + 16 * part; k0 < s1; k0 += step) {
    R w[4][4], a[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int q = 0; q < 4; q++) w[q][u] = nw[q][u];
    if (k0 + step < s1) load_omega(k0 + step, nw);
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int r = k0 + 4 * u + kq;
      const int rel = min(r, s1 - 1) - s0;
      const int idx = idx0 + ((stokes && !NEAR) ? rel >> 1 : rel);
      DofGeo rg;
      rg.x = tab[0][idx];
      rg.y = tab[1][idx];
      rg.n1 = tab[2][idx];
      rg.n2 = tab[3][idx];
      rg.w = tab[4][idx];
      rg.k = tab[5][idx];
      rg.c = stokes ? (NEAR ? rowdofs[rel] & 1 : rel & 1) : 0;
      const double e = TG ? kernel_entry(PDE, rg, cg, NEAR) : kernel_entry(PDE, cg, rg, NEAR);
      a[u] = (cv && r < s1) ? static_cast<R>(e) : R(0);
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int q = 0; q < 4; q++) acc[q] = mfma16x16x4(w[q][u], a[u], acc[q]);
  }
}

constexpr int kRidWideWaves = 8;

// LDS of the wide sketch: the row table, then Y, then the rows of R (a
// raw buffer sized for two QR waves), and the QR's small arrays.
template <typename R>
struct RidWideLds {
  static constexpr int TR = kRidMaxNearW + kNumProxy;
  static constexpr int YMAX = kRidRowsW * (2 * 64 + 1);  // two QR waves' Y
  static constexpr int kBuf = (YMAX * (int)sizeof(R) + 7) / 8 > 6 * TR
                                  ? (YMAX * (int)sizeof(R) + 7) / 8
                                  : 6 * TR;
  double buf[kBuf];
  int rowdofs[kRidMaxNearW];
  R vbuf[4][kRidRowsW];  // two QR waves: each wave's candidate Householder
                         // vector, by step parity
  int spiv[64];
  double kbest[4];  // two QR waves: each wave's pivot key, by step parity
  R hs[4][2];       // two QR waves: each candidate's beta and tau
  int nred0;        // two QR waves: redundant columns of wave 0
};

template <typename R, int PDE, int QW>
__device__ __forceinline__ void rid_wide_box(const DevGeom& g, const DevTree& t,
                                             const DevState& s, int start, int level,
                                             double tol_d, int b, RidWideLds<R>& sh) {
  constexpr int kTiles = 4 * QW, kParts = kRidWideWaves / kTiles;
  constexpr int NT = 64 * kRidWideWaves;
  constexpr int LM = kRidRowsW, YLD = 64 * QW + 1, TR = RidWideLds<R>::TR;
  using V4 = typename Mfma16<R>::V4;
  double (*tab)[TR] = reinterpret_cast<double (*)[TR]>(sh.buf);
  R* ybuf = reinterpret_cast<R*>(sh.buf);
  int* rowdofs = sh.rowdofs;
  int* spiv = sh.spiv;
  double* kbest = sh.kbest;
  const int n = s.act_cnt[b], m = s.m_rows[b], l = s.sk_w[b];
  const int tid = threadIdx.x, lane = tid & 63, wv = wave_id();
  const int P = proxy_rows(g, level) / 2;
  const int cnt = (m - 2 * P) / 2;
  const int* act = s.act_list + (int64_t)b * s.ncap;
  if (wv == 0)
    for_each_row_dof(g, t, s, b, level >= 3, [&](int pos, int dof) { rowdofs[pos] = dof; });
  if (level >= 3 && tid < kNumProxy) {
    double px, py, n1, n2;
    proxy_point(t.cx[b], t.cy[b], t.side[b] * kRadiusRatio, tid, &px, &py, &n1, &n2);
    tab[0][cnt + tid] = px;
    tab[1][cnt + tid] = py;
    tab[2][cnt + tid] = n1;
    tab[3][cnt + tid] = n2;
    tab[4][cnt + tid] = proxy_weight(t, b);
    tab[5][cnt + tid] = 0;
  }
  __syncthreads();
  for (int i = tid; i < cnt; i += NT) {
    const DofGeo d = dof_geo(g, rowdofs[i]);
    tab[0][i] = d.x;
    tab[1][i] = d.y;
    tab[2][i] = d.n1;
    tab[3][i] = d.n2;
    tab[4][i] = d.w;
    tab[5][i] = d.k;
  }
  __syncthreads();

  // Sketch: wave wv takes tile wv % kTiles, row part wv / kTiles.
  const int kq = lane >> 4;
  const int tile = wv % kTiles, part = wv / kTiles;
  V4 acc[4];
#pragma unroll
  for (int q = 0; q < 4; q++) acc[q] = V4{0, 0, 0, 0};
  const int c = 16 * tile + (lane & 15);
  if (16 * tile < n) {
    const bool cv = c < n;
    const DofGeo cg = dof_geo(g, act[cv ? c : n - 1]);
    rid_segment_w<R, PDE, true, true>(tab, rowdofs, s.omega_w, 0, cnt, 0, part, kParts, cg, cv,
                                      acc);
    rid_segment_w<R, PDE, false, true>(tab, rowdofs, s.omega_w, cnt, 2 * cnt, 0, part, kParts,
                                       cg, cv, acc);
    if (P > 0) {
      rid_segment_w<R, PDE, true, false>(tab, rowdofs, s.omega_w, 2 * cnt, 2 * cnt + P, cnt,
                                         part, kParts, cg, cv, acc);
      rid_segment_w<R, PDE, false, false>(tab, rowdofs, s.omega_w, 2 * cnt + P, m, cnt, part,
                                          kParts, cg, cv, acc);
    }
  }
  __syncthreads();  // the row table becomes Y
  for (int p = 0; p < kParts; p++) {  // the row parts added in order
    if (part == p) {
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          R* y = &ybuf[(16 * q + Mfma16<R>::row(kq, v)) * YLD + c];
          *y = p == 0 ? acc[q][v] : *y + acc[q][v];
        }
    }
    __syncthreads();
  }
  if (wv >= QW) return;

  // Pivoted QR of the sketch: wave wv holds column col = 64 wv + lane.
  const int col = 64 * wv + lane;
  R y[LM];
#pragma unroll
  for (int r = 0; r < LM; r++) y[r] = ybuf[r * YLD + col];
  R* rr = ybuf;  // rows of R, rr[j * YLD + c]
  const R tol = static_cast<R>(tol_d);
  bool av = col < n;
  R tl;
  {
    R p4[4] = {0, 0, 0, 0};
#pragma unroll
    for (int r = 1; r < LM; r++) p4[r & 3] += y[r] * y[r];
    tl = (p4[0] + p4[1]) + (p4[2] + p4[3]);
  }
  int k = -1;
  R thresh = 0;
  const int jlim = min(n, l - kRidOver);
  int j = 0;
  for (; j < jlim; j++) {
    double best = wave_max(av ? pivot_key<R, 7>(y[0] * y[0] + tl, col) : -1.0);
    int p, pw, pl;
    R beta = 0, tau = 0, sc = 0;
    const R* vb = nullptr;
    if (QW == 2) {
      // Each wave publishes its own best column's Householder vector before
      // the step's one barrier; after it both read the winner's (by step
      // parity: a wave a step ahead writes the other slot).
      const int par = j & 1;
      if (best >= 0) {
        const int ql = pivot_key_pos<7>(best) & 63;
        const R alpha = lane_value(y[0], ql), xt = lane_value(tl, ql);
        R b0, t0, s0;
        householder_scalars(alpha, xt, &b0, &t0, &s0);
        if (lane == ql) {
          R* w = sh.vbuf[2 * par + wv];
          w[0] = R(1);
#pragma unroll
          for (int r = 1; r < LM; r++) w[r] = y[r] * s0;
          sh.hs[2 * par + wv][0] = b0;
          sh.hs[2 * par + wv][1] = t0;
        }
      }
      if (lane == 0) kbest[2 * par + wv] = best;
      __syncthreads();
      best = fmax(kbest[2 * par], kbest[2 * par + 1]);
      p = pivot_key_pos<7>(best);  // column index
      pw = p >> 6;
      pl = p & 63;
      beta = sh.hs[2 * par + pw][0];
      tau = sh.hs[2 * par + pw][1];
      vb = sh.vbuf[2 * par + pw];
      if (wv == 0 && lane == 0) spiv[j] = p;
    } else {
      p = pivot_key_pos<7>(best);  // column index
      pw = p >> 6;
      pl = p & 63;
      const R alpha = lane_value(y[0], pl), xt = lane_value(tl, pl);
      householder_scalars(alpha, xt, &beta, &tau, &sc);
    }
    if (j == 0) {
      thresh = fabs(tol * beta);
    } else if (2 * fabs(beta) * sqrt(static_cast<R>(l) / static_cast<R>(l - j)) < thresh) {
      k = j;
      break;
    }
    if (col == p) av = false;
    R t4[4] = {0, 0, 0, 0};
    if (QW == 1) {
      // One wave: the pivot column reaches every lane through readlane
      // (twice: kept between the dot and the update it would not fit the
      // registers beside y); the Householder vector is (1, sc yp[1..]).
      if (lane == pl) spiv[j] = p;
      R d4[4] = {0, 0, 0, 0};
#pragma unroll
      for (int r = 1; r < LM; r++) d4[r & 3] += lane_value(y[r], pl) * y[r];
      const R w = tau * (y[0] + sc * ((d4[0] + d4[1]) + (d4[2] + d4[3])));
      const R ws = w * sc;
      rr[j * YLD + col] = (col == p) ? beta : y[0] - w;
      int pl2 = pl;
      __asm__ volatile("" : "+s"(pl2));
#pragma unroll
      for (int r = 0; r < LM - 1; r++) {
        y[r] = y[r + 1] - ws * lane_value(y[r + 1], pl2);
        if (r > 0) t4[r & 3] += y[r] * y[r];
      }
    } else {
      R d4[4] = {0, 0, 0, 0};
#pragma unroll
      for (int r = 0; r < LM; r++) d4[r & 3] += vb[r] * y[r];
      const R w = tau * ((d4[0] + d4[1]) + (d4[2] + d4[3]));
      rr[j * YLD + col] = (col == p) ? beta : y[0] - w;
      // the vector is read again below rather than kept in registers beside y
      __asm__ volatile("" ::: "memory");
#pragma unroll
      for (int r = 0; r < LM - 1; r++) {
        y[r] = y[r + 1] - w * vb[r + 1];
        if (r > 0) t4[r & 3] += y[r] * y[r];
      }
    }
    y[LM - 1] = 0;
    tl = (t4[0] + t4[1]) + (t4[2] + t4[3]);
    if (QW == 1) wave_sync();
  }
  const bool regq = s.qkind[b] == kQrcpReg || s.qkind[b] == kQrcpStack;
  if (k < 0) {
    if (j < n) {  // no cut within the sketch: the full QRCP takes the box
      if (regq && tid == 0) {
        const int pos = atomicAdd(list_slot(s, level, kLstFallback), 1);
        s.lists[(int64_t)kLstFallback * s.nbox + start + pos] = b;
      }
      return;
    }
    k = n;
  }
  // Every wave's rows of R are in LDS; wave 0's redundant count for wave 1.
  const uint64_t mk = __ballot(av);
  if (QW == 2) {
    if (wv == 0 && lane == 0) sh.nred0 = __popcll(mk);
    __syncthreads();
  }

  // T = R11^{-1} R12 by a column sweep in registers, R11 read from LDS.
  constexpr int KMAX = LM - kRidOver;
  R x[KMAX];
#pragma unroll
  for (int i = 0; i < KMAX; i++) x[i] = i < k ? rr[i * YLD + col] : R(0);
#pragma unroll
  for (int q = KMAX - 1; q >= 0; q--) {
    if (q < k) {
      const int pq = spiv[q];
      const R tq = x[q] * recip_any(rr[q * YLD + pq]);
      x[q] = tq;
#pragma unroll
      for (int i = 0; i < q; i++) x[i] -= rr[i * YLD + pq] * tq;
    }
  }
  const int cp = (wv == 1 ? sh.nred0 : 0) + __popcll(mk & ((1ull << lane) - 1));  // redundant index
  R* F = reinterpret_cast<R*>(s.fac_pool) + s.fac_off[b];
  int* out = s.out_list + (int64_t)b * s.ncap;
  int* rl = s.red_list + (int64_t)b * s.ncap;
  int* pm = s.perm + (int64_t)b * s.ncap;
  if (av) {
#pragma unroll
    for (int i = 0; i < KMAX; i++)
      if (i < k) F[i + (int64_t)k * cp] = x[i];
    pm[k + cp] = col;
    rl[cp] = act[col];
  }
  if (wv == 0 && lane < k) {
    const int cc = spiv[lane];
    pm[lane] = cc;
    out[lane] = act[cc];
  }
  if (tid == 0) {
    s.rank[b] = k;
    s.out_cnt[b] = k;
    s.rid_done[b] = 1;
  }
}

// One launch for both forms: boxes with n <= 64 run the one-wave QR,
// the others the two-wave QR.
template <typename R, int PDE>
__global__ __launch_bounds__(64 * kRidWideWaves, 1) void k_rid_wide(
    DevGeom g, DevTree t, DevState s, int start, int level, double tol_d) {
  __shared__ RidWideLds<R> sh;
  if (!pools_ok(s)) return;
  if ((int)blockIdx.x >= list_count(s, level, kLstRid)) return;
  const int b = list_at(s, start, kLstRid, blockIdx.x);
  if (s.sk_w[b] != kRidRowsW) return;  // a narrow sketch (k_rid)
  if (s.act_cnt[b] > 64) rid_wide_box<R, PDE, 2>(g, t, s, start, level, tol_d, b, sh);
  else rid_wide_box<R, PDE, 1>(g, t, s, start, level, tol_d, b, sh);
}

// Gaussian table for the sketches: z = sqrt(-2 ln u1) cos(2 pi u2) (Box and
// Muller 1958) from two 53-bit uniforms of a SplitMix64 hash of the index.
__device__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__global__ __launch_bounds__(256) void k_gaussian(double* out, int64_t n, uint64_t seed) {
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const uint64_t h1 = splitmix64(seed ^ (2 * (uint64_t)i));
    const uint64_t h2 = splitmix64(seed ^ (2 * (uint64_t)i + 1));
    const double u1 = ((h1 >> 11) + 1) * (1.0 / 9007199254740992.0);  // (0, 1]
    const double u2 = (h2 >> 11) * (1.0 / 9007199254740992.0);        // [0, 1)
    out[i] = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
  }
}

// General form (any box the register form does not take): the matrix lives
// in LDS when it fits, else in the box's scratch slot; each wave owns a
// strided set of columns.
template <typename R, int NT>
__device__ __forceinline__ void qrcp_general_box(const DevState& s, int start, int level, int b, double tol_d) {
  constexpr int NW = NT / 64;
  constexpr int kLdsElems = 131072 / sizeof(R);
  __shared__ R smat[kLdsElems];
  __shared__ R vn1[kMaxBoxDofs];
  __shared__ R vn2[kMaxBoxDofs];
  __shared__ int col[kMaxBoxDofs];
  __shared__ R sh_tau;
  __shared__ int sh_stop;

  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int n = s.act_cnt[b], m = s.m_rows[b];
  const int* act = s.act_list + (int64_t)b * s.ncap;
  int* out = s.out_list + (int64_t)b * s.ncap;
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

template <typename R, int NT>
__global__ __launch_bounds__(NT) void k_qrcp(DevState s, int start, int level, double tol_d) {
  if (!pools_ok(s)) return;
  const int nl = list_count(s, level, kLstLds);
  for (int li = blockIdx.x; li < nl; li += gridDim.x) {
    const int b = list_at(s, start, kLstLds, li);
    if (s.rid_done[b]) continue;  // the wide sketch took it
    qrcp_general_box<R, NT>(s, start, level, b, tol_d);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Gauss-Jordan inversion with partial pivoting.  Step j pivots on the row
// with the largest |entry| of column j among rows not yet pivoted (first
// index on ties, LAPACK dgetrf's choice) without moving rows: with pivot
// row prow[j] at step j the result M satisfies
//   X^{-1}(j, prow[jj]) = M(prow[j], jj).
// Steps alternate between two n x n buffers (ld n) starting in `a`, so M
// lands in `a` for even n and in `b` for odd n.  All NT threads take part;
// every wave picks the pivot itself, so a step needs one barrier.
// Returns 0, or j + 1 for the first zero pivot.  n <= 2048.
// ---------------------------------------------------------------------------

template <typename R, int NT>
__device__ int gj_invert(R* a, R* b, int n, int* prow) {
  const int tid = threadIdx.x, lane = tid & 63;
  uint32_t used = 0;  // bit u: row lane + 64 u already pivoted
  int info = 0;
  R* cur = a;
  R* nxt = b;
  const int i0 = tid % n, c0 = tid / n, di = NT % n, dc = NT / n;
  for (int j = 0; j < n; j++) {
    R best = -1;
    int bi = 0x7fffffff;
    for (int u = 0; lane + 64 * u < n; u++) {
      if ((used >> u) & 1) continue;
      const int i = lane + 64 * u;
      const R v = fabs(cur[i + (int64_t)n * j]);
      const R key = (v == v) ? v : R(-0.5);  // NaN rows after all others
      if (key > best) {
        best = key;
        bi = i;
      }
    }
    const int p = wave_argmax<R>(best, bi);
    if ((p & 63) == lane) used |= 1u << (p >> 6);
    if (tid == 0) prow[j] = p;
    const R piv = cur[p + (int64_t)n * j];
    if (piv == R(0) && info == 0) info = j + 1;
    const R inv = R(1) / piv;
    int i = i0, c = c0;
    while (c < n) {
      const int64_t idx = i + (int64_t)n * c;
      const R rp = ((c == j) ? R(1) : cur[p + (int64_t)n * c]) * inv;
      nxt[idx] = (i == p) ? rp : ((c == j) ? R(0) : cur[idx]) - cur[i + (int64_t)n * j] * rp;
      i += di;
      c += dc;
      if (i >= n) {
        i -= n;
        c++;
      }
    }
    __syncthreads();
    R* t = cur;
    cur = nxt;
    nxt = t;
  }
  return info;
}

// Gauss-Jordan inverse of X (r x r, r <= 16 U, in LDS or global memory,
// ld r) with the matrix in the registers of 256 threads on a 16 x 16 grid,
// thread (tr, tc) holding rows tr + 16 u and columns tc + 16 v (U x U
// entries).  Per step:
//   * every thread reads its rows of the pivot column (kept in LDS by its
//     owners) and the pivot row is the largest packed key (same rule as
//     gj_invert) over the 16 lanes of its DPP row, which together hold all
//     16 U rows, so every lane finds it without a block barrier;
//   * the pivot and the pivot row's entries that a thread needs (columns
//     tc + 16 v) are held by lane tr = p % 16 of the same 16-lane group,
//     read with a lane permute (no LDS store, no barrier);
//   * each entry takes one FMA; column j restarts from zero and the pivot
//     row becomes the scaled pivot row;
//   * the owners of column j + 1 publish it, then one block barrier.
// M (as gj_invert) is written back to X.  colbuf holds 2 x 16 U entries.
// Returns 0, or j + 1 for the first zero pivot.
template <typename R, int U = 4>
__device__ __forceinline__ int gj_invert_regs(R* X, int r, int* prow, R* colbuf) {
  constexpr int RM = 16 * U;
  const int tid = threadIdx.x;
  const int tr = tid & 15, tc = tid >> 4;
  R a[U][U];
#pragma unroll
  for (int u = 0; u < U; u++)
#pragma unroll
    for (int v = 0; v < U; v++) {
      const int i = tr + 16 * u, c = tc + 16 * v;
      a[u][v] = (i < r && c < r) ? X[i + r * c] : R(0);
    }
  if (tc == 0) {
#pragma unroll
    for (int u = 0; u < U; u++)
      if (tr + 16 * u < r) colbuf[tr + 16 * u] = a[u][0];
  }
  __syncthreads();
  unsigned used = 0;  // bit u: row tr + 16 u already pivoted
  int info = 0;
  for (int j = 0; j < r; j++) {
    const R* col = colbuf + RM * (j & 1);
    R* coln = colbuf + RM * ((j + 1) & 1);
    R cv[U];
    double key = -1.0;
    unsigned cand = 0;
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int row = tr + 16 * u;
      cv[u] = row < r ? col[row] : R(0);
      const bool c = row < r && !((used >> u) & 1);
      cand |= c ? (1u << u) : 0u;
      const R mag = fabs(cv[u]);
      key = fmax(key, (c && mag == mag) ? pivot_key(mag, row) : -1.0);
    }
    key = fmax(key, dpp_f64<dpp::kQuadSwap1, 0xF>(key));
    key = fmax(key, dpp_f64<dpp::kQuadSwap2, 0xF>(key));
    key = fmax(key, dpp_f64<dpp::kRowHalfMirror, 0xF>(key));
    key = fmax(key, dpp_f64<dpp::kRowMirror, 0xF>(key));
    int p;
    if (key >= 0) {
      p = pivot_key_pos(key);
    } else {  // NaN column: the first row not yet pivoted
      int f = cand ? tr + 16 * (__ffs(cand) - 1) : 0x7fffffff;
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) f = min(f, __shfl_xor(f, o, 16));
      p = f;
    }
    p = __builtin_amdgcn_readfirstlane(p);  // the same in every lane
    const int up = p >> 4, pr = p & 15;
    if (tr == pr) used |= 1u << up;
    if (tid == 0) prow[j] = p;
    // Lane pr of this 16-lane row holds row p's entries of the row's
    // columns, and cv[up] = the pivot.
    R pc = cv[0];
#pragma unroll
    for (int u = 1; u < U; u++) pc = (u == up) ? cv[u] : pc;
    const int src = (threadIdx.x & 48) | pr;
    const R piv = __shfl(pc, src, 64);
    if (piv == R(0) && info == 0) info = j + 1;
    const R inv = R(1) / piv;
    R rv[U];
#pragma unroll
    for (int v = 0; v < U; v++) {
      const int c = tc + 16 * v;
      R pv = a[0][v];
#pragma unroll
      for (int u = 1; u < U; u++) pv = (u == up) ? a[u][v] : pv;
      pv = __shfl(pv, src, 64);
      rv[v] = c < r ? ((c == j) ? R(1) : pv) * inv : R(0);
    }
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int v = 0; v < U; v++) {
        const bool colj = tc + 16 * v == j, rowp = tr + 16 * u == p;
        const R base = colj ? R(0) : a[u][v];
        const R val = base - cv[u] * rv[v];
        a[u][v] = rowp ? rv[v] : val;
      }
    if (j + 1 < r && tc == ((j + 1) & 15)) {
      const int vn = (j + 1) >> 4;
#pragma unroll
      for (int u = 0; u < U; u++) {
        R x = a[u][0];
#pragma unroll
        for (int v = 1; v < U; v++) x = (v == vn) ? a[u][v] : x;
        if (tr + 16 * u < r) coln[tr + 16 * u] = x;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < U; u++)
#pragma unroll
    for (int v = 0; v < U; v++) {
      const int i = tr + 16 * u, c = tc + 16 * v;
      if (i < r && c < r) X[i + r * c] = a[u][v];
    }
  __syncthreads();
  return info;
}

// Inverts X (r x r) in place through the scratch buffer W (>= r^2): the
// steps start in W when r is odd, so M ends in X.
template <typename R, int NT>
__device__ int gj_invert_into(R* X, R* W, int r, int* prow) {
  if (r & 1) {
    for (int idx = threadIdx.x; idx < r * r; idx += NT) W[idx] = X[idx];
    __syncthreads();
    return gj_invert<R, NT>(W, X, r, prow);
  }
  return gj_invert<R, NT>(X, W, r, prow);
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

// ---------------------------------------------------------------------------
// Decouple one box (skel_factorization.cpp:77-128): with s/r the skeleton
// and redundant positions and K_BN = K(act, act) - (children's carried
// Schur blocks),
//   K_rs = K(r,s) - T^T K(s,s)
//   K_sr = K(s,r) - K(s,s) T
//   X_rr = K(r,r) - T^T K(s,r) - K_rs T
//   L = K_sr X_rr^{-1},  U = X_rr^{-1} K_rs,  schur = K_sr X_rr^{-1} K_rs
// and the box's carried block = (children's update on s,s) + schur.  X_rr
// is inverted by Gauss-Jordan (the reference LU-factors it and solves): the
// factor slot keeps X_rr^{-1}, so L, U, the carried block and the solve's
// diagonal step are plain products.  Uncompressed boxes carry their
// children's blocks unchanged.
// ---------------------------------------------------------------------------

// P (n x n, skeleton positions first) -> K_rs, K_sr, X_rr.
// sum_l a[sa l] b[sb l], l < len, in four partial sums (a dependent FP64
// FMA chain costs its latency per term; global operands add a load
// latency per term besides).
template <typename R>
__device__ inline R dot2(const R* a, int64_t sa, const R* b, int64_t sb, int len) {
  R s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  int l = 0;
  for (; l + 4 <= len; l += 4) {
    s0 += a[sa * l] * b[sb * l];
    s1 += a[sa * (l + 1)] * b[sb * (l + 1)];
    s2 += a[sa * (l + 2)] * b[sb * (l + 2)];
    s3 += a[sa * (l + 3)] * b[sb * (l + 3)];
  }
  for (; l < len; l++) s0 += a[sa * l] * b[sb * l];
  return (s0 + s1) + (s2 + s3);
}

template <typename R, int NT>
__device__ void decouple_blocks(const R* P, const R* T, int n, int k, R* Krs, R* Ksr, R* X) {
  const int tid = threadIdx.x, r = n - k;
  for (int idx = tid; idx < r * k; idx += NT) {
    const int i = idx % r, a = idx / r;
    Krs[idx] = P[(k + i) + (int64_t)n * a] - dot2(T + k * i, 1, P + (int64_t)n * a, 1, k);
  }
  for (int idx = tid; idx < k * r; idx += NT) {
    const int a = idx % k, j = idx / k;
    Ksr[idx] = P[a + (int64_t)n * (k + j)] - dot2(P + a, n, T + k * j, 1, k);
  }
  __syncthreads();
  for (int idx = tid; idx < r * r; idx += NT) {
    const int i = idx % r, j = idx / r;
    X[idx] = P[(k + i) + (int64_t)n * (k + j)] - dot2(T + k * i, 1, P + (int64_t)n * (k + j), 1, k) -
             dot2(Krs + i, r, T + k * j, 1, k);
  }
  __syncthreads();
}

// From M / prow of gj_invert on X_rr: U, L, X_rr^{-1} and the carried block
// into the factor slot.  W has room for r k entries; uss, when given, holds
// the children's updates on the skeleton block (k x k), else upd forms them.
template <typename R, int NT, typename Upd>
__device__ void decouple_tail(const R* M, const int* prow, const R* Krs, const R* Ksr,
                              R* W, int r, int k, R* F, const FacLayout& fl,
                              const Upd& upd, const short* sperm, const R* uss = nullptr) {
  const int tid = threadIdx.x;
  for (int idx = tid; idx < r * k; idx += NT) {  // U = X^{-1} K_rs
    const int j = idx % r, a = idx / r;
    const R* mrow = M + prow[j];
    const R* kc = Krs + (int64_t)r * a;
    R s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int jj = 0;
    for (; jj + 4 <= r; jj += 4) {
      s0 += mrow[(int64_t)r * jj] * kc[prow[jj]];
      s1 += mrow[(int64_t)r * (jj + 1)] * kc[prow[jj + 1]];
      s2 += mrow[(int64_t)r * (jj + 2)] * kc[prow[jj + 2]];
      s3 += mrow[(int64_t)r * (jj + 3)] * kc[prow[jj + 3]];
    }
    for (; jj < r; jj++) s0 += mrow[(int64_t)r * jj] * kc[prow[jj]];
    const R acc = (s0 + s1) + (s2 + s3);
    W[idx] = acc;
    F[fl.U + idx] = acc;
  }
  for (int idx = tid; idx < k * r; idx += NT) {  // L = K_sr X^{-1}
    const int a = idx % k, jj = idx / k;
    const R* mcol = M + (int64_t)r * jj;
    R s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int j = 0;
    for (; j + 4 <= r; j += 4) {
      s0 += Ksr[a + (int64_t)k * j] * mcol[prow[j]];
      s1 += Ksr[a + (int64_t)k * (j + 1)] * mcol[prow[j + 1]];
      s2 += Ksr[a + (int64_t)k * (j + 2)] * mcol[prow[j + 2]];
      s3 += Ksr[a + (int64_t)k * (j + 3)] * mcol[prow[j + 3]];
    }
    for (; j < r; j++) s0 += Ksr[a + (int64_t)k * j] * mcol[prow[j]];
    F[fl.L + a + (int64_t)k * prow[jj]] = (s0 + s1) + (s2 + s3);
  }
  for (int idx = tid; idx < r * r; idx += NT) {
    const int j = idx % r, jj = idx / r;
    F[fl.X + j + (int64_t)r * prow[jj]] = M[prow[j] + (int64_t)r * jj];
  }
  __syncthreads();
  for (int idx = tid; idx < k * k; idx += NT) {  // carried block
    const int a = idx % k, c = idx / k;
    const R u = uss ? uss[idx] : upd(sperm[a], sperm[c]);
    F[fl.S + idx] = u + dot2(Ksr + a, k, W + (int64_t)r * c, 1, r);
  }
}

// Compressed boxes with NMIN < |active| <= NMAX (<= 64), every operand in
// LDS.  Only the skeleton rows of K_BN are stored (K(s,s), K(s,r)); the
// redundant rows are evaluated where K_rs and X_rr need them, which keeps
// the buffer at n^2 + 2 k r <= 1.5 NMAX^2 entries: three boxes per CU for
// NMAX = 64, five for NMAX = 44.
template <typename R, int NMIN, int NMAX>
__global__ __launch_bounds__(256) void k_decouple_small(DevGeom g, DevTree t,
                                                        DevState s, int start) {
  constexpr int NT = 256;
  // >= n^2 + 2 k r (k r <= n^2 / 4); the two size classes let the boxes
  // with n <= kSmallBox1 run at a higher occupancy
  constexpr int kBuf = NMAX * NMAX + NMAX * NMAX / 2;
  static_assert(NMAX * sizeof(DofGeo) >= 2 * 64 * sizeof(R), "pivot columns fit in geo");
  __shared__ R buf[kBuf];
  __shared__ DofGeo geo[NMAX];
  __shared__ signed char pch[NMAX];
  __shared__ short ploc[NMAX];
  __shared__ short sperm[NMAX];
  __shared__ int prow[NMAX];
  __shared__ const R* chS[4];
  __shared__ int chN[4];

  const int b = start + blockIdx.x;
  const int tid = threadIdx.x;
  if (!pools_ok(s) || box_clean(s, b)) return;
  const int n = s.act_cnt[b], k = s.rank[b];
  if (n <= NMIN || n > NMAX || k == 0) return;  // the other class's or k_decouple's box
  load_child_blocks<R>(t, s, b, n, chS, chN, pch, ploc, NT);
  const int* act = s.act_list + (int64_t)b * s.ncap;
  for (int i = tid; i < n; i += NT) {
    const int q = s.perm[(int64_t)b * s.ncap + i];
    sperm[i] = (short)q;
    geo[i] = dof_geo(g, act[q]);
  }
  __syncthreads();
  ChildUpdates<R> upd;
  upd.S = chS;
  upd.N = chN;
  upd.pch = pch;
  upd.ploc = ploc;
  const int r = n - k;
  const FacLayout fl = FacLayout::make(k, r);
  R* F = reinterpret_cast<R*>(s.fac_pool) + s.fac_off[b];
  // K_BN entry at permuted positions (skeleton first).
  auto kbn = [&](int i, int c) -> R {
    return static_cast<R>(kernel_entry(g.pde, geo[i], geo[c])) - upd(sperm[i], sperm[c]);
  };
  R* Pss = buf;          // k x k
  R* Psr = Pss + k * k;  // k x r
  R* T = Psr + k * r;    // k x r
  R* X = T + k * r;      // r x r
  R* Krs = X + r * r;    // r x k
  R* Ksr = Krs + r * k;  // k x r
  for (int idx = tid; idx < k * r; idx += NT) T[idx] = F[fl.T + idx];
  for (int idx = tid; idx < k * n; idx += NT) {
    const int a = idx % k, c = idx / k;  // Pss and Psr are contiguous (ld k)
    Pss[idx] = kbn(a, c);
  }
  __syncthreads();
  for (int idx = tid; idx < r * k; idx += NT) {  // K_rs = K(r,s) - T^T K(s,s)
    const int i = idx % r, a = idx / r;
    Krs[idx] = kbn(k + i, a) - dot2(T + k * i, 1, Pss + k * a, 1, k);
  }
  for (int idx = tid; idx < k * r; idx += NT) {  // K_sr = K(s,r) - K(s,s) T
    const int a = idx % k, j = idx / k;
    Ksr[idx] = Psr[idx] - dot2(Pss + a, k, T + k * j, 1, k);
  }
  __syncthreads();
  for (int idx = tid; idx < r * r; idx += NT) {  // X_rr = K(r,r) - T^T K(s,r) - K_rs T
    const int i = idx % r, j = idx / r;
    X[idx] = kbn(k + i, k + j) - dot2(T + k * i, 1, Psr + k * j, 1, k) -
             dot2(Krs + i, r, T + k * j, 1, k);
  }
  __syncthreads();
  if (s.fwd_pool) {
    R* Xf = reinterpret_cast<R*>(s.fwd_pool) + s.fac_off[b];
    for (int idx = tid; idx < r * r; idx += NT) Xf[idx] = X[idx];
  }
  // The DOF geometry is dead: its space holds the pivot columns.
  R* colbuf = reinterpret_cast<R*>(geo);
  const int info = gj_invert_regs<R>(X, r, prow, colbuf);
  if (info != 0 && tid == 0) atomicOr(s.err, kErrSingular);
  decouple_tail<R, NT>(X, prow, Krs, Ksr, buf, r, k, F, fl, upd, sperm);
  if (tid == 0) s.upd_kind[b] = 1;
}

// Every other box: uncompressed boxes pass their children's blocks on;
// compressed boxes with |active| > 64 keep P and X_rr in LDS when they fit
// and in the box's scratch slot otherwise.  Few boxes of a level are
// these, so a small grid walks the level (the others are skipped after
// two loads) instead of one workgroup per box.
constexpr int kDecGeoMax = 96;

template <typename R, int NT>
__global__ __launch_bounds__(NT) void k_decouple(DevGeom g, DevTree t,
                                                 DevState s, int start, int count) {
  // 128 KB: P, X_rr, K_rs and K_sr of a box up to 90 DOFs (FP64) stay in
  // LDS; the Schur products are then LDS-bound, not bound by L2 latency.
  constexpr int kLdsElems = 131072 / sizeof(R);
  __shared__ R sbuf[kLdsElems];
  __shared__ R gjcol[2 * 16 * 8];  // the register Gauss-Jordan's column buffers
  __shared__ DofGeo geo[kDecGeoMax];  // the box's DOFs, permuted, when n fits
  __shared__ signed char pch[kMaxBoxDofs];
  __shared__ short ploc[kMaxBoxDofs];
  __shared__ short sperm[kMaxBoxDofs];
  __shared__ int prow[kMaxBoxDofs];
  __shared__ const R* chS[4];
  __shared__ int chN[4];

  const int tid = threadIdx.x;
  if (!pools_ok(s)) return;
  for (int bi = blockIdx.x; bi < count; bi += gridDim.x) {
  const int b = start + bi;
  if (box_clean(s, b)) continue;
  const int n = s.act_cnt[b];
  if (n == 0) {
    if (tid == 0) s.upd_kind[b] = 0;
    continue;
  }
  if (n <= kSmallBox && s.rank[b] > 0) continue;  // k_decouple_small's box
  __syncthreads();  // the previous box's LDS
  load_child_blocks<R>(t, s, b, n, chS, chN, pch, ploc, NT);
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
    continue;
  }
  const int r = n - k;
  const FacLayout fl = FacLayout::make(k, r);
  const int* act = s.act_list + (int64_t)b * s.ncap;
  const bool lgeo = n <= kDecGeoMax;
  for (int i = tid; i < n; i += NT) {
    const int q = s.perm[(int64_t)b * s.ncap + i];
    sperm[i] = (short)q;
    if (lgeo) geo[i] = dof_geo(g, act[q]);
  }
  __syncthreads();
  // n^2 + r^2 + 2 r k <= 2 n^2 entries.
  R* P = (2 * (int64_t)n * n <= kLdsElems)
             ? sbuf
             : reinterpret_cast<R*>(s.dec_pool) + s.dec_off[b];
  R* X = P + (int64_t)n * n;
  R* Krs = X + (int64_t)r * r;
  R* Ksr = Krs + (int64_t)r * k;
  // T (k x r) and the children's updates on the skeleton block (k x k)
  // join them in LDS when they fit: the Schur products then read no
  // global memory.
  const bool stage = P == sbuf && (int64_t)n * n + (int64_t)r * r + 3 * (int64_t)r * k +
                                          (int64_t)k * k <= kLdsElems;
  R* Tl = Ksr + (int64_t)r * k;
  R* Uss = stage ? Tl + (int64_t)r * k : nullptr;
  if (stage)
    for (int idx = tid; idx < r * k; idx += NT) Tl[idx] = F[fl.T + idx];
  for (int64_t idx = tid; idx < (int64_t)n * n; idx += NT) {
    const int i = (int)(idx % n), c = (int)(idx / n);
    const int ia = sperm[i], ic = sperm[c];
    const R u = upd(ia, ic);
    const double kv = lgeo ? kernel_entry(g.pde, geo[i], geo[c]) : kernel_entry(g, act[ia], act[ic]);
    P[idx] = static_cast<R>(kv) - u;
    if (Uss && i < k && c < k) Uss[i + k * c] = u;
  }
  __syncthreads();
  decouple_blocks<R, NT>(P, stage ? Tl : F + fl.T, n, k, Krs, Ksr, X);
  if (s.fwd_pool) {
    R* Xf = reinterpret_cast<R*>(s.fwd_pool) + s.fac_off[b];
    for (int64_t idx = tid; idx < (int64_t)r * r; idx += NT) Xf[idx] = X[idx];
    __syncthreads();
  }
  // X_rr^{-1}: in registers for r <= 128, else step by step.
  int info;
  if (NT == 256 && r <= 64) info = gj_invert_regs<R, 4>(X, r, prow, gjcol);
  else if (NT == 256 && r <= 128) info = gj_invert_regs<R, 8>(X, r, prow, gjcol);
  else info = gj_invert_into<R, NT>(X, P, r, prow);
  if (info != 0 && tid == 0) atomicOr(s.err, kErrSingular);
  decouple_tail<R, NT>(X, prow, Krs, Ksr, P, r, k, F, fl, upd, sperm, Uss);
  if (tid == 0) s.upd_kind[b] = 1;
  }
}

// ---------------------------------------------------------------------------
// Root: K(allskel, allskel) - carried blocks of the level-1 boxes, inverted.
// ---------------------------------------------------------------------------

// Assembly of root block blockIdx.x: gridDim.y workgroups share its n^2
// entries (one workgroup was latency-bound on the kernel's gathers).
constexpr int kRootSplit = 64;

template <typename R, int NT>
__global__ __launch_bounds__(NT) void k_root(DevGeom g, DevTree t, DevState s) {
  __shared__ signed char pch[kMaxBoxDofs];
  __shared__ short ploc[kMaxBoxDofs];
  __shared__ const R* chS[4];
  __shared__ int chN[4];
  const int b = blockIdx.x;  // root of problem b
  const int tid = threadIdx.x;
  const bool lead = blockIdx.y == 0 && tid == 0;
  if (lead) {  // the root is never compressed
    s.rank[b] = 0;
    s.m_rows[b] = 0;
    s.out_cnt[b] = s.act_cnt[b];
  }
  if (b == 0 && lead && s.plan_skip != nullptr) {
    // The launch plan skipped kernels whose lists it expected empty: any
    // that is not makes the host run the factorization again in full.
    for (int lv = 1; lv < t.n_levels; lv++)
      if (s.plan_skip[lv] & ~plan_skippable(s, lv)) {
        atomicOr(s.err, kErrPlan);
        break;
      }
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
  const int* act = s.act_list + (int64_t)b * s.ncap;
  R* A = root_block<R>(s, b);
  R* Af = s.root_fwd ? reinterpret_cast<R*>(s.root_fwd) + (int64_t)s.ncap * s.ncap * b : nullptr;
  const int64_t nn = (int64_t)n * n, per = (nn + gridDim.y - 1) / gridDim.y;
  const int64_t e0 = per * blockIdx.y, e1 = min(nn, e0 + per);
  for (int64_t idx = e0 + tid; idx < e1; idx += NT) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    const R v = static_cast<R>(kernel_entry(g, act[i], act[j])) - upd(i, j);
    A[idx] = v;
    if (Af) Af[idx] = v;
  }
}

// Gauss-Jordan inverse of the root block with the matrix in registers,
// blocked over panels of kRootPB columns: 1024 threads on a 32 x 32 grid,
// thread (tr, tc) holding rows tr + 32 u and columns tc + 32 v (n <= 192;
// 6 x 6 entries each), so that one wave holds every row of its 12 columns.
// The block is bordered by the identity up to n8, a multiple of kRootPB
// (its pivots stay in the block: border rows hold zeros in its columns and
// lose every tie to a block row), and rows and columns past n8 are zero, so
// no step or entry needs a bound.
//
// Wave 0 runs each panel's pivot steps on the panel's columns alone (they
// depend on nothing else): pivot search as in gj_invert, scaling,
// elimination; it records each step's pivot row p_k, 1 / pivot and the
// column as it was before the step (C[m][i]).  Applying panel P:
//  a. in every wave, the owners of its pivot rows copy them to LDS and one
//     lane per column forms the scaled pivot rows the steps would have
//     made, R_k = (row p_k - sum_{m<k} C[m][p_k] R_m) / pivot_k;
//  b. every thread applies the steps to its entries in step order,
//     a -= C[m][i] R_m, a pivot row restarting from R_k at its own step;
//     panel entries take wave 0's result.
// With one step of lookahead: after (a), the owners of panel P+1's columns
// apply P to a copy of those columns and publish it; after a barrier wave 0
// eliminates panel P+1 from that copy while every wave (wave 0 after it)
// applies P to all its entries; a barrier ends the panel.  Per entry these are the
// operations of the unblocked elimination in its order, so the inverse is
// the same.  The inverse, unpermuted, goes after the assembled matrix.
constexpr int kRootTR = 32, kRootTC = 32, kRootRU = 6, kRootCV = 6;
constexpr int kRootPB = 8;

template <typename R>
__global__ __launch_bounds__(kRootTR * kRootTC) void k_root_inv(DevState s, int factor_lu) {
  constexpr int NMAX = kRootRegMax, PB = kRootPB, NS = NMAX / 64;
  static_assert(kRootTC % PB == 0 && NMAX == kRootTR * kRootRU && NMAX == kRootTC * kRootCV &&
                NMAX % 64 == 0 && NMAX % PB == 0 && NMAX <= 256 && PB * PB <= 64,
                "root tiling");
  __shared__ R pan[2][PB][NMAX];   // panel columns, in and out of the elimination
  __shared__ R cpre[2][PB][NMAX];  // column m before step m (zero where row i
                                   // is pivot at a step >= m); by panel parity
  __shared__ R cpiv[2][PB][PB];    // C[m][p_k], m < k
  __shared__ R pinv[2][PB];
  __shared__ R rowk[PB][NMAX];     // pivot rows, then R_k (each wave its columns)
  __shared__ int pstep[NMAX];      // elimination step that pivots on row i
  __shared__ int prow[NMAX];
  __shared__ int invrow[NMAX];
  __shared__ R park[kRootRU * kRootCV][64];
  if (!factor_lu || !pools_ok(s)) return;
  const int n = s.act_cnt[blockIdx.x];
  if (n == 0 || n > NMAX) return;
  const int n8 = (n + PB - 1) / PB * PB;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int tr = tid % kRootTR, tc = tid / kRootTR;
  R* A = root_block<R>(s, blockIdx.x);
  R a[kRootRU][kRootCV];
#pragma unroll
  for (int u = 0; u < kRootRU; u++)
#pragma unroll
    for (int v = 0; v < kRootCV; v++) {
      const int i = tr + kRootTR * u, c = tc + kRootTC * v;
      a[u][v] = (i < n && c < n) ? A[i + (int64_t)n * c] : (i == c && i < n8) ? R(1) : R(0);
    }
  if (tid < NMAX) pstep[tid] = NMAX;
  // wave 0: bit q set once row lane + 64 q is a pivot row (rows past n8
  // never are)
  uint32_t used = 0;
#pragma unroll
  for (int q = 0; q < NS; q++)
    if (lane + 64 * q >= n8) used |= 1u << q;
  int info = 0;

  // The thread owns panel j0's column in its entries a[.][vj] when
  // tc in [tc0, tc0 + PB) (the same for both halves of a wave).
  auto owns = [&](int j0) { return tc >= j0 % kRootTC && tc < j0 % kRootTC + PB; };
  auto publish = [&](int j0, int pb) {
    const int vj = j0 / kRootTC;
#pragma unroll
    for (int v = 0; v < kRootCV; v++) {
      if (v != vj) continue;
#pragma unroll
      for (int u = 0; u < kRootRU; u++) pan[pb][tc - j0 % kRootTC][tr + kRootTR * u] = a[u][v];
    }
  };
  // wave 0, its own entries parked in LDS meanwhile (the panel and the
  // matrix do not fit its registers)
  auto eliminate = [&](int j0, int pb) {
#pragma unroll
    for (int u = 0; u < kRootRU; u++)
#pragma unroll
      for (int v = 0; v < kRootCV; v++) park[u * kRootCV + v][lane] = a[u][v];
    R x[NS][PB];
#pragma unroll
    for (int q = 0; q < NS; q++)
#pragma unroll
      for (int k = 0; k < PB; k++) x[q][k] = pan[pb][k][lane + 64 * q];
    // pivot keys of column k over the rows not yet pivots (the next
    // column's are formed at the end of each step, beside its updates)
    uint32_t cand = 0;
    auto column_best = [&](int k) {
      // the lane index is formed here each time: kept across the steps,
      // the rows' key bits were spilled
      int l = lane;
      __asm__ volatile("" : "+v"(l));
      double key = -1.0;
      cand = 0;
#pragma unroll
      for (int q = 0; q < NS; q++) {
        const bool ok = !((used >> q) & 1);
        cand |= (ok ? 1u : 0u) << q;
        const R v = fabs(x[q][k]);
        const double kq = pivot_key<R, 8>(v, l + 64 * q);
        key = fmax(key, (ok && v == v) ? kq : -1.0);
      }
      return wave_max(key);
    };
    double best = column_best(0);
#pragma unroll
    for (int k = 0; k < PB; k++) {
      int p;
      if (best >= 0) {
        p = pivot_key_pos<8>(best);
      } else {  // NaN column: the first row not yet pivoted
        p = 0x7fffffff;
#pragma unroll
        for (int q = 0; q < NS; q++) {
          const uint64_t m = __ballot((cand >> q) & 1);
          if (m && p == 0x7fffffff) p = 64 * q + __ffsll((unsigned long long)m) - 1;
        }
      }
      const int pq = p >> 6, pl = p & 63;
      // the pivot row's panel entries, in every lane
      R xp[PB];
#pragma unroll
      for (int k2 = 0; k2 < PB; k2++) {
        R v = x[0][k2];
#pragma unroll
        for (int q = 1; q < NS; q++)
          if (pq == q) v = x[q][k2];
        xp[k2] = lane_value(v, pl);
      }
      const R piv = xp[k];
      if (piv == R(0) && info == 0) info = j0 + k + 1;
      const R inv = R(1) / piv;
      if (lane == 0) {
        pinv[pb][k] = inv;
        prow[j0 + k] = p;
        pstep[p] = j0 + k;
      }
#pragma unroll
      for (int q = 0; q < NS; q++) cpre[pb][k][lane + 64 * q] = x[q][k];
      R rv[PB];
#pragma unroll
      for (int k2 = 0; k2 < PB; k2++) rv[k2] = (k2 == k ? R(1) : xp[k2]) * inv;
#pragma unroll
      for (int q = 0; q < NS; q++) {
        const R c = x[q][k];
        x[q][k] = R(0);
#pragma unroll
        for (int k2 = 0; k2 < PB; k2++) x[q][k2] -= c * rv[k2];
      }
      if (lane == pl) {
        used |= 1u << pq;
#pragma unroll
        for (int q = 0; q < NS; q++)
          if (q == pq) {
#pragma unroll
            for (int k2 = 0; k2 < PB; k2++) x[q][k2] = rv[k2];
          }
      }
      if (k + 1 < PB) best = column_best(k + 1);
      __builtin_amdgcn_sched_barrier(0);  // one step's registers at a time
    }
#pragma unroll
    for (int q = 0; q < NS; q++)
#pragma unroll
      for (int k = 0; k < PB; k++) pan[pb][k][lane + 64 * q] = x[q][k];
    // C at the pivot rows, one (step k, step m <= k) pair per lane: the
    // steps before a row's own step go to cpiv (they form R_k), cpre
    // keeps only the steps after it
    if (lane < PB * PB) {
      const int k = lane / PB, m = lane % PB;
      if (m <= k) {
        const int pk = prow[j0 + k];
        if (m < k) cpiv[pb][k][m] = cpre[pb][m][pk];
        cpre[pb][m][pk] = R(0);
      }
    }
#pragma unroll
    for (int u = 0; u < kRootRU; u++)
#pragma unroll
      for (int v = 0; v < kRootCV; v++) a[u][v] = park[u * kRootCV + v][lane];
  };

  // prologue: panel 0 eliminated
  if (owns(0)) publish(0, 0);
  __syncthreads();
  if (wave == 0) eliminate(0, 0);
  __syncthreads();
  for (int j0 = 0, pb = 0; j0 < n8; j0 += PB, pb ^= 1) {
    const int j1 = j0 + PB, pn = pb ^ 1;
    const bool nxt = j1 < n8;
    // a. the wave's pivot rows of panel j0, then R_k for the wave's columns
    int kk[kRootRU];
#pragma unroll
    for (int u = 0; u < kRootRU; u++) {
      const int q = pstep[tr + kRootTR * u] - j0;
      kk[u] = (q >= 0 && q < PB) ? q : -1;
      if (kk[u] >= 0) {
#pragma unroll
        for (int v = 0; v < kRootCV; v++) rowk[kk[u]][tc + kRootTC * v] = a[u][v];
      }
    }
    wave_sync();
    if (lane < 2 * kRootCV) {
      const int c = (tc & ~1) + (lane & 1) + kRootTC * (lane >> 1);
      R r[PB];
#pragma unroll
      for (int k = 0; k < PB; k++) {
        R t = rowk[k][c];
#pragma unroll
        for (int m = 0; m < k; m++) t -= cpiv[pb][k][m] * r[m];
        r[k] = t * pinv[pb][k];
        rowk[k][c] = r[k];
      }
    }
    wave_sync();
    // lookahead: panel j1's columns take panel j0's steps (as in b, which
    // repeats them on the thread's entries) and are published
    if (nxt && owns(j1)) {
      const int v1 = j1 / kRootTC;
#pragma unroll
      for (int v = 0; v < kRootCV; v++) {
        if (v != v1) continue;
        const int c = tc + kRootTC * v;
        R col[kRootRU];
#pragma unroll
        for (int u = 0; u < kRootRU; u++) col[u] = kk[u] >= 0 ? rowk[kk[u]][c] : a[u][v];
#pragma unroll
        for (int m = 0; m < PB; m++) {
          const R rv = rowk[m][c];
#pragma unroll
          for (int u = 0; u < kRootRU; u++) col[u] -= cpre[pb][m][tr + kRootTR * u] * rv;
        }
#pragma unroll
        for (int u = 0; u < kRootRU; u++) pan[pn][tc - j1 % kRootTC][tr + kRootTR * u] = col[u];
      }
    }
    __syncthreads();
    if (wave == 0 && nxt) eliminate(j1, pn);
    // b. panel j0's steps on the thread's entries (the pivot steps of panel
    // j1 are after j0's in pstep: kk formed again, not kept across)
#pragma unroll
    for (int u = 0; u < kRootRU; u++) {
      const int q = pstep[tr + kRootTR * u] - j0;
      kk[u] = (q >= 0 && q < PB) ? q : -1;
    }
#pragma unroll
    for (int u = 0; u < kRootRU; u++) {
      if (kk[u] < 0) continue;
#pragma unroll
      for (int v = 0; v < kRootCV; v++) a[u][v] = rowk[kk[u]][tc + kRootTC * v];
    }
#pragma unroll
    for (int m = 0; m < PB; m++) {
      R cv[kRootRU], rv[kRootCV];
#pragma unroll
      for (int u = 0; u < kRootRU; u++) cv[u] = cpre[pb][m][tr + kRootTR * u];
#pragma unroll
      for (int v = 0; v < kRootCV; v++) rv[v] = rowk[m][tc + kRootTC * v];
#pragma unroll
      for (int u = 0; u < kRootRU; u++)
#pragma unroll
        for (int v = 0; v < kRootCV; v++) a[u][v] -= cv[u] * rv[v];
    }
    if (owns(j0)) {
      const int vj = j0 / kRootTC;
#pragma unroll
      for (int v = 0; v < kRootCV; v++) {
        if (v != vj) continue;
#pragma unroll
        for (int u = 0; u < kRootRU; u++) a[u][v] = pan[pb][tc - j0 % kRootTC][tr + kRootTR * u];
      }
    }
    __syncthreads();
  }
  if (tid < n8) invrow[prow[tid]] = tid;
  __syncthreads();
  R* B = A + (int64_t)n * n;
#pragma unroll
  for (int u = 0; u < kRootRU; u++)
#pragma unroll
    for (int v = 0; v < kRootCV; v++) {
      const int i = tr + kRootTR * u, c = tc + kRootTC * v;
      if (i < n && c < n) B[invrow[i] + (int64_t)n * prow[c]] = a[u][v];
    }
  if (info != 0 && tid == 0) atomicOr(s.err, kErrSingular);
}
// ---------------------------------------------------------------------------
// Solve sweeps (skel_factorization.cpp:507-592).  One wave per box; the
// box's skeleton and redundant entries are gathered to LDS, updated and
// scattered back (boxes of one level touch disjoint DOFs).
// ---------------------------------------------------------------------------

// sum_j p[stride j] v[j], j < len: eight loads in flight, four partial
// sums (a sweep's per-lane loop is otherwise one load latency per term).
template <typename R>
__device__ inline double dot_ld(const R* p, int64_t stride, const double* v, int len) {
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  int j = 0;
  for (; j + 8 <= len; j += 8) {
    R q[8];
#pragma unroll
    for (int u = 0; u < 8; u++) q[u] = p[stride * (j + u)];
    a0 += (double)q[0] * v[j] + (double)q[4] * v[j + 4];
    a1 += (double)q[1] * v[j + 1] + (double)q[5] * v[j + 5];
    a2 += (double)q[2] * v[j + 2] + (double)q[6] * v[j + 6];
    a3 += (double)q[3] * v[j + 3] + (double)q[7] * v[j + 7];
  }
  for (; j < len; j++) a0 += (double)p[stride * j] * v[j];
  return (a0 + a1) + (a2 + a3);
}

// x <- X^{-1} x with an explicit inverse (r x r, ld r): one wave, x and
// tmp in LDS.
template <typename R>
__device__ void wave_inv_apply(const R* Xi, int r, double* x, double* tmp) {
  const int lane = threadIdx.x & 63;
  for (int i = lane; i < r; i += 64) tmp[i] = dot_ld(Xi + i, r, x, r);
  wave_sync();
  for (int i = lane; i < r; i += 64) x[i] = tmp[i];
  wave_sync();
}

template <typename R>
__global__ __launch_bounds__(64) void k_sweep_up(DevState s, int start,
                                                 SolveVec v, int use_Ut,
                                                 int diag) {
  __shared__ double xs[kMaxBoxDofs];
  __shared__ double xr[kMaxBoxDofs];
  __shared__ double tmp[kMaxBoxDofs];
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
    for (int i = lane; i < r; i += 64) xr[i] -= dot_ld(F + fl.T + (int64_t)k * i, 1, xs, k);
    wave_sync();
    for (int a = lane; a < k; a += 64) {
      if (use_Ut) xs[a] -= dot_ld(F + fl.U + (int64_t)r * a, 1, xr, r);
      else xs[a] -= dot_ld(F + fl.L + a, k, xr, r);
    }
    wave_sync();
    if (diag) wave_inv_apply(F + fl.X, r, xr, tmp);
    for (int a = lane; a < k; a += 64) x[sk[a]] = xs[a];
    for (int i = lane; i < r; i += 64) x[rd[i]] = xr[i];
    wave_sync();
  }
}

template <typename R>
__global__ __launch_bounds__(64) void k_diag(DevState s, int start, SolveVec v) {
  __shared__ double xr[kMaxBoxDofs];
  __shared__ double tmp[kMaxBoxDofs];
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
    wave_inv_apply(F + fl.X, r, xr, tmp);
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
    for (int i = lane; i < r; i += 64) xr[i] -= dot_ld(F + fl.U + i, r, xs, k);
    wave_sync();
    for (int a = lane; a < k; a += 64) xs[a] -= dot_ld(F + fl.T + a, k, xr, r);
    wave_sync();
    for (int a = lane; a < k; a += 64) x[sk[a]] = xs[a];
    for (int i = lane; i < r; i += 64) x[rd[i]] = xr[i];
    wave_sync();
  }
}

// Blocked in-place Gauss-Jordan inverse of a root block with
// kRootRegMax < n <= kRootBlkMax, one workgroup per root, thread t owning
// row t.  Per panel of 16 columns J:
//   1. the panel's 16 elimination steps run on the n x 16 slab in
//      registers (pivot row: largest |value| among the rows not yet
//      pivoted, as gj_invert; one barrier per step);
//   2. the other columns take all 16 steps at once: with W the final slab
//      and P the panel's pivot rows, the steps multiply every other column
//      x by G = I + sum_k (W_k - e_{p_k}) e_{p_k}^T, so
//      x_i += sum_k W_ik x_{p_k} (old values of the pivot rows), and a
//      pivot row's own value is replaced rather than added to.
// The result equals the column-by-column algorithm up to rounding order.
// The inverse, unpermuted, goes after the matrix, as in k_root_inv.
constexpr int kRootBlkB = 16;

template <typename R>
__global__ __launch_bounds__(kRootBlkMax) void k_root_inv_blk(DevState s, int factor_lu) {
  constexpr int B = kRootBlkB, NW = kRootBlkMax / 64;
  __shared__ R rowv[2][NW][B];
  __shared__ double keys[2][NW];
  __shared__ R Rl[kRootBlkMax * B];
  __shared__ int prow[kRootBlkMax];
  __shared__ int invrow[kRootBlkMax];
  __shared__ int pan[B];
  if (!factor_lu || !pools_ok(s)) return;
  const int root = blockIdx.x;
  const int n = s.act_cnt[root];
  if (n <= kRootRegMax) return;
  if (n > kRootBlkMax) {
    // Gauss-Jordan inverse (the reference LU-factors the root block) in
    // global memory for blocks too large for the register forms; the
    // inverse, unpermuted, goes after the assembled matrix
    R* A = root_block<R>(s, root);
    R* B = A + (int64_t)n * n;
    int* gp = root_pivots(s, root);
    const int info = gj_invert_into<R, kRootBlkMax>(A, B, n, gp);
    if (info != 0 && threadIdx.x == 0) atomicOr(s.err, kErrSingular);
    for (int64_t idx = threadIdx.x; idx < (int64_t)n * n; idx += kRootBlkMax) {
      const int j = (int)(idx % n), jj = (int)(idx / n);
      B[j + (int64_t)n * gp[jj]] = A[gp[j] + (int64_t)n * jj];
    }
    return;
  }
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int i = tid;  // own row
  const bool live = i < n;
  R* A = root_block<R>(s, root);
  bool used = !live;
  int info = 0;
  for (int j0 = 0; j0 < n; j0 += B) {
    const int bw = min(B, n - j0);
    R w[B];
#pragma unroll
    for (int k = 0; k < B; k++) w[k] = (live && k < bw) ? A[i + (int64_t)n * (j0 + k)] : R(0);
    bool inpan = false;
#pragma unroll
    for (int k = 0; k < B; k++) {
      if (k < bw) {
        R v = fabs(w[k]);
        if (!isfinite(v)) v = R(0);  // every unpivoted row stays a candidate
        const double key = used ? -1.0 : pivot_key<R, 9>(v, i);
        const double wbest = wave_max(key);
        const int buf = k & 1;
        if (!used && key == wbest) {
#pragma unroll
          for (int c = 0; c < B; c++) rowv[buf][wv][c] = w[c];
        }
        if (lane == 0) keys[buf][wv] = wbest;
        __syncthreads();
        double best = keys[buf][0];
        int bwv = 0;
#pragma unroll
        for (int q = 1; q < NW; q++)
          if (keys[buf][q] > best) {
            best = keys[buf][q];
            bwv = q;
          }
        const int p = pivot_key_pos<9>(best);
        R rv[B];
#pragma unroll
        for (int c = 0; c < B; c++) rv[c] = rowv[buf][bwv][c];
        const R piv = rv[k];
        if (piv == R(0) && info == 0) info = j0 + k + 1;
        const R inv = R(1) / piv;
#pragma unroll
        for (int c = 0; c < B; c++) rv[c] = (c == k) ? inv : rv[c] * inv;
        const R cv = w[k];
        w[k] = R(0);
#pragma unroll
        for (int c = 0; c < B; c++) w[c] -= cv * rv[c];
        if (i == p) {
#pragma unroll
          for (int c = 0; c < B; c++) w[c] = rv[c];
          used = true;
          inpan = true;
        }
        if (tid == 0) {
          prow[j0 + k] = p;
          pan[k] = p;
        }
      }
    }
    __syncthreads();
    // old values of the panel's pivot rows in every other column
    for (int idx = tid; idx < n * B; idx += kRootBlkMax) {
      const int c = idx / B, k = idx % B;
      const bool skip = k >= bw || (c >= j0 && c < j0 + bw);
      Rl[idx] = skip ? R(0) : A[pan[k] + (int64_t)n * c];
    }
    __syncthreads();
    // Own row, 16 columns per chunk; the next chunk's loads are issued
    // before the current chunk's arithmetic so one load latency is exposed
    // per panel rather than per chunk.
    constexpr int CH = 16;
    auto load_chunk = [&](int c0, R* x) {
#pragma unroll
      for (int q = 0; q < CH; q++) {
        const int c = c0 + q;
        const bool other = c < n && (c < j0 || c >= j0 + bw);
        x[q] = (live && other && !inpan) ? A[i + (int64_t)n * c] : R(0);
      }
    };
    R x[CH], xn[CH];
    load_chunk(0, x);
    for (int c0 = 0; c0 < n; c0 += CH) {
      if (c0 + CH < n) load_chunk(c0 + CH, xn);
#pragma unroll
      for (int q = 0; q < CH; q++) {
        const int c = c0 + q;
        if (c < n && (c < j0 || c >= j0 + bw)) {
          R acc = x[q];
#pragma unroll
          for (int k = 0; k < B; k++) acc += w[k] * Rl[c * B + k];
          if (live) A[i + (int64_t)n * c] = acc;
        }
      }
#pragma unroll
      for (int q = 0; q < CH; q++) x[q] = xn[q];
    }
#pragma unroll
    for (int k = 0; k < B; k++)
      if (live && k < bw) A[i + (int64_t)n * (j0 + k)] = w[k];
    __syncthreads();
  }
  if (tid < n) invrow[prow[tid]] = tid;
  __syncthreads();
  R* Bm = A + (int64_t)n * n;
  for (int64_t idx = tid; idx < (int64_t)n * n; idx += kRootBlkMax) {
    const int r = (int)(idx % n), c = (int)(idx / n);
    Bm[invrow[r] + (int64_t)n * prow[c]] = A[idx];
  }
  if (info != 0 && tid == 0) atomicOr(s.err, kErrSingular);
}

// x_root <- (root inverse) x_root: 256 threads, one row each per pass.
template <typename R>
__global__ __launch_bounds__(256) void k_root_solve(DevState s, SolveVec v) {
  __shared__ double xa[kMaxBoxDofs];
  const int tid = threadIdx.x;
  const int n = s.act_cnt[blockIdx.x];
  const int* act = s.act_list + (int64_t)blockIdx.x * s.ncap;
  const R* Ai = root_block<R>(s, blockIdx.x) + (int64_t)n * n;
  for (int c = 0; c < v.nrhs; c++) {
    double* x = v.x + v.ldx * c;
    for (int i = tid; i < n; i += 256) xa[i] = x[act[i]];
    __syncthreads();
    double y[kMaxBoxDofs / 256];
#pragma unroll
    for (int q = 0; q < kMaxBoxDofs / 256; q++) {
      const int i = tid + 256 * q;
      y[q] = i < n ? dot_ld(Ai + i, n, xa, n) : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kMaxBoxDofs / 256; q++) {
      const int i = tid + 256 * q;
      if (i < n) x[act[i]] = y[q];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Compressed forward apply b = A x (SkelFactorization::sparse_matvec,
// skel_factorization.cpp:429-504): the solve's steps undone in reverse
// order.  Per compressed box, levels leaves first:
//   b_s += T b_r;  b_r += U b_s;  b_r = X_rr b_r
// (X_rr may follow each box at once: no later box touches its redundant
// entries); then every root b = A_root b; then, root side first,
//   b_s += L b_r;  b_r += T^T b_s.
// One wave per box.
// ---------------------------------------------------------------------------

template <typename R>
__global__ __launch_bounds__(64) void k_fwd_up(DevState s, int start, double* x) {
  __shared__ double xs[kMaxBoxDofs];
  __shared__ double xr[kMaxBoxDofs];
  const int b = start + blockIdx.x;
  const int k = s.rank[b];
  if (k == 0) return;
  const int lane = threadIdx.x;
  const int r = s.act_cnt[b] - k;
  const FacLayout fl = FacLayout::make(k, r);
  const R* F = reinterpret_cast<const R*>(s.fac_pool) + s.fac_off[b];
  const R* Xf = reinterpret_cast<const R*>(s.fwd_pool) + s.fac_off[b];
  const int* sk = s.out_list + (int64_t)b * s.ncap;
  const int* rd = s.red_list + (int64_t)b * s.ncap;
  for (int a = lane; a < k; a += 64) xs[a] = x[sk[a]];
  for (int i = lane; i < r; i += 64) xr[i] = x[rd[i]];
  wave_sync();
  for (int a = lane; a < k; a += 64) xs[a] += dot_ld(F + fl.T + a, k, xr, r);
  wave_sync();
  for (int i = lane; i < r; i += 64) xr[i] += dot_ld(F + fl.U + i, r, xs, k);
  wave_sync();
  double y[kMaxBoxDofs / 64];
#pragma unroll
  for (int q = 0; q < kMaxBoxDofs / 64; q++) {
    const int i = lane + 64 * q;
    y[q] = i < r ? dot_ld(Xf + i, r, xr, r) : 0.0;
  }
  for (int a = lane; a < k; a += 64) x[sk[a]] = xs[a];
#pragma unroll
  for (int q = 0; q < kMaxBoxDofs / 64; q++) {
    const int i = lane + 64 * q;
    if (i < r) x[rd[i]] = y[q];
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_fwd_root(DevState s, double* x) {
  __shared__ double xa[kMaxBoxDofs];
  const int tid = threadIdx.x;
  const int n = s.act_cnt[blockIdx.x];
  const int* act = s.act_list + (int64_t)blockIdx.x * s.ncap;
  const R* A = reinterpret_cast<const R*>(s.root_fwd) + (int64_t)s.ncap * s.ncap * blockIdx.x;
  for (int i = tid; i < n; i += 256) xa[i] = x[act[i]];
  __syncthreads();
  double y[kMaxBoxDofs / 256];
#pragma unroll
  for (int q = 0; q < kMaxBoxDofs / 256; q++) {
    const int i = tid + 256 * q;
    y[q] = i < n ? dot_ld(A + i, n, xa, n) : 0.0;
  }
#pragma unroll
  for (int q = 0; q < kMaxBoxDofs / 256; q++) {
    const int i = tid + 256 * q;
    if (i < n) x[act[i]] = y[q];
  }
}

template <typename R>
__global__ __launch_bounds__(64) void k_fwd_down(DevState s, int start, double* x) {
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
  for (int a = lane; a < k; a += 64) xs[a] = x[sk[a]];
  for (int i = lane; i < r; i += 64) xr[i] = x[rd[i]];
  wave_sync();
  for (int a = lane; a < k; a += 64) xs[a] += dot_ld(F + fl.L + a, k, xr, r);
  wave_sync();
  for (int i = lane; i < r; i += 64) xr[i] += dot_ld(F + fl.T + (int64_t)k * i, 1, xs, k);
  wave_sync();
  for (int a = lane; a < k; a += 64) x[sk[a]] = xs[a];
  for (int i = lane; i < r; i += 64) x[rd[i]] = xr[i];
}

}  // namespace

void launch_reset(const DevState& s, hipStream_t st) {
  k_reset<<<1, 64, 0, st>>>(s);
}

void launch_gaussian(double* out, int64_t n, uint64_t seed, hipStream_t st) {
  k_gaussian<<<256, 256, 0, st>>>(out, n, seed);
}

template <typename R, int PDE>
void launch_rid_wide(const DevGeom& g, const DevTree& t, const DevState& s, int start,
                     int level, int count, double tol, hipStream_t st) {
  k_rid_wide<R, PDE><<<count, 64 * kRidWideWaves, 0, st>>>(g, t, s, start, level, tol);
}

template <typename R>
void launch_factor_stage(int stage, const DevGeom& g, const DevTree& t,
                         const DevState& s, int level, int start, int count,
                         double tol, const LaunchCaps& caps, hipStream_t st, int skip,
                         hipStream_t st_big) {
  if (count <= 0) return;
  switch (stage) {
    case kStageActive:
      k_active<<<count, 64, 0, st>>>(t, s, start);
      break;
    case kStageRows:
      k_rows<<<count, 64, 0, st>>>(g, t, s, start, level, tol);
      break;
    case kStageScan:
      if (count >= kScanMulti) {
        const int nblk = (count + 255) / 256;
        k_scan_blocks<<<nblk, 256, 0, st>>>(s, start, count);
        k_scan_totals<<<1, 1024, 0, st>>>(s, nblk, level);
        k_scan_place<<<nblk, 256, 0, st>>>(s, start, count, level, nblk);
      } else {
        k_scan<<<1, 1024, 0, st>>>(s, start, count, level);
      }
      break;
    case kStageAssemble:
      if (!(skip & kSkipAssemble))
        k_assemble<R><<<min(count, caps.assemble), 256, 0, st>>>(g, t, s, start, level);
      break;
    case kStageRid:
      if (s.rid_mode != 0) {
        if (skip & kSkipRid) {
          // no 16- or 32-row sketch on this level (launch plan)
        } else if (g.pde == kPdeStokes) {
          k_rid<R, kPdeStokes><<<count, 256, 0, st>>>(g, t, s, start, level, tol);
        } else if (g.pde == kPdeLaplaceNeumann) {
          k_rid<R, kPdeLaplaceNeumann><<<count, 256, 0, st>>>(g, t, s, start, level, tol);
        } else {
          k_rid<R, kPdeLaplace><<<count, 256, 0, st>>>(g, t, s, start, level, tol);
        }
        if (level <= kRidWideTop && !(skip & kSkipRidWide)) {
          if (g.pde == kPdeStokes)
            launch_rid_wide<R, kPdeStokes>(g, t, s, start, level, count, tol, st);
          else if (g.pde == kPdeLaplaceNeumann)
            launch_rid_wide<R, kPdeLaplaceNeumann>(g, t, s, start, level, count, tol, st);
          else
            launch_rid_wide<R, kPdeLaplace>(g, t, s, start, level, count, tol, st);
        }
      }
      break;
    case kStageQrcp: {
      // A register-QRCP kernel whose grid covers its whole list (known from
      // the level's box count) runs one box per workgroup, without the
      // grid-stride loop and the registers it keeps live (no spills).
      const int jobs_max = count * kMaxChunks;
      if (skip & kSkipRegQrcp) {
        // no register-QRCP box on this level (launch plan)
      } else if (s.rid_mode != 0) {
        // Only the boxes without a successful sketch reach the full QRCP.
        k_rowcomp<R, 4><<<min(jobs_max, caps.rowcomp4), kXThreads, 0, st>>>(g, t, s, start,
                                                                            level);
        if (count <= caps.qrcp_id) {
          k_qrcp_id<R, 4, 8, 8, true, false><<<count, 512, 0, st>>>(g, t, s, start, level, tol);
          k_qrcp_id<R, 4, 8, 16, true, false><<<count, kXThreads, 0, st>>>(g, t, s, start,
                                                                           level, tol);
        } else {
          const int grid = caps.qrcp_id;
          k_qrcp_id<R, 4, 8, 8, true><<<grid, 512, 0, st>>>(g, t, s, start, level, tol);
          k_qrcp_id<R, 4, 8, 16, true><<<grid, kXThreads, 0, st>>>(g, t, s, start, level, tol);
        }
      } else {
        k_rowcomp<R, 4><<<min(2 * count, caps.rowcomp4), kXThreads, 0, st>>>(g, t, s, start,
                                                                            level);
        k_qrcp_id<R, 4, 8, 8, false><<<count, 512, 0, st>>>(g, t, s, start, level, tol);
        k_qrcp_id<R, 4, 8, 16, false><<<count, kXThreads, 0, st>>>(g, t, s, start, level, tol);
      }
      if (!(skip & kSkipReg8Qrcp)) {
        k_rowcomp<R, 8><<<min(2 * count, caps.rowcomp8), kXThreads, 0, st>>>(g, t, s, start,
                                                                            level);
        if (count <= caps.qrcp_id8)
          k_qrcp_id<R, 8, 4, 16, true, false><<<count, kXThreads, 0, st>>>(g, t, s, start,
                                                                           level, tol);
        else
          k_qrcp_id<R, 8, 4, 16, true><<<caps.qrcp_id8, kXThreads, 0, st>>>(g, t, s, start,
                                                                            level, tol);
      }
      if (!(skip & kSkipAssemble))
        k_qrcp<R, 512><<<min(count, caps.qrcp), 512, 0, st>>>(s, start, level, tol);
      break;
    }
    case kStageDecouple:
      if (count >= caps.dsweep) {  // one wave per box with |active| <= 64
        launch_decouple_sweep<R>(g, t, s, start, count, st);
      } else if (count >= caps.dsplit) {
        k_decouple_small<R, 0, kSmallBox1><<<count, 256, 0, st>>>(g, t, s, start);
        k_decouple_small<R, kSmallBox1, kSmallBox><<<count, 256, 0, st>>>(g, t, s, start);
      } else {
        k_decouple_small<R, 0, kSmallBox><<<count, 256, 0, st>>>(g, t, s, start);
      }
      // the boxes above the small-box bound on st_big when given (their
      // own stream: the two kernels touch different boxes), there with
      // 1024 threads (level 1's two 84-DOF boxes: the P fill 47k -> 16k
      // cycles, 12 % less in all)
      if (st_big)
        k_decouple<R, 1024><<<min(count, 512), 1024, 0, st_big>>>(g, t, s, start, count);
      else
        k_decouple<R, 256><<<min(count, 512), 256, 0, st>>>(g, t, s, start, count);
      break;
    default:
      break;
  }
}

template <typename R>
void launch_root(const DevGeom& g, const DevTree& t, const DevState& s,
                 int n_roots, bool factor_lu, hipStream_t st) {
  k_active<<<n_roots, 64, 0, st>>>(t, s, 0);
  k_root<R, 256><<<dim3(n_roots, kRootSplit), 256, 0, st>>>(g, t, s);
  k_root_inv<R><<<n_roots, kRootTR * kRootTC, 0, st>>>(s, factor_lu ? 1 : 0);
  k_root_inv_blk<R><<<n_roots, kRootBlkMax, 0, st>>>(s, factor_lu ? 1 : 0);
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
void launch_root_solve(const DevState& s, int n_roots, SolveVec v, hipStream_t st) {
  k_root_solve<R><<<n_roots, 256, 0, st>>>(s, v);
}

template <typename R>
void launch_fwd_up(const DevState& s, int start, int count, double* x, hipStream_t st) {
  if (count > 0) k_fwd_up<R><<<count, 64, 0, st>>>(s, start, x);
}
template <typename R>
void launch_fwd_root(const DevState& s, int n_roots, double* x, hipStream_t st) {
  k_fwd_root<R><<<n_roots, 256, 0, st>>>(s, x);
}
template <typename R>
void launch_fwd_down(const DevState& s, int start, int count, double* x, hipStream_t st) {
  if (count > 0) k_fwd_down<R><<<count, 64, 0, st>>>(s, start, x);
}

#define IEGPU_INSTANTIATE(R)                                                   \
  template void launch_factor_stage<R>(int, const DevGeom&, const DevTree&,   \
                                       const DevState&, int, int, int, double, \
                                       const LaunchCaps&, hipStream_t, int,    \
                                       hipStream_t);                           \
  template void launch_root<R>(const DevGeom&, const DevTree&,                \
                               const DevState&, int, bool, hipStream_t);       \
  template void launch_sweep_up<R>(const DevState&, int, int, SolveVec, bool, \
                                   bool, hipStream_t);                         \
  template void launch_diag<R>(const DevState&, int, int, SolveVec,           \
                               hipStream_t);                                   \
  template void launch_sweep_down<R>(const DevState&, int, int, SolveVec,     \
                                     hipStream_t);                             \
  template void launch_root_solve<R>(const DevState&, int, SolveVec, hipStream_t); \
  template void launch_fwd_up<R>(const DevState&, int, int, double*, hipStream_t); \
  template void launch_fwd_root<R>(const DevState&, int, double*, hipStream_t);    \
  template void launch_fwd_down<R>(const DevState&, int, int, double*, hipStream_t);

IEGPU_INSTANTIATE(double)
IEGPU_INSTANTIATE(float)

}  // namespace iegpu
