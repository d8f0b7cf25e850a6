// Device-side data layout and helpers for the RS factor + solve kernels
// (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace iegpu {

constexpr int kNumProxy = 128;          // ie_solver/kernel/kernel.h:11
constexpr double kRadiusRatio = 1.5;    // ie_solver/kernel/kernel.h:12
constexpr int kMinCompress = 16;        // skel_factorization.h:11
constexpr int kMaxBoxDofs = 1024;       // hard cap on |active| of one box
constexpr int kWave = 64;

enum : int { kPdeLaplace = 0, kPdeLaplaceNeumann = 1, kPdeStokes = 2 };

// Error bits raised by device code (read back once after a factorization).
enum : int {
  kErrBoxCap = 1,      // a box's active list exceeded the per-box capacity
  kErrIdPool = 2,      // ID-matrix scratch pool exhausted
  kErrFacPool = 4,     // factor pool exhausted
  kErrSingular = 8,    // exact zero pivot in an LU
  kErrRowPool = 16,    // row-index pool exhausted
  kErrDecPool = 32,    // decouple operand pool exhausted
  kErrPlan = 64,       // a list whose kernels the launch plan skipped was not empty
};

// Launch plan (DevState::plan_skip, per level): the ID-stage kernels not
// launched because their lists were empty in the last factorization of the
// same inputs.  k_root checks every skipped list and raises kErrPlan.
enum : int {
  kSkipAssemble = 1,   // k_assemble, k_qrcp (general QRCP list)
  kSkipRegQrcp = 2,    // k_rowcomp<4>, k_qrcp_id n <= 32 and n <= 64 shapes
  kSkipReg8Qrcp = 4,   // k_rowcomp<8>, k_qrcp_id<8, 4, 16>
  kSkipRid = 8,        // k_rid (no 16- or 32-row sketch on the level)
  kSkipRidWide = 16,   // k_rid_wide (no 64-row sketch on the level)
};

// How a box's interpolative decomposition runs.
enum QrcpKind : int {
  kQrcpNone = 0,    // not compressed (too few DOFs or no ID rows)
  kQrcpReg = 1,     // register QRCP on the ID matrix (m <= 512, n <= 64)
  kQrcpStack = 2,   // row-compressed in chunks first, then register QRCP
  kQrcpLds = 3,     // general QRCP (LDS or the box's scratch slot)
  kQrcpReg8 = 4,    // wide register QRCP (m <= 256, n <= 128)
  kQrcpStack8 = 5,  // row-compressed, then wide register QRCP
};
constexpr int kRegMaxRows = 512, kRegMaxCols = 64;
constexpr int kRegMaxRows8 = 256, kRegMaxCols8 = 128;
constexpr int kMaxChunks = 8;
constexpr int kMaxLevels = 64;  // tree depth bound (per-level job counters)
constexpr int kSmallBox = 64;     // decouple with every operand in LDS
constexpr int kSmallBox1 = 44;    // smaller LDS class of that decouple
constexpr int kRootRegMax = 192;  // root blocks inverted in registers
constexpr int kRootBlkMax = 512;  // root blocks inverted by the blocked kernel

// Per-level lists of the boxes each QRCP kernel takes, built by prefix
// sums (k_scan) so that a kernel's workgroups map onto exactly its boxes.
enum : int {
  kLstReg = 0,      // register QRCP (kQrcpReg, kQrcpStack), n > 32
  kLstLds = 1,      // general QRCP (kQrcpLds)
  kLstReg8 = 2,     // wide register QRCP (kQrcpReg8, kQrcpStack8)
  kLstRegHalf = 3,  // register QRCP, n <= 32 (8-wave workgroups)
  kLstRid = 4,      // randomized ID (sketch width > 0)
  kLstFallback = 5, // randomized mode: register-QRCP boxes without a successful sketch
  kNumLists = 6,
};
constexpr int kHalfCols = 32;

// Randomized ID: a Gaussian sketch of at most kRidRows rows (a multiple of
// 16 per box); a cut must show within kRidRows - kRidOver pivots, else the
// box goes through the full QRCP.  Boxes with at most kRidMaxNear row DOFs
// and 64 columns take it.
constexpr int kRidRows = 32;
constexpr int kRidOver = 6;
constexpr int kRidMaxNear = 256;
constexpr int kRidTableRows = 2 * kRidMaxNear + 4 * kNumProxy;  // Omega rows
// The wide form for the few large boxes of the top levels (level <=
// kRidWideTop): a sketch of kRidRowsW rows, up to kRidWideCols columns and
// kRidMaxNearW row DOFs, with its own Gaussian table.
constexpr int kRidRowsW = 64;
constexpr int kRidWideCols = 128;
constexpr int kRidMaxNearW = 1280;  // the level-1/2 row sets of small problems in a batch
constexpr int kRidWideTop = 4;
constexpr int kRidTableRowsW = 2 * kRidMaxNearW + 4 * kNumProxy;

// Rank model: an MLP kFeat -> kHid (tanh) -> kHid (tanh) -> 1 on per-box
// features, predicting the ID rank.  Parameter blob (floats): feature
// means and scales (kFeat each), W1 [kFeat][kHid], b1, W2 [kHid][kHid], b2,
// w3 [kHid], b3, then the target mean and scale.
constexpr int kFeat = 9;
constexpr int kHid = 32;
enum : int {
  kMlpMu = 0,
  kMlpSd = kMlpMu + kFeat,
  kMlpW1 = kMlpSd + kFeat,
  kMlpB1 = kMlpW1 + kFeat * kHid,
  kMlpW2 = kMlpB1 + kHid,
  kMlpB2 = kMlpW2 + kHid * kHid,
  kMlpW3 = kMlpB2 + kHid,
  kMlpB3 = kMlpW3 + kHid,
  kMlpYMu = kMlpB3 + 1,
  kMlpYSd = kMlpYMu + 1,
  kMlpLen = kMlpYSd + 1,
  kMlpParams = kMlpYMu - kMlpW1,  // trained parameters (W1 .. b3)
};

// Row chunks of a tall ID matrix, at most `rows` rows each.
__host__ __device__ inline int id_chunks(int m, int rows) {
  return (m + rows - 1) / rows;
}

struct DevGeom {
  const double* pts;   // 2 per point
  const double* nrm;   // 2 per point
  const double* curv;  // 1 per point
  const double* wts;   // 1 per point
  int n_points;
  int sol_dim;
  int pde;
  // Independent problems (a batch): problem p owns points
  // [prob_off[p], prob_off[p+1]); n_prob == 1 for a single boundary.
  int n_prob;
  const int* prob_off;
};

// Static tree, slot-indexed (level-major).
struct DevTree {
  int n_boxes;
  int n_levels;
  const int* level;
  const int* leaf;
  const int* child;  // 4 per slot
  const double* cx;
  const double* cy;
  const double* side;
  const int* orig_off;
  const int* orig_cnt;
  const int* orig_dofs;
  const int* src_off;
  const int* src_cnt;
  const int* src;
};

// Per-box state produced by the factorization.
struct DevState {
  int ncap;            // per-box list capacity
  int* act_cnt;        // |active|
  int* act_list;       // ncap per box
  int* chpos;          // 4 per box: position of each child's block in active
  int* out_cnt;        // skeleton count if compressed, else |active|
  int* out_list;       // ncap per box: skeleton (pivot order) or active
  int* red_list;       // ncap per box: redundant DOFs
  int* perm;           // ncap per box: pivot order (positions into active)
  int* rank;           // k (0: not compressed)
  int* upd_kind;       // 0: no carried update, 1: carried update stored
  int* m_rows;         // ID matrix height
  int* qkind;          // ID route of the box (QrcpKind)
  int nbox;            // box slots
  int* lists;          // kNumLists x nbox: a level's list L starts at its first slot
  int* lcnt;           // kMaxLevels x kNumLists x 2: list lengths (second word spare)
  int* jobs;           // (box, chunk) row-compression jobs, kMaxChunks per box slot
  int* njobs;          // per level: number of row-compression jobs
  int64_t* id_sz;      // scratch doubles/floats needed (per box, level-local)
  int64_t* row_sz;
  int64_t* fac_sz;
  int64_t* id_off;
  int64_t* row_off;
  int64_t* fac_off;
  // Operands of a general decouple (boxes above kSmallBox DOFs whose blocks
  // do not fit LDS): level-local, on the decouple chain only.
  int64_t* dec_sz;
  int64_t* dec_off;
  void* dec_pool;
  int64_t dec_cap;
  void* id_pool;
  int* row_pool;
  void* fac_pool;
  int64_t id_cap;
  int64_t row_cap;
  int64_t fac_cap;
  int64_t* fac_top;    // running factor-pool offset across levels
  int64_t* scan_blk;   // a level's per-block scan totals (16 words per 256
                       // boxes, then the factor-pool base)
  int* err;            // [0] error bits, [1] / [2] / [3] decouple / ID / row
                       // pool size a level needed on overflow (1024 elements)
  void* root_mat;      // per root: 2 ncap^2 elements (matrix + inverse)
  int* root_piv;       // per root: ncap + 64
  // Kept for the compressed forward apply when requested (else null): X_rr
  // of every compressed box at its factor-slot offset, and each root's
  // block (ncap^2 elements per root), as formed before their inversion.
  void* fwd_pool;
  void* root_fwd;
  // Randomized ID (rid_mode 0: off; 1: sketch width rid_width for every
  // eligible box; 2: width from the rank model).
  int rid_mode;
  int rid_width;
  int* sk_w;           // per box: sketch rows (0: full QRCP)
  int* rid_done;       // per box: 1 once the randomized ID succeeded
  const double* omega; // kRidTableRows x kRidRows Gaussian table (row-major)
  const double* omega_w;  // kRidTableRowsW x kRidRowsW table of the wide form
  const float* rank_model;  // kMlpLen floats (rid_mode 2)
  int rid_margin;           // ranks added to the prediction before sizing
  float* feat;              // kFeat per box, written by k_rows when non-null
  // Incremental re-factorization after a geometry update (null: every box).
  // A box with dirty[b] == 0 keeps its lists, rank and factor slot from the
  // previous factorization; the per-level kernels skip it and k_scan gives
  // it no scratch, no list entry and no new factor slot.
  const int* dirty;
  // Factor-pool elements in use before an incremental run (0 otherwise): a
  // re-factored box keeps its slot when that slot lies below this mark and
  // its new factors fit it.
  int64_t kept_top;
  // Per level: kSkip* bits of the launch plan in use (null: every kernel).
  const int* plan_skip;
};

// False once a pool overflow was reported: the level's remaining kernels
// stop and the host repeats the factorization with larger pools.
__device__ inline bool pools_ok(const DevState& s) {
  return (s.err[0] & (kErrIdPool | kErrRowPool | kErrFacPool | kErrBoxCap | kErrDecPool)) == 0;
}

__device__ inline bool box_clean(const DevState& s, int b) {
  return s.dirty != nullptr && s.dirty[b] == 0;
}

__device__ inline int* list_slot(const DevState& s, int level, int L) {
  return s.lcnt + 2 * (level * kNumLists + L);
}
// Row-compression jobs (chunks) of box b.
__device__ inline int row_jobs(const DevState& s, int b) {
  const int kind = s.qkind[b];
  if (kind == kQrcpStack) return id_chunks(s.m_rows[b], kRegMaxRows);
  if (kind == kQrcpStack8) return id_chunks(s.m_rows[b], kRegMaxRows8);
  return 0;
}

__device__ inline int list_count(const DevState& s, int level, int L) {
  return list_slot(s, level, L)[0];
}

// The kSkip* bits a level could have taken: the kernels whose lists are
// empty (in the randomized form the register shapes take only the fallback
// list).
__device__ inline int plan_skippable(const DevState& s, int level) {
  const int reg = s.rid_mode != 0 ? list_count(s, level, kLstFallback)
                                  : list_count(s, level, kLstReg) +
                                        list_count(s, level, kLstRegHalf);
  // kLstRid's second slot int counts the level's 64-row sketches.
  const int wide = list_slot(s, level, kLstRid)[1];
  int m = 0;
  if (list_count(s, level, kLstLds) == 0) m |= kSkipAssemble;
  if (reg == 0) m |= kSkipRegQrcp;
  if (list_count(s, level, kLstReg8) == 0) m |= kSkipReg8Qrcp;
  if (list_count(s, level, kLstRid) == wide) m |= kSkipRid;
  if (wide == 0) m |= kSkipRidWide;
  return m;
}
__device__ inline int list_at(const DevState& s, int start, int L, int idx) {
  return s.lists[(int64_t)L * s.nbox + start + idx];
}

// Offsets of the factor blocks inside one compressed box's factor slot.
// Layout: T (k x r) | L (k x r) | U (r x k) | X_rr^{-1} (r x r) | r spare |
//         carried Schur block (k x k).  Column-major throughout.
struct FacLayout {
  int64_t T, L, U, X, piv, S;
  __host__ __device__ static FacLayout make(int k, int r) {
    FacLayout f;
    f.T = 0;
    f.L = (int64_t)k * r;
    f.U = 2 * (int64_t)k * r;
    f.X = 3 * (int64_t)k * r;
    f.piv = f.X + (int64_t)r * r;
    f.S = f.piv + r;
    return f;
  }
};

// Upper bound of a box's factor slot from |active| alone:
// 3kr + r^2 + r + k^2 <= n^2 + n^2/4 + n for k + r = n.
__host__ __device__ inline int64_t fac_slot_size(int n) {
  const int64_t nn = (int64_t)n * n;
  int64_t s = nn + nn / 4 + n + 8;
  return (s + 7) & ~int64_t(7);
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline int wave_sum_int(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---------------------------------------------------------------------------
// Wave64 reductions with DPP (gfx9 data-parallel primitives): pairs, quads,
// half rows and rows by quad_perm / row mirrors, then rows combined by
// row_bcast15 / row_bcast31, leaving the total in lane 63, which readlane
// broadcasts.  No LDS round trip, unlike shuffles through ds_bpermute.
// ---------------------------------------------------------------------------

namespace dpp {
constexpr int kQuadSwap1 = 0xB1;     // quad_perm [1,0,3,2]
constexpr int kQuadSwap2 = 0x4E;     // quad_perm [2,3,0,1]
constexpr int kRowHalfMirror = 0x141;
constexpr int kRowMirror = 0x140;
constexpr int kRowBcast15 = 0x142;
constexpr int kRowBcast31 = 0x143;
constexpr int kQuadSwap3 = 0x1B;     // quad_perm [3,2,1,0]
constexpr int kQuadBitrev = 0xD8;    // quad_perm [0,2,1,3]
constexpr int kRowRor4 = 0x124;
constexpr int kRowRor8 = 0x128;
}  // namespace dpp

template <int CTRL, int ROWS>
__device__ inline double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int CTRL, int ROWS>
__device__ inline float dpp_f32(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false));
}
template <int CTRL, int ROWS>
__device__ inline int dpp_i32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xF, false);
}

// Index of the calling wave in its workgroup, as a wave-uniform (scalar)
// value: the compiler cannot tell that threadIdx.x / 64 is uniform.
__device__ inline int wave_id() {
  return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}

// Value of `v` in lane `lane` (wave-uniform index), for every lane.
__device__ inline double lane_value(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline float lane_value(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ inline int lane_value(int v, int lane) {
  return __builtin_amdgcn_readlane(v, lane);
}

// Sum over the 64 lanes, returned in every lane.
__device__ inline double wave_total(double v) {
  v += dpp_f64<dpp::kQuadSwap1, 0xF>(v);
  v += dpp_f64<dpp::kQuadSwap2, 0xF>(v);
  v += dpp_f64<dpp::kRowHalfMirror, 0xF>(v);
  v += dpp_f64<dpp::kRowMirror, 0xF>(v);
  v += dpp_f64<dpp::kRowBcast15, 0xA>(v);
  v += dpp_f64<dpp::kRowBcast31, 0xC>(v);
  return lane_value(v, 63);
}
__device__ inline float wave_total(float v) {
  v += dpp_f32<dpp::kQuadSwap1, 0xF>(v);
  v += dpp_f32<dpp::kQuadSwap2, 0xF>(v);
  v += dpp_f32<dpp::kRowHalfMirror, 0xF>(v);
  v += dpp_f32<dpp::kRowMirror, 0xF>(v);
  v += dpp_f32<dpp::kRowBcast15, 0xA>(v);
  v += dpp_f32<dpp::kRowBcast31, 0xC>(v);
  return lane_value(v, 63);
}

// (value, index) with the largest value over the wave; ties go to the
// smaller index (the first maximum, as LAPACK idamax).  Lanes with nothing
// to offer pass value -1 and index INT_MAX.
template <typename R>
__device__ inline void argmax_step(R& v, int& i, R ov, int oi) {
  if (ov > v || (ov == v && oi < i)) {
    v = ov;
    i = oi;
  }
}
template <typename R, int CTRL, int ROWS>
__device__ inline R dpp_any(R v) {
  if constexpr (sizeof(R) == 8) return dpp_f64<CTRL, ROWS>(v);
  else return dpp_f32<CTRL, ROWS>(v);
}
template <typename R>
__device__ inline int wave_argmax(R v, int i) {
  // Lanes outside the row mask of a bcast step read 0 from the DPP move;
  // mask them back out so they cannot win.
  argmax_step(v, i, dpp_any<R, dpp::kQuadSwap1, 0xF>(v), dpp_i32<dpp::kQuadSwap1, 0xF>(i));
  argmax_step(v, i, dpp_any<R, dpp::kQuadSwap2, 0xF>(v), dpp_i32<dpp::kQuadSwap2, 0xF>(i));
  argmax_step(v, i, dpp_any<R, dpp::kRowHalfMirror, 0xF>(v), dpp_i32<dpp::kRowHalfMirror, 0xF>(i));
  argmax_step(v, i, dpp_any<R, dpp::kRowMirror, 0xF>(v), dpp_i32<dpp::kRowMirror, 0xF>(i));
  const int row = (threadIdx.x & 63) >> 4;
  {
    const R ov = dpp_any<R, dpp::kRowBcast15, 0xA>(v);
    const int oi = dpp_i32<dpp::kRowBcast15, 0xA>(i);
    if (row & 1) argmax_step(v, i, ov, oi);
  }
  {
    const R ov = dpp_any<R, dpp::kRowBcast31, 0xC>(v);
    const int oi = dpp_i32<dpp::kRowBcast31, 0xC>(i);
    if (row & 2) argmax_step(v, i, ov, oi);
  }
  return lane_value(i, 63);
}

// ---------------------------------------------------------------------------
// Four-column interleave.  A wave holding four columns ("classes" 0..3) of
// a matrix, row i on lane i % 64, keeps class sl ^ quad_class(l) in slot sl
// of lane l.  The four column sums then need two pair exchanges followed by
// one reduction, instead of four separate wave reductions: after the
// exchanges lane l carries only class quad_class(l).
// ---------------------------------------------------------------------------

__device__ inline int quad_class(int lane) {
  return ((lane & 1) << 1) | ((lane >> 1) & 1);
}

template <typename R>
__device__ inline R shfl_xor_any(R v, int o) {
  return __shfl_xor(v, o, 64);
}

// Per-lane partials p[sl] (slot layout above) -> the wave total of class
// quad_class(lane), in every lane.
template <typename R>
__device__ inline R wave_total4(const R (&p)[4]) {
  const R s0 = p[0] + dpp_any<R, dpp::kQuadSwap1, 0xF>(p[2]);
  const R s1 = p[1] + dpp_any<R, dpp::kQuadSwap1, 0xF>(p[3]);
  R t = s0 + dpp_any<R, dpp::kQuadSwap2, 0xF>(s1);
  t += dpp_any<R, dpp::kRowRor4, 0xF>(t);
  t += dpp_any<R, dpp::kRowRor8, 0xF>(t);
  t += shfl_xor_any(t, 16);
  t += shfl_xor_any(t, 32);
  return t;
}

// Spreads class totals (lane l holding class quad_class(l)) back to the
// slot layout: out[sl] = total of class sl ^ quad_class(lane).
template <typename R>
__device__ inline void spread4(R t, R (&out)[4]) {
  out[0] = t;
  out[1] = dpp_any<R, dpp::kQuadSwap2, 0xF>(t);
  out[2] = dpp_any<R, dpp::kQuadSwap1, 0xF>(t);
  out[3] = dpp_any<R, dpp::kQuadSwap3, 0xF>(t);
}

// Class totals -> lane l of every quad holds class (l & 3).
template <typename R>
__device__ inline R class_by_lane(R t) {
  return dpp_any<R, dpp::kQuadBitrev, 0xF>(t);
}

// Eight-column interleave (same scheme, three exchange steps): slot sl of
// lane l carries class sl ^ oct_class(l), the bit-reversed (l & 7).  The
// third exchange takes the lane 4 away within the row (row_ror:4), whose
// low three bits differ in bit 2 only.
__device__ inline int oct_class(int lane) {
  return ((lane & 1) << 2) | (lane & 2) | ((lane >> 2) & 1);
}

template <typename R>
__device__ inline R wave_total8(const R (&p)[8]) {
  R s[4];
#pragma unroll
  for (int i = 0; i < 4; i++) s[i] = p[i] + dpp_any<R, dpp::kQuadSwap1, 0xF>(p[i + 4]);
  const R u0 = s[0] + dpp_any<R, dpp::kQuadSwap2, 0xF>(s[2]);
  const R u1 = s[1] + dpp_any<R, dpp::kQuadSwap2, 0xF>(s[3]);
  R t = u0 + dpp_any<R, dpp::kRowRor4, 0xF>(u1);
  t += dpp_any<R, dpp::kRowRor8, 0xF>(t);
  t += shfl_xor_any(t, 16);
  t += shfl_xor_any(t, 32);
  return t;
}

template <typename R>
__device__ inline void spread8(R t, R (&out)[8]) {
  const R r4 = dpp_any<R, dpp::kRowRor4, 0xF>(t);
  out[0] = t;
  out[1] = r4;
  out[2] = dpp_any<R, dpp::kQuadSwap2, 0xF>(t);
  out[3] = dpp_any<R, dpp::kQuadSwap2, 0xF>(r4);
  out[4] = dpp_any<R, dpp::kQuadSwap1, 0xF>(t);
  out[5] = dpp_any<R, dpp::kQuadSwap1, 0xF>(r4);
  out[6] = dpp_any<R, dpp::kQuadSwap3, 0xF>(t);
  out[7] = dpp_any<R, dpp::kQuadSwap3, 0xF>(r4);
}

template <typename R>
__device__ inline R class_by_lane8(R t) {
  const int lane = threadIdx.x & 63;
  return __shfl(t, (lane & ~7) | oct_class(lane & 7), 64);
}

// The interleave for CPW = 4 or 8 columns per wave.
template <int CPW>
__device__ inline int slot_class(int lane) {
  if constexpr (CPW == 4) return quad_class(lane);
  else return oct_class(lane);
}
template <typename R, int CPW>
__device__ inline R wave_total_cols(const R (&p)[CPW]) {
  if constexpr (CPW == 4) return wave_total4(p);
  else return wave_total8(p);
}
template <typename R, int CPW>
__device__ inline void spread_cols(R t, R (&out)[CPW]) {
  if constexpr (CPW == 4) spread4(t, out);
  else spread8(t, out);
}
template <typename R, int CPW>
__device__ inline R class_by_lane_cols(R t) {
  if constexpr (CPW == 4) return class_by_lane(t);
  else return class_by_lane8(t);
}

// Pivot key: a magnitude v >= 0 with its position in the low BITS bits,
// ordered as a non-negative double.  The largest key is the largest
// magnitude, ties (equal to 2^(BITS-52) relative in FP64) going to the
// smaller position, as LAPACK's first-maximum rule.  Unavailable entries
// carry -1.
template <typename R, int BITS = 7>
__device__ inline double pivot_key(R v, int pos) {
  constexpr uint64_t kMask = (uint64_t(1) << BITS) - 1;
  uint64_t b;
  if constexpr (sizeof(R) == 8)
    b = (uint64_t)__double_as_longlong(v) & ~kMask;
  else
    b = (uint64_t)__float_as_uint(v) << 32;
  return __longlong_as_double((long long)(b | (kMask - (uint64_t)pos)));
}
template <int BITS = 7>
__device__ inline int pivot_key_pos(double key) {
  constexpr int kMask = (1 << BITS) - 1;
  return kMask - (int)(__double_as_longlong(key) & kMask);
}

// DPP move that leaves lanes outside the row mask with their own value.
template <int CTRL, int ROWS>
__device__ inline double dpp_keep_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int l0 = (int)(b & 0xffffffffll), h0 = (int)(b >> 32);
  const int lo = __builtin_amdgcn_update_dpp(l0, l0, CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(h0, h0, CTRL, ROWS, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Maximum over the wave, in every lane.
__device__ inline double wave_max(double v) {
  v = fmax(v, dpp_f64<dpp::kQuadSwap1, 0xF>(v));
  v = fmax(v, dpp_f64<dpp::kQuadSwap2, 0xF>(v));
  v = fmax(v, dpp_f64<dpp::kRowHalfMirror, 0xF>(v));
  v = fmax(v, dpp_f64<dpp::kRowMirror, 0xF>(v));
  v = fmax(v, dpp_keep_f64<dpp::kRowBcast15, 0xA>(v));
  v = fmax(v, dpp_keep_f64<dpp::kRowBcast31, 0xC>(v));
  return lane_value(v, 63);
}

// Geometry of one DOF as a kernel target or source: its point, normal,
// quadrature weight, curvature and (Stokes) velocity component.
struct DofGeo {
  double x, y, n1, n2, w, k;
  int c;
};

__device__ inline DofGeo dof_geo(const DevGeom& g, int dof) {
  const bool stokes = g.pde == kPdeStokes;
  const int p = stokes ? dof >> 1 : dof;
  DofGeo d;
  d.x = g.pts[2 * p];
  d.y = g.pts[2 * p + 1];
  d.n1 = g.nrm[2 * p];
  d.n2 = g.nrm[2 * p + 1];
  d.w = g.wts[p];
  d.k = g.curv[p];
  d.c = stokes ? dof & 1 : 0;
  return d;
}

// 1 / d from the hardware reciprocal refined by one Newton step (within an
// ulp or two of the rounded quotient; the kernel entries take it in place
// of a full IEEE division).
__device__ inline double recip(double d) {
  const double y = __builtin_amdgcn_rcp(d);
  return fma(y, fma(-d, y, 1.0), y);
}

// Laplace double layer, ie_solver/kernel/kernel.cpp:123-165.  The
// coincident-point (diagonal) value applies only between boundary DOFs.
__device__ inline double laplace_entry(const DofGeo& t, const DofGeo& s, bool diag_ok) {
  const double scale = 1.0 / (2 * M_PI);
  if (diag_ok && t.x == s.x && t.y == s.y) return 0.5 + 0.5 * s.k * s.w * scale;
  const double r0 = t.x - s.x, r1 = t.y - s.y;
  return -s.w * scale * (r0 * s.n1 + r1 * s.n2) * recip(r0 * r0 + r1 * r1);
}

// Stokes velocity double layer (2 DOFs per point, DOF d -> point d/2,
// component d%2), ie_solver/kernel/kernel.cpp:212-290.
__device__ inline double stokes_entry(const DofGeo& t, const DofGeo& s, bool diag_ok) {
  const double scale = 1.0 / M_PI;
  const int a = t.c, b = s.c;
  const double tna = a ? t.n2 : t.n1;
  const double snb = b ? s.n2 : s.n1;
  if (diag_ok && t.x == s.x && t.y == s.y) {
    const double pot = -0.5 * s.k * s.w * scale;
    if (a == 0 && b == 0) return -0.5 + pot * s.n2 * s.n2 + s.w * tna * s.n1;
    if (a == 1 && b == 1) return -0.5 + pot * s.n1 * s.n1 + s.w * tna * s.n2;
    return -pot * s.n1 * s.n2 + s.w * t.n2 * s.n1;
  }
  const double r0 = t.x - s.x, r1 = t.y - s.y;
  const double d2 = r0 * r0 + r1 * r1;
  const double i2 = recip(d2);
  const double pot = s.w * scale * (r0 * s.n1 + r1 * s.n2) * (i2 * i2);
  const double ra = a ? r1 : r0, rb = b ? r1 : r0;
  return pot * ra * rb + s.w * tna * snb;
}

// Laplace Neumann (adjoint double layer with the target's normal, plus
// the rank-one term w_s that fixes the constant), ie_solver/kernel/
// kernel.cpp:80-92 and :168-209 (IMPROVE_CONDITION = 0).
__device__ inline double laplace_neumann_entry(const DofGeo& t, const DofGeo& s, bool diag_ok) {
  const double scale = 1.0 / (2 * M_PI);
  if (diag_ok && t.x == s.x && t.y == s.y) return s.w - 0.5 + 0.5 * s.k * s.w * scale;
  const double r0 = t.x - s.x, r1 = t.y - s.y;
  return s.w + s.w * scale * (r0 * t.n1 + r1 * t.n2) * recip(r0 * r0 + r1 * r1);
}

__device__ inline double kernel_entry(int pde, const DofGeo& t, const DofGeo& s,
                                      bool diag_ok = true) {
  if (pde == kPdeStokes) return stokes_entry(t, s, diag_ok);
  if (pde == kPdeLaplaceNeumann) return laplace_neumann_entry(t, s, diag_ok);
  return laplace_entry(t, s, diag_ok);
}
__device__ inline double kernel_entry(const DevGeom& g, int t, int s) {
  return kernel_entry(g.pde, dof_geo(g, t), dof_geo(g, s));
}

// Proxy point p of the circle radius r about (cx, cy):
// ie_solver/kernel/kernel.cpp:430-441.
__device__ inline void proxy_point(double cx, double cy, double r, int p,
                                   double* px, double* py, double* n1,
                                   double* n2) {
  const double ang = 2 * M_PI * p * (1.0 / kNumProxy);
  double sn, cs;
  sincos(ang, &sn, &cs);
  *px = cx + r * cs;
  *py = cy + r * sn;
  *n1 = cs;
  *n2 = sn;
}

// One row of a box's ID matrix (ie_solver/kernel/kernel.cpp:293-423 and
// 459-591): rows [0, cnt) are K(row DOF, box DOF), [cnt, 2 cnt)
// K(box DOF, row DOF), then the proxy block: for Laplace P = 128 rows
// box -> proxy point and P rows proxy point -> box; for Stokes the two
// components of each proxy point interleave (2P rows each way).  A proxy
// point acts as a DOF with the circle's outward normal and weight
// 2 pi r / 128.  `q` is the row side, `target` whether it is the target.
struct IdRow {
  DofGeo q;
  bool target;
  bool diag_ok;
};

__device__ inline IdRow id_row(const DevGeom& g, const int* rows, int cnt,
                               const double (*pxy)[kNumProxy], double pw, int i) {
  IdRow r;
  if (i < 2 * cnt) {
    const bool first = i < cnt;
    r.q = dof_geo(g, rows[first ? i : i - cnt]);
    r.target = first;
    r.diag_ok = true;
    return r;
  }
  const bool stokes = g.pde == kPdeStokes;
  const int P = stokes ? 2 * kNumProxy : kNumProxy;
  const int row = i - 2 * cnt;
  r.target = row < P;
  const int rr = r.target ? row : row - P;
  const int p = stokes ? rr >> 1 : rr;
  r.q.x = pxy[0][p];
  r.q.y = pxy[1][p];
  r.q.n1 = pxy[2][p];
  r.q.n2 = pxy[3][p];
  r.q.w = pw;
  r.q.k = 0;
  r.q.c = stokes ? rr & 1 : 0;
  r.diag_ok = false;
  return r;
}

__device__ inline double id_entry(int pde, const IdRow& r, const DofGeo& col) {
  return r.target ? kernel_entry(pde, r.q, col, r.diag_ok)
                  : kernel_entry(pde, col, r.q, r.diag_ok);
}

// Inner-circle test of ie_solver/kernel/kernel.cpp:384-394, evaluated
// without fused multiply-adds so that the comparison sees the same
// rounded distance as the host.
__device__ inline bool in_proxy_circle(const DevGeom& g, int dof, double cx,
                                       double cy, double side) {
#pragma clang fp contract(off)
  const int p = dof / g.sol_dim;
  const double dx = cx - g.pts[2 * p], dy = cy - g.pts[2 * p + 1];
  const double d = sqrt(dx * dx + dy * dy);
  return d < kRadiusRatio * side;
}

// DOF list a source box contributes to ID rows: a leaf's original DOFs,
// otherwise its current active DOFs (quadtree.cpp:1014-1022).
struct DofList {
  const int* p;
  int n;
};
__device__ inline DofList source_dofs(const DevTree& t, const DevState& s,
                                      int box) {
  DofList d;
  if (t.leaf[box]) {
    d.p = t.orig_dofs + t.orig_off[box];
    d.n = t.orig_cnt[box];
  } else {
    d.p = s.act_list + (int64_t)box * s.ncap;
    d.n = s.act_cnt[box];
  }
  return d;
}

// Carried Schur block of box c in its factor slot (null: none).
template <typename R>
__device__ const R* carried_block(const DevState& s, int c) {
  if (!s.upd_kind[c]) return nullptr;
  const R* F = reinterpret_cast<const R*>(s.fac_pool) + s.fac_off[c];
  const int k = s.rank[c];
  if (k == 0) return F;
  return F + FacLayout::make(k, s.act_cnt[c] - k).S;
}

// Dense block (matrix, then its inverse) and pivots of root r (slot r).
template <typename R>
__device__ inline R* root_block(const DevState& s, int r) {
  return reinterpret_cast<R*>(s.root_mat) + 2 * (int64_t)s.ncap * s.ncap * r;
}
__device__ inline int* root_pivots(const DevState& s, int r) {
  return s.root_piv + (int64_t)(s.ncap + 64) * r;
}

}  // namespace iegpu
