// C ABI (include/iegpu.h): context, device memory, and the level-by-level
// launch sequence of the RS factor + solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/iegpu.h"
#include "rs_device.hpp"
#include "rs_holes.hpp"
#include "rs_eval.hpp"
#include "rs_fmm.hpp"
#include "rs_kernels.hpp"
#include "rs_matvec.hpp"
#include "rs_mlp.hpp"
#include "tree.hpp"

using namespace iegpu;

static_assert(IEGPU_RANK_FEATURES == kFeat && IEGPU_RANK_MODEL_LEN == kMlpLen &&
                  IEGPU_RANK_PARAMS == kMlpParams,
              "rank model layout of iegpu.h");
static_assert(IEGPU_NUM_LAUNCH_CAPS == kNumCaps, "launch caps of iegpu.h");

namespace {

struct HipError {
  hipError_t e;
  const char* what;
};

#define HIP_TRY(expr)                                   \
  do {                                                  \
    hipError_t _e = (expr);                             \
    if (_e != hipSuccess) throw HipError{_e, #expr};    \
  } while (0)

// Device array of n elements in an allocation of cap >= n: shrinking keeps
// the allocation, growing past it allocates anew (a quarter more than asked
// when it grows again) and drops the contents.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  size_t cap = 0;
  DevBuf() {}
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    cap = 0;
  }
  void alloc(size_t count) {
    if (p && count <= cap) {
      n = count;
      return;
    }
    const size_t want = p ? std::max(count, cap + cap / 4) : count;
    release();
    if (count == 0) return;
    hipError_t e = hipMalloc(&p, want * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      throw HipError{e, "hipMalloc"};
    }
    n = count;
    cap = want;
  }
  void upload(const T* h, size_t count, hipStream_t st) {
    alloc(count);
    if (count) HIP_TRY(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, st));
  }
  void upload(const std::vector<T>& v, hipStream_t st) { upload(v.data(), v.size(), st); }
  // Elements [off, off + count) of an array of the same size.
  void upload_range(const T* h, size_t off, size_t count, hipStream_t st) {
    if (off + count > n) throw HipError{hipErrorInvalidValue, "upload_range"};
    if (count)
      HIP_TRY(hipMemcpyAsync(p + off, h + off, count * sizeof(T), hipMemcpyHostToDevice, st));
  }
};

// Rows of b renumbered: new row r = old row map[r], zero for map[r] < 0
// (rows of `row` elements; map on the device).  The gather goes through
// tmp, then back into b, which keeps its allocation when it is large enough.
template <typename T>
void remap_rows(DevBuf<T>& b, DevBuf<unsigned char>& tmp, const int* d_map, int n_rows,
                int old_rows, size_t row, hipStream_t st) {
  static_assert(sizeof(T) % 4 == 0, "32-bit words");
  if (!b.p) return;
  if (b.n != (size_t)old_rows * row) throw HipError{hipErrorInvalidValue, "remap_rows"};
  const size_t count = (size_t)n_rows * row, bytes = count * sizeof(T);
  if (bytes > tmp.cap || count > b.cap) HIP_TRY(hipStreamSynchronize(st));  // before a free
  tmp.alloc(std::max(tmp.n, bytes));
  launch_gather_rows(tmp.p, b.p, d_map, n_rows, (int)(row * sizeof(T) / 4), st);
  if (count > b.cap) HIP_TRY(hipStreamSynchronize(st));
  b.alloc(count);
  if (bytes) HIP_TRY(hipMemcpyAsync(b.p, tmp.p, bytes, hipMemcpyDeviceToDevice, st));
}

// Profile table rows: the factor stages of rs_kernels.hpp, then these.
enum : int {
  kProfRoot = kNumFactorStages,  // root active list + dense block + LU
  kProfSweepUp,                  // upward sweep (+ X_rr^{-1}) per level
  kProfRootSolve,                // root / bordered solve
  kProfSweepDown,                // downward sweep per level
  kProfBorder,                   // bordered-system factor (holes)
  kProfStages
};
constexpr int kProfLevels = 64;

// Per-(stage, level) device time from HIP events recorded around each
// launch on the launching stream; resolved when read.
struct Profiler {
  bool on = false;
  struct Mark {
    int stage, level;
    hipEvent_t a, b;
  };
  std::vector<Mark> pending;
  std::vector<hipEvent_t> spare;
  double ms[kProfStages][kProfLevels] = {};
  int launches[kProfStages][kProfLevels] = {};

  hipEvent_t event() {
    if (!spare.empty()) {
      hipEvent_t e = spare.back();
      spare.pop_back();
      return e;
    }
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    return e;
  }
  void resolve() {
    for (const Mark& m : pending) {
      HIP_TRY(hipEventSynchronize(m.b));
      float t = 0;
      HIP_TRY(hipEventElapsedTime(&t, m.a, m.b));
      ms[m.stage][m.level] += t;
      launches[m.stage][m.level] += 1;
      spare.push_back(m.a);
      spare.push_back(m.b);
    }
    pending.clear();
  }
  void reset() {
    resolve();
    std::memset(ms, 0, sizeof(ms));
    std::memset(launches, 0, sizeof(launches));
  }
  void release() {
    for (const Mark& m : pending) {
      (void)hipEventDestroy(m.a);
      (void)hipEventDestroy(m.b);
    }
    pending.clear();
    for (hipEvent_t e : spare) (void)hipEventDestroy(e);
    spare.clear();
  }
  template <class F>
  void run(int stage, int level, hipStream_t st, F&& f) {
    if (!on) {
      f();
      return;
    }
    Mark m{stage, std::min(std::max(level, 0), kProfLevels - 1), event(), event()};
    HIP_TRY(hipEventRecord(m.a, st));
    f();
    HIP_TRY(hipEventRecord(m.b, st));
    pending.push_back(m);
  }
};

}  // namespace

struct iegpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // Decouple chain of the factorization (iegpu_set_overlap): level l's ID
  // stages on `stream` overlap level l+1's decouple on `stream2`.
  hipStream_t stream2 = nullptr;
  bool overlap = true;
  std::vector<hipEvent_t> evs;  // fork, one per level, join
  std::string err;
  Profiler prof;

  // geometry
  int pde = kPdeLaplace;
  int sol_dim = 1;
  int n_points = 0;
  std::vector<double> h_pts, h_nrm, h_curv, h_wts;
  DevBuf<double> d_pts, d_nrm, d_curv, d_wts;
  // batch of independent problems: point offsets (n_prob + 1)
  int n_prob = 1;
  std::vector<int> h_poff;
  DevBuf<int> d_poff;

  // tree
  bool have_tree = false;
  HostTree tree;
  DevBuf<int> t_level, t_leaf, t_child, t_orig_off, t_orig_cnt, t_orig_dofs,
      t_src_off, t_src_cnt, t_src;
  DevBuf<double> t_cx, t_cy, t_side;

  // holes
  int n_holes = 0;
  int hole_dim = 0;
  std::vector<double> h_hole_c, h_hole_r;
  DevBuf<double> d_U, d_PsiT;  // ndof x hole_dim each
  DevBuf<double> d_mU, d_mPsiT, d_DinvC, d_gram, d_partial, d_Smat, d_tail,
      d_ybuf, d_alpha;
  DevBuf<int> d_Spiv, d_mask;
  DevBuf<double> d_hole_c;  // hole centres (2 per hole)
  // interior evaluation buffers
  DevBuf<double> e_tg, e_mu, e_alpha, e_partial, e_out;
  DevBuf<int> e_mask;
  // the device address of the allocation-free evaluation's static scratch
  double* e_scratch = nullptr;

  // factor state
  int ncap = 0;
  int precision = IEGPU_FP64;
  bool factored = false;
  DevBuf<int> s_act_cnt, s_act_list, s_chpos, s_out_cnt, s_out_list,
      s_red_list, s_perm, s_rank, s_upd_kind, s_m_rows, s_row_pool, s_err,
      s_root_piv, s_qkind, s_jobs, s_njobs, s_lists, s_lcnt, s_sk_w, s_rid_done;
  DevBuf<int64_t> s_id_sz, s_row_sz, s_fac_sz, s_id_off, s_row_off, s_fac_off,
      s_fac_top, s_dec_sz, s_dec_off, s_scan_blk;
  DevBuf<unsigned char> s_id_pool, s_fac_pool, s_root_mat, s_dec_pool;
  // X_rr and root blocks kept for the compressed forward apply
  bool keep_operator = false;
  DevBuf<unsigned char> s_fwd_pool, s_root_fwd;
  int64_t id_cap = 0, row_cap = 0, fac_cap = 0, dec_cap = 0;
  // randomized ID
  int rid_mode = 0, rid_width = 0;
  DevBuf<double> d_omega, d_omega_w;
  // factorization replayed as a captured HIP graph (iegpu_set_graphs)
  bool use_graphs = false;
  hipGraphExec_t fgraph = nullptr;
  DevState fgraph_state{};
  LaunchCaps caps, fgraph_caps;
  bool fgraph_overlap = true;
  const double* fgraph_fb = nullptr;
  double* fgraph_fx = nullptr;
  double fgraph_tol = 0;
  int fgraph_prec = -1;
  // rank model
  bool have_model = false, record_features = false;
  int rid_margin = 0;
  DevBuf<float> d_model, d_feat, m_X, m_y, m_work;
  DevBuf<double> d_work;  // solve vector for host-pointer solves
  // iterative refinement: b, x, residual / correction (ndof), border parts,
  // reduction partials
  DevBuf<double> r_b, r_x, r_r, r_a, r_ra, r_part;
  // FP64 operator form (iegpu_set_matvec) and the FMM's tree, translation
  // tables and expansions
  int mv_method = IEGPU_MATVEC_AUTO, fmm_order = 48, fmm_leaf = 64;
  int mv_last = 0;
  bool fmm_ready = false;
  FmmHostTree fmm;
  DevBuf<int> f_int;
  DevBuf<double> f_dbl, f_tab, f_exp;
  FmmDev fdev{};

  double t_factor_ms = 0, t_solve_ms = 0, t_root_ms = 0;
  // Timing events of the factorization and the solve (created on first
  // use) and the pinned read-back of the factorization's error words and
  // factor-pool top.
  hipEvent_t ev_f0 = nullptr, ev_f1 = nullptr, ev_s0 = nullptr, ev_s1 = nullptr;
  int* h_rb = nullptr;
  double fac_bytes = 0;

  // geometry update (iegpu_update_geometry): boxes to re-factor per slot,
  // and whether the next factorization may keep the other boxes' factors
  std::vector<int> h_dirty;
  DevBuf<int> s_dirty;
  // how the last update changed the tree: leaves split, boxes merged into
  // leaves, and per slot the box's creation id before (-1: new; empty: ids
  // unchanged)
  int upd_split = 0, upd_merged = 0;
  bool upd_rebuilt = false;
  std::vector<int> upd_old_id;
  DevBuf<int> remap_map;              // old slot per new slot (device)
  DevBuf<unsigned char> remap_tmp;    // staging of the renumbered rows
  bool update_pending = false;  // dirty flags valid for the next factor
  bool inc_run = false;         // the factorization being enqueued is incremental
  bool last_incremental = false;
  int n_dirty = 0;
  double last_tol = 0;  // id_tol and precision of the stored factorization
  int last_prec = -1;
  int64_t kept_top = 0;  // factor-pool elements in use after the last factor
  int64_t full_top = 0;  // the same after the last full factorization

  // Launch plan: per level, the kSkip* bits of the ID-stage kernels whose
  // lists were empty in the last full factorization of the same inputs
  // (plan_epoch counts every change of geometry, tree, holes, border, ID
  // method and rank model).  k_root verifies each skipped list; the host
  // repeats the factorization without the plan on a miss.
  uint64_t plan_epoch = 0;
  bool plan_valid = false, plan_on = false;
  int plan_mode = 1;        // iegpu_set_launch_plan
  bool plan_used = false;   // the last factorization ran with the plan
  int plan_misses = 0;      // factorizations repeated because a skipped list had boxes
  struct PlanKey {
    uint64_t epoch = 0;
    double tol = 0;
    int prec = -1, rid_mode = 0, rid_width = 0, rid_margin = 0, hole_dim = 0, levels = 0;
    LaunchCaps caps;
    bool operator==(const PlanKey& o) const {
      return epoch == o.epoch && tol == o.tol && prec == o.prec && rid_mode == o.rid_mode &&
             rid_width == o.rid_width && rid_margin == o.rid_margin &&
             hole_dim == o.hole_dim && levels == o.levels &&
             std::memcmp(&caps, &o.caps, sizeof(LaunchCaps)) == 0;
    }
  } plan_key;
  std::vector<int> plan_skip;
  DevBuf<int> d_plan_skip;

  int ndof() const { return n_points * sol_dim; }
  size_t elem() const { return precision == IEGPU_FP32 ? 4 : 8; }

  DevGeom geom() const {
    DevGeom g;
    g.pts = d_pts.p;
    g.nrm = d_nrm.p;
    g.curv = d_curv.p;
    g.wts = d_wts.p;
    g.n_points = n_points;
    g.sol_dim = sol_dim;
    g.pde = pde;
    g.n_prob = n_prob;
    g.prob_off = d_poff.p;
    return g;
  }
  DevTree dtree() const {
    DevTree t;
    t.n_boxes = tree.n_boxes;
    t.n_levels = tree.n_levels;
    t.level = t_level.p;
    t.leaf = t_leaf.p;
    t.child = t_child.p;
    t.cx = t_cx.p;
    t.cy = t_cy.p;
    t.side = t_side.p;
    t.orig_off = t_orig_off.p;
    t.orig_cnt = t_orig_cnt.p;
    t.orig_dofs = t_orig_dofs.p;
    t.src_off = t_src_off.p;
    t.src_cnt = t_src_cnt.p;
    t.src = t_src.p;
    return t;
  }
  DevState dstate() {
    DevState s{};  // zeroed padding: states are compared bytewise
    s.ncap = ncap;
    s.act_cnt = s_act_cnt.p;
    s.act_list = s_act_list.p;
    s.chpos = s_chpos.p;
    s.out_cnt = s_out_cnt.p;
    s.out_list = s_out_list.p;
    s.red_list = s_red_list.p;
    s.perm = s_perm.p;
    s.rank = s_rank.p;
    s.upd_kind = s_upd_kind.p;
    s.m_rows = s_m_rows.p;
    s.qkind = s_qkind.p;
    s.nbox = tree.n_boxes;
    s.lists = s_lists.p;
    s.lcnt = s_lcnt.p;
    s.jobs = s_jobs.p;
    s.njobs = s_njobs.p;
    s.id_sz = s_id_sz.p;
    s.row_sz = s_row_sz.p;
    s.fac_sz = s_fac_sz.p;
    s.id_off = s_id_off.p;
    s.row_off = s_row_off.p;
    s.fac_off = s_fac_off.p;
    s.dec_sz = s_dec_sz.p;
    s.dec_off = s_dec_off.p;
    s.dec_pool = s_dec_pool.p;
    s.dec_cap = dec_cap;
    s.id_pool = s_id_pool.p;
    s.row_pool = s_row_pool.p;
    s.fac_pool = s_fac_pool.p;
    s.id_cap = id_cap;
    s.row_cap = row_cap;
    s.fac_cap = fac_cap;
    s.fac_top = s_fac_top.p;
    s.scan_blk = s_scan_blk.p;
    s.err = s_err.p;
    s.root_mat = s_root_mat.p;
    s.root_piv = s_root_piv.p;
    s.fwd_pool = keep_operator ? s_fwd_pool.p : nullptr;
    s.root_fwd = keep_operator ? s_root_fwd.p : nullptr;
    s.rid_mode = rid_mode;
    s.rid_width = rid_width;
    s.sk_w = s_sk_w.p;
    s.rid_done = s_rid_done.p;
    s.omega = d_omega.p;
    s.omega_w = d_omega_w.p;
    s.rank_model = d_model.p;
    s.rid_margin = rid_margin;
    s.feat = record_features ? d_feat.p : nullptr;
    s.dirty = inc_run ? s_dirty.p : nullptr;
    s.kept_top = inc_run ? kept_top : 0;
    s.plan_skip = plan_on ? d_plan_skip.p : nullptr;
    return s;
  }
};

namespace {

int fail(iegpu_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int from_hip(iegpu_ctx* c, const HipError& e) {
  const bool oom = e.e == hipErrorOutOfMemory || e.e == hipErrorMemoryAllocation;
  std::string m = std::string(e.what) + ": " + hipGetErrorString(e.e);
  return fail(c, oom ? IEGPU_ERR_OOM : IEGPU_ERR_HIP, m);
}

// U / Psi^T on the boundary points: ie_solver/linear_solve.cpp:20-112.
void build_border(iegpu_ctx* c, const double* centers, const double* radii,
                  std::vector<double>* U, std::vector<double>* PsiT) {
  const int np = c->n_points, h = c->n_holes;
  const int64_t nd = c->ndof();
  if (c->pde == kPdeLaplaceNeumann) {
    // linear_solve.cpp:61-64,106-109: no border for Laplace Neumann; the
    // multiply connected problem is solved like a simply connected one.
    c->hole_dim = 0;
    U->clear();
    PsiT->clear();
    return;
  }
  if (c->pde == kPdeLaplace) {
    c->hole_dim = h;
    U->assign(nd * h, 0.0);
    PsiT->assign(nd * h, 0.0);
    for (int i = 0; i < np; i++) {
      const double x0 = c->h_pts[2 * i], x1 = c->h_pts[2 * i + 1];
      for (int k = 0; k < h; k++) {
        const double r0 = x0 - centers[2 * k], r1 = x1 - centers[2 * k + 1];
        (*U)[i + nd * k] = std::log(std::sqrt(r0 * r0 + r1 * r1));
      }
      for (int k = 0; k < h; k++) {
        const double r0 = x0 - centers[2 * k], r1 = x1 - centers[2 * k + 1];
        if (std::sqrt(r0 * r0 + r1 * r1) < radii[k] + 1e-8) {
          (*PsiT)[i + nd * k] = c->h_wts[i];
          break;
        }
      }
    }
    return;
  }
  c->hole_dim = 3 * h;
  const int hd = 3 * h;
  U->assign(nd * hd, 0.0);
  PsiT->assign(nd * hd, 0.0);
  const double scale = 1.0 / (4 * M_PI);
  for (int i = 0; i < np; i++) {
    const double x0 = c->h_pts[2 * i], x1 = c->h_pts[2 * i + 1];
    const int64_t r0i = 2 * i, r1i = 2 * i + 1;
    for (int k = 0; k < h; k++) {
      const double r0 = x0 - centers[2 * k], r1 = x1 - centers[2 * k + 1];
      const double rn = std::sqrt(r0 * r0 + r1 * r1);
      const double rn2 = rn * rn;
      (*U)[r0i + nd * (3 * k)] = scale * (std::log(1 / rn) + (1.0 / rn2) * r0 * r0);
      (*U)[r1i + nd * (3 * k)] = scale * ((1.0 / rn2) * r1 * r0);
      (*U)[r0i + nd * (3 * k + 1)] = scale * ((1.0 / rn2) * r0 * r1);
      (*U)[r1i + nd * (3 * k + 1)] = scale * (std::log(1 / rn) + (1.0 / rn2) * r1 * r1);
      (*U)[r0i + nd * (3 * k + 2)] = r1 * (scale / rn2);
      (*U)[r1i + nd * (3 * k + 2)] = -r0 * (scale / rn2);
    }
    for (int k = 0; k < h; k++) {
      const double r0 = x0 - centers[2 * k], r1 = x1 - centers[2 * k + 1];
      if (std::sqrt(r0 * r0 + r1 * r1) < radii[k] + 1e-8) {
        const double w = c->h_wts[i];
        (*PsiT)[r0i + nd * (3 * k)] = w;
        (*PsiT)[r1i + nd * (3 * k + 1)] = w;
        (*PsiT)[r0i + nd * (3 * k + 2)] = w * x1;
        (*PsiT)[r1i + nd * (3 * k + 2)] = -w * x0;
        break;
      }
    }
  }
}

// Grows a device buffer to `bytes`, keeping its contents.
void grow_keep(DevBuf<unsigned char>& b, size_t bytes, hipStream_t st) {
  if (bytes <= b.n) return;
  unsigned char* p = nullptr;
  HIP_TRY(hipMalloc(&p, bytes));
  if (b.p && b.n) {
    const hipError_t e = hipMemcpyAsync(p, b.p, b.n, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) HIP_TRY(hipStreamSynchronize(st));
    if (e != hipSuccess) {
      (void)hipFree(p);
      throw HipError{e, "hipMemcpyAsync"};
    }
  }
  b.release();
  b.p = p;
  b.n = bytes;
  b.cap = bytes;
}

void alloc_state(iegpu_ctx* c) {
  const int nb = c->tree.n_boxes;
  const int64_t nl = (int64_t)nb * c->ncap;
  hipStream_t st = c->stream;
  (void)st;
  c->s_act_cnt.alloc(nb);
  c->s_act_list.alloc(nl);
  c->s_chpos.alloc(4 * (size_t)nb);
  c->s_out_cnt.alloc(nb);
  c->s_out_list.alloc(nl);
  c->s_red_list.alloc(nl);
  c->s_perm.alloc(nl);
  c->s_rank.alloc(nb);
  c->s_upd_kind.alloc(nb);
  c->s_m_rows.alloc(nb);
  c->s_qkind.alloc(nb);
  c->s_sk_w.alloc(nb);
  if (c->record_features && c->d_feat.n != (size_t)nb * kFeat) {
    c->d_feat.alloc((size_t)nb * kFeat);
    HIP_TRY(hipMemset(c->d_feat.p, 0, sizeof(float) * nb * kFeat));
  }
  c->s_rid_done.alloc(nb);
  c->s_lists.alloc((size_t)kNumLists * nb);
  c->s_lcnt.alloc((size_t)2 * kMaxLevels * kNumLists);
  c->s_jobs.alloc((size_t)2 * kMaxChunks * nb);
  c->s_njobs.alloc(kMaxLevels);
  c->s_scan_blk.alloc(((size_t)nb / 256 + 2) * 16);
  c->s_id_sz.alloc(nb);
  c->s_row_sz.alloc(nb);
  c->s_fac_sz.alloc(nb);
  c->s_id_off.alloc(nb);
  c->s_row_off.alloc(nb);
  c->s_fac_off.alloc(nb);
  c->s_dec_sz.alloc(nb);
  c->s_dec_off.alloc(nb);
  c->s_fac_top.alloc(1);
  c->s_err.alloc(4);
  const size_t nr = c->tree.n_roots;
  c->s_root_piv.alloc(nr * (c->ncap + 64));
  c->s_root_mat.alloc(nr * 2 * c->ncap * c->ncap * 8);  // matrix + inverse per root
  c->s_id_pool.alloc((size_t)2 * c->id_cap * c->elem());  // odd and even levels
  c->s_row_pool.alloc((size_t)c->row_cap);
  c->s_fac_pool.alloc((size_t)c->fac_cap * c->elem());
  c->s_dec_pool.alloc((size_t)c->dec_cap * c->elem());
  if (c->keep_operator) {
    c->s_fwd_pool.alloc((size_t)c->fac_cap * c->elem());
    c->s_root_fwd.alloc(nr * c->ncap * c->ncap * 8);
  }
}

// The ID of level l needs only the skeletons of level l+1, not its Schur
// complements, so with overlap on the decouple stage of every level runs
// on a second stream, after its level's ID and the decouple of the level
// below: level l's ID proceeds while level l+1 is decoupled.  The root
// waits for both chains.  Per-stage profiling keeps one stream.
// With fb / fx (device, ndof) the solve x = A^{-1} b is fused in: each
// level's upward sweep follows that level's decouple on the decouple chain,
// and the root solve and the downward sweeps follow the root.
template <typename R>
void enqueue_factor(iegpu_ctx* c, double tol, const double* fb = nullptr,
                    double* fx = nullptr) {
  DevGeom g = c->geom();
  DevTree t = c->dtree();
  DevState s = c->dstate();
  hipStream_t st = c->stream;
  const HostTree& T = c->tree;
  const SolveVec fv{fx, (int64_t)c->ndof(), 1};
  if (fx && fb != fx) launch_copy(fx, fb, c->ndof(), st);
  const bool split = c->overlap && !c->prof.on && c->stream2 != nullptr;
  hipStream_t st2 = split ? c->stream2 : st;
  // evs[1 + lv]: level lv's ID done (the decouple chain's first work waits
  // for the deepest level's, which follows everything enqueued before it);
  // evs[n_levels + 1]: the decouple chain done (without a level 1);
  // evs[n_levels + 2] and evs[n_levels + 3]: level 1's decouple forked to
  // the ID stream (idle from there to the root) for its boxes above 64
  // DOFs, and the ID stream's join of the decouple chain's end, so level
  // 1's two decouple kernels run side by side.  The two chains share no
  // scratch: the ID pool
  // (one half per level parity) belongs to the ID chain, the decouple pool
  // to the decouple chain.  The events order the two streams of one device
  // only, so they release at device scope (no system-scope cache
  // write-back: 0.05 ms per step at N = 262,144).
  const size_t n_ev = (size_t)T.n_levels + 4;
  if (split && c->evs.size() < n_ev) {
    for (size_t i = c->evs.size(); i < n_ev; i++) {
      hipEvent_t e;
      HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming | hipEventDisableSystemFence));
      c->evs.push_back(e);
    }
  }
  launch_reset(s, st);
  bool joined = false;
  for (int lv = T.n_levels - 1; lv >= 1; lv--) {
    const int start = T.level_start[lv];
    const int count = T.level_start[lv + 1] - start;
    const int skip = c->plan_on ? c->plan_skip[lv] : 0;
    for (int stage = 0; stage < kStageDecouple; stage++)
      c->prof.run(stage, lv, st, [&] {
        launch_factor_stage<R>(stage, g, t, s, lv, start, count, tol, c->caps, st, skip);
      });
    if (split) {
      HIP_TRY(hipEventRecord(c->evs[1 + lv], st));
      HIP_TRY(hipStreamWaitEvent(st2, c->evs[1 + lv], 0));
    }
    const bool fork = split && lv == 1;
    if (fork) {
      HIP_TRY(hipEventRecord(c->evs[T.n_levels + 2], st2));
      HIP_TRY(hipStreamWaitEvent(st, c->evs[T.n_levels + 2], 0));
    }
    c->prof.run(kStageDecouple, lv, st2, [&] {
      launch_factor_stage<R>(kStageDecouple, g, t, s, lv, start, count, tol, c->caps, st2, 0,
                             fork ? st : nullptr);
    });
    // After the fork the ID stream joins the decouple chain's last work and
    // runs level 1's upward sweep itself (no hop back and forth).
    hipStream_t sw = st2;
    if (fork) {
      HIP_TRY(hipEventRecord(c->evs[T.n_levels + 3], st2));
      HIP_TRY(hipStreamWaitEvent(st, c->evs[T.n_levels + 3], 0));
      sw = st;
      joined = true;
    }
    if (fx)
      c->prof.run(kProfSweepUp, lv, sw,
                  [&] { launch_sweep_up<R>(s, start, count, fv, false, true, sw); });
  }
  if (split && !joined) {
    HIP_TRY(hipEventRecord(c->evs[T.n_levels + 1], st2));
    HIP_TRY(hipStreamWaitEvent(st, c->evs[T.n_levels + 1], 0));
  }
  c->prof.run(kProfRoot, 0, st,
              [&] { launch_root<R>(g, t, s, T.n_roots, c->hole_dim == 0, st); });
  if (fx) {
    c->prof.run(kProfRootSolve, 0, st, [&] { launch_root_solve<R>(s, T.n_roots, fv, st); });
    for (int lv = 1; lv < T.n_levels; lv++) {
      const int start = T.level_start[lv];
      const int count = T.level_start[lv + 1] - start;
      c->prof.run(kProfSweepDown, lv, st,
                  [&] { launch_sweep_down<R>(s, start, count, fv, st); });
    }
  }
}

template <typename R>
void enqueue_border_factor(iegpu_ctx* c) {
  DevState s = c->dstate();
  hipStream_t st = c->stream;
  const HostTree& T = c->tree;
  const int64_t nd = c->ndof();
  const int h = c->hole_dim;
  c->prof.run(kProfBorder, 0, st, [&] {
    launch_copy(c->d_mU.p, c->d_U.p, nd * h, st);
    launch_copy(c->d_mPsiT.p, c->d_PsiT.p, nd * h, st);
    SolveVec vu{c->d_mU.p, nd, h};
    SolveVec vp{c->d_mPsiT.p, nd, h};
    for (int lv = T.n_levels - 1; lv >= 1; lv--) {
      const int start = T.level_start[lv];
      const int count = T.level_start[lv + 1] - start;
      launch_sweep_up<R>(s, start, count, vu, false, false, st);
      launch_sweep_up<R>(s, start, count, vp, true, false, st);
    }
    launch_copy(c->d_DinvC.p, c->d_mU.p, nd * h, st);
    SolveVec vd{c->d_DinvC.p, nd, h};
    for (int lv = T.n_levels - 1; lv >= 1; lv--) {
      const int start = T.level_start[lv];
      const int count = T.level_start[lv + 1] - start;
      launch_diag<R>(s, start, count, vd, st);
    }
    launch_mark_skel(s, c->d_mask.p, nd, st);
    launch_border_gram(c->d_mPsiT.p, c->d_DinvC.p, c->d_mask.p, nd, h,
                       c->d_partial.p, c->d_gram.p, st);
    launch_border_factor<R>(s, c->d_mU.p, c->d_mPsiT.p, c->d_gram.p, nd, h,
                            c->d_Smat.p, c->d_Spiv.p, st);
  });
}

// x (ndof) <- M^{-1} x in place; with holes alpha (hole_dim) is written and
// g (hole_dim, device, may be NULL = 0) is the border right-hand side.
template <typename R>
void enqueue_solve(iegpu_ctx* c, double* x, double* alpha, hipStream_t st,
                   const double* g = nullptr, int nrhs = 1) {
  DevState s = c->dstate();
  const HostTree& T = c->tree;
  const int64_t nd = c->ndof();
  SolveVec v{x, nd, nrhs};
  const bool holes = c->hole_dim > 0;
  for (int lv = T.n_levels - 1; lv >= 1; lv--) {
    const int start = T.level_start[lv];
    const int count = T.level_start[lv + 1] - start;
    c->prof.run(kProfSweepUp, lv, st,
                [&] { launch_sweep_up<R>(s, start, count, v, false, true, st); });
  }
  c->prof.run(kProfRootSolve, 0, st, [&] {
    if (!holes) {
      launch_root_solve<R>(s, T.n_roots, v, st);
    } else {
      const int h = c->hole_dim;
      launch_border_gemv(c->d_mPsiT.p, x, c->d_mask.p, nd, h, c->d_partial.p,
                         c->d_tail.p, -1.0, st);
      launch_border_solve(s, c->d_Smat.p, c->d_Spiv.p, h, x, c->d_tail.p, g,
                          alpha, c->d_ybuf.p, st);
      launch_red_update(x, c->d_DinvC.p, c->d_mask.p, alpha, nd, h, st);
    }
  });
  for (int lv = 1; lv < T.n_levels; lv++) {
    const int start = T.level_start[lv];
    const int count = T.level_start[lv + 1] - start;
    c->prof.run(kProfSweepDown, lv, st,
                [&] { launch_sweep_down<R>(s, start, count, v, st); });
  }
}

}  // namespace

extern "C" {

int iegpu_create(iegpu_ctx** out, int device) {
  if (!out) return IEGPU_ERR_INVALID;
  *out = nullptr;
  if (device == -1) {  // host-only context: geometry and tree, no factorization
    iegpu_ctx* c = new iegpu_ctx();
    c->device = -1;
    *out = c;
    return IEGPU_OK;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return IEGPU_ERR_NO_DEVICE;
  if (device < 0 || device >= ndev) return IEGPU_ERR_INVALID;
  std::unique_ptr<iegpu_ctx> c(new iegpu_ctx());
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) return IEGPU_ERR_HIP;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return IEGPU_ERR_HIP;
  if (hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    return IEGPU_ERR_HIP;
  }
  // The static scratch of iegpu_evaluate_noalloc on this device.
  if (eval_scratch_address(device, &c->e_scratch) != 0) {
    (void)hipStreamDestroy(c->stream);
    (void)hipStreamDestroy(c->stream2);
    return IEGPU_ERR_HIP;
  }
  *out = c.release();
  return IEGPU_OK;
}

int iegpu_destroy(iegpu_ctx* c) {
  if (!c) return IEGPU_ERR_INVALID;
  if (c->device < 0) {
    delete c;
    return IEGPU_OK;
  }
  (void)hipSetDevice(c->device);
  if (c->stream) {
    (void)hipStreamSynchronize(c->stream);
  }
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  c->prof.release();
  if (c->fgraph) (void)hipGraphExecDestroy(c->fgraph);
  for (hipEvent_t e : c->evs) (void)hipEventDestroy(e);
  for (hipEvent_t e : {c->ev_f0, c->ev_f1, c->ev_s0, c->ev_s1})
    if (e) (void)hipEventDestroy(e);
  if (c->h_rb) (void)hipHostFree(c->h_rb);
  hipStream_t st = c->stream, st2 = c->stream2;
  delete c;
  if (st) (void)hipStreamDestroy(st);
  if (st2) (void)hipStreamDestroy(st2);
  return IEGPU_OK;
}

const char* iegpu_last_error(const iegpu_ctx* c) {
  return c ? c->err.c_str() : "null context";
}

namespace {
int set_geometry(iegpu_ctx* c, int pde, int n_problems, const int* point_off,
                 const double* points, const double* normals,
                 const double* curvatures, const double* weights) {
  const int n_points = point_off[n_problems];
  if (pde != kPdeLaplace && pde != kPdeLaplaceNeumann && pde != kPdeStokes)
    return fail(c, IEGPU_ERR_INVALID, "pde must be LAPLACE (0), LAPLACE_NEUMANN (1) or STOKES (2)");
  if (n_points <= 0 || !points || !normals || !curvatures || !weights)
    return fail(c, IEGPU_ERR_INVALID, "empty boundary or null array");
  for (int i = 0; i < 2 * n_points; i++)
    if (!std::isfinite(points[i]) || !std::isfinite(normals[i]))
      return fail(c, IEGPU_ERR_INVALID, "non-finite point or normal");
  for (int i = 0; i < n_points; i++)
    if (!std::isfinite(curvatures[i]) || !std::isfinite(weights[i]))
      return fail(c, IEGPU_ERR_INVALID, "non-finite curvature or weight");
  try {
    c->pde = pde;
    c->sol_dim = (pde == kPdeStokes) ? 2 : 1;
    c->n_points = n_points;
    c->h_pts.assign(points, points + 2 * n_points);
    c->h_nrm.assign(normals, normals + 2 * n_points);
    c->h_curv.assign(curvatures, curvatures + n_points);
    c->h_wts.assign(weights, weights + n_points);
    c->n_prob = n_problems;
    c->h_poff.assign(point_off, point_off + n_problems + 1);
    if (c->device >= 0) {
      HIP_TRY(hipSetDevice(c->device));
      c->d_pts.upload(c->h_pts, c->stream);
      c->d_nrm.upload(c->h_nrm, c->stream);
      c->d_curv.upload(c->h_curv, c->stream);
      c->d_wts.upload(c->h_wts, c->stream);
      c->d_poff.upload(c->h_poff, c->stream);
      HIP_TRY(hipStreamSynchronize(c->stream));
    }
    c->have_tree = false;
    c->factored = false;
    c->update_pending = false;
    c->fmm_ready = false;
    c->plan_epoch++;
    c->n_holes = 0;
    c->hole_dim = 0;
  } catch (const HipError& e) {
    return from_hip(c, e);
  }
  return IEGPU_OK;
}
}  // namespace

int iegpu_set_geometry(iegpu_ctx* c, int pde, int n_points,
                       const double* points, const double* normals,
                       const double* curvatures, const double* weights) {
  if (!c) return IEGPU_ERR_INVALID;
  const int off[2] = {0, n_points};
  return set_geometry(c, pde, 1, off, points, normals, curvatures, weights);
}

int iegpu_set_geometry_batch(iegpu_ctx* c, int pde, int n_problems,
                             const int* point_offsets, const double* points,
                             const double* normals, const double* curvatures,
                             const double* weights) {
  if (!c) return IEGPU_ERR_INVALID;
  if (n_problems < 1 || !point_offsets)
    return fail(c, IEGPU_ERR_INVALID, "n_problems must be >= 1 with point offsets");
  if (point_offsets[0] != 0) return fail(c, IEGPU_ERR_INVALID, "point_offsets[0] must be 0");
  for (int p = 0; p < n_problems; p++)
    if (point_offsets[p + 1] <= point_offsets[p])
      return fail(c, IEGPU_ERR_INVALID, "every problem needs at least one point");
  return set_geometry(c, pde, n_problems, point_offsets, points, normals, curvatures,
                      weights);
}

namespace {
void upload_tree(iegpu_ctx* c);
void upload_border(iegpu_ctx* c);
bool use_fmm(const iegpu_ctx* c);
void prepare_fmm(iegpu_ctx* c);
std::vector<int> mark_update(const HostTree& T, const std::vector<int>& seeds);
void remap_state(iegpu_ctx* c, const std::vector<int>& old_slot, int old_nb);
}  // namespace

int iegpu_update_geometry(iegpu_ctx* c, int n_points, const double* points,
                          const double* normals, const double* curvatures,
                          const double* weights, int* tree_kept) {
  if (!c) return IEGPU_ERR_INVALID;
  if (tree_kept) *tree_kept = 0;
  if (!c->have_tree) return fail(c, IEGPU_ERR_INVALID, "build_tree first");
  if (n_points != c->n_points)
    return fail(c, IEGPU_ERR_INVALID, "an update keeps the number of points");
  if (!points || !normals || !curvatures || !weights)
    return fail(c, IEGPU_ERR_INVALID, "null array");
  try {
    // Changed nodes; only they can be non-finite (the stored data is finite).
    std::vector<int> changed;
    for (int i = 0; i < n_points; i++)
      if (points[2 * i] != c->h_pts[2 * i] || points[2 * i + 1] != c->h_pts[2 * i + 1] ||
          normals[2 * i] != c->h_nrm[2 * i] || normals[2 * i + 1] != c->h_nrm[2 * i + 1] ||
          curvatures[i] != c->h_curv[i] || weights[i] != c->h_wts[i])
        changed.push_back(i);
    for (int i : changed) {
      if (!std::isfinite(points[2 * i]) || !std::isfinite(points[2 * i + 1]) ||
          !std::isfinite(normals[2 * i]) || !std::isfinite(normals[2 * i + 1]))
        return fail(c, IEGPU_ERR_INVALID, "non-finite point or normal");
      if (!std::isfinite(curvatures[i]) || !std::isfinite(weights[i]))
        return fail(c, IEGPU_ERR_INVALID, "non-finite curvature or weight");
    }
    // The moved points go into the kept cells, boxes split and merged as
    // QuadTree::perturb does; a point outside its root cell rebuilds the
    // tree for the new nodes.
    const int old_nb = c->tree.n_boxes;
    TreeUpdate tu;
    const bool keep = c->tree.update(c->h_pts.data(), points, changed, &tu);
    if (!keep) {
      HostTree nt;
      nt.build_forest(points, c->h_poff.data(), c->n_prob, c->sol_dim, c->tree.leaf_cap);
      if (nt.n_levels > kMaxLevels)
        return fail(c, IEGPU_ERR_INVALID, "quadtree deeper than 64 levels");
      c->tree = std::move(nt);
    }
    const bool had_factor = c->factored;
    // Host copies: only the span of changed nodes
    const int lo = changed.empty() ? 0 : changed.front();
    const int cnt = changed.empty() ? 0 : changed.back() + 1 - lo;
    std::copy(points + 2 * (size_t)lo, points + 2 * (size_t)(lo + cnt), c->h_pts.begin() + 2 * (size_t)lo);
    std::copy(normals + 2 * (size_t)lo, normals + 2 * (size_t)(lo + cnt), c->h_nrm.begin() + 2 * (size_t)lo);
    std::copy(curvatures + lo, curvatures + lo + cnt, c->h_curv.begin() + lo);
    std::copy(weights + lo, weights + lo + cnt, c->h_wts.begin() + lo);
    c->factored = false;
    c->update_pending = false;
    c->fmm_ready = false;
    c->plan_epoch++;
    c->h_dirty = keep ? mark_update(c->tree, tu.seeds) : std::vector<int>(c->tree.n_boxes, 1);
    c->n_dirty = 0;
    for (int v : c->h_dirty) c->n_dirty += v;
    c->upd_split = tu.n_split;
    c->upd_merged = tu.n_merged;
    c->upd_rebuilt = !keep;
    c->upd_old_id.swap(tu.old_id);
    if (c->device >= 0) {
      HIP_TRY(hipSetDevice(c->device));
      hipStream_t st = c->stream;
      if (c->fgraph) {
        HIP_TRY(hipGraphExecDestroy(c->fgraph));
        c->fgraph = nullptr;
      }
      c->d_pts.upload_range(c->h_pts.data(), 2 * (size_t)lo, 2 * (size_t)cnt, st);
      c->d_nrm.upload_range(c->h_nrm.data(), 2 * (size_t)lo, 2 * (size_t)cnt, st);
      c->d_curv.upload_range(c->h_curv.data(), lo, cnt, st);
      c->d_wts.upload_range(c->h_wts.data(), lo, cnt, st);
      if (keep) c->s_dirty.upload(c->h_dirty, st);
      if (!keep || tu.restructured) {
        upload_tree(c);
      } else if (tu.lists_changed) {
        const HostTree& T = c->tree;
        c->t_orig_off.upload(T.orig_off, st);
        c->t_orig_cnt.upload(T.orig_cnt, st);
        c->t_orig_dofs.upload(T.orig_dofs, st);
      }
      if (keep && tu.restructured && had_factor) remap_state(c, tu.old_slot, old_nb);
      HIP_TRY(hipStreamSynchronize(st));
      if (c->n_holes > 0) upload_border(c);
      c->update_pending = keep && had_factor;
    }
    if (tree_kept) *tree_kept = keep ? 1 : 0;
  } catch (const HipError& e) {
    return from_hip(c, e);
  } catch (const std::exception& e) {
    return fail(c, IEGPU_ERR_INVALID, e.what());
  }
  return IEGPU_OK;
}

int iegpu_update_geometry_batch(iegpu_ctx* c, int n_problems, const int* point_offsets,
                                const double* points, const double* normals,
                                const double* curvatures, const double* weights, int* tree_kept) {
  if (!c) return IEGPU_ERR_INVALID;
  if (tree_kept) *tree_kept = 0;
  if (!c->have_tree) return fail(c, IEGPU_ERR_INVALID, "build_tree first");
  if (n_problems != c->n_prob || !point_offsets)
    return fail(c, IEGPU_ERR_INVALID, "an update keeps the number of problems");
  for (int p = 0; p <= n_problems; p++)
    if (point_offsets[p] != c->h_poff[p])
      return fail(c, IEGPU_ERR_INVALID,
                  "an update keeps every problem's number of points (problem " +
                      std::to_string(p > 0 ? p - 1 : 0) + " differs)");
  return iegpu_update_geometry(c, point_offsets[n_problems], points, normals, curvatures,
                               weights, tree_kept);
}

int iegpu_discard_update(iegpu_ctx* c) {
  if (!c) return IEGPU_ERR_INVALID;
  c->update_pending = false;
  return IEGPU_OK;
}

int iegpu_update_info(const iegpu_ctx* c, int* n_dirty, int* last_incremental, int* dirty) {
  if (!c) return IEGPU_ERR_INVALID;
  if (n_dirty) *n_dirty = c->n_dirty;
  if (last_incremental) *last_incremental = c->last_incremental ? 1 : 0;
  if (dirty && c->have_tree) {
    const HostTree& T = c->tree;
    for (int id = 0; id < T.n_boxes; id++)
      dirty[id] = (int)c->h_dirty.size() == T.n_boxes ? c->h_dirty[T.slot_of_id[id]] : 1;
  }
  return IEGPU_OK;
}

int iegpu_update_tree_info(const iegpu_ctx* c, int* n_split, int* n_merged, int* old_id) {
  if (!c) return IEGPU_ERR_INVALID;
  if (n_split) *n_split = c->upd_split;
  if (n_merged) *n_merged = c->upd_merged;
  if (old_id && c->have_tree) {
    const HostTree& T = c->tree;
    for (int id = 0; id < T.n_boxes; id++) {
      const int s = T.slot_of_id[id];
      old_id[id] = c->upd_rebuilt ? -1
                   : c->upd_old_id.empty() ? id
                   : c->upd_old_id[s];
    }
  }
  return IEGPU_OK;
}

int iegpu_build_tree(iegpu_ctx* c, int leaf_cap) {
  if (!c) return IEGPU_ERR_INVALID;
  c->plan_epoch++;
  if (c->n_points == 0) return fail(c, IEGPU_ERR_INVALID, "set_geometry first");
  if (leaf_cap < c->sol_dim || leaf_cap > kMaxBoxDofs / 2)
    return fail(c, IEGPU_ERR_INVALID, "leaf_cap out of range");
  try {
    c->tree = HostTree();
    c->tree.build_forest(c->h_pts.data(), c->h_poff.data(), c->n_prob, c->sol_dim, leaf_cap);
    c->upd_split = c->upd_merged = 0;
    c->upd_rebuilt = false;
    c->upd_old_id.clear();
    if (c->tree.n_levels > kMaxLevels)
      return fail(c, IEGPU_ERR_INVALID, "quadtree deeper than 64 levels");
    c->have_tree = true;
    c->factored = false;
    c->update_pending = false;
    if (c->device < 0) return IEGPU_OK;
    upload_tree(c);
    if (use_fmm(c)) prepare_fmm(c);
  } catch (const HipError& e) {
    return from_hip(c, e);
  } catch (const std::exception& e) {
    return fail(c, IEGPU_ERR_INVALID, e.what());
  }
  return IEGPU_OK;
}

namespace {
void upload_tree(iegpu_ctx* c) {
    HIP_TRY(hipSetDevice(c->device));
    const HostTree& T = c->tree;
    hipStream_t st = c->stream;
    c->t_level.upload(T.level, st);
    c->t_leaf.upload(T.leaf, st);
    c->t_child.upload(T.child, st);
    c->t_cx.upload(T.cx, st);
    c->t_cy.upload(T.cy, st);
    c->t_side.upload(T.side, st);
    c->t_orig_off.upload(T.orig_off, st);
    c->t_orig_cnt.upload(T.orig_cnt, st);
    c->t_orig_dofs.upload(T.orig_dofs, st);
    c->t_src_off.upload(T.src_off, st);
    c->t_src_cnt.upload(T.src_cnt, st);
    if (T.src.empty()) {
      const std::vector<int> dummy(1, 0);
      c->t_src.upload(dummy, st);
    } else {
      c->t_src.upload(T.src, st);
    }
    HIP_TRY(hipStreamSynchronize(st));
}

// Boxes to re-factor after a geometry update.  The walks start from the
// seeds: the leaves that lost, gained or hold a changed point, following
// QuadTree::perturb (quadtree.cpp:680-697: the leaf of every added point,
// and for deleted points the leaf they leave, is passed to
// mark_neighbors_and_parents), and after a split or merge every new box and
// every box whose leaf flag or row sources changed.  A walk follows
// mark_neighbors_and_parents (quadtree.cpp:435-461: the box, its
// neighbours, then the same for its parent, up to the root).  A box's ID
// rows come from its row-source boxes' active lists, which change when a
// source holds changed nodes or re-factors a child, and the sources are
// not always among the neighbours (all boxes of the level at levels 1 and
// 2): such a box is marked too (the reference keeps its old ID, which then
// misses the moved nodes: residuals of 2e-5 at tol 1e-10 on a bumped
// circle).  Then every marked box's parent is
// marked (a box re-factored under a kept parent would leave the parent's
// lists stale; the reference's rule gives this by itself on the tested
// boundaries), and every root is re-factored.
std::vector<int> mark_update(const HostTree& T, const std::vector<int>& seeds) {
  std::vector<int> dirty(T.n_boxes, 0);
  std::vector<char> seen(T.n_boxes, 0), walk(T.n_boxes, 0);
  for (int s : seeds) {
    // one walk per seed: the walk from a box depends on the box alone
    if (s < 0 || seen[s]) continue;
    seen[s] = 1;
    for (; s >= 0; s = T.parent[s]) {
      dirty[s] = 1;
      walk[s] = 1;
      for (int t = 0; t < T.nbr_cnt[s]; t++) dirty[T.nbr[T.nbr_off[s] + t]] = 1;
    }
  }
  // Bottom up: a box's active list changes when it is a leaf holding
  // changed nodes or when a child is re-factored (new skeleton); a box
  // whose ID rows come from such a box is re-factored.
  std::vector<char> chg(T.n_boxes, 0);
  for (int s = 0; s < T.n_boxes; s++) chg[s] = T.leaf[s] && walk[s];
  for (int lv = T.n_levels - 1; lv >= 1; lv--) {
    for (int s = T.level_start[lv]; s < T.level_start[lv + 1]; s++)
      for (int q = 0; q < 4; q++) {
        const int ch = T.child[4 * (size_t)s + q];
        if (ch >= 0 && dirty[ch]) chg[s] = 1;
      }
    for (int s = T.level_start[lv]; s < T.level_start[lv + 1]; s++) {
      if (chg[s]) dirty[s] = 1;
      for (int u = 0; u < T.src_cnt[s]; u++)
        if (chg[T.src[T.src_off[s] + u]]) dirty[s] = 1;
    }
  }
  for (int lv = T.n_levels - 1; lv >= 1; lv--)
    for (int s = T.level_start[lv]; s < T.level_start[lv + 1]; s++)
      if (dirty[s]) dirty[T.parent[s]] = 1;
  for (int r = 0; r < T.n_roots; r++) dirty[r] = 1;
  return dirty;
}

// After boxes were split or merged: the kept boxes' factor state moves to
// their new slots (new boxes get zeroed rows; every one of them is marked
// for re-factoring).  Level-local scratch (lists, jobs) is not kept.
void remap_state(iegpu_ctx* c, const std::vector<int>& old_slot, int old_nb) {
  hipStream_t st = c->stream;
  const int nb = (int)old_slot.size();
  DevBuf<unsigned char>& tmp = c->remap_tmp;
  c->remap_map.upload(old_slot, st);
  const int* map = c->remap_map.p;
  for (DevBuf<int>* b : {&c->s_act_cnt, &c->s_out_cnt, &c->s_rank, &c->s_upd_kind, &c->s_m_rows,
                         &c->s_qkind, &c->s_sk_w, &c->s_rid_done})
    remap_rows(*b, tmp, map, nb, old_nb, 1, st);
  remap_rows(c->s_chpos, tmp, map, nb, old_nb, 4, st);
  for (DevBuf<int>* b : {&c->s_act_list, &c->s_out_list, &c->s_red_list, &c->s_perm})
    remap_rows(*b, tmp, map, nb, old_nb, (size_t)c->ncap, st);
  for (DevBuf<int64_t>* b : {&c->s_id_sz, &c->s_row_sz, &c->s_fac_sz, &c->s_id_off,
                             &c->s_row_off, &c->s_fac_off, &c->s_dec_sz, &c->s_dec_off})
    remap_rows(*b, tmp, map, nb, old_nb, 1, st);
  if (c->d_feat.p && c->d_feat.n == (size_t)old_nb * kFeat)
    remap_rows(c->d_feat, tmp, map, nb, old_nb, kFeat, st);
}
}  // namespace

int iegpu_set_holes(iegpu_ctx* c, int n_holes, const double* centers,
                    const double* radii) {
  if (c) c->plan_epoch++;
  if (!c) return IEGPU_ERR_INVALID;
  if (c->n_points == 0) return fail(c, IEGPU_ERR_INVALID, "set_geometry first");
  if (n_holes < 0 || (n_holes > 0 && (!centers || !radii)))
    return fail(c, IEGPU_ERR_INVALID, "bad hole arrays");
  if (n_holes > 64) return fail(c, IEGPU_ERR_INVALID, "at most 64 holes");
  if (n_holes > 0 && c->n_prob > 1)
    return fail(c, IEGPU_ERR_INVALID, "holes are not supported in a batch of problems");
  if (c->device < 0) return fail(c, IEGPU_ERR_NO_DEVICE, "host-only context");
  try {
    HIP_TRY(hipSetDevice(c->device));
    c->n_holes = n_holes;
    // The holes enter only the border blocks, which every factorization
    // rebuilds: a pending geometry update stays incremental.
    c->factored = false;
    if (n_holes == 0) {
      c->hole_dim = 0;
      return IEGPU_OK;
    }
    c->h_hole_c.assign(centers, centers + 2 * n_holes);
    c->h_hole_r.assign(radii, radii + n_holes);
    upload_border(c);
  } catch (const HipError& e) {
    return from_hip(c, e);
  }
  return IEGPU_OK;
}

namespace {
// Border blocks of the holes' extra unknowns (they depend on the boundary
// geometry, so a geometry update rebuilds them).
void upload_border(iegpu_ctx* c) {
    const double* centers = c->h_hole_c.data();
    const double* radii = c->h_hole_r.data();
    const int n_holes = c->n_holes;
    std::vector<double> U, PsiT;
    build_border(c, centers, radii, &U, &PsiT);
    const int64_t nd = c->ndof();
    const int h = c->hole_dim;
    hipStream_t st = c->stream;
    c->d_U.upload(U, st);
    c->d_PsiT.upload(PsiT, st);
    c->d_mU.alloc(nd * h);
    c->d_mPsiT.alloc(nd * h);
    c->d_DinvC.alloc(nd * h);
    c->d_gram.alloc((size_t)h * h);
    c->d_partial.alloc((size_t)border_partial_blocks(nd) * h * h);
    c->d_tail.alloc(h);
    c->d_alpha.alloc(h);
    c->d_mask.alloc(nd);
    c->d_hole_c.upload(centers, 2 * (size_t)n_holes, st);
    HIP_TRY(hipStreamSynchronize(st));
}
}  // namespace

int iegpu_set_id_method(iegpu_ctx* c, int method, int sketch_rows) {
  if (!c) return IEGPU_ERR_INVALID;
  c->plan_epoch++;
  c->update_pending = false;  // the kept factors would mix two ID methods
  if (method == IEGPU_ID_QRCP) {
    c->rid_mode = 0;
    c->rid_width = 0;
    return IEGPU_OK;
  }
  if (method != IEGPU_ID_RANDOMIZED)
    return fail(c, IEGPU_ERR_INVALID, "unknown ID method");
  if (sketch_rows == 0) {
    if (!c->have_model) return fail(c, IEGPU_ERR_INVALID, "sketch_rows 0 needs a rank model");
    c->rid_mode = 2;
    c->rid_width = 0;
    return IEGPU_OK;
  }
  if (sketch_rows < 16 || sketch_rows > kRidRows || sketch_rows % 16 != 0)
    return fail(c, IEGPU_ERR_INVALID, "sketch_rows must be 0, 16 or 32");
  c->rid_mode = 1;
  c->rid_width = sketch_rows;
  return IEGPU_OK;
}

int iegpu_set_keep_operator(iegpu_ctx* c, int on) {
  if (!c) return IEGPU_ERR_INVALID;
  c->update_pending = false;  // kept boxes would lack their forward-apply blocks
  c->keep_operator = on != 0;
  c->factored = false;
  return IEGPU_OK;
}

int iegpu_sparse_matvec(iegpu_ctx* c, const double* x, double* b) {
  if (!c) return IEGPU_ERR_INVALID;
  if (!c->factored) return fail(c, IEGPU_ERR_INVALID, "factor first");
  if (!c->keep_operator)
    return fail(c, IEGPU_ERR_INVALID, "factor with iegpu_set_keep_operator(ctx, 1) first");
  if (!x || !b) return fail(c, IEGPU_ERR_INVALID, "null vector");
  try {
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const int64_t nd = c->ndof();
    c->d_work.alloc(nd);
    HIP_TRY(hipMemcpyAsync(c->d_work.p, x, nd * sizeof(double), hipMemcpyHostToDevice, st));
    const DevState s = c->dstate();
    const HostTree& T = c->tree;
    double* v = c->d_work.p;
    const bool f32 = c->precision == IEGPU_FP32;
    for (int lv = T.n_levels - 1; lv >= 1; lv--) {
      const int start = T.level_start[lv], count = T.level_start[lv + 1] - start;
      if (f32) launch_fwd_up<float>(s, start, count, v, st);
      else launch_fwd_up<double>(s, start, count, v, st);
    }
    if (f32) launch_fwd_root<float>(s, T.n_roots, v, st);
    else launch_fwd_root<double>(s, T.n_roots, v, st);
    for (int lv = 1; lv < T.n_levels; lv++) {
      const int start = T.level_start[lv], count = T.level_start[lv + 1] - start;
      if (f32) launch_fwd_down<float>(s, start, count, v, st);
      else launch_fwd_down<double>(s, start, count, v, st);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(b, v, nd * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
  } catch (const HipError& e) {
    return from_hip(c, e);
  }
  return IEGPU_OK;
}

int iegpu_set_overlap(iegpu_ctx* c, int on) {
  if (!c) return IEGPU_ERR_INVALID;
  c->overlap = on != 0;
  return IEGPU_OK;
}

int iegpu_set_graphs(iegpu_ctx* c, int on) {
  if (!c) return IEGPU_ERR_INVALID;
  c->use_graphs = on != 0;
  return IEGPU_OK;
}

int iegpu_record_features(iegpu_ctx* c, int on) {
  if (!c) return IEGPU_ERR_INVALID;
  c->record_features = on != 0;
  return IEGPU_OK;
}

int iegpu_box_features(iegpu_ctx* c, float* feat, int cap) {
  if (!c) return IEGPU_ERR_INVALID;
  if (!c->factored || !c->d_feat.p)
    return fail(c, IEGPU_ERR_INVALID, "no factorization recorded features");
  const HostTree& T = c->tree;
  const int64_t need = (int64_t)T.n_boxes * kFeat;
  if (!feat || cap < need) return fail(c, IEGPU_ERR_INVALID, "feature buffer too small");
  std::vector<float> dev(need);
  try {
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemcpy(dev.data(), c->d_feat.p, sizeof(float) * need, hipMemcpyDeviceToHost));
  } catch (const HipError& e) {
    return from_hip(c, e);
  }
  // device rows are level-major slots; the ABI's boxes are creation ids
  for (int id = 0; id < T.n_boxes; id++)
    for (int q = 0; q < kFeat; q++)
      feat[(int64_t)id * kFeat + q] = dev[(int64_t)T.slot_of_id[id] * kFeat + q];
  return IEGPU_OK;
}

int iegpu_rank_model_init(iegpu_ctx* c, const float* X, const float* y, int n, unsigned seed,
                          float* model) {
  if (!c) return IEGPU_ERR_INVALID;
  if (!X || !y || !model || n < 2) return fail(c, IEGPU_ERR_INVALID, "bad model arguments");
  std::vector<double> mu(kFeat, 0.0), sd(kFeat, 0.0);
  double ym = 0, ys = 0;
  for (int i = 0; i < n; i++) {
    for (int q = 0; q < kFeat; q++) mu[q] += X[(int64_t)i * kFeat + q];
    ym += y[i];
  }
  for (int q = 0; q < kFeat; q++) mu[q] /= n;
  ym /= n;
  for (int i = 0; i < n; i++) {
    for (int q = 0; q < kFeat; q++) {
      const double d = X[(int64_t)i * kFeat + q] - mu[q];
      sd[q] += d * d;
    }
    ys += (y[i] - ym) * (y[i] - ym);
  }
  for (int q = 0; q < kFeat; q++) {
    const double v = std::sqrt(sd[q] / n);
    model[kMlpMu + q] = (float)mu[q];
    model[kMlpSd + q] = (float)(v > 1e-6 ? v : 1.0);  // constant features pass through
  }
  model[kMlpYMu] = (float)ym;
  model[kMlpYSd] = (float)(std::sqrt(ys / n) > 1e-6 ? std::sqrt(ys / n) : 1.0);
  // Glorot-uniform weights from a 32-bit xorshift stream; zero biases.
  uint32_t st = seed ? seed : 0x9E3779B9u;
  auto uni = [&st]() {
    st ^= st << 13;
    st ^= st >> 17;
    st ^= st << 5;
    return (st >> 8) * (1.0 / 16777216.0);
  };
  auto fill = [&](int off, int count, int fan_in, int fan_out) {
    const double a = std::sqrt(6.0 / (fan_in + fan_out));
    for (int i = 0; i < count; i++) model[off + i] = (float)((2 * uni() - 1) * a);
  };
  fill(kMlpW1, kFeat * kHid, kFeat, kHid);
  fill(kMlpW2, kHid * kHid, kHid, kHid);
  fill(kMlpW3, kHid, kHid, 1);
  for (int j = 0; j < kHid; j++) model[kMlpB1 + j] = model[kMlpB2 + j] = 0;
  model[kMlpB3] = 0;
  return IEGPU_OK;
}

namespace {
int upload_samples(iegpu_ctx* c, const float* X, const float* y, int n, const float* model) {
  c->m_X.upload(X, (size_t)n * kFeat, c->stream);
  c->m_y.upload(y, (size_t)n, c->stream);
  c->m_work.alloc(kMlpLen + kMlpParams + 8);
  HIP_TRY(hipMemcpyAsync(c->m_work.p, model, sizeof(float) * kMlpLen, hipMemcpyHostToDevice,
                         c->stream));
  return IEGPU_OK;
}
}  // namespace

int iegpu_rank_model_train(iegpu_ctx* c, const float* X, const float* y, int n, float* model,
                           int steps, float lr, unsigned seed, float* loss) {
  if (!c) return IEGPU_ERR_INVALID;
  if (c->device < 0) return fail(c, IEGPU_ERR_NO_DEVICE, "host-only context");
  if (!X || !y || !model || n < 1 || steps < 1 || !(lr > 0))
    return fail(c, IEGPU_ERR_INVALID, "bad training arguments");
  try {
    HIP_TRY(hipSetDevice(c->device));
    upload_samples(c, X, y, n, model);
    float* dl = c->m_work.p + kMlpLen + kMlpParams;
    launch_mlp_train(c->m_X.p, c->m_y.p, n, c->m_work.p, steps, lr, seed, dl, c->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(model, c->m_work.p, sizeof(float) * kMlpLen, hipMemcpyDeviceToHost,
                           c->stream));
    float l = 0;
    HIP_TRY(hipMemcpyAsync(&l, dl, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (loss) *loss = l;
  } catch (const HipError& e) {
    return from_hip(c, e);
  }
  return IEGPU_OK;
}

int iegpu_rank_model_grad(iegpu_ctx* c, const float* model, const float* X, const float* y,
                          int n, float* grad, float* loss) {
  if (!c) return IEGPU_ERR_INVALID;
  if (c->device < 0) return fail(c, IEGPU_ERR_NO_DEVICE, "host-only context");
  if (!X || !y || !model || !grad || n < 1) return fail(c, IEGPU_ERR_INVALID, "bad arguments");
  try {
    HIP_TRY(hipSetDevice(c->device));
    upload_samples(c, X, y, n, model);
    float* g = c->m_work.p + kMlpLen;
    float* dl = g + kMlpParams;
    launch_mlp_grad(c->m_X.p, c->m_y.p, n, c->m_work.p, g, dl, c->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(grad, g, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, c->stream));
    float l = 0;
    HIP_TRY(hipMemcpyAsync(&l, dl, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (loss) *loss = l;
  } catch (const HipError& e) {
    return from_hip(c, e);
  }
  return IEGPU_OK;
}

int iegpu_set_rank_model(iegpu_ctx* c, const float* model, int margin) {
  if (!c) return IEGPU_ERR_INVALID;
  c->plan_epoch++;
  c->update_pending = false;  // the kept factors would mix two ID methods
  if (c->device < 0) return fail(c, IEGPU_ERR_NO_DEVICE, "host-only context");
  if (!model || margin < 0 || margin > 16) return fail(c, IEGPU_ERR_INVALID, "bad rank model");
  for (int i = 0; i < kMlpLen; i++)
    if (!std::isfinite(model[i])) return fail(c, IEGPU_ERR_INVALID, "rank model not finite");
  for (int q = 0; q < kFeat; q++)
    if (!(model[kMlpSd + q] > 0)) return fail(c, IEGPU_ERR_INVALID, "rank model scale <= 0");
  try {
    HIP_TRY(hipSetDevice(c->device));
    c->d_model.upload(model, kMlpLen, c->stream);
    HIP_TRY(hipStreamSynchronize(c->stream));
  } catch (const HipError& e) {
    return from_hip(c, e);
  }
  c->have_model = true;
  c->rid_margin = margin;
  return IEGPU_OK;
}

int iegpu_id_stats(iegpu_ctx* c, int* n_sketched, int* n_fallback, double* sketch_flops) {
  if (!c) return IEGPU_ERR_INVALID;
  if (!c->factored) return fail(c, IEGPU_ERR_INVALID, "factor first");
  int ns = 0, nf = 0;
  double fl = 0;
  if (c->rid_mode != 0) {
    try {
      HIP_TRY(hipSetDevice(c->device));
      const int lo = c->tree.level_start[1], nb = c->tree.n_boxes;
      std::vector<int> w(nb), d(nb), m(nb), n(nb);
      HIP_TRY(hipMemcpy(w.data(), c->s_sk_w.p, sizeof(int) * nb, hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(d.data(), c->s_rid_done.p, sizeof(int) * nb, hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(m.data(), c->s_m_rows.p, sizeof(int) * nb, hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(n.data(), c->s_act_cnt.p, sizeof(int) * nb, hipMemcpyDeviceToHost));
      for (int b = lo; b < nb; b++) {
        if (w[b] <= 0) continue;
        if (d[b]) ns++;
        else nf++;
        fl += 2.0 * w[b] * m[b] * (16 * ((n[b] + 15) / 16));
      }
    } catch (const HipError& e) {
      return from_hip(c, e);
    }
  }
  if (n_sketched) *n_sketched = ns;
  if (n_fallback) *n_fallback = nf;
  if (sketch_flops) *sketch_flops = fl;
  return IEGPU_OK;
}

namespace {
int factor_impl(iegpu_ctx* c, double id_tol, int precision, const double* fb, double* fx);
}

int iegpu_factor(iegpu_ctx* c, double id_tol, int precision) {
  return factor_impl(c, id_tol, precision, nullptr, nullptr);
}

int iegpu_factor_solve_device(iegpu_ctx* c, double id_tol, int precision, const double* d_b,
                              double* d_x) {
  if (!c) return IEGPU_ERR_INVALID;
  if (!d_b || !d_x) return fail(c, IEGPU_ERR_INVALID, "null vector");
  if (d_b == d_x)  // a workspace retry restarts from b
    return fail(c, IEGPU_ERR_INVALID, "b and x must be distinct buffers");
  if (c->hole_dim > 0)
    return fail(c, IEGPU_ERR_INVALID, "with holes, factor and solve separately");
  return factor_impl(c, id_tol, precision, d_b, d_x);
}

namespace {
int factor_impl(iegpu_ctx* c, double id_tol, int precision, const double* fb, double* fx) {
  if (!c) return IEGPU_ERR_INVALID;
  if (!c->have_tree) return fail(c, IEGPU_ERR_INVALID, "build_tree first");
  if (c->device < 0) return fail(c, IEGPU_ERR_NO_DEVICE, "host-only context");
  if (!(id_tol > 0) || !std::isfinite(id_tol))
    return fail(c, IEGPU_ERR_INVALID, "id_tol must be > 0");
  if (precision != IEGPU_FP64 && precision != IEGPU_FP32)
    return fail(c, IEGPU_ERR_INVALID, "precision must be FP64 or FP32");
  try {
    HIP_TRY(hipSetDevice(c->device));
    const int64_t nd = c->ndof();
    if (c->precision != precision) {
      c->s_id_pool.release();
      c->s_fac_pool.release();
      c->s_dec_pool.release();
    }
    c->precision = precision;
    if (c->ncap == 0) c->ncap = 256;
    // Register-QRCP boxes need no ID scratch; the pool holds the few
    // row-compressed or general-QRCP boxes of one level (grown to the
    // measured need if a level overflows it).
    if (c->id_cap == 0) c->id_cap = std::max<int64_t>(nd * 16, 1 << 22);
    if (c->row_cap == 0) c->row_cap = std::max<int64_t>(nd * 8, 1 << 16);
    if (c->fac_cap == 0) c->fac_cap = std::max<int64_t>(nd * 160, 1 << 22);
    if (c->dec_cap == 0) c->dec_cap = std::max<int64_t>(nd * 8, 1 << 20);
    if (c->rid_mode != 0 && !c->d_omega.p) {
      const int64_t no = (int64_t)kRidTableRows * kRidRows;
      c->d_omega.alloc(no);
      launch_gaussian(c->d_omega.p, no, 0x5EEDull, c->stream);
      const int64_t nw = (int64_t)kRidTableRowsW * kRidRowsW;
      c->d_omega_w.alloc(nw);
      launch_gaussian(c->d_omega_w.p, nw, 0x5EEDull, c->stream);
    }
    // After iegpu_update_geometry on the same tree, only the marked boxes
    // are re-factored (same tolerance and precision as the kept factors).
    // Boxes whose new factors outgrow their kept slot take new ones past
    // the pool top; once the top passes twice a full factorization's, the
    // next factorization is a full one, which packs the pool again.
    c->inc_run = c->update_pending && c->last_tol == id_tol && c->last_prec == precision &&
                 c->kept_top <= 2 * c->full_top;
    c->update_pending = false;
    iegpu_ctx::PlanKey key;
    key.epoch = c->plan_epoch;
    key.tol = id_tol;
    key.prec = precision;
    key.rid_mode = c->rid_mode;
    key.rid_width = c->rid_width;
    key.rid_margin = c->rid_margin;
    key.hole_dim = c->hole_dim;
    key.levels = c->tree.n_levels;
    key.caps = c->caps;
    const bool plan_ok = !c->inc_run && !c->use_graphs && c->plan_mode != 0;
    c->plan_on = plan_ok && c->plan_valid && key == c->plan_key;
    if (plan_ok && c->plan_mode == 2) {  // every list claimed empty (a test of the check)
      c->plan_skip.assign(kMaxLevels, kSkipAssemble | kSkipRegQrcp | kSkipReg8Qrcp | kSkipRid |
                                          kSkipRidWide);
      c->d_plan_skip.upload(c->plan_skip, c->stream);
      c->plan_on = true;
      c->plan_mode = 1;
    }
    c->plan_used = c->plan_on;
    if (!c->ev_f0) HIP_TRY(hipEventCreate(&c->ev_f0));
    if (!c->ev_f1) HIP_TRY(hipEventCreate(&c->ev_f1));
    if (!c->h_rb) HIP_TRY(hipHostMalloc((void**)&c->h_rb, 8 * sizeof(int), hipHostMallocDefault));
    hipEvent_t e0 = c->ev_f0, e1 = c->ev_f1;
    int errbits = 0;
    for (int attempt = 0; attempt < 8; attempt++) {
      alloc_state(c);
      if (c->hole_dim > 0) {
        const int ns_max = c->ncap + c->hole_dim;
        c->d_Smat.alloc((size_t)ns_max * ns_max);
        c->d_Spiv.alloc(ns_max);
        c->d_ybuf.alloc(ns_max);
      }
      if (c->inc_run)  // new slots go past the kept ones (a failed attempt moved the top)
        HIP_TRY(hipMemcpyAsync(c->s_fac_top.p, &c->kept_top, sizeof(int64_t),
                               hipMemcpyHostToDevice, c->stream));
      HIP_TRY(hipEventRecord(e0, c->stream));
      if (c->use_graphs && !c->prof.on && c->hole_dim == 0 && !c->inc_run) {
        // Capture the launch sequence once per state; replay it after.
        const DevState cur = c->dstate();
        if (!c->fgraph || std::memcmp(&cur, &c->fgraph_state, sizeof(DevState)) != 0 ||
            std::memcmp(&c->caps, &c->fgraph_caps, sizeof(LaunchCaps)) != 0 ||
            c->overlap != c->fgraph_overlap ||
            c->fgraph_tol != id_tol || c->fgraph_prec != precision || c->fgraph_fb != fb ||
            c->fgraph_fx != fx) {
          if (c->fgraph) HIP_TRY(hipGraphExecDestroy(c->fgraph));
          c->fgraph = nullptr;
          HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
          if (precision == IEGPU_FP32) enqueue_factor<float>(c, id_tol, fb, fx);
          else enqueue_factor<double>(c, id_tol, fb, fx);
          hipGraph_t gr = nullptr;
          HIP_TRY(hipStreamEndCapture(c->stream, &gr));
          const hipError_t ie = hipGraphInstantiate(&c->fgraph, gr, nullptr, nullptr, 0);
          (void)hipGraphDestroy(gr);
          HIP_TRY(ie);
          c->fgraph_state = cur;
          c->fgraph_caps = c->caps;
          c->fgraph_fb = fb;
          c->fgraph_fx = fx;
          c->fgraph_overlap = c->overlap;
          c->fgraph_tol = id_tol;
          c->fgraph_prec = precision;
        }
        HIP_TRY(hipGraphLaunch(c->fgraph, c->stream));
      } else if (precision == IEGPU_FP32) {
        enqueue_factor<float>(c, id_tol, fb, fx);
      } else {
        enqueue_factor<double>(c, id_tol, fb, fx);
      }
      HIP_TRY(hipGetLastError());
      // One read-back: the error words and the factor-pool top.
      if (c->hole_dim == 0) HIP_TRY(hipEventRecord(e1, c->stream));
      int* h_err = c->h_rb;
      HIP_TRY(hipMemcpyAsync(h_err, c->s_err.p, 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipMemcpyAsync(h_err + 4, c->s_fac_top.p, sizeof(int64_t), hipMemcpyDeviceToHost,
                             c->stream));
      HIP_TRY(hipStreamSynchronize(c->stream));
      errbits = h_err[0];
      if (errbits & kErrPlan) {  // a skipped list had boxes: again, every kernel
        c->plan_valid = false;
        c->plan_on = false;
        c->plan_used = false;
        c->plan_misses++;
        continue;
      }
      const int grow = errbits & (kErrBoxCap | kErrIdPool | kErrRowPool | kErrFacPool | kErrDecPool);
      if (!grow) {
        if (c->hole_dim > 0) {
          if (precision == IEGPU_FP32) enqueue_border_factor<float>(c);
          else enqueue_border_factor<double>(c);
          HIP_TRY(hipGetLastError());
          HIP_TRY(hipMemcpyAsync(h_err, c->s_err.p, 4 * sizeof(int), hipMemcpyDeviceToHost,
                                 c->stream));
          HIP_TRY(hipEventRecord(e1, c->stream));
          HIP_TRY(hipStreamSynchronize(c->stream));
          errbits = h_err[0];
        }
        break;
      }
      // A larger box capacity changes every per-box list's stride, so the
      // kept boxes are lost; the ID and row pools are scratch, and the
      // factor pool grows with its contents.  A pool overflow leaves the
      // level's boxes uncompressed, which can overflow their parents' box
      // capacity as well: in an incremental run that bit counts only when
      // it comes alone.
      bool box_cap = (errbits & kErrBoxCap) != 0;
      if (c->inc_run && (errbits & (kErrIdPool | kErrRowPool | kErrFacPool | kErrDecPool)))
        box_cap = false;
      if (box_cap) c->inc_run = false;
      if (box_cap) {
        if (c->ncap >= kMaxBoxDofs) {
          return fail(c, IEGPU_ERR_CAPACITY,
                      "a box's active DOF count exceeds " + std::to_string(kMaxBoxDofs));
        }
        c->ncap = std::min(kMaxBoxDofs, c->ncap * 2);
      }
      if (errbits & kErrIdPool)
        c->id_cap = std::max<int64_t>(2 * c->id_cap, ((int64_t)h_err[2] << 10) * 5 / 4);
      if (errbits & kErrRowPool)
        c->row_cap = std::max<int64_t>(2 * c->row_cap, ((int64_t)h_err[3] << 10) * 5 / 4);
      if (errbits & kErrDecPool)
        c->dec_cap = std::max<int64_t>(2 * c->dec_cap, ((int64_t)h_err[1] << 10) * 5 / 4);
      if (errbits & kErrFacPool) {
        c->fac_cap *= 2;
        if (c->inc_run) {
          grow_keep(c->s_fac_pool, (size_t)c->fac_cap * c->elem(), c->stream);
          if (c->keep_operator)
            grow_keep(c->s_fwd_pool, (size_t)c->fac_cap * c->elem(), c->stream);
        }
      }
      if (attempt == 7) {
        return fail(c, IEGPU_ERR_CAPACITY, "workspace growth did not converge");
      }
    }
    if (plan_ok && !c->plan_on && !(errbits & kErrSingular)) {
      // Record the plan from this factorization's per-level list counts.
      std::vector<int> lc((size_t)2 * kMaxLevels * kNumLists);
      HIP_TRY(hipMemcpy(lc.data(), c->s_lcnt.p, sizeof(int) * lc.size(), hipMemcpyDeviceToHost));
      c->plan_skip.assign(kMaxLevels, 0);
      for (int lv = 1; lv < c->tree.n_levels; lv++) {
        auto cnt = [&](int L) { return lc[(size_t)2 * (lv * kNumLists + L)]; };
        const int reg = c->rid_mode != 0 ? cnt(kLstFallback) : cnt(kLstReg) + cnt(kLstRegHalf);
        const int wide = lc[(size_t)2 * (lv * kNumLists + kLstRid) + 1];
        int m = 0;
        if (cnt(kLstLds) == 0) m |= kSkipAssemble;
        if (reg == 0) m |= kSkipRegQrcp;
        if (cnt(kLstReg8) == 0) m |= kSkipReg8Qrcp;
        if (cnt(kLstRid) == wide) m |= kSkipRid;
        if (wide == 0) m |= kSkipRidWide;
        c->plan_skip[lv] = m;
      }
      c->d_plan_skip.upload(c->plan_skip, c->stream);
      HIP_TRY(hipStreamSynchronize(c->stream));
      c->plan_key = key;
      c->plan_valid = true;
    }
    c->plan_on = false;
    c->last_incremental = c->inc_run;
    c->inc_run = false;
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    c->t_factor_ms = ms;
    int64_t top = 0;
    std::memcpy(&top, c->h_rb + 4, sizeof(top));
    c->fac_bytes = (double)top * c->elem();
    c->kept_top = top;
    if (!c->last_incremental) c->full_top = top;
    if (errbits & kErrSingular) {
      c->factored = false;
      return fail(c, IEGPU_ERR_SINGULAR, "exact zero pivot in an LU factorization");
    }
    if (c->hole_dim > 0) {
      // skel_factorization.cpp:209-212: every DOF is a skeleton DOF.
      int nroot = 0;
      HIP_TRY(hipMemcpy(&nroot, c->s_act_cnt.p, sizeof(int), hipMemcpyDeviceToHost));
      if (nroot == nd) {
        c->factored = false;
        return fail(c, IEGPU_ERR_NO_COMPRESSION, "No compression possible");
      }
    }
    c->factored = true;
    c->last_tol = id_tol;
    c->last_prec = precision;
  } catch (const HipError& e) {
    c->factored = false;
    c->inc_run = false;
    return from_hip(c, e);
  }
  return IEGPU_OK;
}

}  // namespace

int iegpu_solve_device(iegpu_ctx* c, const double* d_b, double* d_mu,
                       double* d_alpha, void* stream) {
  if (!c) return IEGPU_ERR_INVALID;
  if (!c->factored) return fail(c, IEGPU_ERR_INVALID, "factor first");
  if (!d_b || !d_mu) return fail(c, IEGPU_ERR_INVALID, "null vector");
  if (c->hole_dim > 0 && !d_alpha)
    return fail(c, IEGPU_ERR_INVALID, "alpha required with holes");
  try {
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    const int64_t nd = c->ndof();
    if (d_b != d_mu) launch_copy(d_mu, d_b, nd, st);
    if (c->precision == IEGPU_FP32) enqueue_solve<float>(c, d_mu, d_alpha, st);
    else enqueue_solve<double>(c, d_mu, d_alpha, st);
    HIP_TRY(hipGetLastError());
  } catch (const HipError& e) {
    return from_hip(c, e);
  }
  return IEGPU_OK;
}

int iegpu_solve(iegpu_ctx* c, const double* b, double* mu, double* alpha) {
  if (!c) return IEGPU_ERR_INVALID;
  if (!c->factored) return fail(c, IEGPU_ERR_INVALID, "factor first");
  if (!b || !mu) return fail(c, IEGPU_ERR_INVALID, "null vector");
  if (c->hole_dim > 0 && !alpha)
    return fail(c, IEGPU_ERR_INVALID, "alpha required with holes");
  try {
    HIP_TRY(hipSetDevice(c->device));
    const int64_t nd = c->ndof();
    c->d_work.alloc(nd);
    HIP_TRY(hipMemcpyAsync(c->d_work.p, b, nd * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (!c->ev_s0) HIP_TRY(hipEventCreate(&c->ev_s0));
    if (!c->ev_s1) HIP_TRY(hipEventCreate(&c->ev_s1));
    hipEvent_t e0 = c->ev_s0, e1 = c->ev_s1;
    HIP_TRY(hipEventRecord(e0, c->stream));
    double* da = c->hole_dim > 0 ? c->d_alpha.p : nullptr;
    if (c->precision == IEGPU_FP32) enqueue_solve<float>(c, c->d_work.p, da, c->stream);
    else enqueue_solve<double>(c, c->d_work.p, da, c->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(e1, c->stream));
    HIP_TRY(hipMemcpyAsync(mu, c->d_work.p, nd * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (da) HIP_TRY(hipMemcpyAsync(alpha, da, c->hole_dim * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    c->t_solve_ms = ms;
  } catch (const HipError& e) {
    return from_hip(c, e);
  }
  return IEGPU_OK;
}

int iegpu_solve_nrhs(iegpu_ctx* c, int nrhs, const double* b, double* mu, double* alpha) {
  if (!c) return IEGPU_ERR_INVALID;
  if (!c->factored) return fail(c, IEGPU_ERR_INVALID, "factor first");
  if (nrhs < 0 || (nrhs > 0 && (!b || !mu)))
    return fail(c, IEGPU_ERR_INVALID, "bad right-hand sides");
  const int h = c->hole_dim;
  if (h > 0 && nrhs > 0 && !alpha) return fail(c, IEGPU_ERR_INVALID, "alpha required with holes");
  if (nrhs == 0) return IEGPU_OK;
  try {
    HIP_TRY(hipSetDevice(c->device));
    const int64_t nd = c->ndof();
    hipStream_t st = c->stream;
    if (h > 0) {
      // the bordered solve takes one column at a time
      c->d_work.alloc(nd);
      for (int q = 0; q < nrhs; q++) {
        HIP_TRY(hipMemcpyAsync(c->d_work.p, b + nd * q, nd * sizeof(double),
                               hipMemcpyHostToDevice, st));
        if (c->precision == IEGPU_FP32) enqueue_solve<float>(c, c->d_work.p, c->d_alpha.p, st);
        else enqueue_solve<double>(c, c->d_work.p, c->d_alpha.p, st);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(mu + nd * q, c->d_work.p, nd * sizeof(double),
                               hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(alpha + (int64_t)h * q, c->d_alpha.p, h * sizeof(double),
                               hipMemcpyDeviceToHost, st));
      }
    } else {
      // every column through each sweep kernel in one pass
      c->d_work.alloc(nd * nrhs);
      HIP_TRY(hipMemcpyAsync(c->d_work.p, b, nd * nrhs * sizeof(double), hipMemcpyHostToDevice,
                             st));
      if (c->precision == IEGPU_FP32) enqueue_solve<float>(c, c->d_work.p, nullptr, st, nullptr, nrhs);
      else enqueue_solve<double>(c, c->d_work.p, nullptr, st, nullptr, nrhs);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(mu, c->d_work.p, nd * nrhs * sizeof(double), hipMemcpyDeviceToHost,
                             st));
    }
    HIP_TRY(hipStreamSynchronize(st));
  } catch (const HipError& e) {
    return from_hip(c, e);
  }
  return IEGPU_OK;
}

int iegpu_set_border(iegpu_ctx* c, int hole_dim, const double* U, const double* Psi) {
  if (!c) return IEGPU_ERR_INVALID;
  c->plan_epoch++;
  if (c->device < 0) return fail(c, IEGPU_ERR_NO_DEVICE, "host-only context");
  if (c->n_holes == 0) return fail(c, IEGPU_ERR_INVALID, "set_holes first");
  if (hole_dim != c->hole_dim || !U || !Psi)
    return fail(c, IEGPU_ERR_INVALID,
                "U must be ndof x " + std::to_string(c->hole_dim) + " and Psi its transpose's shape");
  const int64_t nd = c->ndof();
  const int h = hole_dim;
  for (int64_t i = 0; i < nd * h; i++)
    if (!std::isfinite(U[i]) || !std::isfinite(Psi[i]))
      return fail(c, IEGPU_ERR_INVALID, "non-finite border entry");
  try {
    HIP_TRY(hipSetDevice(c->device));
    std::vector<double> PsiT(nd * h);
    for (int k = 0; k < h; k++)  // Psi is h x ndof column-major
      for (int64_t i = 0; i < nd; i++) PsiT[i + nd * k] = Psi[k + (int64_t)h * i];
    c->d_U.upload(U, nd * h, c->stream);
    c->d_PsiT.upload(PsiT, c->stream);
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->factored = false;
  } catch (const HipError& e) {
    return from_hip(c, e);
  }
  return IEGPU_OK;
}

int iegpu_box_factors(iegpu_ctx* c, int box, int* k_out, int* r_out, double* T, double* L,
                      double* U, double* Xinv, double* S) {
  if (!c) return IEGPU_ERR_INVALID;
  if (!c->factored) return fail(c, IEGPU_ERR_INVALID, "factor first");
  if (box < 0 || box >= c->tree.n_boxes) return fail(c, IEGPU_ERR_INVALID, "bad box");
  try {
    HIP_TRY(hipSetDevice(c->device));
    const int slot = c->tree.slot_of_id[box];
    int n = 0, k = 0, kind = 0;
    int64_t off = 0;
    HIP_TRY(hipMemcpy(&n, c->s_act_cnt.p + slot, sizeof(int), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(&k, c->s_rank.p + slot, sizeof(int), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(&kind, c->s_upd_kind.p + slot, sizeof(int), hipMemcpyDeviceToHost));
    const bool root = slot < c->tree.n_roots;
    const int r = (k > 0 && !root) ? n - k : 0;
    if (k_out) *k_out = root ? 0 : k;
    if (r_out) *r_out = r;
    if (r == 0 || root) return IEGPU_OK;
    HIP_TRY(hipMemcpy(&off, c->s_fac_off.p + slot, sizeof(int64_t), hipMemcpyDeviceToHost));
    const FacLayout fl = FacLayout::make(k, r);
    const int64_t len = fl.S + (int64_t)k * k;
    std::vector<double> h(len);
    if (c->precision == IEGPU_FP32) {
      std::vector<float> f(len);
      HIP_TRY(hipMemcpy(f.data(), reinterpret_cast<float*>(c->s_fac_pool.p) + off,
                        len * sizeof(float), hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < len; i++) h[i] = f[i];
    } else {
      HIP_TRY(hipMemcpy(h.data(), reinterpret_cast<double*>(c->s_fac_pool.p) + off,
                        len * sizeof(double), hipMemcpyDeviceToHost));
    }
    auto put = [&](double* dst, int64_t at, int64_t cnt) {
      if (dst) std::copy(h.begin() + at, h.begin() + at + cnt, dst);
    };
    put(T, fl.T, (int64_t)k * r);
    put(L, fl.L, (int64_t)k * r);
    put(U, fl.U, (int64_t)r * k);
    put(Xinv, fl.X, (int64_t)r * r);
    if (S) {
      if (kind) put(S, fl.S, (int64_t)k * k);
      else std::fill(S, S + (int64_t)k * k, 0.0);
    }
  } catch (const HipError& e) {
    return from_hip(c, e);
  }
  return IEGPU_OK;
}

int iegpu_tree_info(const iegpu_ctx* c, int* n_levels, int* n_boxes, int* ndof,
                    int* hole_dim) {
  if (!c) return IEGPU_ERR_INVALID;
  if (n_levels) *n_levels = c->have_tree ? c->tree.n_levels : 0;
  if (n_boxes) *n_boxes = c->have_tree ? c->tree.n_boxes : 0;
  if (ndof) *ndof = c->ndof();
  if (hole_dim) *hole_dim = c->hole_dim;
  return IEGPU_OK;
}

int iegpu_box_info(iegpu_ctx* c, int* level, int* parent, int* is_leaf,
                   int* children, int* n_active, int* rank, int* m_rows) {
  if (!c) return IEGPU_ERR_INVALID;
  if (!c->have_tree) return fail(c, IEGPU_ERR_INVALID, "build_tree first");
  const HostTree& T = c->tree;
  const int nb = T.n_boxes;
  std::vector<int> act(nb, 0), rk(nb, 0), mr(nb, 0);
  if (c->factored && (n_active || rank || m_rows)) {
    try {
      HIP_TRY(hipSetDevice(c->device));
      HIP_TRY(hipMemcpy(act.data(), c->s_act_cnt.p, nb * sizeof(int), hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(rk.data(), c->s_rank.p, nb * sizeof(int), hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(mr.data(), c->s_m_rows.p, nb * sizeof(int), hipMemcpyDeviceToHost));
    } catch (const HipError& e) {
      return from_hip(c, e);
    }
  }
  for (int id = 0; id < nb; id++) {
    const int s = T.slot_of_id[id];
    if (level) level[id] = T.level[s];
    if (parent) parent[id] = T.parent[s] >= 0 ? T.box_id[T.parent[s]] : -1;
    if (is_leaf) is_leaf[id] = T.leaf[s];
    if (children)
      for (int q = 0; q < 4; q++) {
        const int ch = T.child[4 * s + q];
        children[4 * id + q] = ch >= 0 ? T.box_id[ch] : -1;
      }
    if (n_active) n_active[id] = act[s];
    if (rank) rank[id] = rk[s];
    if (m_rows) m_rows[id] = mr[s];
  }
  return IEGPU_OK;
}

int iegpu_box_list(iegpu_ctx* c, int box, int which, int* out, int cap) {
  if (!c) return IEGPU_ERR_INVALID;
  if (!c->have_tree) return -fail(c, IEGPU_ERR_INVALID, "build_tree first");
  const HostTree& T = c->tree;
  std::vector<int> v;
  if (which == 5) {
    if (box < 0 || box >= T.n_levels) return -fail(c, IEGPU_ERR_INVALID, "bad level");
    for (int s = T.level_start[box]; s < T.level_start[box + 1]; s++) v.push_back(T.box_id[s]);
  } else {
    if (box < 0 || box >= T.n_boxes) return -fail(c, IEGPU_ERR_INVALID, "bad box");
    const int s = T.slot_of_id[box];
    if (which == 0) {
      v = T.dofs_below(s);
    } else if (which == 1) {
      for (int t = 0; t < T.nbr_cnt[s]; t++) v.push_back(T.box_id[T.nbr[T.nbr_off[s] + t]]);
    } else if (which >= 2 && which <= 4) {
      if (!c->factored) return -fail(c, IEGPU_ERR_INVALID, "factor first");
      try {
        HIP_TRY(hipSetDevice(c->device));
        int n = 0, k = 0;
        HIP_TRY(hipMemcpy(&n, c->s_act_cnt.p + s, sizeof(int), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&k, c->s_rank.p + s, sizeof(int), hipMemcpyDeviceToHost));
        const int64_t base = (int64_t)s * c->ncap;
        if (which == 2) {
          v.resize(n);
          if (n) HIP_TRY(hipMemcpy(v.data(), c->s_act_list.p + base, n * sizeof(int), hipMemcpyDeviceToHost));
        } else if (which == 3) {
          v.resize(k);
          if (k) HIP_TRY(hipMemcpy(v.data(), c->s_out_list.p + base, k * sizeof(int), hipMemcpyDeviceToHost));
        } else {
          const int r = k ? n - k : 0;
          v.resize(r);
          if (r) HIP_TRY(hipMemcpy(v.data(), c->s_red_list.p + base, r * sizeof(int), hipMemcpyDeviceToHost));
        }
      } catch (const HipError& e) {
        return -from_hip(c, e);
      }
    } else {
      return -fail(c, IEGPU_ERR_INVALID, "bad list selector");
    }
  }
  if (out)
    for (int i = 0; i < (int)v.size() && i < cap; i++) out[i] = v[i];
  return (int)v.size();
}

int iegpu_last_timings(const iegpu_ctx* c, double* t, int nt) {
  if (!c || !t) return IEGPU_ERR_INVALID;
  const double v[4] = {c->t_factor_ms, c->t_solve_ms, c->t_root_ms, c->fac_bytes};
  for (int i = 0; i < nt && i < 4; i++) t[i] = v[i];
  return IEGPU_OK;
}

namespace {
// The FMM applies the operator of a single boundary; batches are summed
// directly (each problem's block is small).
bool use_fmm(const iegpu_ctx* c) {
  if (c->n_prob != 1) return false;
  if (c->mv_method == IEGPU_MATVEC_FMM) return true;
  if (c->mv_method == IEGPU_MATVEC_DIRECT) return false;
  return c->n_points >= IEGPU_FMM_AUTO_POINTS;
}

// Builds the FMM tree of the current geometry, uploads it with the
// translation tables and sorts the geometry on the device.
void prepare_fmm(iegpu_ctx* c) {
  if (c->fmm_ready) return;
  hipStream_t st = c->stream;
  FmmHostTree& T = c->fmm;
  T = FmmHostTree();
  T.build(c->h_pts.data(), c->n_points, c->fmm_leaf);
  std::vector<int> iv;
  auto put = [&](const std::vector<int>& v) {
    const size_t o = iv.size();
    iv.insert(iv.end(), v.begin(), v.end());
    return o;
  };
  const size_t o_par = put(T.parent), o_coff = put(T.child_off), o_ccnt = put(T.child_cnt),
               o_vlo = put(T.vl_off), o_vl = put(T.vl), o_pto = put(T.pt_off),
               o_nbo = put(T.nb_off), o_nb = put(T.nb), o_perm = put(T.perm);
  iv.push_back(0);  // keeps every offset inside the buffer
  c->f_int.upload(iv, st);
  const int64_t nb = T.n_boxes(), np = c->n_points;
  c->f_dbl.alloc(2 * nb + 6 * np + 2 * np);
  HIP_TRY(hipMemcpyAsync(c->f_dbl.p, T.center.data(), sizeof(double) * 2 * nb,
                         hipMemcpyHostToDevice, st));
  if (!c->f_tab.p) {
    std::vector<double> a, b, d;
    fmm_tables(a, b, d);
    a.insert(a.end(), b.begin(), b.end());
    a.insert(a.end(), d.begin(), d.end());
    c->f_tab.upload(a, st);
    HIP_TRY(hipStreamSynchronize(st));
  }
  const int ne = fmm_n_exp(c->pde);
  const int64_t per = nb * ne * c->fmm_order * 2;
  c->f_exp.alloc(2 * per);
  FmmDev& f = c->fdev;
  f.order = c->fmm_order;
  f.depth = T.depth;
  f.n_leaves = T.n_leaves();
  f.first_leaf = T.first_leaf();
  f.side = T.side;
  f.center = c->f_dbl.p;
  double* sgeo = c->f_dbl.p + 2 * nb;
  f.sgeo = sgeo;
  f.xs = sgeo + 6 * np;
  const int* ib = c->f_int.p;
  f.parent = ib + o_par;
  f.child_off = ib + o_coff;
  f.child_cnt = ib + o_ccnt;
  f.vl_off = ib + o_vlo;
  f.vl = ib + o_vl;
  f.pt_off = ib + o_pto;
  f.nb_off = ib + o_nbo;
  f.nb = ib + o_nb;
  f.perm = ib + o_perm;
  f.mexp = c->f_exp.p;
  f.lexp = c->f_exp.p + per;
  const int64_t tab = int64_t(kFmmMaxOrder) * kFmmMaxOrder;
  f.t_m2m = c->f_tab.p;
  f.t_m2l = c->f_tab.p + tab;
  f.t_l2l = c->f_tab.p + 2 * tab;
  launch_fmm_gather(c->geom(), f, sgeo, st);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  c->fmm_ready = true;
}

// y = K x on the context's stream, in the form iegpu_set_matvec selects;
// r_part holds the reduction partials.
void enqueue_matvec(iegpu_ctx* c, const double* x, double* y, hipStream_t st) {
  if (use_fmm(c)) {
    prepare_fmm(c);
    launch_fmm_matvec(c->geom(), c->fdev, c->fmm.level_off, x, y, c->r_part.p, st);
    c->mv_last = IEGPU_MATVEC_FMM;
  } else {
    launch_matvec(c->geom(), x, y, c->r_part.p, st);
    c->mv_last = IEGPU_MATVEC_DIRECT;
  }
}
}  // namespace

int iegpu_set_launch_plan(iegpu_ctx* c, int mode) {
  if (!c) return IEGPU_ERR_INVALID;
  if (mode < 0 || mode > 2) return fail(c, IEGPU_ERR_INVALID, "mode must be 0, 1 or 2");
  c->plan_mode = mode;
  if (mode == 0) c->plan_valid = false;
  return IEGPU_OK;
}

int iegpu_launch_plan_info(const iegpu_ctx* c, int* info, int n) {
  if (!c || (n > 0 && !info)) return IEGPU_ERR_INVALID;
  int skipped = 0;
  if (c->plan_valid)
    for (int m : c->plan_skip) skipped += __builtin_popcount(m);
  const int v[3] = {c->plan_used ? 1 : 0, c->plan_misses, skipped};
  for (int i = 0; i < n && i < 3; i++) info[i] = v[i];
  return IEGPU_OK;
}

int iegpu_set_matvec(iegpu_ctx* c, int method, int order, int leaf_points) {
  if (!c) return IEGPU_ERR_INVALID;
  if (method < IEGPU_MATVEC_AUTO || method > IEGPU_MATVEC_FMM)
    return fail(c, IEGPU_ERR_INVALID, "method must be AUTO (0), DIRECT (1) or FMM (2)");
  if (order == 0) order = 48;
  if (!fmm_order_supported(order))
    return fail(c, IEGPU_ERR_INVALID, "FMM order must be 32, 48 or 64");
  if (leaf_points == 0) leaf_points = 64;
  if (leaf_points < 1 || leaf_points > 4096)
    return fail(c, IEGPU_ERR_INVALID, "leaf_points must be in 1..4096");
  if (order != c->fmm_order || leaf_points != c->fmm_leaf) c->fmm_ready = false;
  c->mv_method = method;
  c->fmm_order = order;
  c->fmm_leaf = leaf_points;
  return IEGPU_OK;
}

int iegpu_matvec_info(const iegpu_ctx* c, int* info, int n) {
  if (!c || (n > 0 && !info)) return IEGPU_ERR_INVALID;
  const bool t = c->fmm_ready;
  const int v[4] = {c->mv_last, t ? c->fmm.depth : 0, t ? c->fmm.n_boxes() : 0,
                    t ? c->fmm.n_leaves() : 0};
  for (int i = 0; i < n && i < 4; i++) info[i] = v[i];
  return IEGPU_OK;
}

int iegpu_solve_refined(iegpu_ctx* c, const double* b, double* x, double* alpha,
                        double rtol, int max_iter, int* iters, double* relres) {
  if (!c) return IEGPU_ERR_INVALID;
  if (!c->factored) return fail(c, IEGPU_ERR_INVALID, "factor first");
  if (!b || !x || max_iter < 0 || !(rtol >= 0))
    return fail(c, IEGPU_ERR_INVALID, "bad refinement arguments");
  const int h = c->hole_dim;
  if (h > 0 && !alpha) return fail(c, IEGPU_ERR_INVALID, "alpha required with holes");
  try {
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const int64_t nd = c->ndof();
    const int nres = residual_blocks();
    c->r_b.upload(b, nd, st);
    c->r_x.alloc(nd);
    c->r_r.alloc(nd);
    c->r_part.alloc(std::max<int64_t>(2 * nres, matvec_partials(c->geom())));
    if (h > 0) {
      c->r_a.alloc(h);
      c->r_ra.alloc(h);
    }
    std::vector<double> part(2 * nres);
    auto solve = [&](double* v, double* a, const double* g) {
      if (c->precision == IEGPU_FP32) enqueue_solve<float>(c, v, a, st, g);
      else enqueue_solve<double>(c, v, a, st, g);
    };
    // x0 = M^{-1} b
    launch_copy(c->r_x.p, c->r_b.p, nd, st);
    solve(c->r_x.p, h ? c->r_a.p : nullptr, nullptr);
    int it = 0;
    double rel = 0;
    for (;; it++) {
      // r = b - A x  (A bordered by [U; Psi, -I] with holes)
      enqueue_matvec(c, c->r_x.p, c->r_r.p, st);
      if (h > 0)
        launch_border_apply(c->d_U.p, c->d_PsiT.p, c->r_x.p, c->r_a.p, nd, h, c->r_r.p,
                            c->r_ra.p, st);
      launch_residual(c->r_b.p, c->r_r.p, nd, c->r_part.p, st);
      HIP_TRY(hipMemcpyAsync(part.data(), c->r_part.p, sizeof(double) * 2 * nres,
                             hipMemcpyDeviceToHost, st));
      std::vector<double> ra(h);
      if (h > 0) {
        // border residual 0 - (Psi x - alpha)
        HIP_TRY(hipMemcpyAsync(ra.data(), c->r_ra.p, sizeof(double) * h,
                               hipMemcpyDeviceToHost, st));
      }
      HIP_TRY(hipStreamSynchronize(st));
      double rr = 0, bb = 0;
      for (int i = 0; i < nres; i++) {
        rr += part[2 * i];
        bb += part[2 * i + 1];
      }
      for (int i = 0; i < h; i++) {
        ra[i] = -ra[i];
        rr += ra[i] * ra[i];
      }
      rel = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
      if (rel <= rtol || it >= max_iter) break;
      // x += M^{-1} r
      if (h > 0) HIP_TRY(hipMemcpyAsync(c->r_ra.p, ra.data(), sizeof(double) * h,
                                        hipMemcpyHostToDevice, st));
      solve(c->r_r.p, h ? c->d_alpha.p : nullptr, h ? c->r_ra.p : nullptr);
      launch_axpy(c->r_x.p, c->r_r.p, nd, st);
      if (h > 0) launch_axpy(c->r_a.p, c->d_alpha.p, h, st);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(x, c->r_x.p, sizeof(double) * nd, hipMemcpyDeviceToHost, st));
    if (h > 0) HIP_TRY(hipMemcpyAsync(alpha, c->r_a.p, sizeof(double) * h, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (iters) *iters = it;
    if (relres) *relres = rel;
  } catch (const HipError& e) {
    return from_hip(c, e);
  }
  return IEGPU_OK;
}

int iegpu_apply(iegpu_ctx* c, const double* x, const double* alpha, double* y,
                double* z) {
  if (!c) return IEGPU_ERR_INVALID;
  if (c->n_points == 0) return fail(c, IEGPU_ERR_INVALID, "set_geometry first");
  if (c->device < 0) return fail(c, IEGPU_ERR_NO_DEVICE, "host-only context");
  if (!x || !y) return fail(c, IEGPU_ERR_INVALID, "null vector");
  const int h = c->hole_dim;
  try {
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const int64_t nd = c->ndof();
    c->r_x.upload(x, nd, st);
    c->r_r.alloc(nd);
    c->r_part.alloc(std::max<int64_t>(2 * residual_blocks(), matvec_partials(c->geom())));
    enqueue_matvec(c, c->r_x.p, c->r_r.p, st);
    if (h > 0 && alpha && z) {
      c->r_a.upload(alpha, h, st);
      c->r_ra.alloc(h);
      launch_border_apply(c->d_U.p, c->d_PsiT.p, c->r_x.p, c->r_a.p, nd, h, c->r_r.p,
                          c->r_ra.p, st);
      HIP_TRY(hipMemcpyAsync(z, c->r_ra.p, sizeof(double) * h, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(y, c->r_r.p, sizeof(double) * nd, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
  } catch (const HipError& e) {
    return from_hip(c, e);
  }
  return IEGPU_OK;
}

int iegpu_evaluate(iegpu_ctx* c, int n_targets, const double* targets,
                   const int* in_domain, const double* mu, const double* alpha,
                   double* out) {
  if (!c) return IEGPU_ERR_INVALID;
  if (c->n_points == 0) return fail(c, IEGPU_ERR_INVALID, "set_geometry first");
  if (c->device < 0) return fail(c, IEGPU_ERR_NO_DEVICE, "host-only context");
  if (n_targets < 0 || (n_targets > 0 && (!targets || !mu || !out)))
    return fail(c, IEGPU_ERR_INVALID, "bad evaluation arrays");
  if (c->n_prob > 1)
    return fail(c, IEGPU_ERR_INVALID, "interior evaluation takes one problem, not a batch");
  if (n_targets == 0) return IEGPU_OK;
  try {
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const int sd = c->sol_dim;
    c->e_tg.upload(targets, 2 * (size_t)n_targets, st);
    c->e_mu.upload(mu, (size_t)c->ndof(), st);
    const int* dmask = nullptr;
    if (in_domain) {
      c->e_mask.upload(in_domain, n_targets, st);
      dmask = c->e_mask.p;
    }
    const double* dalpha = nullptr;
    const bool holes = c->hole_dim > 0 && alpha;
    if (holes) {
      c->e_alpha.upload(alpha, c->hole_dim, st);
      dalpha = c->e_alpha.p;
    }
    c->e_partial.alloc((size_t)eval_chunks(c->n_points) * sd * n_targets);
    c->e_out.alloc((size_t)sd * n_targets);
    launch_eval(c->geom(), n_targets, c->e_tg.p, dmask, c->e_mu.p,
                holes ? c->n_holes : 0, c->d_hole_c.p, dalpha, c->e_partial.p,
                c->e_out.p, st);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, c->e_out.p, sizeof(double) * sd * n_targets,
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
  } catch (const HipError& e) {
    return from_hip(c, e);
  }
  return IEGPU_OK;
}

int iegpu_evaluate_noalloc(iegpu_ctx* c, int n_targets, const double* d_targets,
                           const int* d_in_domain, const double* d_mu, const double* d_alpha,
                           double* d_out, void* stream) {
  if (!c) return IEGPU_ERR_INVALID;
  if (c->n_points == 0) return fail(c, IEGPU_ERR_INVALID, "set_geometry first");
  if (c->device < 0) return fail(c, IEGPU_ERR_NO_DEVICE, "host-only context");
  if (n_targets < 0 || (n_targets > 0 && (!d_targets || !d_mu || !d_out)))
    return fail(c, IEGPU_ERR_INVALID, "bad evaluation arrays");
  if (c->n_prob > 1)
    return fail(c, IEGPU_ERR_INVALID, "interior evaluation takes one problem, not a batch");
  if (n_targets == 0) return IEGPU_OK;
  const int64_t need = eval_scratch_need(c->n_points, c->sol_dim, n_targets);
  if (need > kEvalScratchElems)
    return fail(c, IEGPU_ERR_CAPACITY,
                "evaluation needs " + std::to_string(need) + " partial sums; the static scratch "
                "holds " + std::to_string(kEvalScratchElems) + " (fewer targets per call)");
  if (!c->e_scratch) return fail(c, IEGPU_ERR_HIP, "no scratch address for this device");
  const hipError_t e0 = hipSetDevice(c->device);
  if (e0 != hipSuccess) return fail(c, IEGPU_ERR_HIP, hipGetErrorString(e0));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
  const bool holes = c->hole_dim > 0 && d_alpha;
  launch_eval(c->geom(), n_targets, d_targets, d_in_domain, d_mu, holes ? c->n_holes : 0,
              c->d_hole_c.p, holes ? d_alpha : nullptr, c->e_scratch, d_out, st);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, IEGPU_ERR_HIP, hipGetErrorString(e));
  return IEGPU_OK;
}

int iegpu_set_launch_caps(iegpu_ctx* c, const int* caps, int n) {
  if (!c) return IEGPU_ERR_INVALID;
  if (!caps || n != kNumCaps) return fail(c, IEGPU_ERR_INVALID, "need IEGPU_NUM_LAUNCH_CAPS caps");
  for (int i = 0; i < n; i++)
    if (caps[i] < 1 || caps[i] > (1 << 20))
      return fail(c, IEGPU_ERR_INVALID, "launch caps must be in [1, 2^20]");
  LaunchCaps& L = c->caps;
  L.assemble = caps[0];
  L.qrcp_id = caps[1];
  L.rowcomp4 = caps[2];
  L.rowcomp8 = caps[3];
  L.qrcp_id8 = caps[4];
  L.qrcp = caps[5];
  L.dsplit = caps[6];
  L.dsweep = caps[7];
  return IEGPU_OK;
}

int iegpu_get_launch_caps(const iegpu_ctx* c, int* caps, int n) {
  if (!c || !caps || n != kNumCaps) return IEGPU_ERR_INVALID;
  const LaunchCaps& L = c->caps;
  const int v[kNumCaps] = {L.assemble, L.qrcp_id, L.rowcomp4, L.rowcomp8, L.qrcp_id8, L.qrcp,
                           L.dsplit, L.dsweep};
  for (int i = 0; i < n; i++) caps[i] = v[i];
  return IEGPU_OK;
}

int iegpu_autotune(iegpu_ctx* c, double id_tol, int precision, int reps, int* caps_out,
                   double* ms_out) {
  if (!c) return IEGPU_ERR_INVALID;
  if (reps < 1 || reps > 64) return fail(c, IEGPU_ERR_INVALID, "reps must be in [1, 64]");
  // Coordinate descent over the caps: each candidate's time is the best of
  // `reps` factorizations after one untimed run (which also re-captures
  // the graph when graphs are on).  A candidate replaces the current value
  // only if it is at least 1 % faster, so timing noise does not move it.
  static const int kCand[kNumCaps][4] = {{64, 256, 1024, 4096}, {64, 256, 1024, 4096},
                                         {64, 256, 1024, 4096}, {64, 256, 1024, 4096},
                                         {64, 256, 1024, 4096}, {64, 256, 1024, 4096},
                                         {1, 2048, 8192, 1 << 20}, {1, 2048, 4096, 1 << 20}};
  auto measure = [&](double* best) -> int {
    *best = 1e300;
    for (int r = 0; r <= reps; r++) {
      const int st = iegpu_factor(c, id_tol, precision);
      if (st != IEGPU_OK) return st;
      if (r > 0) *best = std::min(*best, c->t_factor_ms);
    }
    return IEGPU_OK;
  };
  int cur[kNumCaps];
  iegpu_get_launch_caps(c, cur, kNumCaps);
  double best = 0;
  int st = measure(&best);
  if (st != IEGPU_OK) return st;
  for (int k = 0; k < kNumCaps; k++) {
    int pick = cur[k];
    for (int v : kCand[k]) {
      if (v == cur[k]) continue;
      int trial[kNumCaps];
      std::memcpy(trial, cur, sizeof(trial));
      trial[k] = v;
      iegpu_set_launch_caps(c, trial, kNumCaps);
      double t = 0;
      st = measure(&t);
      if (st != IEGPU_OK) return st;
      if (t < 0.99 * best) {
        best = t;
        pick = v;
      }
    }
    cur[k] = pick;
    iegpu_set_launch_caps(c, cur, kNumCaps);
  }
  st = iegpu_factor(c, id_tol, precision);  // leave a factorization with the chosen caps
  if (st != IEGPU_OK) return st;
  if (caps_out) std::memcpy(caps_out, cur, sizeof(cur));
  if (ms_out) *ms_out = best;
  return IEGPU_OK;
}

int iegpu_set_profiling(iegpu_ctx* c, int on) {
  if (!c) return IEGPU_ERR_INVALID;
  if (c->device < 0) return fail(c, IEGPU_ERR_NO_DEVICE, "host-only context");
  try {
    HIP_TRY(hipSetDevice(c->device));
    c->prof.reset();
    c->prof.on = on != 0;
  } catch (const HipError& e) {
    return from_hip(c, e);
  }
  return IEGPU_OK;
}

int iegpu_profile(iegpu_ctx* c, int stage, int level, double* ms,
                  int* launches) {
  if (!c || !ms || !launches) return IEGPU_ERR_INVALID;
  if (stage < 0 || stage >= kProfStages || level >= kProfLevels)
    return fail(c, IEGPU_ERR_INVALID, "bad stage or level");
  if (c->device < 0) return fail(c, IEGPU_ERR_NO_DEVICE, "host-only context");
  try {
    HIP_TRY(hipSetDevice(c->device));
    c->prof.resolve();
  } catch (const HipError& e) {
    return from_hip(c, e);
  }
  double t = 0;
  int n = 0;
  for (int l = 0; l < kProfLevels; l++) {
    if (level >= 0 && l != level) continue;
    t += c->prof.ms[stage][l];
    n += c->prof.launches[stage][l];
  }
  *ms = t;
  *launches = n;
  return IEGPU_OK;
}

}  // extern "C"
