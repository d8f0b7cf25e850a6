// FP64 matrix-free apply of the second-kind boundary operator, y = K x,
// with the entries of ie_solver/kernel/kernel.cpp:123-165 (Laplace),
// :168-209 (Laplace Neumann) and :212-290 (Stokes), used for the residual
// of iterative refinement.
//
// One thread per target point; sources staged through LDS in tiles of 256
// (pre-scaled by the weights), every wave reading the same source
// (broadcast).  The reciprocal of |r|^2 is v_rcp_f64 refined by one Newton
// step.  For Stokes the w n_t n_s^T part of every entry sums to
// n_t (sum_j w_j n_j . x_j), a global scalar computed first.
//
// A batch of independent problems has a block-diagonal operator: a target
// sums only over its own problem's points, and a block of targets walks
// only the source tiles of the problems it touches.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "rs_device.hpp"
#include "rs_matvec.hpp"

namespace iegpu {

namespace {

constexpr int kTile = 256;

__device__ inline double rcp_f64(double d) {
  double y = __builtin_amdgcn_rcp(d);
  const double e = fma(-d, y, 1.0);
  return fma(y, e, y);
}

// c_p = sum_j w_j (n_j . x_j) over Stokes DOF pairs, or sum_j w_j x_j for
// Laplace Neumann (its rank-one term), over problem p's points, in one
// fixed order: chunks of kWnChunk points, each reduced by 256 threads
// (thread t sums points t, t + 256, ... of the chunk, then a tree), the
// chunk sums added in chunk order.  A single boundary spreads its chunks
// over workgroups (partial[1 + c]) and k_wn_final adds them into
// partial[0]; a batch takes one workgroup per problem, walking its chunks.
// Both give the same bits for the same problem.
constexpr int kWnChunk = 4096;

__device__ inline double wn_chunk_sum(const DevGeom& g, const double* x, int64_t lo,
                                      int64_t hi, double* red) {
  const bool stokes = g.pde == kPdeStokes;
  double acc = 0;
  for (int64_t j = lo + threadIdx.x; j < hi; j += 256)
    acc += stokes ? g.wts[j] * (g.nrm[2 * j] * x[2 * j] + g.nrm[2 * j + 1] * x[2 * j + 1])
                  : g.wts[j] * x[j];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  const double v = red[0];
  __syncthreads();
  return v;
}

// One chunk of a single boundary per workgroup.
__global__ __launch_bounds__(256) void k_wn_chunk(DevGeom g, const double* x,
                                                  double* partial) {
  __shared__ double red[256];
  const int64_t lo = (int64_t)blockIdx.x * kWnChunk;
  const int64_t hi = lo + kWnChunk < g.n_points ? lo + kWnChunk : g.n_points;
  const double v = wn_chunk_sum(g, x, lo, hi, red);
  if (threadIdx.x == 0) partial[1 + blockIdx.x] = v;
}

__global__ void k_wn_final(double* partial, int nchunks) {
  if (threadIdx.x != 0) return;
  double acc = 0;
  for (int c = 0; c < nchunks; c++) acc += partial[1 + c];
  partial[0] = acc;
}

// One problem of a batch per workgroup.
__global__ __launch_bounds__(256) void k_wn_dot(DevGeom g, const double* x,
                                                double* partial) {
  __shared__ double red[256];
  const int64_t lo = g.prob_off[blockIdx.x], hi = g.prob_off[blockIdx.x + 1];
  double acc = 0;
  for (int64_t c = lo; c < hi; c += kWnChunk)
    acc += wn_chunk_sum(g, x, c, c + kWnChunk < hi ? c + kWnChunk : hi, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

int wn_chunks(int n) { return (n + kWnChunk - 1) / kWnChunk; }

// Problem owning point i: the largest p with prob_off[p] <= i.
__device__ inline int problem_of(const DevGeom& g, int i) {
  int lo = 0, hi = g.n_prob;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (g.prob_off[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(kTile) void k_apply(DevGeom g, const double* x,
                                                 const double* wn, double* y) {
  __shared__ double sx[kTile], sy[kTile], sa[kTile], sb[kTile], sm0[kTile],
      sm1[kTile], sd0[kTile], sd1[kTile];
  const int i = blockIdx.x * kTile + threadIdx.x;
  const bool live = i < g.n_points;
  const double xi = live ? g.pts[2 * i] : 0, yi = live ? g.pts[2 * i + 1] : 0;
  const bool stokes = g.pde == kPdeStokes;
  const bool neumann = g.pde == kPdeLaplaceNeumann;
  const double tn0 = live ? g.nrm[2 * i] : 0, tn1 = live ? g.nrm[2 * i + 1] : 0;
  // source range of this block (blo, bhi) and of this target (plo, phi)
  int blo = 0, bhi = g.n_points, plo = 0, phi = g.n_points, prob = 0;
  if (g.n_prob > 1) {
    const int first = blockIdx.x * kTile;
    const int last = min(first + kTile, g.n_points) - 1;
    blo = g.prob_off[problem_of(g, first)];
    bhi = g.prob_off[problem_of(g, last) + 1];
    prob = problem_of(g, live ? i : last);
    plo = g.prob_off[prob];
    phi = g.prob_off[prob + 1];
  }
  double a0 = 0, a1 = 0;
  for (int base = blo; base < bhi; base += kTile) {
    const int j = base + threadIdx.x;
    if (j < bhi) {
      const double w = g.wts[j], kap = g.curv[j];
      const double n0 = g.nrm[2 * j], n1 = g.nrm[2 * j + 1];
      sx[threadIdx.x] = g.pts[2 * j];
      sy[threadIdx.x] = g.pts[2 * j + 1];
      // sd: the source's term at a target on the same point (its own node
      // or a coincident one), whose entry is the diagonal value of
      // kernel.cpp:149-157,245-263 for this source (rank-one parts aside).
      if (stokes) {
        sa[threadIdx.x] = w * (1.0 / M_PI) * n0;
        sb[threadIdx.x] = w * (1.0 / M_PI) * n1;
        const double x0 = x[2 * j], x1 = x[2 * j + 1];
        sm0[threadIdx.x] = x0;
        sm1[threadIdx.x] = x1;
        const double pot = -0.5 * kap * w * (1.0 / M_PI);
        const double tx = -n1 * x0 + n0 * x1;
        sd0[threadIdx.x] = -0.5 * x0 - pot * n1 * tx;
        sd1[threadIdx.x] = -0.5 * x1 + pot * n0 * tx;
      } else if (neumann) {
        sa[threadIdx.x] = w * (1.0 / (2 * M_PI)) * x[j];
        sd0[threadIdx.x] = (-0.5 + 0.5 * kap * w * (1.0 / (2 * M_PI))) * x[j];
      } else {
        const double s = -w * (1.0 / (2 * M_PI)) * x[j];
        sa[threadIdx.x] = s * n0;
        sb[threadIdx.x] = s * n1;
        sd0[threadIdx.x] = (0.5 + 0.5 * kap * w * (1.0 / (2 * M_PI))) * x[j];
      }
    }
    __syncthreads();
    const int q0 = max(0, plo - base), q1 = min(min(kTile, bhi - base), phi - base);
    if (stokes) {
      for (int q = q0; q < q1; q++) {
        const double r0 = xi - sx[q], r1 = yi - sy[q];
        const double d2 = fma(r0, r0, r1 * r1);
        if (d2 == 0) {  // same point: the source's diagonal value
          a0 += sd0[q];
          a1 += sd1[q];
          continue;
        }
        const double inv = rcp_f64(d2);
        const double pot = fma(r0, sa[q], r1 * sb[q]) * inv * inv;
        const double proj = fma(r0, sm0[q], r1 * sm1[q]) * pot;
        a0 = fma(r0, proj, a0);
        a1 = fma(r1, proj, a1);
      }
    } else if (neumann) {
      for (int q = q0; q < q1; q++) {
        const double r0 = xi - sx[q], r1 = yi - sy[q];
        const double d2 = fma(r0, r0, r1 * r1);
        if (d2 == 0) {
          a0 += sd0[q];
          continue;
        }
        a0 = fma(fma(r0, tn0, r1 * tn1) * sa[q], rcp_f64(d2), a0);
      }
    } else {
      for (int q = q0; q < q1; q++) {
        const double r0 = xi - sx[q], r1 = yi - sy[q];
        const double d2 = fma(r0, r0, r1 * r1);
        if (d2 == 0) {
          a0 += sd0[q];
          continue;
        }
        a0 = fma(fma(r0, sa[q], r1 * sb[q]), rcp_f64(d2), a0);
      }
    }
    __syncthreads();
  }
  if (!live) return;
  if (stokes) {  // plus the n n^T part of every entry (kernel.cpp:245-263)
    const double c = wn[prob];
    y[2 * i] = a0 + tn0 * c;
    y[2 * i + 1] = a1 + tn1 * c;
  } else if (neumann) {
    // sum_j w_j x_j (every entry's rank-one part, the diagonal's too)
    y[i] = a0 + wn[prob];
  } else {
    y[i] = a0;
  }
}

// y[i] += sum_c U[i + ld c] a[c]  and  z[c] = sum_i P[i + ld c] x[i] - a[c].
__global__ void k_border_apply_u(double* y, const double* U, const double* a,
                                 int64_t ndof, int h) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ndof;
       i += (int64_t)gridDim.x * blockDim.x) {
    double acc = y[i];
    for (int c = 0; c < h; c++) acc += U[i + ndof * c] * a[c];
    y[i] = acc;
  }
}

__global__ __launch_bounds__(256) void k_border_apply_psi(const double* P,
                                                          const double* x,
                                                          const double* a,
                                                          int64_t ndof, int h,
                                                          double* z) {
  __shared__ double red[256];
  const int c = blockIdx.x;
  double acc = 0;
  for (int64_t i = threadIdx.x; i < ndof; i += 256) acc += P[i + ndof * c] * x[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) z[c] = red[0] - a[c];
}

// r = b - y (in place on y); partial sums of r^2 and b^2 per block.
__global__ __launch_bounds__(256) void k_residual(const double* b, double* y,
                                                  int64_t n, double* partial) {
  __shared__ double rr[256], bb[256];
  double a = 0, c = 0;
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double bi = b ? b[i] : 0.0;
    const double r = bi - y[i];
    y[i] = r;
    a += r * r;
    c += bi * bi;
  }
  rr[threadIdx.x] = a;
  bb[threadIdx.x] = c;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      rr[threadIdx.x] += rr[threadIdx.x + o];
      bb[threadIdx.x] += bb[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = rr[0];
    partial[2 * blockIdx.x + 1] = bb[0];
  }
}

__global__ void k_axpy(double* y, const double* x, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] += x[i];
}

}  // namespace

int residual_blocks() { return 256; }

void launch_residual(const double* b, double* y, int64_t n, double* partial,
                     hipStream_t st) {
  k_residual<<<256, 256, 0, st>>>(b, y, n, partial);
}

void launch_axpy(double* y, const double* x, int64_t n, hipStream_t st) {
  k_axpy<<<512, 256, 0, st>>>(y, x, n);
}

int matvec_partials(const DevGeom& g) {
  const int single = 1 + wn_chunks(g.n_points);
  return g.n_prob > single ? g.n_prob : single;
}

void launch_wn_dot(const DevGeom& g, const double* x, double* partial, hipStream_t st) {
  if (g.n_prob == 1) {
    const int nc = wn_chunks(g.n_points);
    k_wn_chunk<<<nc, 256, 0, st>>>(g, x, partial);
    k_wn_final<<<1, 64, 0, st>>>(partial, nc);
  } else {
    k_wn_dot<<<g.n_prob, 256, 0, st>>>(g, x, partial);
  }
}

void launch_matvec(const DevGeom& g, const double* x, double* y, double* partial,
                   hipStream_t st) {
  if (g.pde == kPdeStokes || g.pde == kPdeLaplaceNeumann) launch_wn_dot(g, x, partial, st);
  k_apply<<<(g.n_points + kTile - 1) / kTile, kTile, 0, st>>>(g, x, partial, y);
}

void launch_border_apply(const double* U, const double* PsiT, const double* x,
                         const double* alpha, int64_t ndof, int h, double* y,
                         double* z, hipStream_t st) {
  k_border_apply_u<<<512, 256, 0, st>>>(y, U, alpha, ndof, h);
  k_border_apply_psi<<<h, 256, 0, st>>>(PsiT, x, alpha, ndof, h, z);
}

}  // namespace iegpu
