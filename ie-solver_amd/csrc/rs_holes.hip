// Bordered system for multiply connected domains
// (skel_factorization.cpp:188-305 factor side, :596-777 solve side).
//
// With the upward-swept border matrices mU (ndof x h) and mPsi^T
// (ndof x h) and D the block diagonal of the boxes' X_rr:
//   S = [ A_ss            mU(skel,:)                          ]
//       [ mPsi^T(skel,:)^T  -I - mPsi^T(red,:)^T D^{-1} mU(red,:) ]
// "skel" is the root's active list and "red" every other DOF.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "rs_device.hpp"
#include "rs_holes.hpp"

namespace iegpu {

namespace {

__global__ void k_fill_int(int* p, int64_t n, int v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_mark_skel(const DevState s, int* mask) {
  const int n = s.act_cnt[0];
  for (int i = threadIdx.x; i < n; i += blockDim.x) mask[s.act_list[i]] = 1;
}

__global__ void k_copy(double* dst, const double* src, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// partial[blk][a + h*c] = sum over redundant rows i of chunk blk of
// P[i,a] * C[i,c]   (P, C: ndof x h, ld ndof).
__global__ __launch_bounds__(256) void k_gram_partial(const double* P,
                                                      const double* C,
                                                      const int* mask,
                                                      int64_t ndof, int h,
                                                      int64_t chunk,
                                                      double* partial) {
  const int64_t lo = blockIdx.x * chunk;
  const int64_t hi = lo + chunk < ndof ? lo + chunk : ndof;
  for (int pc = threadIdx.x; pc < h * h; pc += blockDim.x) {
    const int a = pc % h, c = pc / h;
    double acc = 0;
    for (int64_t i = lo; i < hi; i++)
      if (!mask[i]) acc += P[i + ndof * a] * C[i + ndof * c];
    partial[blockIdx.x * (int64_t)h * h + pc] = acc;
  }
}

// partial[blk][a] = sum over redundant rows i of P[i,a] * x[i].
__global__ __launch_bounds__(64) void k_gemv_partial(const double* P,
                                                     const double* x,
                                                     const int* mask,
                                                     int64_t ndof, int h,
                                                     int64_t chunk,
                                                     double* partial) {
  const int64_t lo = blockIdx.x * chunk;
  const int64_t hi = lo + chunk < ndof ? lo + chunk : ndof;
  for (int a = threadIdx.x; a < h; a += blockDim.x) {
    double acc = 0;
    for (int64_t i = lo; i < hi; i++)
      if (!mask[i]) acc += P[i + ndof * a] * x[i];
    partial[blockIdx.x * (int64_t)h + a] = acc;
  }
}

// out[j] = sum_blk partial[blk][j] (fixed order: deterministic).
__global__ void k_reduce_partials(const double* partial, int nblk, int len,
                                  double* out, double sign) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < len;
       j += gridDim.x * blockDim.x) {
    double acc = 0;
    for (int b = 0; b < nblk; b++) acc += partial[(int64_t)b * len + j];
    out[j] = sign * acc;
  }
}

template <typename R>
__device__ int lu_inplace_1024(R* X, int n, int ld, int* ipiv) {
  __shared__ int sh_p;
  const int tid = threadIdx.x, lane = tid & 63;
  constexpr int NT = 1024;
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
    if (p != j)
      for (int c = tid; c < n; c += NT) {
        const R t = X[j + (int64_t)ld * c];
        X[j + (int64_t)ld * c] = X[p + (int64_t)ld * c];
        X[p + (int64_t)ld * c] = t;
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
    for (int64_t idx = tid; idx < (int64_t)len * len; idx += NT) {
      const int i = j + 1 + (int)(idx % len), c = j + 1 + (int)(idx / len);
      X[i + (int64_t)ld * c] -= X[i + (int64_t)ld * j] * X[j + (int64_t)ld * c];
    }
    __syncthreads();
  }
  return info;
}

// Assemble S (ns + h square, ld ns + h) and factor it (one workgroup).
template <typename R>
__global__ __launch_bounds__(1024) void k_border_factor(DevState s,
                                                        const double* mU,
                                                        const double* mPsiT,
                                                        const double* gram,
                                                        int64_t ndof, int h,
                                                        double* Smat,
                                                        int* Spiv) {
  const int ns = s.act_cnt[0];
  const int N = ns + h;
  const int* act = s.act_list;
  const R* A = reinterpret_cast<const R*>(s.root_mat);
  for (int64_t idx = threadIdx.x; idx < (int64_t)N * N; idx += blockDim.x) {
    const int i = (int)(idx % N), j = (int)(idx / N);
    double v;
    if (i < ns && j < ns) v = (double)A[i + (int64_t)ns * j];
    else if (i >= ns && j < ns) v = mPsiT[act[j] + ndof * (i - ns)];
    else if (i < ns) v = mU[act[i] + ndof * (j - ns)];
    else v = (i == j ? -1.0 : 0.0) - gram[(i - ns) + h * (j - ns)];
    Smat[idx] = v;
  }
  __syncthreads();
  const int info = lu_inplace_1024<double>(Smat, N, N, Spiv);
  if (info != 0 && threadIdx.x == 0) atomicOr(s.err, kErrSingular);
}

// y = S^{-1} [x(skel); g + tail]; then x(skel) = y_head, alpha = y_tail.
// g is the right-hand side of the border rows (NULL: zero, as in the
// reference's [f; 0]).
__global__ __launch_bounds__(64) void k_border_solve(DevState s,
                                                     const double* Smat,
                                                     const int* Spiv, int h,
                                                     double* x,
                                                     const double* tail,
                                                     const double* g,
                                                     double* alpha,
                                                     double* ybuf) {
  const int lane = threadIdx.x;
  const int ns = s.act_cnt[0];
  const int N = ns + h;
  const int* act = s.act_list;
  for (int i = lane; i < ns; i += 64) ybuf[i] = x[act[i]];
  for (int i = lane; i < h; i += 64) ybuf[ns + i] = tail[i] + (g ? g[i] : 0.0);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (lane == 0)
    for (int i = 0; i < N; i++) {
      const int p = Spiv[i];
      if (p != i) {
        const double t = ybuf[i];
        ybuf[i] = ybuf[p];
        ybuf[p] = t;
      }
    }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int l = 0; l < N; l++) {
    const double yl = ybuf[l];
    for (int i = l + 1 + lane; i < N; i += 64) ybuf[i] -= Smat[i + (int64_t)N * l] * yl;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  for (int i = N - 1; i >= 0; i--) {
    if (lane == 0) ybuf[i] /= Smat[i + (int64_t)N * i];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const double yi = ybuf[i];
    for (int l = lane; l < i; l += 64) ybuf[l] -= Smat[l + (int64_t)N * i] * yi;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  for (int i = lane; i < ns; i += 64) x[act[i]] = ybuf[i];
  for (int i = lane; i < h; i += 64) alpha[i] = ybuf[ns + i];
}

// x(red) -= C(red,:) alpha.
__global__ void k_red_update(double* x, const double* C, const int* mask,
                             const double* alpha, int64_t ndof, int h) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ndof;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (mask[i]) continue;
    double acc = x[i];
    for (int c = 0; c < h; c++) acc -= C[i + ndof * c] * alpha[c];
    x[i] = acc;
  }
}

constexpr int64_t kChunk = 4096;

}  // namespace

void launch_mark_skel(const DevState& s, int* mask, int64_t ndof,
                      hipStream_t st) {
  k_fill_int<<<256, 256, 0, st>>>(mask, ndof, 0);
  k_mark_skel<<<1, 1024, 0, st>>>(s, mask);
}

void launch_copy(double* dst, const double* src, int64_t n, hipStream_t st) {
  k_copy<<<512, 256, 0, st>>>(dst, src, n);
}

// dst row r = src row map[r] (rows of row_words 32-bit words); rows with
// map[r] < 0 are zeroed.
__global__ void k_gather_rows(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src,
                              const int* __restrict__ map, int n_rows, int row_words) {
  const int64_t total = (int64_t)n_rows * row_words;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(i / row_words);
    const int o = map[r];
    dst[i] = o >= 0 ? src[(int64_t)o * row_words + (i - (int64_t)r * row_words)] : 0u;
  }
}

void launch_gather_rows(void* dst, const void* src, const int* map, int n_rows, int row_words,
                        hipStream_t st) {
  if (n_rows <= 0 || row_words <= 0) return;
  const int64_t total = (int64_t)n_rows * row_words;
  const int blocks = (int)std::min<int64_t>((total + 255) / 256, 8192);
  k_gather_rows<<<blocks, 256, 0, st>>>(static_cast<uint32_t*>(dst),
                                        static_cast<const uint32_t*>(src), map, n_rows,
                                        row_words);
}

int64_t border_partial_blocks(int64_t ndof) { return (ndof + kChunk - 1) / kChunk; }

void launch_border_gram(const double* P, const double* C, const int* mask,
                        int64_t ndof, int h, double* partial, double* out,
                        hipStream_t st) {
  const int nb = (int)border_partial_blocks(ndof);
  k_gram_partial<<<nb, 256, 0, st>>>(P, C, mask, ndof, h, kChunk, partial);
  k_reduce_partials<<<(h * h + 255) / 256, 256, 0, st>>>(partial, nb, h * h,
                                                        out, 1.0);
}

void launch_border_gemv(const double* P, const double* x, const int* mask,
                        int64_t ndof, int h, double* partial, double* out,
                        double sign, hipStream_t st) {
  const int nb = (int)border_partial_blocks(ndof);
  k_gemv_partial<<<nb, 64, 0, st>>>(P, x, mask, ndof, h, kChunk, partial);
  k_reduce_partials<<<1, 64, 0, st>>>(partial, nb, h, out, sign);
}

template <typename R>
void launch_border_factor(const DevState& s, const double* mU,
                          const double* mPsiT, const double* gram,
                          int64_t ndof, int h, double* Smat, int* Spiv,
                          hipStream_t st) {
  k_border_factor<R><<<1, 1024, 0, st>>>(s, mU, mPsiT, gram, ndof, h, Smat, Spiv);
}

void launch_border_solve(const DevState& s, const double* Smat,
                         const int* Spiv, int h, double* x, const double* tail,
                         const double* g, double* alpha, double* ybuf,
                         hipStream_t st) {
  k_border_solve<<<1, 64, 0, st>>>(s, Smat, Spiv, h, x, tail, g, alpha, ybuf);
}

void launch_red_update(double* x, const double* C, const int* mask,
                       const double* alpha, int64_t ndof, int h,
                       hipStream_t st) {
  k_red_update<<<512, 256, 0, st>>>(x, C, mask, alpha, ndof, h);
}

template void launch_border_factor<double>(const DevState&, const double*,
                                           const double*, const double*,
                                           int64_t, int, double*, int*,
                                           hipStream_t);
template void launch_border_factor<float>(const DevState&, const double*,
                                          const double*, const double*,
                                          int64_t, int, double*, int*,
                                          hipStream_t);

}  // namespace iegpu
