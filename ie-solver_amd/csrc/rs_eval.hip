// Interior evaluation u = K_domain mu + U_forward alpha, matrix-free
// (ie_solver/initialization.cpp:11-135, linear_solve.cpp:140-154,188-266).
//
// Laplace: u(x) = sum_j -(w_j / 2 pi) ((x - y_j) . n_j) / |x - y_j|^2 mu_j.
// Stokes:  u_a(x) = sum_j (w_j / pi) ((x - y_j) . n_j) / |x - y_j|^4
//                   r_a (r_0 mu_2j + r_1 mu_2j+1)   (no n n^T term off the
//                   boundary, kernel.cpp:29-63 with tgt.is_boundary = false).
// Targets outside the domain give 0 (initialization.cpp:43-49).
//
// Grid: (target blocks of 256) x (source chunks); each thread accumulates
// one target over a source chunk staged in LDS, then a second kernel sums
// the chunks in a fixed order and adds the hole terms.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "rs_device.hpp"
#include "rs_eval.hpp"

namespace iegpu {

// Partial sums of the allocation-free evaluation (iegpu_evaluate_noalloc):
// one copy per GPU, held until the process ends.  Host code uses the
// address hipGetSymbolAddress returns for the current device, never this
// name (which is a host-side stand-in, not a device address).
static __device__ __align__(256) double scratch[kEvalScratchElems];

namespace {

constexpr int kTile = 256;

__global__ __launch_bounds__(kTile) void k_eval_partial(DevGeom g, int nt,
                                                        const double* tg,
                                                        const double* mu,
                                                        int64_t chunk,
                                                        double* partial) {
  __shared__ double sx[kTile], sy[kTile], snx[kTile], sny[kTile], sw[kTile],
      sm0[kTile], sm1[kTile];
  const int t = blockIdx.x * kTile + threadIdx.x;
  const int sd = g.sol_dim;
  const double x0 = t < nt ? tg[2 * t] : 0, x1 = t < nt ? tg[2 * t + 1] : 0;
  const int64_t lo = blockIdx.y * chunk;
  const int64_t hi = lo + chunk < g.n_points ? lo + chunk : g.n_points;
  double a0 = 0, a1 = 0;
  for (int64_t base = lo; base < hi; base += kTile) {
    const int64_t j = base + threadIdx.x;
    if (j < hi) {
      sx[threadIdx.x] = g.pts[2 * j];
      sy[threadIdx.x] = g.pts[2 * j + 1];
      snx[threadIdx.x] = g.nrm[2 * j];
      sny[threadIdx.x] = g.nrm[2 * j + 1];
      sw[threadIdx.x] = g.wts[j];
      sm0[threadIdx.x] = mu[sd * j];
      sm1[threadIdx.x] = sd == 2 ? mu[2 * j + 1] : 0.0;
    }
    __syncthreads();
    const int cnt = (int)((hi - base) < kTile ? (hi - base) : kTile);
    if (g.pde == kPdeStokes) {
      for (int q = 0; q < cnt; q++) {
        const double r0 = x0 - sx[q], r1 = x1 - sy[q];
        const double d2 = r0 * r0 + r1 * r1;
        const double pot = sw[q] * (1.0 / M_PI) * (r0 * snx[q] + r1 * sny[q]) / (d2 * d2);
        const double proj = r0 * sm0[q] + r1 * sm1[q];
        a0 += pot * r0 * proj;
        a1 += pot * r1 * proj;
      }
    } else if (g.pde == kPdeLaplaceNeumann) {
      // single layer w (1/2 pi) log|r| (kernel.cpp:95-102)
      for (int q = 0; q < cnt; q++) {
        const double r0 = x0 - sx[q], r1 = x1 - sy[q];
        a0 += sw[q] * (1.0 / (2 * M_PI)) * log(sqrt(r0 * r0 + r1 * r1)) * sm0[q];
      }
    } else {
      for (int q = 0; q < cnt; q++) {
        const double r0 = x0 - sx[q], r1 = x1 - sy[q];
        a0 += -sw[q] * (1.0 / (2 * M_PI)) * (r0 * snx[q] + r1 * sny[q]) /
              (r0 * r0 + r1 * r1) * sm0[q];
      }
    }
    __syncthreads();
  }
  if (t < nt) {
    partial[(int64_t)blockIdx.y * sd * nt + (int64_t)sd * t] = a0;
    if (sd == 2) partial[(int64_t)blockIdx.y * sd * nt + (int64_t)sd * t + 1] = a1;
  }
}

// Sum chunks in order, add U_forward alpha (linear_solve.cpp:20-67 on the
// targets), zero targets outside the domain.
__global__ void k_eval_finish(DevGeom g, int nt, const double* tg,
                              const int* in_domain, const double* partial,
                              int nchunks, int n_holes, const double* hc,
                              const double* alpha, double* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int sd = g.sol_dim;
  double a0 = 0, a1 = 0;
  for (int c = 0; c < nchunks; c++) {
    a0 += partial[(int64_t)c * sd * nt + (int64_t)sd * t];
    if (sd == 2) a1 += partial[(int64_t)c * sd * nt + (int64_t)sd * t + 1];
  }
  const double x0 = tg[2 * t], x1 = tg[2 * t + 1];
  for (int k = 0; k < n_holes && alpha; k++) {
    const double r0 = x0 - hc[2 * k], r1 = x1 - hc[2 * k + 1];
    const double rn = sqrt(r0 * r0 + r1 * r1);
    if (g.pde == kPdeStokes) {
      const double s = 1.0 / (4 * M_PI), rn2 = rn * rn;
      const double* al = alpha + 3 * k;
      a0 += s * (log(1 / rn) + r0 * r0 / rn2) * al[0] + s * (r0 * r1 / rn2) * al[1] +
            r1 * (s / rn2) * al[2];
      a1 += s * (r1 * r0 / rn2) * al[0] + s * (log(1 / rn) + r1 * r1 / rn2) * al[1] -
            r0 * (s / rn2) * al[2];
    } else {
      a0 += log(rn) * alpha[k];
    }
  }
  const bool inside = in_domain ? in_domain[t] != 0 : true;
  out[(int64_t)sd * t] = inside ? a0 : 0.0;
  if (sd == 2) out[(int64_t)sd * t + 1] = inside ? a1 : 0.0;
}

}  // namespace

int eval_scratch_address(int device, double** ptr) {
  static std::mutex mu;
  static double* addr[kEvalMaxDevices] = {};
  if (device < 0 || device >= kEvalMaxDevices || !ptr) return (int)hipErrorInvalidValue;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    // the device must be current: each GPU holds its own copy
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return (int)e;
    if (cur != device) return (int)hipErrorInvalidDevice;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    addr[device] = static_cast<double*>(p);
  }
  *ptr = addr[device];
  return 0;
}

int64_t eval_scratch_need(int64_t n_points, int sol_dim, int64_t n_targets) {
  return (int64_t)eval_chunks(n_points) * sol_dim * n_targets;
}

int eval_chunks(int64_t n_points) {
  int64_t c = (n_points + 8191) / 8192;
  return (int)(c < 1 ? 1 : (c > 64 ? 64 : c));
}

void launch_eval(const DevGeom& g, int nt, const double* tg, const int* in_domain,
                 const double* mu, int n_holes, const double* hole_centers,
                 const double* alpha, double* partial, double* out,
                 hipStream_t st) {
  if (nt <= 0) return;
  const int nc = eval_chunks(g.n_points);
  const int64_t chunk = (g.n_points + nc - 1) / nc;
  dim3 grid((nt + kTile - 1) / kTile, nc);
  k_eval_partial<<<grid, kTile, 0, st>>>(g, nt, tg, mu, chunk, partial);
  k_eval_finish<<<(nt + 255) / 256, 256, 0, st>>>(g, nt, tg, in_domain, partial, nc,
                                                  n_holes, hole_centers, alpha, out);
}

}  // namespace iegpu
