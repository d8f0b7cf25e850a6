// FP64 fast multipole apply of the second-kind boundary operator, y = K x,
// with the entries of ie_solver/kernel/kernel.cpp:123-165 (Laplace),
// :168-209 (Laplace Neumann) and :212-290 (Stokes): the far field through
// complex (Cauchy) expansions on a uniform-depth quadtree (fmm_tree.hpp),
// the adjacent leaves summed directly with the same real-valued entries,
// diagonal rule and rank-one terms as the direct apply (rs_matvec.hip).
//
// Complex forms of the far-field entries (t target, y source, r = t - y,
// n the unit normal and w the quadrature weight of the source, all as
// complex numbers):
//   Laplace DLP   Re( sum q/r ),       q = -(w x / 2 pi) n
//   Neumann       Re( n_t sum q/r ),   q = w x / 2 pi
//   Stokes        S1 + conj(S2) + (t - o) conj(D3) - conj(D4), with
//                 S1 = sum c1/r, S2 = sum c2/r, D3 = sum c1/r^2,
//                 D4 = sum c4/r^2,  c1 = (w / 4 pi) n mu,
//                 c2 = (w / 2 pi) Re(conj(n) mu), c4 = c1 conj(y - o),
//                 mu = x0 + i x1, for any origin o.
// The Stokes origin o is each expansion's own box centre: with a fixed
// origin, (t - o) conj(D3) and conj(D4) are each |t - o| / |r| times larger
// than their difference (2^depth at the nearest far boxes), and rounding
// grows with the tree depth.  Moving the origin from o to o' adds
// conj(o - o') times the c1 expansion to the c4 expansion, so every
// translation first shifts the c4 coefficients by the c1 ones.
// Expansions are scaled by the box side s and centred at the box centre c
// (u = (z - c) / s):
//   multipole  phi(z) = (1/s) sum_k M_k / u^(k+1),   M_k = sum q u_y^k
//   local      phi(z) = (1/s) sum_l L_l u^l
//   M2M        M'_k = sum_{m<=k} C(k,m) 2^-m d^(k-m) M_m      (d = (c - c')/s')
//   M2L        L_l  = (-1)^l w^(l+1) sum_k C(k+l,l) w^k M_k   (w = s/(c_t - c_s))
//   L2L        L'_m = 2^-(m+1) sum_{l>=m} C(l,m) d^(l-m) L_l   (d = (c' - c)/s)
// One lane per coefficient: one wave per leaf, four per box in the
// translations.  Interaction lists hold boxes at least one box apart, so
// the truncation error of order P falls like 0.55^P or faster (P = 48:
// about 1e-15 relative on smooth curves).

#include <hip/hip_runtime.h>

#include <cstdint>

#include "rs_fmm.hpp"
#include "rs_matvec.hpp"

namespace iegpu {

namespace {

struct cplx {
  double r, i;
};

__device__ inline cplx cmul(cplx a, cplx b) {
  return {fma(a.r, b.r, -a.i * b.i), fma(a.r, b.i, a.i * b.r)};
}
__device__ inline cplx cfma(cplx a, cplx b, cplx c) {  // a b + c
  return {fma(a.r, b.r, fma(-a.i, b.i, c.r)), fma(a.r, b.i, fma(a.i, b.r, c.i))};
}
__device__ inline cplx cscale(double s, cplx a) { return {s * a.r, s * a.i}; }
__device__ inline cplx crecip(cplx a) {
  const double d = 1.0 / fma(a.r, a.r, a.i * a.i);
  return {a.r * d, -a.i * d};
}
// a^k for 0 <= k < 64 by binary powering.
__device__ inline cplx cpow(cplx a, int k) {
  cplx p = {1.0, 0.0};
#pragma unroll
  for (int bit = 0; bit < 6; bit++) {
    const cplx q = cmul(p, a);
    p.r = (k >> bit) & 1 ? q.r : p.r;
    p.i = (k >> bit) & 1 ? q.i : p.i;
    a = cmul(a, a);
  }
  return p;
}
__device__ inline cplx ld(const double* p) { return {p[0], p[1]}; }

__device__ inline double rcp_f64(double d) {
  double y = __builtin_amdgcn_rcp(d);
  const double e = fma(-d, y, 1.0);
  return fma(y, e, y);
}

template <int PDE>
struct Pde {
  static constexpr int NE = PDE == kPdeStokes ? 3 : 1;
  static constexpr int SD = PDE == kPdeStokes ? 2 : 1;
};

__global__ void k_fmm_gather(DevGeom g, const int* perm, int n, double* sgeo) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const int o = perm[j];
    double* d = sgeo + 6 * (size_t)j;
    d[0] = g.pts[2 * o];
    d[1] = g.pts[2 * o + 1];
    d[2] = g.nrm[2 * o];
    d[3] = g.nrm[2 * o + 1];
    d[4] = g.wts[o];
    d[5] = g.curv[o];
  }
}

// Leaf multipoles from the nodes' charges (staged in LDS, 64 nodes at a
// time); also the sorted density.
template <int P, int PDE>
__global__ __launch_bounds__(64) void k_fmm_p2m(FmmDev f, const double* __restrict__ x) {
  constexpr int NE = Pde<PDE>::NE;
  __shared__ double su[64][2], sq[64][NE][2];
  const int leaf = blockIdx.x, lane = threadIdx.x;
  const int b = f.first_leaf + leaf;
  const int p0 = f.pt_off[leaf], p1 = f.pt_off[leaf + 1];
  const double inv_s = 1.0 / ldexp(f.side, -f.depth);
  const double cx = f.center[2 * b], cy = f.center[2 * b + 1];
  const int k = lane < P ? lane : 0;
  cplx acc[NE];
#pragma unroll
  for (int e = 0; e < NE; e++) acc[e] = {0.0, 0.0};
  for (int base = p0; base < p1; base += 64) {
    const int j = base + lane;
    if (j < p1) {
      const double* gj = f.sgeo + 6 * (size_t)j;
      const double px = gj[0], py = gj[1], n0 = gj[2], n1 = gj[3], w = gj[4];
      const int o = f.perm[j];
      su[lane][0] = (px - cx) * inv_s;
      su[lane][1] = (py - cy) * inv_s;
      cplx q[NE];
      if (PDE == kPdeStokes) {
        const double x0 = x[2 * o], x1 = x[2 * o + 1];
        f.xs[2 * j] = x0;
        f.xs[2 * j + 1] = x1;
        const double a = w * (0.25 / M_PI);
        q[0] = {a * (n0 * x0 - n1 * x1), a * (n0 * x1 + n1 * x0)};
        q[1] = {w * (0.5 / M_PI) * (n0 * x0 + n1 * x1), 0.0};
        q[NE - 1] = cmul(q[0], {px - cx, -(py - cy)});
      } else if (PDE == kPdeLaplaceNeumann) {
        const double xj = x[o];
        f.xs[j] = xj;
        q[0] = {w * (0.5 / M_PI) * xj, 0.0};
      } else {
        const double xj = x[o];
        f.xs[j] = xj;
        const double a = -w * (0.5 / M_PI) * xj;
        q[0] = {a * n0, a * n1};
      }
#pragma unroll
      for (int e = 0; e < NE; e++) {
        sq[lane][e][0] = q[e].r;
        sq[lane][e][1] = q[e].i;
      }
    }
    __syncthreads();
    const int cnt = min(64, p1 - base);
    for (int t = 0; t < cnt; t++) {
      const cplx pk = cpow({su[t][0], su[t][1]}, k);
#pragma unroll
      for (int e = 0; e < NE; e++) acc[e] = cfma({sq[t][e][0], sq[t][e][1]}, pk, acc[e]);
    }
    __syncthreads();
  }
  if (lane < P) {
#pragma unroll
    for (int e = 0; e < NE; e++) {
      double* m = f.mexp + 2 * (((size_t)b * NE + e) * P + lane);
      m[0] = acc[e].r;
      m[1] = acc[e].i;
    }
  }
}

// Multipoles of level `level`'s boxes from their children: four waves per
// box, wave w translating child w (its coefficients staged in LDS with the
// Stokes c4 part already moved to the parent's centre); the four partial
// sums are added in wave order.
template <int P, int NE>
__global__ __launch_bounds__(256) void k_fmm_m2m(FmmDev f, int first, int level) {
  __shared__ double sm[4][NE][P][2];
  const int b = first + blockIdx.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int k = lane < P ? lane : P - 1;
  const double inv_s = 1.0 / ldexp(f.side, -level);
  const double cx = f.center[2 * b], cy = f.center[2 * b + 1];
  const int c0 = f.child_off[b], nc = f.child_cnt[b];
  const bool mine = wv < nc;
  const int c = c0 + (mine ? wv : 0);
  const double dx = f.center[2 * c] - cx, dy = f.center[2 * c + 1] - cy;
  if (mine && lane < P) {
    const double* mc = f.mexp + 2 * ((size_t)c * NE * P);
    cplx m[NE];
#pragma unroll
    for (int e = 0; e < NE; e++) m[e] = ld(mc + 2 * (e * P + lane));
    if (NE == 3) m[NE - 1] = cfma({dx, -dy}, m[0], m[NE - 1]);  // conj(c_child - c_parent)
#pragma unroll
    for (int e = 0; e < NE; e++) {
      sm[wv][e][lane][0] = m[e].r;
      sm[wv][e][lane][1] = m[e].i;
    }
  }
  cplx acc[NE];
#pragma unroll
  for (int e = 0; e < NE; e++) acc[e] = {0.0, 0.0};
  __syncthreads();
  if (mine) {
    double tm[P];  // C(k, m) 2^-m, 0 for m > k
#pragma unroll
    for (int m = 0; m < P; m++) tm[m] = f.t_m2m[m * kFmmMaxOrder + k];
    const cplx d = {dx * inv_s, dy * inv_s};
    cplx pw = {1.0, 0.0};
#pragma unroll
    for (int m = P - 1; m >= 0; m--) {
      const cplx tp = cscale(tm[m], pw);
#pragma unroll
      for (int e = 0; e < NE; e++) acc[e] = cfma(tp, {sm[wv][e][m][0], sm[wv][e][m][1]}, acc[e]);
      const cplx q = cmul(pw, d);
      pw.r = m <= k ? q.r : pw.r;
      pw.i = m <= k ? q.i : pw.i;
    }
  }
  __syncthreads();
  if (lane < P) {
#pragma unroll
    for (int e = 0; e < NE; e++) {
      sm[wv][e][lane][0] = acc[e].r;
      sm[wv][e][lane][1] = acc[e].i;
    }
  }
  __syncthreads();
  if (wv == 0 && lane < P) {
#pragma unroll
    for (int e = 0; e < NE; e++) {
      double r = sm[0][e][lane][0], i = sm[0][e][lane][1];
      for (int w = 1; w < nc; w++) {
        r += sm[w][e][lane][0];
        i += sm[w][e][lane][1];
      }
      double* m = f.mexp + 2 * (((size_t)b * NE + e) * P + lane);
      m[0] = r;
      m[1] = i;
    }
  }
}

// A source box's multipole with its Stokes c4 coefficients moved to the
// target's centre (lane < P: coefficient `lane`).
template <int P, int NE>
__device__ inline void load_source(const FmmDev& f, int src, int lane, double cx, double cy,
                                   cplx (&m)[NE]) {
  const double* ms = f.mexp + 2 * ((size_t)src * NE * P);
#pragma unroll
  for (int e = 0; e < NE; e++) m[e] = ld(ms + 2 * (e * P + lane));
  if (NE == 3)  // conj(c_source - c_target)
    m[NE - 1] = cfma({f.center[2 * src] - cx, -(f.center[2 * src + 1] - cy)}, m[0], m[NE - 1]);
}

// Locals of level `level`'s boxes: the parent's local shifted in (L2L,
// wave 0) plus the interaction list's multipoles (M2L): W waves per box,
// wave w taking sources w, w + W, ... (the next one loaded while the
// current one is summed); the partial sums are added in wave order.  Four
// waves cut the latency of a box on the upper levels, where there are too
// few boxes to fill the chip; one wave per box wastes none on the lower.
template <int P, int NE, int W>
__global__ __launch_bounds__(64 * W) void k_fmm_down(FmmDev f, int first, int level) {
  __shared__ double sa[W][NE][P][2];
  const int b = first + blockIdx.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int j = lane < P ? lane : P - 1;
  const double s = ldexp(f.side, -level);
  const double cx = f.center[2 * b], cy = f.center[2 * b + 1];
  cplx acc[NE];
#pragma unroll
  for (int e = 0; e < NE; e++) acc[e] = {0.0, 0.0};
  if (level > 2 && wv == 0) {
    const int p = f.parent[b];
    const double dx = cx - f.center[2 * p], dy = cy - f.center[2 * p + 1];
    const double inv_sp = 0.5 / s;
    const cplx d = {dx * inv_sp, dy * inv_sp};
    const double* lp = f.lexp + 2 * ((size_t)p * NE * P);
    cplx l[NE];
#pragma unroll
    for (int e = 0; e < NE; e++) l[e] = ld(lp + 2 * (e * P + j));
    if (NE == 3) l[NE - 1] = cfma({-dx, dy}, l[0], l[NE - 1]);  // conj(c_parent - c_child)
    if (lane < P) {
#pragma unroll
      for (int e = 0; e < NE; e++) {
        sa[0][e][lane][0] = l[e].r;
        sa[0][e][lane][1] = l[e].i;
      }
    }
    // one wave reads what it wrote: a wave-level barrier orders the LDS access
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    cplx pw = {1.0, 0.0};
    // the table is read in the loop (once per box) so that its registers
    // stay free for the M2L table below
#pragma unroll 4
    for (int q = 0; q < P; q++) {
      const cplx tp = cscale(f.t_l2l[q * kFmmMaxOrder + j], pw);  // C(q, j) 2^-(j+1), 0 for q < j
#pragma unroll
      for (int e = 0; e < NE; e++) acc[e] = cfma(tp, {sa[0][e][q][0], sa[0][e][q][1]}, acc[e]);
      const cplx t = cmul(pw, d);
      pw.r = q >= j ? t.r : pw.r;
      pw.i = q >= j ? t.i : pw.i;
    }
  }
  __syncthreads();
  double tc[P];  // C(k + j, j), k = 0..P-1
#pragma unroll
  for (int k = 0; k < P; k++) tc[k] = f.t_m2l[k * kFmmMaxOrder + j];
  const double sign = (j & 1) ? -1.0 : 1.0;
  const int v0 = f.vl_off[b], v1 = f.vl_off[b + 1];
  const int rounds = (v1 - v0 + W - 1) / W;
  cplx nxt[NE];
#pragma unroll
  for (int e = 0; e < NE; e++) nxt[e] = {0.0, 0.0};
  if (v0 + wv < v1 && lane < P) load_source<P, NE>(f, f.vl[v0 + wv], lane, cx, cy, nxt);
  for (int it = 0; it < rounds; it++) {
    const int v = v0 + W * it + wv;
    const bool live = v < v1;
    const int src = f.vl[live ? v : v0];
    cplx cur[NE];
#pragma unroll
    for (int e = 0; e < NE; e++) cur[e] = nxt[e];
    if (v + W < v1 && lane < P) load_source<P, NE>(f, f.vl[v + W], lane, cx, cy, nxt);
    const cplx w = crecip({(cx - f.center[2 * src]) / s, (cy - f.center[2 * src + 1]) / s});
    const cplx wj = cpow(w, j);
    if (lane < P) {
#pragma unroll
      for (int e = 0; e < NE; e++) {
        const cplx a = cmul(cur[e], wj);
        sa[wv][e][lane][0] = a.r;
        sa[wv][e][lane][1] = a.i;
      }
    }
    __syncthreads();
    if (live) {
      cplx sum[NE];
#pragma unroll
      for (int e = 0; e < NE; e++) sum[e] = {0.0, 0.0};
#pragma unroll
      for (int k = 0; k < P; k++) {
#pragma unroll
        for (int e = 0; e < NE; e++) {
          sum[e].r = fma(tc[k], sa[wv][e][k][0], sum[e].r);
          sum[e].i = fma(tc[k], sa[wv][e][k][1], sum[e].i);
        }
      }
      const cplx g = cscale(sign, cmul(wj, w));
#pragma unroll
      for (int e = 0; e < NE; e++) acc[e] = cfma(g, sum[e], acc[e]);
    }
    __syncthreads();
  }
  if (W > 1 && lane < P) {
#pragma unroll
    for (int e = 0; e < NE; e++) {
      sa[wv][e][lane][0] = acc[e].r;
      sa[wv][e][lane][1] = acc[e].i;
    }
  }
  if (W > 1) __syncthreads();
  if (wv == 0 && lane < P) {
#pragma unroll
    for (int e = 0; e < NE; e++) {
      double r = W > 1 ? sa[0][e][lane][0] : acc[e].r, i = W > 1 ? sa[0][e][lane][1] : acc[e].i;
      for (int w = 1; w < W; w++) {
        r += sa[w][e][lane][0];
        i += sa[w][e][lane][1];
      }
      double* m = f.lexp + 2 * (((size_t)b * NE + e) * P + lane);
      m[0] = r;
      m[1] = i;
    }
  }
}

// Leaf targets: local expansion (value and derivative by Horner) plus the
// adjacent leaves' nodes summed directly; written in the caller's order.
template <int P, int PDE>
__global__ __launch_bounds__(64) void k_fmm_eval(FmmDev f, const double* wn, double* y) {
  constexpr int NE = Pde<PDE>::NE;
  __shared__ double sx[64], sy[64], sa[64], sb[64], sm0[64], sm1[64], sd0[64], sd1[64];
  __shared__ double lb[NE * P * 2];  // the leaf's local coefficients
  const bool stokes = PDE == kPdeStokes, neumann = PDE == kPdeLaplaceNeumann;
  const int leaf = blockIdx.x, lane = threadIdx.x;
  const int b = f.first_leaf + leaf;
  const int p0 = f.pt_off[leaf], p1 = f.pt_off[leaf + 1];
  const double s = ldexp(f.side, -f.depth), inv_s = 1.0 / s;
  const double cx = f.center[2 * b], cy = f.center[2 * b + 1];
  for (int t = lane; t < NE * P * 2; t += 64) lb[t] = f.lexp[(size_t)b * NE * P * 2 + t];
  __syncthreads();
  const int q0 = f.nb_off[leaf], q1 = f.nb_off[leaf + 1];
  for (int base = p0; base < p1; base += 64) {
    const int i = base + lane;
    const bool live = i < p1;
    const double* gi = f.sgeo + 6 * (size_t)(live ? i : p0);
    const double xi = gi[0], yi = gi[1], tn0 = gi[2], tn1 = gi[3];
    const cplx u = {(xi - cx) * inv_s, (yi - cy) * inv_s};
    // far field
    double a0, a1 = 0;
    {
      cplx v0 = {0.0, 0.0}, dv0 = {0.0, 0.0};
#pragma unroll 8
      for (int l = P - 1; l >= 0; l--) {
        const cplx c = ld(lb + 2 * l);
        dv0 = l > 0 ? cfma(dv0, u, cscale(double(l), c)) : dv0;
        v0 = cfma(v0, u, c);
      }
      const cplx phi = cscale(inv_s, v0);
      if (stokes) {
        cplx v1 = {0.0, 0.0}, dv2 = {0.0, 0.0};
#pragma unroll 8
        for (int l = P - 1; l >= 0; l--) {
          v1 = cfma(v1, u, ld(lb + 2 * (P + l)));
          if (l > 0) dv2 = cfma(dv2, u, cscale(double(l), ld(lb + 2 * (2 * P + l))));
        }
        const double is2 = inv_s * inv_s;
        const cplx S2 = cscale(inv_s, v1);
        const cplx D3 = cscale(-is2, dv0), D4 = cscale(-is2, dv2);
        const double tr = xi - cx, ti = yi - cy;
        a0 = phi.r + S2.r + tr * D3.r + ti * D3.i - D4.r;
        a1 = phi.i - S2.i + ti * D3.r - tr * D3.i + D4.i;
      } else if (neumann) {
        a0 = tn0 * phi.r - tn1 * phi.i;
      } else {
        a0 = phi.r;
      }
    }
    // adjacent leaves, directly
    for (int q = q0; q < q1; q++) {
      const int nl = f.nb[q];
      const int s0 = f.pt_off[nl], s1 = f.pt_off[nl + 1];
      for (int sb0 = s0; sb0 < s1; sb0 += 64) {
        const int jj = sb0 + lane;
        if (jj < s1) {
          const double* gj = f.sgeo + 6 * (size_t)jj;
          const double n0 = gj[2], n1 = gj[3], w = gj[4], kap = gj[5];
          sx[lane] = gj[0];
          sy[lane] = gj[1];
          // sd: the source's term at a target on the same point, the
          // diagonal value of kernel.cpp:149-157,245-263 (rank-one parts aside)
          if (stokes) {
            const double x0 = f.xs[2 * jj], x1 = f.xs[2 * jj + 1];
            sa[lane] = w * (1.0 / M_PI) * n0;
            sb[lane] = w * (1.0 / M_PI) * n1;
            sm0[lane] = x0;
            sm1[lane] = x1;
            const double pot = -0.5 * kap * w * (1.0 / M_PI);
            const double tx = -n1 * x0 + n0 * x1;
            sd0[lane] = -0.5 * x0 - pot * n1 * tx;
            sd1[lane] = -0.5 * x1 + pot * n0 * tx;
          } else if (neumann) {
            const double xj = f.xs[jj];
            sa[lane] = w * (1.0 / (2 * M_PI)) * xj;
            sd0[lane] = (-0.5 + 0.5 * kap * w * (1.0 / (2 * M_PI))) * xj;
          } else {
            const double xj = f.xs[jj];
            const double sc = -w * (1.0 / (2 * M_PI)) * xj;
            sa[lane] = sc * n0;
            sb[lane] = sc * n1;
            sd0[lane] = (0.5 + 0.5 * kap * w * (1.0 / (2 * M_PI))) * xj;
          }
        }
        __syncthreads();
        const int cnt = min(64, s1 - sb0);
        for (int t = 0; t < cnt; t++) {
          const double r0 = xi - sx[t], r1 = yi - sy[t];
          const double d2 = fma(r0, r0, r1 * r1);
          if (d2 == 0) {
            a0 += sd0[t];
            if (stokes) a1 += sd1[t];
            continue;
          }
          const double inv = rcp_f64(d2);
          if (stokes) {
            const double pot = fma(r0, sa[t], r1 * sb[t]) * inv * inv;
            const double proj = fma(r0, sm0[t], r1 * sm1[t]) * pot;
            a0 = fma(r0, proj, a0);
            a1 = fma(r1, proj, a1);
          } else if (neumann) {
            a0 = fma(fma(r0, tn0, r1 * tn1) * sa[t], inv, a0);
          } else {
            a0 = fma(fma(r0, sa[t], r1 * sb[t]), inv, a0);
          }
        }
        __syncthreads();
      }
    }
    if (live) {
      const int o = f.perm[i];
      if (stokes) {
        const double c = wn[0];
        y[2 * o] = a0 + tn0 * c;
        y[2 * o + 1] = a1 + tn1 * c;
      } else if (neumann) {
        y[o] = a0 + wn[0];
      } else {
        y[o] = a0;
      }
    }
  }
}

// Levels with fewer boxes than this take four waves per box downward.
constexpr int kDownWideBoxes = 1536;

template <int P, int PDE>
void run_fmm(const DevGeom& g, const FmmDev& f, const std::vector<int>& level_off,
             const double* x, double* y, const double* wn, hipStream_t st) {
  constexpr int NE = Pde<PDE>::NE;
  const int L = f.depth;
  k_fmm_p2m<P, PDE><<<f.n_leaves, 64, 0, st>>>(f, x);
  for (int l = L - 1; l >= 2; l--) {
    const int nbx = level_off[l - 1] - level_off[l - 2];
    k_fmm_m2m<P, NE><<<nbx, 256, 0, st>>>(f, level_off[l - 2], l);
  }
  for (int l = 2; l <= L; l++) {
    const int nbx = level_off[l - 1] - level_off[l - 2];
    if (nbx < kDownWideBoxes) k_fmm_down<P, NE, 4><<<nbx, 256, 0, st>>>(f, level_off[l - 2], l);
    else k_fmm_down<P, NE, 1><<<nbx, 64, 0, st>>>(f, level_off[l - 2], l);
  }
  k_fmm_eval<P, PDE><<<f.n_leaves, 64, 0, st>>>(f, wn, y);
}

template <int P>
void run_order(const DevGeom& g, const FmmDev& f, const std::vector<int>& level_off,
               const double* x, double* y, const double* wn, hipStream_t st) {
  if (g.pde == kPdeStokes) run_fmm<P, kPdeStokes>(g, f, level_off, x, y, wn, st);
  else if (g.pde == kPdeLaplaceNeumann) run_fmm<P, kPdeLaplaceNeumann>(g, f, level_off, x, y, wn, st);
  else run_fmm<P, kPdeLaplace>(g, f, level_off, x, y, wn, st);
}

}  // namespace

bool fmm_order_supported(int order) { return order == 32 || order == 48 || order == 64; }

int fmm_n_exp(int pde) { return pde == kPdeStokes ? 3 : 1; }

void launch_fmm_gather(const DevGeom& g, const FmmDev& f, double* sgeo, hipStream_t st) {
  k_fmm_gather<<<512, 256, 0, st>>>(g, f.perm, g.n_points, sgeo);
}

void launch_fmm_matvec(const DevGeom& g, const FmmDev& f,
                       const std::vector<int>& level_off, const double* x,
                       double* y, double* partial, hipStream_t st) {
  if (g.pde == kPdeStokes || g.pde == kPdeLaplaceNeumann) launch_wn_dot(g, x, partial, st);
  if (f.order == 32) run_order<32>(g, f, level_off, x, y, partial, st);
  else if (f.order == 64) run_order<64>(g, f, level_off, x, y, partial, st);
  else run_order<48>(g, f, level_off, x, y, partial, st);
}

}  // namespace iegpu
