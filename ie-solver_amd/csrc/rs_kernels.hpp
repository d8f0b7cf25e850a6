// Launch wrappers for the RS factor + solve kernels (rs_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "rs_device.hpp"

namespace iegpu {

struct SolveVec {
  double* x;      // device, ldx x nrhs column-major
  int64_t ldx;
  int nrhs;
};

void launch_reset(const DevState& s, hipStream_t st);
// Standard normal table (n entries) for the randomized-ID sketches.
void launch_gaussian(double* out, int64_t n, uint64_t seed, hipStream_t st);

// Stages of one tree level of the factorization, in launch order.
enum FactorStage : int {
  kStageActive = 0,   // active lists from the children
  kStageRows = 1,     // ID row counts and slot sizes
  kStageScan = 2,     // slot offsets
  kStageAssemble = 3, // ID matrices
  kStageRid = 4,      // randomized ID (Gaussian sketch + QR of the sketch)
  kStageQrcp = 5,     // pivoted QR, rank, T
  kStageDecouple = 6, // Schur complement, X_rr LU, L, U
  kNumFactorStages = 7
};

// Launch shapes chosen per workload by iegpu_autotune.  The caps are the
// grids of the kernels that walk a device-side box list with a grid
// stride (the host knows only the level's box count, not the list's
// length); workgroups past the list's end exit at once.
struct LaunchCaps {
  int assemble = 1024;   // k_assemble (general-QRCP boxes' ID matrices)
  int qrcp_id = 1024;    // k_qrcp_id register shapes, randomized-ID fallback
  int rowcomp4 = 1024;   // k_rowcomp<4> (box, chunk) jobs
  int rowcomp8 = 1024;   // k_rowcomp<8> (box, chunk) jobs
  int qrcp_id8 = 1024;   // k_qrcp_id<8, 4, 16> (n <= 128 after row compression)
  int qrcp = 1024;       // k_qrcp (general LDS/global QRCP)
  // Levels with at least this many boxes decouple their n <= 64 boxes in
  // two LDS size classes (n <= 44 at five boxes per CU, the rest at
  // three), in two launches; smaller levels use one launch of the larger
  // class, whose single box latency is all they pay.
  int dsplit = 8192;
  // Levels with at least this many boxes decouple their compressed boxes
  // with |active| <= 64 by the register sweep, one wave per box
  // (k_decouple_sweep: more boxes per CU); smaller levels by Gauss-Jordan
  // in 256-thread workgroups (k_decouple_small, LDS classes by dsplit:
  // four waves shorten each box's latency, which is all a small level pays).
  int dsweep = 4096;
};
constexpr int kNumCaps = 8;

// `skip`: kSkip* bits of kernels not to launch (their lists are empty by
// the launch plan, which k_root verifies).  `st_big`: the decouple stage's
// boxes above 64 DOFs go on this stream when given (the caller orders it).
template <typename R>
void launch_factor_stage(int stage, const DevGeom& g, const DevTree& t,
                         const DevState& s, int level, int start, int count,
                         double tol, const LaunchCaps& caps, hipStream_t st,
                         int skip = 0, hipStream_t st_big = nullptr);

// Decouple of one level's compressed boxes with |active| <= 64, one wave
// per box (rs_sweep.hip); the other boxes exit at once.
template <typename R>
void launch_decouple_sweep(const DevGeom& g, const DevTree& t, const DevState& s, int start,
                           int count, hipStream_t st);

// Roots (slots 0 .. n_roots-1, one per problem): active list, dense
// skeleton block, inverse.
template <typename R>
void launch_root(const DevGeom& g, const DevTree& t, const DevState& s,
                 int n_roots, bool factor_lu, hipStream_t st);

// Upward sweep of one level: x_r -= T^T x_s; x_s -= L x_r (or U^T x_r when
// `use_Ut`); then x_r = X_rr^{-1} x_r when `diag`.
template <typename R>
void launch_sweep_up(const DevState& s, int start, int count, SolveVec v,
                     bool use_Ut, bool diag, hipStream_t st);

// Block-diagonal X_rr^{-1} on the redundant rows of one level.
template <typename R>
void launch_diag(const DevState& s, int start, int count, SolveVec v,
                 hipStream_t st);

// Downward sweep of one level: x_r -= U x_s; x_s -= T x_r.
template <typename R>
void launch_sweep_down(const DevState& s, int start, int count, SolveVec v,
                       hipStream_t st);

// x(allskel) = A_root^{-1} x(allskel) for every root.
template <typename R>
void launch_root_solve(const DevState& s, int n_roots, SolveVec v, hipStream_t st);

// Compressed forward apply b = A x in place on x (ndof), from the factors
// and the X_rr / root blocks kept in fwd_pool / root_fwd: per level,
// leaves first, launch_fwd_up; then launch_fwd_root; then per level, root
// side first, launch_fwd_down.
template <typename R>
void launch_fwd_up(const DevState& s, int start, int count, double* x, hipStream_t st);
template <typename R>
void launch_fwd_root(const DevState& s, int n_roots, double* x, hipStream_t st);
template <typename R>
void launch_fwd_down(const DevState& s, int start, int count, double* x, hipStream_t st);

}  // namespace iegpu
