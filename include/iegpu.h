/*
 * iegpu.h -- C ABI of the MI355X (gfx950) recursive-skeletonization (RS)
 * factor + solve for 2-D second-kind boundary integral equations.
 *
 * The reference (jpryan1/ie-solver) exposes this path as a C++ library API
 * with no FFI; each entry point below names the reference interface it
 * replaces (paths relative to the reference root).  Plain pointers and
 * sizes only; host arrays are borrowed for the duration of the call;
 * device buffers are owned by the context.  Each context owns two HIP
 * streams (the second carries the factorization's decouple chain, see
 * iegpu_set_overlap); a context is not thread-safe; separate contexts are
 * independent.
 *
 * Every function returns an iegpu_status (0 = success).  On failure
 * iegpu_last_error() describes the cause.
 */
#ifndef IEGPU_H_
#define IEGPU_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct iegpu_ctx iegpu_ctx;

typedef enum {
  IEGPU_OK = 0,
  IEGPU_ERR_INVALID = 1,        /* bad argument or call order            */
  IEGPU_ERR_NO_COMPRESSION = 2, /* skel_factorization.cpp:209-212         */
  IEGPU_ERR_SINGULAR = 3,       /* exact zero pivot (LAPACK info > 0)     */
  IEGPU_ERR_OOM = 4,            /* device allocation failed               */
  IEGPU_ERR_HIP = 5,            /* HIP runtime error                      */
  IEGPU_ERR_CAPACITY = 6,       /* a box exceeded the supported size      */
  IEGPU_ERR_NO_DEVICE = 7       /* no HIP device available                */
} iegpu_status;

/* ie_solver_config::Pde (ie_solver/io/io.h:19-23). */
enum { IEGPU_PDE_LAPLACE = 0, IEGPU_PDE_LAPLACE_NEUMANN = 1, IEGPU_PDE_STOKES = 2 };

/* Arithmetic of the stored factorization. */
enum { IEGPU_FP64 = 0, IEGPU_FP32 = 1 };

/* Create / destroy a context on HIP device `device` (device = -1: a
 * host-only context that builds trees but cannot factor).  No reference
 * counterpart (the reference holds state in QuadTree/SkelFactorization
 * objects, ie_solver/quadtree/quadtree.h:95-139). */
int iegpu_create(iegpu_ctx** ctx, int device);
int iegpu_destroy(iegpu_ctx* ctx);
const char* iegpu_last_error(const iegpu_ctx* ctx);

/* Boundary discretization: n_points nodes, SoA arrays points[2n],
 * normals[2n], curvatures[n], weights[n]; pde selects the kernel and the
 * DOFs per node (1 Laplace, 2 Stokes).
 * Replaces Kernel(sol_dim, dom_dim, pde, boundary, ...) and the Boundary
 * data contract: ie_solver/kernel/kernel.h:28-35,
 * ie_solver/boundaries/boundary.h:39-41. */
int iegpu_set_geometry(iegpu_ctx* ctx, int pde, int n_points,
                       const double* points, const double* normals,
                       const double* curvatures, const double* weights);

/* A batch of n_problems independent boundaries in one context: problem p
 * owns nodes [point_offsets[p], point_offsets[p+1]) of the concatenated
 * arrays (point_offsets[0] = 0, each problem non-empty).  build_tree makes
 * one quadtree per problem, merged level by level, so every stage of
 * factor and solve covers all problems in one launch per level; vectors
 * are the problems' DOF vectors concatenated.  Holes and interior
 * evaluation are not supported for n_problems > 1.  No counterpart in the
 * reference (one boundary per
 * QuadTree, ie_solver/quadtree/quadtree.h:95-139). */
int iegpu_set_geometry_batch(iegpu_ctx* ctx, int pde, int n_problems,
                             const int* point_offsets, const double* points,
                             const double* normals, const double* curvatures,
                             const double* weights);

/* Replaces QuadTree::perturb (ie_solver/quadtree/quadtree.cpp:520-801)
 * followed by skeletonize's skip of compressed boxes
 * (ie_solver/skel_factorization/skel_factorization.cpp:157): new node
 * data for the same n_points (same order, same problems).  A node whose
 * coordinates changed leaves its leaf and descends through the kept cells
 * to a new one, as the reference inserts added points; then, as there,
 * every leaf that received a node and holds more than the leaf capacity is
 * subdivided (:726-733), and on the old path of every moved node the first
 * box from the root that is not a leaf and holds fewer DOFs than the leaf
 * capacity becomes a leaf, its descendants removed (:689-712, :735-741).
 * Neighbour lists are re-formed where they can change (the reference
 * re-forms all of them, :743-801, with the same result).  Surviving boxes
 * keep their order within a level, new boxes follow; box ids are
 * renumbered (iegpu_update_tree_info maps them).  A node whose point,
 * normal, curvature or weight changed marks its old and its new leaf, and
 * every new box and every box whose leaf flag or ID row sources changed is
 * marked; each marked box marks itself, its neighbours, and the same for
 * every ancestor (mark_neighbors_and_parents, quadtree.cpp:435-461);
 * parents of marked boxes and the roots are marked too.  The next
 * iegpu_factor / iegpu_factor_solve_device with the tolerance and precision
 * of the stored factorization re-factors only the marked boxes and keeps
 * every other box's skeleton and factors (*tree_kept = 1).  A node outside
 * its root cell rebuilds the tree for the new nodes (*tree_kept = 0) and
 * the next factorization covers every box.  The stored factorization is
 * invalid until then.  Holes are kept and their border blocks rebuilt. */
int iegpu_update_geometry(iegpu_ctx* ctx, int n_points, const double* points,
                          const double* normals, const double* curvatures,
                          const double* weights, int* tree_kept);

/* iegpu_update_geometry for a batch (iegpu_set_geometry_batch): the
 * point_offsets (n_problems + 1) must equal the batch's, so that every
 * problem keeps its number of points; IEGPU_ERR_INVALID otherwise.  No
 * reference counterpart (one boundary per QuadTree). */
int iegpu_update_geometry_batch(iegpu_ctx* ctx, int n_problems, const int* point_offsets,
                                const double* points, const double* normals,
                                const double* curvatures, const double* weights,
                                int* tree_kept);

/* Drops a pending geometry update's marks: the next factorization covers
 * every box (a full re-factorization of the kept tree).  No reference
 * counterpart (the reference clears QuadTreeNode::compressed,
 * quadtree.cpp:841-961 via reset). */
int iegpu_discard_update(iegpu_ctx* ctx);

/* After iegpu_update_geometry: the number of marked boxes, whether the last
 * factorization re-factored only them (1) or every box (0), and per box id
 * the marks (dirty: n_boxes ints, may be null).  No reference counterpart
 * (the reference reads QuadTreeNode::compressed). */
int iegpu_update_info(const iegpu_ctx* ctx, int* n_dirty, int* last_incremental,
                      int* dirty);

/* After iegpu_update_geometry: the leaves it subdivided (n_split), the
 * boxes it turned into leaves (n_merged), and per box id the id the box
 * had before (old_id: n_boxes ints, -1 for a new box or after a rebuilt
 * tree; may be null).  No reference counterpart (the reference's
 * QuadTreeNode objects persist across perturb). */
int iegpu_update_tree_info(const iegpu_ctx* ctx, int* n_split, int* n_merged, int* old_id);

/* Build the quadtree (leaf boxes hold at most leaf_cap DOFs) and upload it.
 * Replaces QuadTree::initialize_tree, ie_solver/quadtree/quadtree.h:114-117
 * (quadtree.cpp:39-180; cap = MAX_LEAF_DOFS, quadtree.h:10). */
int iegpu_build_tree(iegpu_ctx* ctx, int leaf_cap);

/* Multiply connected domains: hole centres[2h] and radii[h]; builds the
 * U / Psi border matrices on the boundary points.  Replaces
 * initialize_U_mat / initialize_Psi_mat and the assignment to
 * quadtree->U / quadtree->Psi, ie_solver/linear_solve.cpp:20-112,256-257.
 * n_holes = 0 clears them. */
int iegpu_set_holes(iegpu_ctx* ctx, int n_holes, const double* centers,
                    const double* radii);

/* Recursive-skeletonization factorization with ID tolerance id_tol
 * (> 0) in IEGPU_FP64 or IEGPU_FP32.  Replaces
 * SkelFactorization::skeletonize, skel_factorization.h:55
 * (skel_factorization.cpp:140-310). */
int iegpu_factor(iegpu_ctx* ctx, double id_tol, int precision);

/* iegpu_factor followed by the solve x = A^{-1} b on device buffers (ndof
 * doubles each, distinct), ordered on the context's stream: each level's
 * upward sweep runs right after that level's decouple, beside the next
 * levels' work, and the root solve and downward sweeps after the root.
 * The stored factorization is the same as iegpu_factor's; x equals
 * iegpu_factor + iegpu_solve_device to rounding.  Not with holes.  No
 * reference counterpart (bie_time_trial times skeletonize + solve,
 * ie_solver/linear_solve.cpp:156-185). */
int iegpu_factor_solve_device(iegpu_ctx* ctx, double id_tol, int precision,
                              const double* d_b, double* d_x);

/* Solve with the stored factorization: mu (ndof) from b (ndof); with holes
 * alpha (hole_dim) is also returned (must be non-NULL).  Host pointers.
 * Replaces linear_solve, ie_solver/linear_solve.h:33-35 (solve /
 * multiply_connected_solve, skel_factorization.h:92-94). */
int iegpu_solve(iegpu_ctx* ctx, const double* b, double* mu, double* alpha);

/* As iegpu_solve for nrhs right-hand sides at once: b and mu are ndof x
 * nrhs column-major, alpha (with holes) hole_dim x nrhs.  Without holes
 * every column passes through each sweep kernel together; with holes the
 * bordered solve runs column by column.  Each column equals iegpu_solve of
 * that column.  Replaces solve / multiply_connected_solve on a
 * multi-column b (apply_sweep_matrix sweeps every column,
 * ie_solver/skel_factorization/skel_factorization.cpp:388-406,507-592). */
int iegpu_solve_nrhs(iegpu_ctx* ctx, int nrhs, const double* b, double* mu, double* alpha);

/* The border blocks of a multiply connected problem from the caller: U
 * (ndof x hole_dim, column-major) and Psi (hole_dim x ndof), replacing the
 * ones iegpu_set_holes built from the hole formulas (hole_dim must equal
 * theirs; the hole centres stay for interior evaluation).  A later
 * iegpu_set_holes or geometry update rebuilds them from the formulas.
 * Replaces the assignment to quadtree->U / quadtree->Psi that skeletonize
 * and multiply_connected_solve read, ie_solver/linear_solve.cpp:256-257. */
int iegpu_set_border(iegpu_ctx* ctx, int hole_dim, const double* U, const double* Psi);

/* One box's factors (creation id `box`) after a factorization: its rank k
 * and redundant count r (0 and 0 for uncompressed boxes and roots), and,
 * where the pointer is not NULL, T (k x r), L (k x r), U (r x k), X_rr^{-1}
 * (r x r) and the carried Schur block (k x k), column-major, rows and
 * columns in the order of iegpu_box_list's skeleton (3) and redundant (4)
 * lists.  Mirrors QuadTreeNode::{T, L, U, X_rr, schur_update},
 * ie_solver/quadtree/quadtree.h:55-66 (X_rr is kept inverted here). */
int iegpu_box_factors(iegpu_ctx* ctx, int box, int* k, int* r, double* T, double* L,
                      double* U, double* Xinv, double* S);

/* As iegpu_solve on device buffers (b may alias mu), ordered on `stream`
 * (NULL: the context's stream). */
int iegpu_solve_device(iegpu_ctx* ctx, const double* d_b, double* d_mu,
                       double* d_alpha, void* stream);

/* Iterative refinement in FP64 with the stored factorization (FP64 or
 * FP32) as the preconditioner M: x0 = M^{-1} b, then x += M^{-1}(b - A x)
 * until ||b - A x||_2 <= rtol ||b||_2 or max_iter corrections.  A is the
 * FP64 operator applied matrix-free on the GPU (bordered by U and Psi when
 * holes are set: [K U; Psi -I][x; alpha] = [b; 0]).  Host pointers; iters
 * and relres (nullable) report the corrections made and the final relative
 * residual.  No reference counterpart: the reference solves once with its
 * FP64 factorization (ie_solver/linear_solve.cpp:115-129); this is the
 * mixed-precision form of the same solve. */
int iegpu_solve_refined(iegpu_ctx* ctx, const double* b, double* x,
                        double* alpha, double rtol, int max_iter, int* iters,
                        double* relres);

/* y = K x (ndof) in FP64, matrix-free; with holes and non-NULL alpha, z,
 * also y += U alpha and z = Psi x - alpha.  The forward operator the
 * reference forms densely in its tests (ie_solver_test.cpp:63-93). */
int iegpu_apply(iegpu_ctx* ctx, const double* x, const double* alpha,
                double* y, double* z);

/* How iegpu_apply and iegpu_solve_refined apply the FP64 operator K:
 *   IEGPU_MATVEC_DIRECT  the O(N^2) sum over every source node;
 *   IEGPU_MATVEC_FMM     a fast multipole method: complex expansions of
 *                        `order` coefficients (32, 48 or 64; 0 = 48) on a
 *                        uniform quadtree whose non-empty leaves hold about
 *                        `leaf_points` nodes (0 = 64), the adjacent leaves
 *                        summed directly; O(N), ~1e-15 relative at order 48;
 *   IEGPU_MATVEC_AUTO    (default) the FMM for a single boundary of at
 *                        least IEGPU_FMM_AUTO_POINTS nodes, else direct.
 * A batch of problems always takes the direct sum.  The FMM's tree is built
 * from the geometry on first use (or by iegpu_build_tree) and rebuilt after
 * a geometry change.  No reference counterpart: the reference forms K
 * densely in its tests (ie_solver_test.cpp:63-93). */
#define IEGPU_MATVEC_AUTO 0
#define IEGPU_MATVEC_DIRECT 1
#define IEGPU_MATVEC_FMM 2
#define IEGPU_FMM_AUTO_POINTS 16384
int iegpu_set_matvec(iegpu_ctx* ctx, int method, int order, int leaf_points);

/* info[0..3] = operator form used by the last iegpu_apply /
 * iegpu_solve_refined (IEGPU_MATVEC_DIRECT or IEGPU_MATVEC_FMM, 0 before
 * any), the FMM tree's leaf level, box count and leaf count (0 without a
 * tree). */
int iegpu_matvec_info(const iegpu_ctx* ctx, int* info, int n);

/* Field at n_targets interior points: out (sol_dim per target) =
 * K_domain mu + U_forward alpha, matrix-free on the GPU; targets with
 * in_domain[t] == 0 give 0 (in_domain may be NULL: all inside); alpha may
 * be NULL without holes.  Host pointers.  Replaces
 * Initialization::InitializeDomainKernel and the products of schur_solve,
 * ie_solver/initialization.cpp:11-135, ie_solver/linear_solve.cpp:140-154. */
int iegpu_evaluate(iegpu_ctx* ctx, int n_targets, const double* targets,
                   const int* in_domain, const double* mu, const double* alpha,
                   double* out);

/* As iegpu_evaluate, allocating nothing: device buffers d_targets (2 per
 * target), d_in_domain (nullable), d_mu (ndof), d_alpha (hole_dim, nullable),
 * d_out (sol_dim per target), ordered on `stream` (NULL: the context's
 * stream), no synchronization.  The partial sums go to a static device
 * array of 2^19 doubles per GPU (4 MiB, held until the process ends):
 * eval_chunks * sol_dim * n_targets must fit it (eval_chunks =
 * ceil(n_points / 8192), at most 64; e.g. 16384 targets at 262144 Laplace
 * nodes), else IEGPU_ERR_CAPACITY.  Every call on a GPU shares that GPU's
 * array, so calls that overlap on one GPU (two streams, two contexts)
 * race.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (HIP graph capture, serving loops). */
int iegpu_evaluate_noalloc(iegpu_ctx* ctx, int n_targets, const double* d_targets,
                           const int* d_in_domain, const double* d_mu, const double* d_alpha,
                           double* d_out, void* stream);

/* Sizes: tree levels, boxes, DOFs, hole border dimension. */
int iegpu_tree_info(const iegpu_ctx* ctx, int* n_levels, int* n_boxes,
                    int* ndof, int* hole_dim);

/* Per box, indexed by creation order (QuadTreeNode::id order from the
 * root): level, parent id, leaf flag, children ids[4] (bl, tl, tr, br;
 * -1 none), and after a factorization |active|, rank (0: not compressed)
 * and ID matrix height.  Any pointer may be NULL.  Mirrors the per-node
 * fields of QuadTreeNode, ie_solver/quadtree/quadtree.h:51-82. */
int iegpu_box_info(iegpu_ctx* ctx, int* level, int* parent, int* is_leaf,
                   int* children, int* n_active, int* rank, int* m_rows);

/* One box list (creation id `box`): which = 0 original DOFs, 1 neighbour
 * ids, 2 active DOFs, 3 skeleton DOFs, 4 redundant DOFs; which = 5 lists
 * the box ids of tree level `box` in the reference's level order.
 * Returns the length (>= 0) and copies at most cap entries, or a negative
 * status.  Mirrors InteractionLists, ie_solver/quadtree/quadtree.h:16-49. */
int iegpu_box_list(iegpu_ctx* ctx, int box, int which, int* out, int cap);

/* Device milliseconds of the last factor / solve, split by stage:
 * t[0] factor total, t[1] solve total, t[2] root LU, t[3] bytes of
 * stored factors (as a double). */
int iegpu_last_timings(const iegpu_ctx* ctx, double* t, int nt);

/* Per-stage device timing with HIP events recorded around each launch on
 * the launching stream (stands in for the reference's omp_get_wtime
 * accumulators, skel_factorization.cpp:45-56,155-166).  on != 0 enables
 * and resets the counters.  Stages: 0 active lists, 1 ID row counts,
 * 2 slot offsets, 3 ID-matrix assembly, 4 randomized ID (sketches), 5 full
 * pivoted QR, 6 decouple, 7 root, 8 upward sweep, 9 root / bordered solve,
 * 10 downward sweep, 11 bordered factor.  level = -1 sums all levels. */
int iegpu_set_profiling(iegpu_ctx* ctx, int on);
int iegpu_profile(iegpu_ctx* ctx, int stage, int level, double* ms,
                  int* launches);

/* on != 0: iegpu_factor captures its launch sequence (every level's
 * kernels, the root) into a HIP graph and replays it on later calls while
 * the tolerance, precision, ID method and workspaces are unchanged (a
 * change, or a workspace growth, recaptures).  Not used while per-stage
 * profiling is on or with holes.  No reference counterpart. */
int iegpu_set_graphs(iegpu_ctx* ctx, int on);

/* on != 0: the next iegpu_factor also keeps X_rr of every compressed box
 * and each root block as formed before their inversion (one more factor
 * pool and ncap^2 doubles per root), so that iegpu_sparse_matvec can apply
 * the compressed operator.  Default off.  No reference counterpart (the
 * reference's tree keeps X_rr and allskel_mat, quadtree.h:62-66,104). */
int iegpu_set_keep_operator(iegpu_ctx* ctx, int on);

/* b = A x with the compressed operator of the stored factorization (the
 * skeletonization applied forward; no border with holes), host pointers of
 * ndof doubles.  Replaces SkelFactorization::sparse_matvec,
 * skel_factorization.h:89 (skel_factorization.cpp:429-504).  Needs a
 * factorization made with iegpu_set_keep_operator(ctx, 1). */
int iegpu_sparse_matvec(iegpu_ctx* ctx, const double* x, double* b);

/* on != 0 (default): iegpu_factor runs each level's decouple stage on a
 * second stream of the context, so that the ID stages of level l overlap
 * the decouple of level l+1 (the ID of a level needs only the skeletons
 * below it, not their Schur complements).  Results are identical either
 * way.  Per-stage profiling always uses one stream.  No reference
 * counterpart (the reference runs each level's boxes in one OpenMP loop,
 * skel_factorization.cpp:150-168). */
int iegpu_set_overlap(iegpu_ctx* ctx, int on);

/* Launch plan of the factorization.  After a full factorization the
 * context records, per tree level, which ID-stage kernels (general-QRCP
 * assembly and QRCP, register QRCP with its row compression, the n <= 128
 * register QRCP with its row compression, the 32-row and 64-row sketches)
 * had empty work lists; the next
 * factorization with the same inputs (geometry, tree, holes, border,
 * tolerance, precision, ID method, rank model, launch caps) does not
 * launch those.  The root kernel checks every skipped list on the device;
 * if one is not empty the factorization is repeated with every kernel (and
 * the plan recorded again), so the result never depends on the plan.
 * mode 1 (default): on; 0: off (every kernel on every factorization); 2:
 * the next factorization takes every list as empty (a test of the check;
 * it repeats in full), then back to 1.  Not used with captured graphs or
 * incremental re-factorizations.  No reference counterpart. */
int iegpu_set_launch_plan(iegpu_ctx* ctx, int mode);

/* info[0..2] = whether the last factorization ran with a launch plan (1)
 * or launched every kernel (0), the number of factorizations repeated
 * because a skipped list had work, and the number of skipped kernel
 * groups in the current plan (0 without one). */
int iegpu_launch_plan_info(const iegpu_ctx* ctx, int* info, int n);

/* Launch shapes of the factorization, caps[IEGPU_NUM_LAUNCH_CAPS]:
 * [0..5] workgroup caps of the kernels that walk a device-side box list
 * (the host knows a level's box count, not the list's length): the
 * general-QRCP assembly, the register-QRCP fallback of the randomized ID,
 * the 4- and 8-wave row compressions, the 8-wave register QRCP and the
 * general QRCP (default 1024 each); [6] the box count from which a level
 * decouples its small boxes in two LDS size classes (two launches, higher
 * occupancy for n <= 44) instead of one (default 8192); [7] the box count
 * from which a level decouples its compressed boxes with |active| <= 64 by
 * a register Gauss-Jordan sweep, one wave per box, instead of a
 * Gauss-Jordan inverse by a 256-thread workgroup per box followed by the
 * products (default 4096).  Each in [1, 2^20].  Launch shapes [0..6]
 * change no box's arithmetic; the two decouple forms agree to rounding.
 * No reference counterpart. */
#define IEGPU_NUM_LAUNCH_CAPS 8
int iegpu_set_launch_caps(iegpu_ctx* ctx, const int* caps, int n);
int iegpu_get_launch_caps(const iegpu_ctx* ctx, int* caps, int n);

/* Autotuner: coordinate descent over the launch caps, each candidate
 * (64, 256, 1024, 4096 for the grid caps; 1, 2048, 8192, 2^20 for the
 * decouple split) timed as the best of `reps` factorizations of the
 * current workload at (id_tol, precision) after one untimed run; a value
 * changes only for a gain of at least 1 %.  Leaves the chosen caps set and
 * the workload factored with them; writes the caps (may be NULL) and the
 * best factorization time in ms (may be NULL).  No reference counterpart. */
int iegpu_autotune(iegpu_ctx* ctx, double id_tol, int precision, int reps,
                   int* caps, double* ms);

/* Interpolative decomposition used by the next iegpu_factor.
 * IEGPU_ID_QRCP (default): the reference's full column-pivoted QR of each
 * box's ID matrix with the rank cut |R_ii| < tol |R_00|
 * (ie_solver/ie_mat/ie_mat.cpp:548-588).  IEGPU_ID_RANDOMIZED: boxes with
 * at most 64 active DOFs (and at most 256 near-field row DOFs) take the ID
 * of a Gaussian sketch Omega A of l = sketch_rows rows (16 or 32; 0: per
 * box from the rank model, iegpu_set_rank_model) with the
 * same cut applied to the sketch's R, |R_jj| scaled by 2 sqrt(l / (l - j))
 * (a sketch keeps l - j dimensions at step j); a box whose sketch shows no
 * cut within l - 6 pivots falls back to the full QRCP.  No reference
 * counterpart (the reference always factors the full ID matrix). */
enum { IEGPU_ID_QRCP = 0, IEGPU_ID_RANDOMIZED = 1 };
int iegpu_set_id_method(iegpu_ctx* ctx, int method, int sketch_rows);

/* After a factorization with IEGPU_ID_RANDOMIZED: boxes whose ID came from
 * the sketch, boxes that fell back to the full QRCP (their sketch wasted),
 * and the sketches' matrix-core work in flops, sum of 2 l m n16 over the
 * sketched and fallen-back boxes (l sketch rows, m ID rows, n16 = |active|
 * rounded up to 16).  Any pointer may be NULL. */
int iegpu_id_stats(iegpu_ctx* ctx, int* n_sketched, int* n_fallback, double* sketch_flops);

/* Rank model for the randomized ID (no reference counterpart).  An MLP
 * IEGPU_RANK_FEATURES -> 32 (tanh) -> 32 (tanh) -> 1 predicts each box's
 * ID rank from features the factorization computes on the GPU: log2
 * |active|, log2 ID rows, level, log2(near-field rows + 1), side x mean and
 * max |curvature|, 1 - |mean normal| of the active DOFs, -log10 id_tol,
 * Stokes flag.  The model is a float blob of IEGPU_RANK_MODEL_LEN: feature
 * means and scales, W1 [9][32], b1, W2 [32][32], b2, w3, b3 (the
 * IEGPU_RANK_PARAMS trained parameters), target mean and scale. */
#define IEGPU_RANK_FEATURES 9
#define IEGPU_RANK_MODEL_LEN 1429
#define IEGPU_RANK_PARAMS 1409

/* on != 0: the next factorizations record every box's features. */
int iegpu_record_features(iegpu_ctx* ctx, int on);
/* Features of every box (n_boxes x IEGPU_RANK_FEATURES, box creation
 * order) from the last factorization that recorded them; boxes without an
 * ID (m_rows 0 in iegpu_box_info) hold no meaningful row.  cap = floats
 * available in feat. */
int iegpu_box_features(iegpu_ctx* ctx, float* feat, int cap);
/* A fresh model for samples X (n x IEGPU_RANK_FEATURES) with ranks y:
 * normalization from the samples, weights uniform in +-sqrt(6/(in+out))
 * from seed.  Host only (ctx may be a host-only context). */
int iegpu_rank_model_init(iegpu_ctx* ctx, const float* X, const float* y, int n,
                          unsigned seed, float* model);
/* Trains model in place on the GPU: `steps` Adam steps (learning rate lr)
 * on batches of 64 samples drawn with replacement, loss = mean squared
 * error of the standardized rank; *loss (nullable) receives the mean loss
 * of the last 64 steps. */
int iegpu_rank_model_train(iegpu_ctx* ctx, const float* X, const float* y, int n, float* model,
                           int steps, float lr, unsigned seed, float* loss);
/* Loss over all n samples and its gradient (IEGPU_RANK_PARAMS floats, blob
 * order) computed by the GPU kernels the training uses. */
int iegpu_rank_model_grad(iegpu_ctx* ctx, const float* model, const float* X, const float* y,
                          int n, float* grad, float* loss);
/* Model used by iegpu_set_id_method(ctx, IEGPU_ID_RANDOMIZED, 0): a box's
 * sketch has 16 rows if (predicted rank + margin) <= 10, 32 rows if it is
 * <= 26, and none otherwise (the full QRCP takes the box directly). */
int iegpu_set_rank_model(iegpu_ctx* ctx, const float* model, int margin);

#ifdef __cplusplus
}
#endif

#endif /* IEGPU_H_ */
