/*
 * hodlr.h -- C ABI of the MI355X-native HODLR factor-and-solve library
 *            (libhodlr_amd.so, built from fastsolver2_amd/csrc).
 *
 * Drop-in boundary for the hot path of Charles-Chao-Chen/fastSolver2: each
 * entry point names the reference interface it replaces.  Conventions follow
 * the reference (src/ptr_matrix.cc:29-37, include/lapack_blas.hpp:28-57):
 *   - FP64, column-major, element (r,c) at ptr[r + c*ld];
 *   - the HODLR matrix of size N = leaf * 2^levels has dense leaf blocks K_i
 *     (K is N x leaf, leaf i = rows [i*leaf,(i+1)*leaf)) and every
 *     off-diagonal block at every level equal to U(I,:) V(J,:)^T with global
 *     N x rank matrices U, V (src/tree.cc:59, :189-197);
 *   - the solution overwrites the right-hand side ("b will be overwritten by
 *     x", src/lmatrix.cc:447-448); the plan entry points only READ the leaf
 *     blocks K (their LU factors are kept in the plan), while the
 *     reference-layout hsolve and the batched LAPACK entry points overwrite
 *     their blocks with the factors, as dgesv does.
 * All pointers passed to entry points taking a hipStream_t are DEVICE
 * pointers; work is enqueued on that stream and the call returns without
 * synchronizing unless stated.  Nothing here aborts: every entry point
 * returns a status (0 = success, < 0 = bad argument / runtime error,
 * > 0 = numerically singular block, see hodlr_plan_status()).
 *
 * Threading and streams.  A plan (and a batch) is stream-ordered and not
 * re-entrant: its entry points must not run concurrently with each other on
 * different streams or threads; distinct plans may run concurrently on
 * distinct streams.  Entry points that take no plan (matvec, gradients,
 * operators, batched LAPACK, hsolve) use library scratch kept per (device,
 * stream): calls on different streams never share it, calls on one stream
 * are ordered by that stream.  Calls on ONE stream from several host threads
 * must be serialized by the caller.  Scratch grows stream-ordered (no device
 * synchronization); it cannot grow while its stream is being captured
 * (HODLR_ERR_STATE: make one call outside capture first).
 */
#ifndef HODLR_H
#define HODLR_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define HODLR_OK                 0
#define HODLR_ERR_ARG           -1   /* an argument is out of range          */
#define HODLR_ERR_SHAPE         -2   /* N != leaf * 2^levels, leaf <= rank...  */
#define HODLR_ERR_UNSUPPORTED   -3   /* size beyond what the kernels handle   */
#define HODLR_ERR_HIP           -4   /* a HIP runtime call failed             */
#define HODLR_ERR_NOMEM         -5   /* device allocation failed              */
#define HODLR_ERR_STATE         -6   /* solve before factor, etc.             */
#define HODLR_SINGULAR_LEAF      1   /* zero pivot in a leaf LU               */
#define HODLR_SINGULAR_CORE      2   /* zero pivot in a 2r x 2r core LU       */

/* Human-readable text for a status code (static storage, never NULL). */
const char *hodlr_strerror(int status);
/* Detail of the last HODLR_ERR_HIP / HODLR_ERR_NOMEM on this thread. */
const char *hodlr_last_error_detail(void);
/* Releases the library scratch of `stream` on the current device
 * (stream-ordered free; for callers about to destroy the stream).  No
 * reference counterpart (the reference's tasks leak their scratch,
 * src/tasks/leaf_solve.cc:120-121). */
int hodlr_workspace_release(hipStream_t stream);

/* ---------------------------------------------------------------- plans */
enum hodlr_precision { HODLR_FP64 = 0, HODLR_MIXED_IR = 1 };

typedef struct hodlr_desc {
  int64_t n;          /* N = leaf << levels                                  */
  int     leaf;       /* leaf size m (reference -leaf, solver.cc:234)        */
  int     levels;     /* tree levels L (reference -mtxlvl)                   */
  int     rank;       /* off-diagonal rank r (reference -rank)               */
  int     nrhs_max;   /* largest number of right-hand sides solved at once   */
  int     precision;  /* enum hodlr_precision                                */
  int     device;     /* HIP device ordinal                                  */
  int     flags;      /* 0 or HODLR_PLAN_TRAILING_EMPTY                      */
} hodlr_desc;

/* hodlr_desc.flags: n = leaf * P with P <= 2^levels; the tree keeps 2^levels
 * leaf slots and the last 2^levels - P are empty (no rows).  A system of P
 * leaves, P not a power of two, then needs no padding rows. */
#define HODLR_PLAN_TRAILING_EMPTY 1

typedef struct hodlr_plan_s *hodlr_plan;

/* Replaces HMatrix::HMatrix(nProc, level) + UTree/VTree/KTree::init
 * (include/hmatrix.hpp:12-18, src/hmatrix.cc:22-47): validates the shape and
 * allocates all factor storage on the device (per-level Y_t = A_c^-1 U
 * blocks, leaf/core pivots, core LU factors).  The plan keeps pointers to
 * the K, U, V passed to hodlr_factor; they must stay allocated until the
 * plan is destroyed or factored again. */
int hodlr_plan_create(hodlr_plan *plan, const hodlr_desc *desc);
/* A tree of uneven leaves: 2^levels leaf slots of sizes leaf_sizes[i]
 * (each 0 -- an empty slot -- or above the rank), desc->n their sum
 * (desc->leaf and desc->flags are ignored / must be 0).  Leaf i covers rows
 * [sum_{j<i} leaf_sizes[j], ... + leaf_sizes[i]), its block K_i sits in
 * columns [0, leaf_sizes[i]) of those rows of K (N x max leaf), and the
 * node of leaf range [a, b) has children [a, mid) and [mid, b) of the slot
 * tree, however many rows they hold.  This is the shape of the reference's
 * irregular drivers: shards of 2^k leaves each with their own leaf size
 * (spmd_driver/solver_irregular_part.cc:73-75, heavy shards of 10x leaves),
 * joined by the top levels.  Every entry point taking a plan accepts such
 * a plan; the per-leaf launches run once per run of equal leaf sizes. */
int hodlr_plan_create_uneven(hodlr_plan *plan, const hodlr_desc *desc, const int *leaf_sizes);
/* Replaces HMatrix::destroy (include/hmatrix.hpp:23, declared but never
 * defined in the reference) / LMatrix::clear (src/lmatrix.cc:886-890). */
int hodlr_plan_destroy(hodlr_plan plan);
/* Device bytes the plan owns (factor storage + workspace). */
int64_t hodlr_plan_device_bytes(hodlr_plan plan);

/* Factorization only (leaf LUs, the per-level Y_t blocks, core LUs).
 * K (N x leaf, ld ldk) is read, not modified: the leaf LU factors (FP64, or
 * FP32 for a HODLR_MIXED_IR plan) are stored in the plan, so the same K can
 * serve the refinement residuals.  Leaf blocks whose partial-pivoting LU
 * keeps every diagonal pivot (|multiplier| <= 1 - 2^-20 at every step) are
 * factored without interchanges on a fast path; the others with row
 * interchanges, as dgetrf (the pivots equal dgetrf's either way). */
int hodlr_factor(hodlr_plan plan, const double *K, int64_t ldk,
                 const double *U, int64_t ldu, const double *V, int64_t ldv,
                 hipStream_t stream);

/* Solve with a stored factorization: B (N x nrhs, ld ldb) <- A^-1 B.
 * The factor-reuse form of HMatrix::solve (src/hmatrix.cc:49-96), with
 * nrhs columns in place of the hard-coded single RHS (src/tree.cc:9). */
int hodlr_solve(hodlr_plan plan, double *B, int64_t ldb, int nrhs, hipStream_t stream);

/* Fused factor + solve in one pass: the reference's HMatrix::solve /
 * spmd driver level loop (spmd_driver/solver_irregular_part.cc:104-140),
 * which factors and solves together.  Leaves the factorization in the plan
 * for later hodlr_solve calls. */
int hodlr_factor_solve(hodlr_plan plan, const double *K, int64_t ldk,
                       const double *U, int64_t ldu, const double *V, int64_t ldv,
                       double *B, int64_t ldb, int nrhs, hipStream_t stream);

/* Iterative refinement in FP64 (the HODLR_MIXED_IR plans' second half; also
 * valid for an FP64 plan).  X (N x nrhs) <- A^-1 B refined:
 *   X = solve(B); repeat { R = B - A X (FP64 HODLR matvec with the ORIGINAL
 *   K, U, V); stop if max_j |R_j|_inf / |B_j|_inf <= tol; X += solve(R) }
 * with solve() the plan's stored factorization (FP32 leaf LU for a
 * HODLR_MIXED_IR plan).  At most
 * maxit corrections.  Synchronizes `stream` once per iteration.  Returns
 * HODLR_OK whether or not tol was reached (see stats->converged).  No
 * reference counterpart (the reference solves in FP64 only). */
typedef struct hodlr_ir_stats {
  int    iterations;    /* corrections applied                              */
  int    converged;     /* 1 if the residual reached tol                    */
  double residual;      /* final max_j |b_j - A x_j|_inf / |b_j|_inf        */
  double history[33];   /* residual after 0, 1, ... corrections            */
} hodlr_ir_stats;
int hodlr_solve_ir(hodlr_plan plan, const double *K, int64_t ldk, const double *U, int64_t ldu,
                   const double *V, int64_t ldv, const double *B, int64_t ldb, double *X,
                   int64_t ldx, int nrhs, double tol, int maxit, int flags,
                   hodlr_ir_stats *stats, hipStream_t stream);
/* flags for hodlr_solve_ir: X already holds solve(B) (for instance from
 * hodlr_factor_solve on a copy of B), so the first solve is skipped. */
#define HODLR_IR_X_GIVEN 1

/* Adjoint (transposed) solve with the stored factorization, in place:
 * G <- A^-T G, nrhs columns (any number; processed in chunks).  The
 * operator is the transpose of hodlr_solve's, applied top-down: per level
 * t = 0..L-1 and node, y = S_t^-T [u_c0^T g_c0; u_c1^T g_c1], then
 * g_c0 -= V_c0 y_1, g_c1 -= V_c1 y_0; finally g <- K^-T g per leaf.
 * No reference counterpart (the reference has no gradient path); it
 * transposes the reference's level loop (spmd_driver/solver_irregular_part.cc:104-140). */
int hodlr_solve_adjoint(hodlr_plan plan, double *G, int64_t ldg, int nrhs, hipStream_t stream);

/* Gradients of a scalar loss l(X), X = A^-1 B, with respect to the leaf
 * blocks, U and V, given X and Lambda = A^-T dl/dX (hodlr_solve_adjoint):
 *   dK_i = -Lambda_i X_i^T                          (leaf x leaf, rows of leaf i)
 *   dU_i = -Lambda_i (V^T X - V_i^T X_i)^T          (leaf x rank)
 *   dV_i = -X_i (U^T Lambda - U_i^T Lambda_i)^T     (leaf x rank)
 * (dl/dB = Lambda).  dK is N x leaf (ld lddk), dU, dV are N x rank.  Any of
 * dK, dU, dV may be NULL to skip it.  No reference counterpart. */
int hodlr_grad(int64_t n, int leaf, int rank, int nrhs, const double *U, int64_t ldu,
               const double *V, int64_t ldv, const double *X, int64_t ldx,
               const double *Lambda, int64_t ldl, double *dK, int64_t lddk,
               double *dU, int64_t lddu, double *dV, int64_t lddv, hipStream_t stream);

/* ------------------------------------------------------------- batches */
/* Many independent HODLR systems (uneven N, leaf size, rank, nrhs), one
 * plan each.  hodlr_batch_factor_solve enqueues every system's fused factor
 * + solve (hodlr_factor_solve semantics, per system i: K[i] only read, its
 * leaf LU kept in system i's plan, B[i] overwritten by the solution; a desc
 * may carry HODLR_PLAN_TRAILING_EMPTY for P leaves with P not a power of
 * two).  On a capturable stream (not the
 * legacy null stream) the whole batch is captured once as ONE HIP graph --
 * the stream forks into `lanes` internal streams that run systems
 * concurrently and joins back -- and calls with the same operand pointers
 * replay it.  No reference counterpart (the reference solves one system
 * per run); it extends the driver of spmd_driver/solver_irregular_part.cc
 * to independent systems. */
typedef struct hodlr_batch_s *hodlr_batch;
int hodlr_batch_create(hodlr_batch *batch, int nsys, const hodlr_desc *descs);
int hodlr_batch_destroy(hodlr_batch batch);
int hodlr_batch_size(hodlr_batch batch);
/* The plan of system i (owned by the batch; for hodlr_solve and friends). */
hodlr_plan hodlr_batch_plan(hodlr_batch batch, int i);
int hodlr_batch_factor_solve(hodlr_batch batch, double *const *K, const int64_t *ldk,
                             const double *const *U, const int64_t *ldu,
                             const double *const *V, const int64_t *ldv,
                             double *const *B, const int64_t *ldb, const int *nrhs,
                             hipStream_t stream);
/* Synchronizes; first failing system in *sys and its hodlr_plan_status
 * (*where, *level).  Any of sys, where, level may be NULL. */
int hodlr_batch_status(hodlr_batch batch, hipStream_t stream, int *sys, int *where,
                       int *level);
/* Launch shape knobs: lanes (1..16 concurrent streams) and the leaf-solve
 * column tile (0 = by width, else 1, 2 or 5 MFMA tiles per workgroup). */
int hodlr_batch_set_tuning(hodlr_batch batch, int lanes, int tile);
int hodlr_batch_get_tuning(hodlr_batch batch, int *lanes, int *tile);
/* Autotuner: for every (lanes, tile) pair, captures the batch, replays it
 * once untimed (a graph's first launch uploads it), times the faster of two
 * replays, each started on an idle stream, and keeps the fastest pair
 * (*best_ms).  Every B is overwritten many times: pass scratch copies. */
int hodlr_batch_tune(hodlr_batch batch, double *const *K, const int64_t *ldk,
                     const double *const *U, const int64_t *ldu,
                     const double *const *V, const int64_t *ldv,
                     double *const *B, const int64_t *ldb, const int *nrhs,
                     hipStream_t stream, double *best_ms);

/* -------------------------------------------------------------- pairs */
/* Two sibling shards with different off-diagonal ranks, the system of the
 * reference's spmd_driver/solver_irregular_ranks.cc (rank0 != rank1):
 *   A = [A_0, U_0 W_1^T; U_1 W_0^T, A_1]
 * with A_c the HODLR matrix of shard c (its own leaves, levels and rank
 * r_c; operands K_c, U_c, V_c as for a plan), U_c (n_c x r_c) the same U
 * its own levels use, and W_c (n_c x r_{1-c}) its coupling V of the
 * sibling's rank (the driver's VMatp, :113-117).  The root core is
 * (r0+r1) x (r0+r1), solved as NodeSolveRegionTask (hodlr_node_solve_region).
 * The shards are described as for hodlr_plan_create / _uneven
 * (leaf_sizes_c NULL for a balanced shard); nrhs_max as for a plan.  No
 * mixed precision.  Stream-ordered and not re-entrant, as a plan. */
typedef struct hodlr_pair_s *hodlr_pair;
int hodlr_pair_create(hodlr_pair *pair, const hodlr_desc *shard0, const int *leaf_sizes0,
                      const hodlr_desc *shard1, const int *leaf_sizes1, int nrhs_max);
int hodlr_pair_destroy(hodlr_pair pair);
/* The plan of shard 0 or 1 (owned by the pair; it holds A_c's factors). */
hodlr_plan hodlr_pair_plan(hodlr_pair pair, int shard);
/* Factor and solve: K, U, V, W are two-element arrays (shard 0, shard 1) of
 * device pointers with their leading dimensions; B ((n_0 + n_1) x nrhs, ld
 * ldb, shard 0's rows first) is overwritten by the solution.  The driver's
 * order (solver_irregular_ranks.cc:146-251): each shard solves [B_c | U_c],
 * projects W_c^T [d_c | Y_c], the root core is solved, d_c -= Y_c eta_c. */
int hodlr_pair_factor_solve(hodlr_pair pair, const double *const *K, const int64_t *ldk,
                            const double *const *U, const int64_t *ldu,
                            const double *const *V, const int64_t *ldv,
                            const double *const *W, const int64_t *ldw,
                            double *B, int64_t ldb, int nrhs, hipStream_t stream);
/* Synchronizes; HODLR_OK, or the first failing shard's hodlr_plan_status
 * (*shard = 0 or 1), or HODLR_SINGULAR_CORE for the root core (*shard = -1,
 * *where = 0). */
int hodlr_pair_status(hodlr_pair pair, hipStream_t stream, int *shard, int *where, int *level);

/* Synchronizes `stream` and reports the numerical status of the last factor:
 * HODLR_OK, HODLR_SINGULAR_LEAF (*where = leaf index, *level = levels) or
 * HODLR_SINGULAR_CORE (*where = node index within tree level *level, the
 * root being level 0).  `where` and `level` may be NULL; both are -1 on
 * HODLR_OK.  Replaces the reference's `assert(INFO == 0)`
 * (src/tasks/leaf_solve.cc:90,141). */
int hodlr_plan_status(hodlr_plan plan, hipStream_t stream, int *where, int *level);

/* FP64 HODLR matrix-vector product with the ORIGINAL K (not its LU):
 * Y = alpha * A X + beta * Y, per leaf y_i = K_i x_i + U_i (V^T x - V_i^T x_i).
 * No reference counterpart (the reference has no residual check:
 * tests/test.cc:651-658 is `#if 0`); used for residuals and refinement. */
int hodlr_matvec(int64_t n, int leaf, int rank, int nrhs,
                 const double *K, int64_t ldk, const double *U, int64_t ldu,
                 const double *V, int64_t ldv, const double *X, int64_t ldx,
                 double alpha, double beta, double *Y, int64_t ldy,
                 hipStream_t stream);
/* hodlr_matvec without any device allocation, for callers that cannot
 * call hipMalloc (code under graph capture, tight serving loops).  No
 * reference counterpart.  Same arguments and result as hodlr_matvec; its
 * scratch (the per-leaf V_i^T x partials, (N/leaf + 1) * rank * nrhs
 * doubles) is a fixed device array of 2^20 doubles (8 MiB) that every GPU
 * the library runs on holds until the process ends; larger problems are
 * rejected with HODLR_ERR_UNSUPPORTED before anything is enqueued.  All
 * calls on one GPU share that array: calls that overlap in time on one GPU
 * (different streams, or threads) race and must be serialized by the
 * caller.  No scratch content needs to start at zero. */
int hodlr_matvec_noalloc(int64_t n, int leaf, int rank, int nrhs,
                         const double *K, int64_t ldk, const double *U, int64_t ldu,
                         const double *V, int64_t ldv, const double *X, int64_t ldx,
                         double alpha, double beta, double *Y, int64_t ldy,
                         hipStream_t stream);
/* hodlr_matvec for the tree of a plan (any leaf sizes, see
 * hodlr_plan_create_uneven), on the given operands (not the plan's
 * factors); trans != 0 applies A^T.  No reference counterpart. */
int hodlr_plan_matvec(hodlr_plan plan, int trans, int nrhs, const double *K, int64_t ldk,
                      const double *U, int64_t ldu, const double *V, int64_t ldv,
                      const double *X, int64_t ldx, double alpha, double beta, double *Y,
                      int64_t ldy, hipStream_t stream);
/* The same with A^T: y_i = K_i^T x_i + V_i (U^T x - U_i^T x_i). */
int hodlr_matvec_t(int64_t n, int leaf, int rank, int nrhs,
                   const double *K, int64_t ldk, const double *U, int64_t ldu,
                   const double *V, int64_t ldv, const double *X, int64_t ldx,
                   double alpha, double beta, double *Y, int64_t ldy,
                   hipStream_t stream);

/* ------------------------------------------- reference-layout operators */
/* GPU form of hsolve (src/tasks/leaf_solve.cc:8-10, :73-154), same argument
 * meaning and in-place semantics: X is nrow x (nrhs + rank*log2(nPart))
 * with LD rows, columns [0,nrhs) the right-hand sides d and the following
 * rank-wide blocks the per-level u copies; K (nrow x nrow/nPart, LD) holds
 * the leaf blocks and is overwritten by their LU; V is nrow x rank.
 * Synchronizes `stream` and returns INFO (0 or > 0 singular). */
int hodlr_hsolve_ref(int nrow, int nrhs, int rank, int nPart, int LD,
                     double *K, double *X, const double *V, hipStream_t stream);

/* LMatrix::gemmRed (src/lmatrix.cc:713-761; gemm_reduce.cc:58-76):
 * C[slot] = alpha * sum_{p in slot} A_p^T B_p over nchunk equal row chunks of
 * A (rows x ka) and B (rows x kb), slot = p / (nchunk / nslot); C is
 * (nslot*ka) x kb (ld ldc), slots stacked by rows (beta = 0 as asserted at
 * lmatrix.cc:718). */
int hodlr_gemm_red(int64_t rows, int nchunk, int nslot, double alpha,
                   const double *A, int64_t lda, int ka,
                   const double *B, int64_t ldb, int kb,
                   double *C, int64_t ldc, hipStream_t stream);

/* LMatrix::gemmBro (src/lmatrix.cc:814-865; gemm_broadcast.cc:58-82):
 * D_p += alpha * A_p * E[slot(p)] for row chunks p of A (rows x k) and
 * D (rows x nd); E stacked by slots of k rows (ld lde).  beta = 1 (asserted
 * at lmatrix.cc:820). */
int hodlr_gemm_bro(int64_t rows, int nchunk, int nslot, double alpha,
                   const double *A, int64_t lda, int k,
                   const double *E, int64_t lde, double *D, int64_t ldd, int nd,
                   hipStream_t stream);

/* LMatrix::node_solve (src/lmatrix.cc:534-581; node_solve.cc:60-81): for
 * each of nnode node slots of 2r rows, S = [I, VTu(r:2r,:); VTu(0:r,:), I],
 * the RHS halves of VTd swapped, VTd <- S^-1 VTd.  `info` is a device array
 * of nnode ints receiving each core's LAPACK INFO (0, or the 1-based column
 * of the first zero pivot) -- the reference asserts INFO == 0
 * (node_solve.cc:78).  Asynchronous on `stream`. */
int hodlr_node_solve(int nnode, int r, const double *VTu, int64_t ldu,
                     double *VTd, int64_t ldd, int nrhs, int *info, hipStream_t stream);

/* Static LMatrix::node_solve(VTu0, VTu1, VTd0, VTd1, ...) /
 * NodeSolveRegionTask (src/lmatrix.cc:583-636, src/tasks/node_solve_region.cc:
 * 35-108): the core solve of siblings with different ranks r0 != r1
 * (spmd_driver/solver_irregular_ranks.cc:364-365).  VTu0 is r1 x r0,
 * VTu1 r0 x r1, VTd0 r1 x nrhs, VTd1 r0 x nrhs; per node
 *   S = I_{r0+r1},  S(i, r0+j) = VTu1(i, j),  S(r0+j, i) = VTu0(j, i),
 *   B = [VTd1; VTd0],  B <- S^-1 B (partial pivoting, dgesv),
 *   VTd1 <- B(0:r0, :),  VTd0 <- B(r0:r0+r1, :).
 * Batched over nnode nodes: node k's operands start at element offsets
 * k*su0, k*su1, k*sd0, k*sd1 (the reference solves one node per task).
 * `info` is a device array of nnode ints (LAPACK INFO of each core).
 * Asynchronous on `stream`. */
int hodlr_node_solve_region(int nnode, int r0, int r1, int nrhs,
                            const double *VTu0, int64_t ldu0, int64_t su0,
                            const double *VTu1, int64_t ldu1, int64_t su1,
                            double *VTd0, int64_t ldd0, int64_t sd0,
                            double *VTd1, int64_t ldd1, int64_t sd1,
                            int *info, hipStream_t stream);

/* ------------------------------------------------------ batched LAPACK */
/* Batched LU with partial pivoting (dgetrf semantics, dgetrf.f) of `batch`
 * n x n blocks A_b = A + b*bstride (ld lda) in place.  ipiv (batch x n,
 * 1-based, LAPACK IPIV) and info (batch) are device arrays.  Replaces the
 * dgetrf_ calls behind dgesv_ (include/lapack_blas.hpp:45-57). */
int hodlr_dgetrf_batched(int n, double *A, int64_t lda, int64_t bstride, int batch,
                         int *ipiv, int *info, hipStream_t stream);
/* Out-of-place batched dgetrf (dgetrf.f): LU_b (n x n, ld ldl, LU + b*lbs)
 * <- the LU factors with partial pivoting of K_b (ld ldk, K + b*kbs), K
 * not modified.  ipiv (batch x n, 1-based) and info (batch) are device
 * arrays with dgetrf's meaning.  The leaf factorization of the plans:
 * blocks whose pivots are all diagonal take the no-interchange fast path
 * (n = 128, 256, 384 or 512), the others the pivoted kernels.  Replaces the
 * dgetrf_ call behind dgesv_ (include/lapack_blas.hpp:45-57) for callers
 * that keep their matrix. */
int hodlr_dgetrf_batched_oop(int n, const double *K, int64_t ldk, int64_t kbs, double *LU,
                             int64_t ldl, int64_t lbs, int batch, int *ipiv, int *info,
                             hipStream_t stream);
/* Batched dgesv_ (include/lapack_blas.hpp:45-50): A_b overwritten by LU,
 * B_b = B + b*bbstride (n x nrhs, ld ldb) by the solution.  ipiv (batch x n,
 * 1-based) and info (batch ints, LAPACK INFO per item) are device arrays.
 * Asynchronous on `stream`, like hodlr_dgetrf_batched. */
int hodlr_dgesv_batched(int n, int nrhs, double *A, int64_t lda, int64_t bstride,
                        int batch, int *ipiv, int *info, double *B, int64_t ldb,
                        int64_t bbstride, hipStream_t stream);

/* ------------------------------------------------ workload generators */
/* Device Philox4x64-10 generator: element e of (seed, stream, system) is
 * word e%4 of philox(ctr={e/4, stream, system, 0}, key={seed, 0}).
 * kind 0: uniform (w>>11)*2^-53 * scale; kind 1: Box-Muller normal * scale.
 * Writes an rows x cols column-major block (ld) with element index
 * e = row + col*rows. */
int hodlr_generate(uint64_t seed, uint32_t stream_id, uint32_t system, int kind,
                   int64_t rows, int cols, double scale, double *out, int64_t ld,
                   hipStream_t stream);
/* The reference's generators on the device: Matrix::rand / Vector::rand
 * (src/matrix.cc:250-297, :79-102) with PtrMatrix::rand (src/ptr_matrix.cc:
 * 62-71).  Partition k's seed is the k-th lrand48_r of srand48_r(master)
 * (master standing for the reference's time(NULL)); rows [k*rows_per_part,
 * (k+1)*rows_per_part) of out (npart*rows_per_part x cols, ld) are that
 * seed's drand48_r stream visited rows outer, columns inner, plus offset --
 * bit for bit glibc's values.  With cols = 1 it is Vector::rand(offset).
 * Together with hodlr_dense_blocks it builds the reference-mode K, U, V, D
 * without host data (the reference's has_entry = false leaf generation,
 * src/tasks/dense_block.cc:54-75). */
int hodlr_generate_ref(long master, int npart, int64_t rows_per_part, int cols, double offset,
                       double *out, int64_t ld, hipStream_t stream);
/* DenseBlockTask::cpu_task + LMatrix::init_dense_blocks (src/tasks/
 * dense_block.cc:54-75, src/lmatrix.cc:273-297): K_i = U_i V_i^T +
 * diag(D_i + offset) for every leaf i of K (N x leaf); D has N entries.
 * Products and sums unfused, in dgemm's order: the reference's K bit for
 * bit. */
int hodlr_dense_blocks(int64_t n, int leaf, int rank, const double *U, int64_t ldu,
                       const double *V, int64_t ldv, const double *D, double offset,
                       double *K, int64_t ldk, hipStream_t stream);
/* Adds `shift` to the diagonal of every leaf block of K (N x leaf). */
int hodlr_shift_leaf_diagonal(int64_t n, int leaf, double shift, double *K,
                              int64_t ldk, hipStream_t stream);

/* ---------------------------------------------------- kernel profiling */
/* Kernel classes timed by the profiler. */
enum hodlr_prof_kind {
  HODLR_PROF_GETRF_LEAF = 0, /* batched LU of the leaf blocks              */
  HODLR_PROF_GETRS_LEAF = 1, /* leaf solves K^-1 [U | B]                   */
  HODLR_PROF_TRTRI = 2,      /* inverses of 32x32 diagonal blocks          */
  HODLR_PROF_REDUCE = 3,     /* level reduce V^T [Y | B] (gemmRed)         */
  HODLR_PROF_ASSEMBLE = 4,   /* core assembly S = [I, G1; G0, I]           */
  HODLR_PROF_GETRF_CORE = 5, /* batched LU of the 2r x 2r cores            */
  HODLR_PROF_GETRS_CORE = 6, /* core solves                                */
  HODLR_PROF_UPDATE = 7,     /* level update [Y | B] -= Y eta (gemmBro)    */
  HODLR_PROF_LEAF_LEVEL = 8, /* per-leaf r x r level step (W, G, projections) */
  HODLR_PROF_NKINDS = 9
};
/* No reference counterpart (the reference profiles with Legion Prof,
 * tests/Makefile:119-212).  While enabled, the plan entry points record HIP
 * events around every launch on its stream, with the launch's algorithmic
 * FLOPs and minimum HBM bytes. */
int hodlr_profile_enable(int on);
/* Kinds recorded while enabled: bit k for kind k (default all).  Each
 * recorded launch costs two event records on its stream, so a timing run
 * can keep the events to the kinds it reports. */
int hodlr_profile_mask(unsigned mask);
int hodlr_profile_reset(void);
/* Synchronizes on the recorded events; totals over all launches of `kind`. */
int hodlr_profile_read(int kind, int64_t *count, double *total_ms, double *flops,
                       double *bytes);

/* Device self-test of the FP64 matrix-core lane layout the kernels assume:
 * writes 256 doubles (64 lanes x 4 accumulator elements) to device `out`;
 * element 4*l + v must equal row + 100*(col % 4) with row = l/16 + 4v,
 * col = l%16.  No reference counterpart. */
int hodlr_selftest_mfma(double *out, hipStream_t stream);
/* Same probe for the FP32 matrix core (v_mfma_f32_16x16x4f32) used by the
 * mixed-precision leaf kernels: 256 values (as doubles), element 4*l + v
 * equal to row + 100*(col % 4) with row = 4*(l/16) + v, col = l%16.  No
 * reference counterpart. */
int hodlr_selftest_mfma_f32(double *out, hipStream_t stream);

/* Library version string. */
const char *hodlr_version(void);

#ifdef __cplusplus
}
#endif
#endif /* HODLR_H */
