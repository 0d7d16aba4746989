/*
 * hodlr_oracle.h -- CPU oracle for the HODLR factor-and-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * never by the product library (fastsolver2_amd/).  It restates, in plain C,
 * the reference's CPU algorithm (Charles-Chao-Chen/fastSolver2) and the
 * netlib LAPACK routines that algorithm calls, so that GPU results can be
 * compared with "the reference's own CPU solve" on identical inputs.
 *
 * Storage convention (as the reference, src/ptr_matrix.cc:29-37): FP64,
 * column-major, element (r,c) at ptr[r + c*ld].  All row/column indices are
 * 0-based here; pivot vectors are 1-based as LAPACK's IPIV.
 */
#ifndef HODLR_ORACLE_H
#define HODLR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- POSIX drand48 family (restated LCG; src/ptr_matrix.cc:62-71 uses
 *      srand48_r / drand48_r, src/matrix.cc:250-297 lrand48_r) ------------ */
typedef struct { uint64_t x; } orc_rand48;
void   orc_srand48(orc_rand48 *s, long seed);
double orc_drand48(orc_rand48 *s);
long   orc_lrand48(orc_rand48 *s);

/* Per-partition seeds: srand48(master) then one lrand48 per partition
 * (matrix.cc:250-264, where the master seed is time(NULL)+lrand48()). */
void orc_ref_seeds(long master, int npart, long *seeds);
/* PtrMatrix::rand(seed, offset): rows outer, columns inner (ptr_matrix.cc:62-71). */
void orc_ptr_rand(long seed, int offset, int rows, int cols, int ld, double *a);
/* Matrix::rand() for an (nPart*rows_per_part) x cols matrix, ld = total rows. */
void orc_ref_matrix(long master, int npart, int rows_per_part, int cols, double *a);
/* Vector::rand(offset) (matrix.cc:79-102): per-partition sequential fill + offset. */
void orc_ref_vector(long master, int npart, int rows_per_part, int offset, double *a);
/* DenseBlockTask (dense_block.cc:54-75): K_i = U_i V_i^T + diag(D_i + offset),
 * the generator regenerating U_i, V_i, D_i from per-leaf seeds. K is N x m, ld N. */
void orc_ref_dense_blocks(int npart, int m, int rank, int offset,
                          const long *useeds, const long *vseeds, const long *dseeds,
                          double *K, int64_t ldk);

/* ---- Philox4x64-10 workload generator (Salmon et al., SC'11) ---------- */
void orc_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);
/* Element e of stream (seed, stream, system): block e/4, word e%4 of
 * philox(ctr = {e/4, stream, system, 0}, key = {seed, 0}).
 * uniform: (w >> 11) * 2^-53.  normal: Box-Muller on word pairs (0,1),(2,3). */
void orc_gen_uniform(uint64_t seed, uint32_t stream, uint32_t system,
                     int64_t n, double scale, double *out);
void orc_gen_normal(uint64_t seed, uint32_t stream, uint32_t system,
                    int64_t n, double scale, double *out);

/* ---- netlib LAPACK/BLAS restatements (the netlib_blas Fortran files) ---------------- */
int  orc_idamax(int n, const double *x, int incx);               /* 1-based */
void orc_dgemm(char ta, char tb, int m, int n, int k, double alpha,
               const double *A, int lda, const double *B, int ldb,
               double beta, double *C, int ldc);
void orc_dtrsm_left(char uplo, char trans, char diag, int m, int n,
                    const double *A, int lda, double *B, int ldb);
void orc_dlaswp(int n, double *A, int lda, int k1, int k2, const int *ipiv, int incx);
int  orc_dgetrf2(int m, int n, double *A, int lda, int *ipiv);
int  orc_dgetrf(int m, int n, double *A, int lda, int *ipiv);
int  orc_dgetrs(char trans, int n, int nrhs, const double *A, int lda,
                const int *ipiv, double *B, int ldb);
int  orc_dgesv(int n, int nrhs, double *A, int lda, int *ipiv, double *B, int ldb);

/* ---- BLAS backend used by the HODLR restatement ------------------------
 * Default: the restated netlib routines above.  orc_blas_load() switches
 * dgemm_/dgesv_ to a Fortran-ABI BLAS/LAPACK found at run time (for the CPU
 * baseline; the reference links the system -llapack -lblas,
 * spmd_driver/Makefile:52).  Returns 0 on success. */
int  orc_blas_load(const char *path, const char *symbol_prefix);
void orc_blas_use_netlib(void);
const char *orc_blas_name(void);

/* ---- reference algorithm ---------------------------------------------- */
/* hsolve (src/tasks/leaf_solve.cc:73-154): recursive fused factor+solve of
 * one subtree, in place on X (nrow x (nrhs + rank*log2(nPart))).  Returns
 * 0, or >0 for a singular leaf/core (LAPACK INFO), <0 for bad arguments. */
int orc_hsolve(int nrow, int nrhs, int rank, int nPart, int LD,
               double *K, double *X, const double *V);

/* gemmRed (gemm_reduce.cc:58-76, lmatrix.cc:713-761): for row chunk p of
 * nchunk equal chunks of A (rows x ka) and B (rows x kb):
 *   C[slot(p)] += alpha * A_p^T B_p,  slot = p / (nchunk / nslot),
 * C stacked by slots of ka rows (ld ldc), zeroed first (beta = 0). */
void orc_gemm_red(int rows, int nchunk, int nslot, double alpha,
                  const double *A, int lda, int ka,
                  const double *B, int ldb, int kb,
                  double *C, int ldc);
/* gemmBro (gemm_broadcast.cc:58-82): d_p += alpha * u_p * eta[slot(p)],
 * eta stacked by slots of k rows (ld lde). */
void orc_gemm_bro(int rows, int nchunk, int nslot, double alpha,
                  const double *u, int ldu, int k,
                  const double *eta, int lde, double *d, int ldd, int nd);
/* node_solve (node_solve.cc:60-81): for each of nnode node slots (2r rows):
 * S = [I, VTu(r:2r,:); VTu(0:r,:), I]; swap the RHS halves; dgesv.
 * VTd (2r*nnode x nrhs, ld ldd) is overwritten by eta. */
int orc_node_solve(int nnode, int r, const double *VTu, int ldu,
                   double *VTd, int ldd, int nrhs);
/* NodeSolveRegionTask (node_solve_region.cc:77-107): ranks r0 != r1. */
int orc_node_solve_region(int r0, int r1, int nrhs,
                          const double *VTu0, int ld0, const double *VTu1, int ld1,
                          double *VTd0, int ldd0, double *VTd1, int ldd1);

/* Whole driver (spmd_driver/solver_irregular_part.cc:104-140): leaf tasks at
 * task level T (2^T tasks of 2^(L-T) leaves, each running hsolve), then the
 * level loop gemmRed x2 -> node_solve -> gemmBro for tree levels T-1..0.
 * X is N x (s + r*L) (ld N): cols [0,s) = b, then L copies of U.
 * nthreads > 1 runs tasks / nodes / chunks on pthreads.  Returns INFO. */
int orc_hodlr_solve(int64_t N, int m, int L, int r, int s, int task_level,
                    int nthreads, double *K, double *X, const double *V);

/* Build X = [B | U | U | ... | U] (N x (s + r*L)), init_matrix.cc:57-72. */
void orc_build_x(int64_t N, int r, int s, int L, const double *B, int64_t ldb,
                 const double *U, int64_t ldu, double *X);

/* HODLR matvec y = A x for A = blockdiag(K_i) + (U V^T off the leaf blocks),
 * per leaf y_i = K_i x_i + U_i (V^T x - V_i^T x_i).  nrhs columns. */
void orc_hodlr_matvec(int64_t N, int m, int r, int nrhs,
                      const double *K, int64_t ldk, const double *U, int64_t ldu,
                      const double *V, int64_t ldv, const double *x, int64_t ldx,
                      double *y, int64_t ldy);
/* Dense assembly of the same A (N x N, ld N) for small N. */
void orc_dense_assemble(int64_t N, int m, int r, const double *K, int64_t ldk,
                        const double *U, int64_t ldu, const double *V, int64_t ldv,
                        double *A);

#ifdef __cplusplus
}
#endif
#endif
