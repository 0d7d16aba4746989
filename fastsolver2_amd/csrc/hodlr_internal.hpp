// hodlr_internal.hpp -- launchers shared by the C ABI (hodlr_capi.cpp) and
// the kernels (hodlr_kernels.hip).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

struct hodlr_plan_s;

namespace hodlr {

// Column view of an operand made of two column segments [A | B]: column j
// of the concatenation is read from inA (j < nA) or inB, and written to
// outA / outB.  Either segment may alias its output (in-place update) or
// point elsewhere (the Y_t -> Y_{t-1} update writes to a different level).
// Row / batch offsets are added by the kernels.
struct ColSeg {
  const double* inA;
  double* outA;
  int64_t ldiA, ldoA;
  int nA;
  const double* inB;
  double* outB;
  int64_t ldiB, ldoB;
};

inline ColSeg seg_single(double* X, int64_t ld) {
  ColSeg s;
  s.inA = X; s.outA = X; s.ldiA = ld; s.ldoA = ld; s.nA = 1 << 30;
  s.inB = nullptr; s.outB = nullptr; s.ldiB = 0; s.ldoB = 0;
  return s;
}

// Largest n the batched LU / solve kernels accept (LDS-resident panel).
constexpr int kMaxBlock = 1216;
// Diagonal-block size of the stored triangular inverses.
constexpr int kTri = 32;

__host__ __device__ inline int ntri(int n) { return (n + kTri - 1) / kTri; }
// doubles of diagonal-inverse storage per batch item
__host__ __device__ inline int64_t dinv_elems(int n) { return (int64_t)ntri(n) * 2 * kTri * kTri; }

// Batched LU with partial pivoting; perm receives the row permutation
// (row i of the factored block is row perm[i] of the input), ipiv the
// 1-based LAPACK IPIV interchanges, info the first zero pivot (1-based).
hipError_t launch_getrf(int n, double* A, int64_t lda, int64_t bstride, int batch,
                        int* ipiv, int* perm, int* info, hipStream_t st);
// Leaf factorization out of place: LU_b (ld ldl, stride lbs) <- LU of K_b
// (ld ldk, stride kbs) with pivots, INFO and the stored 32x32 diagonal
// inverses (dinv).  retry holds batch ints of workspace (blocks that left
// the no-interchange fast path).  K is not modified.
hipError_t launch_getrf_leaf(int n, const double* K, int64_t ldk, int64_t kbs, double* LU,
                             int64_t ldl, int64_t lbs, int batch, double* dinv, int* ipiv,
                             int* perm, int* info, int* retry, hipStream_t st);
hipError_t launch_getrf_leaf_f32(int n, const double* K, int64_t ldk, int64_t kbs, float* LU,
                                 int64_t ldl, int64_t lbs, int batch, float* dinv, int* ipiv,
                                 int* perm, int* info, int* retry, hipStream_t st);
// Inverts the 32x32 diagonal blocks of L (unit) and U of each LU.
// FP32 leaf factorization / solve for the mixed-precision path (n <= 512);
// the solve reads and writes FP64 right-hand sides.
hipError_t launch_getrf_f32(int n, float* A, int64_t lda, int64_t bstride, int batch, int* ipiv,
                            int* perm, int* info, hipStream_t st);
hipError_t launch_trtri_diag_f32(int n, const float* LU, int64_t lda, int64_t bstride, int batch,
                                 float* dinv, hipStream_t st);
hipError_t launch_getrs_f32(int n, const float* LU, int64_t lda, int64_t bstride, int batch,
                            const int* perm, const float* dinv, const ColSeg& X, int64_t xbstride,
                            int ncols, bool trans, hipStream_t st);
hipError_t launch_convert_f32(int64_t rows, int cols, const double* src, int64_t lds, float* dst,
                              int64_t ldd, hipStream_t st);
// max |A(:, j)| into out[j]; `part` holds kColmaxParts * cols doubles of
// partial maxima
constexpr int kColmaxParts = 128;
hipError_t launch_colmax(int64_t rows, int cols, const double* A, int64_t lda, double* out,
                         double* part, hipStream_t st);
hipError_t launch_add_cols(int64_t rows, int cols, const double* X, int64_t ldx, double* Y,
                           int64_t ldy, hipStream_t st);
// LU of HODLR cores S = [I, A; B, I] (n2 = 2r): the block structure when
// max|B| <= 1, else the general kernel (per core; `general` is batch ints of
// workspace).  Same factors, pivots and INFO as launch_getrf.
// Cores whose Schur complement keeps every diagonal pivot are factored
// without interchanges and, when r is a multiple of 32, get their 32x32
// diagonal inverses (dinv, dinv_elems(n2) per core) written here; tmask[b]
// is set for every core whose inverses still have to be formed.
// Scratch for cores of rank 64 < r <= 128 (per core: the padded 128 x 128
// Schur complement and its LU, ld 136, the LU's diagonal inverses, pivots,
// INFO and retry flag); null for other ranks (then the general LU runs).
struct CoreSchur {
  double* C = nullptr;
  double* LU = nullptr;
  double* dinv = nullptr;
  int* ipiv = nullptr;
  int* perm = nullptr;
  int* info = nullptr;
  int* retry = nullptr;
};
constexpr int64_t kCoreSchurElems = 136 * 128;  // doubles of C (and of LU) per core
hipError_t launch_getrf_core(int n2, double* S, int64_t lda, int64_t bstride, int batch, int* ipiv,
                             int* perm, int* info, int* general, double* dinv, int* tmask,
                             const CoreSchur* cs, hipStream_t st);
// launch_trtri_diag for the items b with only[b] != 0.
hipError_t launch_trtri_diag_masked(int n, const double* LU, int64_t lda, int64_t bstride,
                                    int batch, double* dinv, const int* only, hipStream_t st);
hipError_t launch_trtri_diag(int n, const double* LU, int64_t lda, int64_t bstride,
                             int batch, double* dinv, hipStream_t st);
// X_b <- U^-1 L^-1 P X_b for batch item b, columns [0, ncols) of X, where
// item b's rows start at element offset b*xbstride of each segment.
hipError_t launch_getrs(int n, const double* LU, int64_t lda, int64_t bstride, int batch,
                        const int* perm, const double* dinv, const ColSeg& X,
                        int64_t xbstride, int ncols, hipStream_t st);
// Transposed solve X_b <- P^T L^-T U^-T X_b (dgetrs TRANS='T').
// Column tile of the register-resident solve for blocks up to 512: 1, 2 or
// 5 MFMA tiles (16, 32, 80 right-hand sides) per workgroup; 0 = chosen by
// the number of columns.  Process-wide (set by the batch autotuner).
void set_getrs_tile(int nct);
int getrs_tile();
// true while per-launch profiling is on (no stream capture then)
bool profiling_enabled();
// A plan's "factored" state, for the batch: a captured factorization holds
// no factors until its graph has been launched.
void plan_mark_unfactored(hodlr_plan_s* p);
void plan_mark_factored(hodlr_plan_s* p);
// The solve restricted to the blocks b with only[b] != 0 (blocks of up to
// 512 rows).
hipError_t launch_getrs_masked(int n, const double* LU, int64_t lda, int64_t bstride, int batch,
                               const int* perm, const double* dinv, const ColSeg& X,
                               int64_t xbstride, int ncols, const int* only, hipStream_t st);
// Core solve with the structured LU of S = [I, A; B, I] (ld 2r, stride 4r^2,
// r = 32 or 64: core_solve_structured); cores flagged in `general` are
// solved by launch_getrs_masked, launched before this.  With PR non-null
// (batch even) every node b also hands T_b = E_b(0:r) + E_b(r:2r) to its
// parent b/2 at the level above: R_p = [T_c1; T_c0] at PR + p*rs (ld 2r,
// nc columns) and, with PS non-null, S_p = [I, T_c1; T_c0, I] at
// PS + p*4r^2 from the first r columns (what launch_assemble_pair forms).
bool core_solve_structured(int r);
hipError_t launch_core_solve_s(int r, const double* S, int batch, const int* perm,
                               const double* dinv, const int* general, double* R, int64_t rs,
                               int nc, double* PS, double* PR, hipStream_t st);
hipError_t launch_getrs_t(int n, const double* LU, int64_t lda, int64_t bstride, int batch,
                          const int* perm, const double* dinv, const ColSeg& X,
                          int64_t xbstride, int ncols, hipStream_t st);

// Split-K reduction over `nchild` consecutive row blocks of n_child rows:
// part[child*cpc + q] (r x ncols, ld r) = V_q^T X_q over chunk q (of at most
// rc rows) of that child.  Returns chunks per child in *cpc.
hipError_t launch_reduce(const double* V, int64_t ldv, int r, const ColSeg& X, int ncols,
                         int64_t n_child, int nchild, int64_t rc, double* part, int* cpc,
                         hipStream_t st);
int64_t reduce_chunks_per_child(int64_t n_child, int64_t rc);

// Per node k (children 2k, 2k+1), summing each child's cpc partials in order:
//   S_k = [I, G_1; G_0, I] (2r x 2r, ld 2r)  with G_c = part cols [gcol, gcol+r)
//   R_k = [H_1; H_0]       (2r x nh, ld 2r)  with H_c = part cols [hcol, hcol+nh)
// (R_k = [H_0; H_1] when swap is false: the adjoint's u_c^T g_c stacking).
// S may be null (solve with a stored core factorization).
// The level loop in node-sum form (kernels in hodlr_kernels.hip): assembly of
// a level from the level below's core solutions, the top-down pass that
// forms G_i for the leaves [i0, i0 + count), and the adjoint's two passes.
hipError_t launch_assemble_pair(const double* Ec, int64_t ecs, int r, int gcol, int hcol, int nh,
                                int nnode, double* S, double* R, int64_t rs, hipStream_t st);
hipError_t launch_path_down(const double* EY, int64_t eys, const double* EB, int64_t ebs, int L,
                            int r, int s, int64_t i0, int64_t count, double* G, hipStream_t st);
hipError_t launch_adj_up(const double* EY, int64_t eys, const double* Hc, int r, int s, int nnode,
                         double* Hn, hipStream_t st);
hipError_t launch_adj_assemble(const double* Hc, const double* Rp, int64_t rps, int r, int s,
                               int nnode, double* R, int64_t rs, hipStream_t st);
hipError_t launch_assemble(const double* part, int cpc, int r, int ncols_part, int gcol,
                           int hcol, int nh, int nnode, double* S, double* R, hipStream_t st,
                           bool swap = true);

// X_c[:, j] += alpha * u_c * eta_s[:, j] for row blocks c of n_child rows,
// slot s = c / csz, eta_s = eta + (s/2)*eta_node_stride + ((s%2)^half_xor)*r
// (ld eta_ld), j < ncols.  HODLR levels use csz = 1 (slot = child); the
// adjoint uses half_xor = 1 (child c takes the other child's half).
hipError_t launch_update(const double* u, int64_t ldu, int r, const double* eta,
                         int64_t eta_ld, int64_t eta_node_stride, int csz, int64_t n_child,
                         int nchild, double alpha, const ColSeg& X, int ncols, hipStream_t st,
                         int half_xor = 0);
// S and swapped R for LMatrix::node_solve from VTu / VTd slots.
hipError_t launch_node_assemble(int nnode, int r, const double* VTu, int64_t ldu,
                                const double* VTd, int64_t ldd, int nrhs, double* S, double* R,
                                hipStream_t st);
// Cores of sibling ranks r0 != r1 ((r0+r1) x (r0+r1)) and R = [VTd1; VTd0]
// for NodeSolveRegionTask; node k's operands at element offsets k*s*.
hipError_t launch_region_assemble(int nnode, int r0, int r1, int nrhs, const double* VTu0,
                                  int64_t ldu0, int64_t su0, const double* VTu1, int64_t ldu1,
                                  int64_t su1, const double* VTd0, int64_t ldd0, int64_t sd0,
                                  const double* VTd1, int64_t ldd1, int64_t sd1, double* S,
                                  double* R, hipStream_t st);
// R (node-major (r0+r1) x nrhs) back into VTd1 (rows [0, r0)) and VTd0.
hipError_t launch_region_scatter(int nnode, int r0, int r1, int nrhs, const double* R,
                                 double* VTd0, int64_t ldd0, int64_t sd0, double* VTd1,
                                 int64_t ldd1, int64_t sd1, hipStream_t st);

// Slot-mapped reduce / broadcast for the LMatrix operators.
hipError_t launch_sum_slots(const double* part, int nchunk, int nslot, int ka, int kb,
                            double alpha, double* C, int64_t ldc, hipStream_t st);

// A run of consecutive leaves of one size m (m = 0: empty leaf slots) in a
// tree of uneven leaves: leaves [first, first + count), rows starting at
// row0, their LU factors at element offset lu_off of the plan's factor
// buffer (leading dimension ldl, stride lbs per leaf) and their diagonal
// inverses at dinv_off.
struct LeafRun {
  int64_t first = 0, count = 0;
  int m = 0;
  int64_t row0 = 0, lu_off = 0, dinv_off = 0, ldl = 0, lbs = 0;
};

// HODLR matvec pieces.
// The matvec on a tree of uneven leaves (runs cover nleaf leaves; work as
// matvec_work_elems(nleaf, 1, rank, nrhs)).
hipError_t launch_matvec_runs(const LeafRun* runs, int nrun, int64_t nleaf, int rank, int nrhs,
                              const double* K, int64_t ldk, const double* U, int64_t ldu,
                              const double* V, int64_t ldv, const double* X, int64_t ldx,
                              double alpha, double beta, double* Y, int64_t ldy, double* work,
                              hipStream_t st, bool trans = false);
hipError_t launch_matvec(int64_t n, int leaf, int rank, int nrhs, const double* K,
                         int64_t ldk, const double* U, int64_t ldu, const double* V,
                         int64_t ldv, const double* X, int64_t ldx, double alpha, double beta,
                         double* Y, int64_t ldy, double* work, hipStream_t st,
                         bool trans = false);
int64_t matvec_work_elems(int64_t n, int leaf, int rank, int nrhs);
// The allocation-free matvec's fixed device scratch (a file-scope device
// array): its size in doubles, and the current device's address of it.
constexpr int64_t kNoAllocScratchElems = int64_t(1) << 20;
hipError_t noalloc_scratch(double** out);
// part (nleaf x r x nrhs) = V_i^T X_i per leaf, tot (r x nrhs) = their sum
// in leaf order.
hipError_t launch_leaf_projections(int64_t n, int leaf, int rank, int nrhs, const double* V,
                                   int64_t ldv, const double* X, int64_t ldx, double* part,
                                   double* tot, hipStream_t st);

// Per leaf i (rows [i*leaf, (i+1)*leaf)), C_i (leaf x nc, ld ldc) =
//   alpha * A_i (leaf x k, ld lda) * B_i^T
// with B_i (nc x k) = B rows of leaf i (ld ldb) when tot is null, else
// B_i = tot - part_i with tot, part_i (nc x k, ld nc) from
// launch_leaf_projections (nc = rank).
hipError_t launch_leaf_outer(int64_t n, int leaf, int nc, int k, double alpha, const double* A,
                             int64_t lda, const double* B, int64_t ldb, const double* part,
                             const double* tot, double* C, int64_t ldc, hipStream_t st);

hipError_t launch_generate(uint64_t seed, uint32_t stream_id, uint32_t system, int kind,
                           int64_t rows, int cols, double scale, double* out, int64_t ld,
                           hipStream_t st);
// Reference-mode generator (drand48 streams per partition from a master
// seed), rows_per_part x cols per partition, stacked by rows.
hipError_t launch_ref_fill(uint64_t master, int npart, int64_t rows, int cols, double offset,
                           double* out, int64_t ld, hipStream_t st);
hipError_t launch_shift_diag(int64_t n, int leaf, double shift, double* K, int64_t ldk,
                             hipStream_t st);
hipError_t launch_dense_blocks(int64_t n, int leaf, int rank, const double* U, int64_t ldu,
                               const double* V, int64_t ldv, const double* D, double offset,
                               double* K, int64_t ldk, hipStream_t st);
hipError_t launch_fill_scale(int64_t rows, int cols, double* A, int64_t ld, int mode, double v,
                             hipStream_t st);
// Batched small GEMM over leaves, see k_bgemm.
struct BGemmArgs {
  int batch = 0, M = 0, K = 0, Nc = 0, csplit = 0, transA = 0, bshift = 0, bflip = 0, ioff = 0;
  const double* A = nullptr;
  int64_t lda = 0, sa = 0;
  const double* B = nullptr;
  int64_t ldb = 0, bs2 = 0, bs1 = 0;
  double alpha1 = 1.0, beta1 = 0.0, alpha2 = 1.0, beta2 = 0.0;
  const double* D1 = nullptr;
  int64_t ldd1 = 0, sd1 = 0;
  double* C1 = nullptr;
  int64_t ldc1 = 0, sc1 = 0;
  const double* D2 = nullptr;
  int64_t ldd2 = 0, sd2 = 0;
  double* C2 = nullptr;
  int64_t ldc2 = 0, sc2 = 0;
};
hipError_t launch_bgemm(const BGemmArgs& g, hipStream_t st);
// Two independent products (no aliasing between them) in one launch when
// their batch and tile shapes agree, else two launches.
hipError_t launch_bgemm2(const BGemmArgs& g0, const BGemmArgs& g1, hipStream_t st);
hipError_t launch_eye(int batch, int r, int64_t sw, double* W, hipStream_t st);
hipError_t launch_selftest_mfma(double* out, hipStream_t st);
hipError_t launch_selftest_mfma_f32(double* out, hipStream_t st);
hipError_t launch_copy_cols(int64_t rows, int cols, const double* src, int64_t lds,
                            double* dst, int64_t ldd, hipStream_t st);

}  // namespace hodlr
