// hodlr_pair.cpp -- two sibling shards with different off-diagonal ranks
// (include/hodlr.h, "pairs"): the single-GPU form of the reference's
// spmd_driver/solver_irregular_ranks.cc.
//
//   A = [A_0,      U_0 W_1^T]      A_c  the HODLR matrix of shard c (own
//       [U_1 W_0^T,      A_1]           plan: leaves, levels, rank r_c)
//                                  U_c  n_c x r_c, the rows its own levels use
//                                  W_c  n_c x r_{1-c} (the reference's VMatp)
//
// Each shard is factored and solved by its own plan on [B_c | U_c]
// (d_c = A_c^-1 B_c, Y_c = A_c^-1 U_c: the leaf task + in-machine level loop
// of the driver, :146-172); the projections W_c^T [d_c | Y_c] are the
// driver's LMatrix::gemm (:181-182); the (r0+r1)^2 root core is
// NodeSolveRegionTask (hodlr_node_solve_region, :218-220); the updates
// d_c -= Y_c eta are LMatrix::gemm_inplace (:250-251).
#include "../../include/hodlr.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <vector>

#include "hodlr_internal.hpp"

struct hodlr_pair_s {
  int device = 0;
  hodlr_plan plan[2] = {nullptr, nullptr};
  int64_t n[2] = {0, 0};
  int r[2] = {0, 0};
  int nrhs_max = 0;
  // per shard: X_c = [d_c | Y_c] (n_c x (nrhs_max + r_c)), T_c = W_c^T X_c
  // (r_{1-c} x (nrhs_max + r_c)); the core's INFO
  double* X[2] = {nullptr, nullptr};
  double* T[2] = {nullptr, nullptr};
  int* info = nullptr;
  bool factored = false;
};

namespace {

void pair_free(hodlr_pair q) {
  for (int c = 0; c < 2; ++c) {
    if (q->plan[c]) (void)hodlr_plan_destroy(q->plan[c]);
    if (q->X[c]) (void)hipFree(q->X[c]);
    if (q->T[c]) (void)hipFree(q->T[c]);
  }
  if (q->info) (void)hipFree(q->info);
}

}  // namespace

extern "C" {

int hodlr_pair_create(hodlr_pair* pair, const hodlr_desc* d0, const int* leaf_sizes0,
                      const hodlr_desc* d1, const int* leaf_sizes1, int nrhs_max) {
  if (!pair || !d0 || !d1 || nrhs_max < 0) return HODLR_ERR_ARG;
  *pair = nullptr;
  if (d0->device != d1->device || d0->precision != HODLR_FP64 || d1->precision != HODLR_FP64)
    return HODLR_ERR_ARG;
  if (d0->rank + d1->rank > hodlr::kMaxBlock) return HODLR_ERR_UNSUPPORTED;
  int prev = 0;
  if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(d0->device) != hipSuccess)
    return HODLR_ERR_ARG;
  hodlr_pair q = new hodlr_pair_s();
  q->device = d0->device;
  q->nrhs_max = nrhs_max;
  const hodlr_desc* d[2] = {d0, d1};
  const int* sz[2] = {leaf_sizes0, leaf_sizes1};
  int rc = HODLR_OK;
  for (int c = 0; c < 2 && !rc; ++c) {
    hodlr_desc dc = *d[c];
    dc.nrhs_max = nrhs_max + dc.rank;  // [B_c | U_c]
    rc = sz[c] ? hodlr_plan_create_uneven(&q->plan[c], &dc, sz[c])
               : hodlr_plan_create(&q->plan[c], &dc);
    if (rc) break;
    q->n[c] = dc.n;
    q->r[c] = dc.rank;
  }
  for (int c = 0; c < 2 && !rc; ++c) {
    const int w = nrhs_max + q->r[c];
    if (hipMalloc(&q->X[c], sizeof(double) * (size_t)q->n[c] * w) != hipSuccess ||
        hipMalloc(&q->T[c], sizeof(double) * (size_t)q->r[1 - c] * w) != hipSuccess)
      rc = HODLR_ERR_NOMEM;
  }
  if (!rc && hipMalloc(&q->info, sizeof(int)) != hipSuccess) rc = HODLR_ERR_NOMEM;
  (void)hipSetDevice(prev);
  if (rc) {
    pair_free(q);
    delete q;
    return rc;
  }
  *pair = q;
  return HODLR_OK;
}

int hodlr_pair_destroy(hodlr_pair q) {
  if (!q) return HODLR_ERR_ARG;
  int prev = 0;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(q->device);
  (void)hipDeviceSynchronize();
  pair_free(q);
  (void)hipSetDevice(prev);
  delete q;
  return HODLR_OK;
}

hodlr_plan hodlr_pair_plan(hodlr_pair q, int shard) {
  return (q && (shard == 0 || shard == 1)) ? q->plan[shard] : nullptr;
}

int hodlr_pair_factor_solve(hodlr_pair q, const double* const* K, const int64_t* ldk,
                            const double* const* U, const int64_t* ldu,
                            const double* const* V, const int64_t* ldv,
                            const double* const* W, const int64_t* ldw, double* B, int64_t ldb,
                            int nrhs, hipStream_t st) {
  if (!q || !K || !ldk || !U || !ldu || !V || !ldv || !W || !ldw || nrhs < 0) return HODLR_ERR_ARG;
  if (nrhs > q->nrhs_max || (nrhs > 0 && (!B || ldb < q->n[0] + q->n[1]))) return HODLR_ERR_ARG;
  for (int c = 0; c < 2; ++c)
    if (!K[c] || !U[c] || !V[c] || !W[c] || ldw[c] < q->n[c] || ldu[c] < q->n[c])
      return HODLR_ERR_ARG;
  int prev = 0;
  if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(q->device) != hipSuccess)
    return HODLR_ERR_HIP;
  q->factored = false;
  int rc = HODLR_OK;
  int64_t row0[2] = {0, q->n[0]};
  for (int c = 0; c < 2 && !rc; ++c) {
    // X_c = [B_c | U_c]; shard solve: X_c <- A_c^-1 X_c = [d_c | Y_c]
    const int64_t n = q->n[c];
    if (nrhs > 0 && hodlr::launch_copy_cols(n, nrhs, B + row0[c], ldb, q->X[c], n, st) != hipSuccess)
      rc = HODLR_ERR_HIP;
    if (!rc && hodlr::launch_copy_cols(n, q->r[c], U[c], ldu[c], q->X[c] + n * nrhs, n, st) !=
                   hipSuccess)
      rc = HODLR_ERR_HIP;
    if (!rc)
      rc = hodlr_factor_solve(q->plan[c], K[c], ldk[c], U[c], ldu[c], V[c], ldv[c], q->X[c], n,
                              nrhs + q->r[c], st);
    // T_c = W_c^T [d_c | Y_c]  (r_{1-c} x (nrhs + r_c))
    if (!rc)
      rc = hodlr_gemm_red(n, 1, 1, 1.0, W[c], ldw[c], q->r[1 - c], q->X[c], n, nrhs + q->r[c],
                          q->T[c], q->r[1 - c], st);
  }
  // root core: [I, W_1^T Y_1; W_0^T Y_0, I] [eta_0; eta_1] = [W_1^T d_1; W_0^T d_0]
  const int r0 = q->r[0], r1 = q->r[1];
  if (!rc)
    rc = hodlr_node_solve_region(1, r0, r1, nrhs, q->T[0] + (int64_t)r1 * nrhs, r1, 0,
                                 q->T[1] + (int64_t)r0 * nrhs, r0, 0, q->T[0], r1, 0, q->T[1], r0,
                                 0, q->info, st);
  // x_0 = d_0 - Y_0 eta_0 (eta_0 in T_1's d columns), x_1 = d_1 - Y_1 eta_1
  for (int c = 0; c < 2 && !rc && nrhs > 0; ++c) {
    const int64_t n = q->n[c];
    rc = hodlr_gemm_bro(n, 1, 1, -1.0, q->X[c] + n * nrhs, n, q->r[c], q->T[1 - c], q->r[c],
                        q->X[c], n, nrhs, st);
    if (!rc && hodlr::launch_copy_cols(n, nrhs, q->X[c], n, B + row0[c], ldb, st) != hipSuccess)
      rc = HODLR_ERR_HIP;
  }
  if (!rc) q->factored = true;
  (void)hipSetDevice(prev);
  return rc;
}

int hodlr_pair_status(hodlr_pair q, hipStream_t st, int* shard, int* where, int* level) {
  if (!q) return HODLR_ERR_ARG;
  if (shard) *shard = -1;
  if (where) *where = -1;
  if (level) *level = -1;
  for (int c = 0; c < 2; ++c) {
    int w = -1, l = -1;
    const int s = hodlr_plan_status(q->plan[c], st, &w, &l);
    if (s != HODLR_OK) {
      if (shard) *shard = c;
      if (where) *where = w;
      if (level) *level = l;
      return s;
    }
  }
  int info = 0;
  int prev = 0;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(q->device);
  const hipError_t e = hipMemcpy(&info, q->info, sizeof(int), hipMemcpyDeviceToHost);
  (void)hipSetDevice(prev);
  if (e != hipSuccess) return HODLR_ERR_HIP;
  if (info) {
    if (where) *where = 0;
    if (level) *level = -1;
    return HODLR_SINGULAR_CORE;
  }
  return HODLR_OK;
}

}  // extern "C"
