// hodlr_capi.cpp -- C ABI (include/hodlr.h) over the gfx950 kernels.
//
// The HODLR solve is level-synchronous: one batched launch per step for all
// leaves, then per tree level t = L-1..0 one reduce, one core assembly, one
// batched core LU + solve and one update launch -- the reference's index
// launches of LeafSolveTask / GemmRedTask / NodeSolveTask / GemmBroTask
// (spmd_driver/solver_irregular_part.cc:114-140) become stream-ordered
// kernels on one device.
//
// Fast path (hodlr_factor / hodlr_solve / hodlr_factor_solve).  The
// reference carries, for every level t, its own copy u_t of U through the
// leaf solve and all deeper levels (UTree column layout, src/tree.cc:64-73,
// filled with identical copies by init_matrix.cc:64-71).  Every copy receives
// exactly the same operations below its level, so one block per level
// suffices: Y_{L-1} = K^-1 U and Y_{t-1} = Y_t - Y_t E_t (E_t the u-part of
// the level-t core solve, per child).  On one leaf i that recursion is a
// right multiplication by r x r matrices, so Y_t,i = Y_{L-1,i} W_t,i with
// W_{L-1} = I and W_{t-1,i} = W_t,i (I - E_t,c(i)); likewise the level
// projections V_i^T [Y_t | B_t] obey P_{t-1,i} = P_t,i - P_t,i(:, 0:r) E_t,c(i)
// and B_final = B_{L-1} - Y_{L-1} G with G_i = sum_t W_t,i E^B_t,c(i).  The
// plan therefore touches N-row data only at the leaves (one solve on [U | B],
// one projection, one final B update); every level is O(leaves * r^2 (r+s))
// work on r x r blocks.  Same arithmetic as the reference up to association
// order.  hodlr_hsolve_ref keeps the reference's full column layout for
// callers that pass it.
#include "../../include/hodlr.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "hodlr_internal.hpp"

using hodlr::ColSeg;

namespace {

thread_local char g_detail[256] = "";

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      snprintf(g_detail, sizeof g_detail, "%s: %s", #expr, hipGetErrorString(e_)); \
      return HODLR_ERR_HIP;                                                        \
    }                                                                              \
  } while (0)

struct DeviceGuard {
  int prev = -1;
  bool ok = true;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (dev >= 0 && dev != prev) ok = (hipSetDevice(dev) == hipSuccess);
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

// Row-chunk length of the split-K level reduce.
int64_t reduce_rc(int64_t n_child) { return n_child < 512 ? n_child : 512; }

// Library workspace for entry points without a plan: one buffer per
// (device, stream).  Calls on different streams never share scratch; calls
// on one stream are ordered by the stream itself.  Growth is stream-ordered
// as well (hipFreeAsync / hipMallocAsync on that stream): no device-wide
// synchronization and no other stream is touched.  A buffer cannot grow
// while its stream is being captured (the allocation would become a graph
// node); the call then fails with HODLR_ERR_STATE and a first call outside
// capture sizes it.
struct Workspace {
  std::mutex mu;
  std::map<std::pair<int, uintptr_t>, std::pair<void*, size_t>> buf;
};
Workspace g_ws;

int workspace(size_t bytes, hipStream_t st, void** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_ws.mu);
  auto& slot = g_ws.buf[{dev, (uintptr_t)st}];
  if (slot.second < bytes) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (st != nullptr) HIP_TRY(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone) {
      snprintf(g_detail, sizeof g_detail,
               "workspace of %zu bytes needed during stream capture; call once outside capture "
               "first", bytes);
      return HODLR_ERR_STATE;
    }
    if (slot.first) {
      HIP_TRY(hipFreeAsync(slot.first, st));
      slot.first = nullptr;
      slot.second = 0;
    }
    const size_t want = bytes + bytes / 4 + 4096;
    if (hipMallocAsync(&slot.first, want, st) != hipSuccess) {
      (void)hipGetLastError();
      snprintf(g_detail, sizeof g_detail, "workspace allocation of %zu bytes failed", want);
      slot.first = nullptr;
      return HODLR_ERR_NOMEM;
    }
    slot.second = want;
  }
  *out = slot.first;
  return HODLR_OK;
}

// Per-launch timing with HIP events on the launch stream, enabled by
// hodlr_profile_enable(); each record carries the launch's algorithmic
// FLOPs and minimum HBM bytes so a caller can place it on the roofline.
struct ProfRec {
  int kind;
  hipEvent_t a, b;
  double flops, bytes;
};
struct Profiler {
  std::mutex mu;
  bool on = false;
  unsigned mask = ~0u;  // kinds recorded while on
  std::vector<ProfRec> recs;
};
Profiler g_prof;

struct ProfScope {
  bool on;
  ProfRec rec;
  hipStream_t st;
  ProfScope(int kind, hipStream_t s, double flops, double bytes)
      : on(g_prof.on && ((g_prof.mask >> kind) & 1u)), st(s) {
    if (!on) return;
    rec.kind = kind; rec.flops = flops; rec.bytes = bytes;
    if (hipEventCreate(&rec.a) != hipSuccess || hipEventCreate(&rec.b) != hipSuccess) {
      on = false;
      return;
    }
    (void)hipEventRecord(rec.a, st);
  }
  ~ProfScope() {
    if (!on) return;
    (void)hipEventRecord(rec.b, st);
    std::lock_guard<std::mutex> lk(g_prof.mu);
    g_prof.recs.push_back(rec);
  }
};

bool is_pow2(int64_t x) { return x > 0 && (x & (x - 1)) == 0; }
int ilog2(int64_t x) {
  int l = 0;
  while ((int64_t(1) << l) < x) ++l;
  return l;
}

// Workspace layout for one batched core step (2^t nodes of 2r x 2r).
struct CoreBufs {
  double* S;
  double* R;
  int* ipiv;
  int* perm;
  int* info;
  double* dinv;
  int* general = nullptr;  // per-core flags of the structured LU (null: general LU)
  int* tmask = nullptr;    // cores whose diagonal inverses are still to be formed
  const hodlr::CoreSchur* schur = nullptr;  // scratch for ranks 64 < r <= 128
  // the level above (null at the root): its cores' S (stride 4r^2) and
  // right-hand sides (the node stride of R), formed by the structured
  // solve's epilogue; `assembled` reports whether it did
  double* PS = nullptr;
  double* PR = nullptr;
  bool* assembled = nullptr;
};

// Core step shared by every level: S (already assembled, n2 x n2 per node)
// is factored in place, R (n2 x nh per node) overwritten by S^-1 R.
double lu_flops(double n) { return 2.0 / 3.0 * n * n * n; }

// R <- S^-1 R for the factored cores of one level (R: n2 x nh per node at
// stride rs, ld n2): the structured solve where the cores were factored by
// the structured LU (their flags in c.general; r = 32 or 64) plus the
// generic solve masked to the cores it flagged general, else the generic
// solve.
int core_solve(int nnode, int n2, const CoreBufs& c, double* R, int64_t rs, int nh,
               hipStream_t st) {
  const int r = n2 / 2;
  ColSeg x = hodlr::seg_single(R, n2);
  if (c.assembled) *c.assembled = false;
  if (c.general && (n2 & 1) == 0 && hodlr::core_solve_structured(r)) {
    // the general cores first: the structured kernel hands every node's
    // solution, theirs included, to the level above
    HIP_TRY(hodlr::launch_getrs_masked(n2, c.S, n2, (int64_t)n2 * n2, nnode, c.perm, c.dinv, x, rs,
                                       nh, c.general, st));
    const bool up = c.PR && nnode >= 2;
    HIP_TRY(hodlr::launch_core_solve_s(r, c.S, nnode, c.perm, c.dinv, c.general, R, rs, nh,
                                       up ? c.PS : nullptr, up ? c.PR : nullptr, st));
    if (c.assembled) *c.assembled = up;
    return HODLR_OK;
  }
  HIP_TRY(hodlr::launch_getrs(n2, c.S, n2, (int64_t)n2 * n2, nnode, c.perm, c.dinv, x, rs, nh, st));
  return HODLR_OK;
}

int core_factor_solve_n(int nnode, int n2, const CoreBufs& c, int nh, hipStream_t st) {
  const double blk = 8.0 * n2 * n2 * nnode;
  {
    ProfScope ps(HODLR_PROF_GETRF_CORE, st, nnode * lu_flops(n2), 2.0 * blk);
    if (c.general)
      HIP_TRY(hodlr::launch_getrf_core(n2, c.S, n2, (int64_t)n2 * n2, nnode, c.ipiv, c.perm, c.info,
                                       c.general, c.dinv, c.tmask, c.schur, st));
    else
      HIP_TRY(hodlr::launch_getrf(n2, c.S, n2, (int64_t)n2 * n2, nnode, c.ipiv, c.perm, c.info,
                                  st));
  }
  // cores of up to 128 rows: the LU kernels formed every inverse themselves
  if (!(c.general && n2 <= 128 && (n2 & 1) == 0)) {
    ProfScope ps(HODLR_PROF_TRTRI, st, 0.0, blk);
    if (c.general)  // only the cores the structured LU left without inverses
      HIP_TRY(hodlr::launch_trtri_diag_masked(n2, c.S, n2, (int64_t)n2 * n2, nnode, c.dinv, c.tmask,
                                              st));
    else
      HIP_TRY(hodlr::launch_trtri_diag(n2, c.S, n2, (int64_t)n2 * n2, nnode, c.dinv, st));
  }
  if (nh > 0) {
    ProfScope ps(HODLR_PROF_GETRS_CORE, st, 2.0 * n2 * n2 * nh * nnode,
                 blk + 16.0 * n2 * nh * nnode);
    return core_solve(nnode, n2, c, c.R, (int64_t)n2 * nh, nh, st);
  }
  return HODLR_OK;
}

// The HODLR cores: 2r x 2r.
int core_factor_solve(int nnode, int r, const CoreBufs& c, int nh, hipStream_t st) {
  return core_factor_solve_n(nnode, 2 * r, c, nh, st);
}

}  // namespace

bool hodlr::profiling_enabled() { return g_prof.on; }

// =====================================================================
// plans
// =====================================================================
struct hodlr_plan_s {
  hodlr_desc d;
  int64_t N = 0;
  int m = 0, L = 0, r = 0, smax = 0;  // m: the largest leaf
  int64_t P = 0;
  // leaves as runs of equal size (one run for a balanced tree of equal
  // leaves; empty leaf slots form runs of size 0)
  std::vector<hodlr::LeafRun> runs;
  // leaf factorization, out of place: leaf i's LU factors in the
  // contiguous block LU + i*lbs (ld ldl = leaf + pad), FP64 or (mixed
  // precision) FP32; the caller's K is only read
  double* LU = nullptr;
  float* LU32 = nullptr;
  int64_t ldl = 0, lbs = 0;
  int* leaf_retry = nullptr;  // blocks that left the no-interchange path
  int* leaf_ipiv = nullptr;
  int* leaf_perm = nullptr;
  int* leaf_info = nullptr;
  double* leaf_dinv = nullptr;
  // Y = Y_{L-1} = K^-1 U (N x r, ld N)
  double* Y = nullptr;
  // the factorization's core solutions E_n = S_n^-1 R_n (2r x nq, ld 2r)
  // of every node, heap order (node n of level t at (2^t - 1 + n)*2r*nq):
  // their first r columns, E^Y, serve the solves
  double* E = nullptr;
  int nq = 0;
  // cores, all levels; level t at offsets *_off[t]
  double* coreS = nullptr;
  int* core_ipiv = nullptr;
  int* core_perm = nullptr;
  int* core_info = nullptr;
  double* core_dinv = nullptr;
  int* core_general = nullptr;  // flags of the structured core LU, every level (heap order)
  int* core_tmask = nullptr;    // cores left without diagonal inverses (largest level)
  hodlr::CoreSchur schur;       // Schur-complement scratch for ranks 64 < r <= 128
  std::vector<int64_t> s_off, p_off, d_off, i_off;
  // workspace: per-leaf projections and accumulators G_i (P x r x ncmax),
  // a solve's core solutions (heap order, 2r x ncmax per node) and the
  // adjoint's per-node sums (heap order, leaves included: r x ncmax)
  double* Pb = nullptr;
  double* G = nullptr;
  double* EB = nullptr;
  double* Hb = nullptr;
  int ncmax = 0;
  // mixed precision: FP32 leaf factors (LU32) and their diagonal-block
  // inverses
  bool mixed = false;
  float* leaf_dinv32 = nullptr;
  // refinement buffers (allocated on first use): residual block, matvec
  // workspace, column norms
  double* ir_buf = nullptr;
  int64_t ir_elems = 0;
  // the correction solve on the residual block, captured as a HIP graph
  // (stream-capture of hodlr_solve; rebuilt when the block or nrhs change,
  // dropped by every factorization)
  hipGraphExec_t ir_exec = nullptr;
  const double* ir_exec_R = nullptr;
  int ir_exec_nrhs = 0;
  // factored operands
  const double* K = nullptr;
  const double* U = nullptr;
  const double* V = nullptr;
  int64_t ldk = 0, ldu = 0, ldv = 0;
  bool factored = false;
  int64_t bytes = 0;
  std::vector<void*> allocs;
};

void hodlr::plan_mark_unfactored(hodlr_plan p) {
  if (p) p->factored = false;
}
void hodlr::plan_mark_factored(hodlr_plan p) {
  if (p) p->factored = true;
}

namespace {

template <typename T>
int palloc(hodlr_plan p, T** out, int64_t count) {
  *out = nullptr;
  if (count <= 0) return HODLR_OK;
  void* ptr = nullptr;
  const size_t bytes = sizeof(T) * (size_t)count;
  if (hipMalloc(&ptr, bytes) != hipSuccess) {
    snprintf(g_detail, sizeof g_detail, "plan allocation of %zu bytes failed", bytes);
    return HODLR_ERR_NOMEM;
  }
  p->allocs.push_back(ptr);
  p->bytes += (int64_t)bytes;
  *out = static_cast<T*>(ptr);
  return HODLR_OK;
}

void drop_graph(hodlr_plan p) {
  if (p->ir_exec) (void)hipGraphExecDestroy(p->ir_exec);
  p->ir_exec = nullptr;
  p->ir_exec_R = nullptr;
  p->ir_exec_nrhs = 0;
}

void plan_free(hodlr_plan p) {
  for (void* a : p->allocs) (void)hipFree(a);
  p->allocs.clear();
  if (p->ir_buf) (void)hipFree(p->ir_buf);
  p->ir_buf = nullptr;
  p->ir_elems = 0;
  drop_graph(p);
}

// The column view x with its rows shifted down by row0.
ColSeg shift_rows(const ColSeg& x, int64_t row0) {
  ColSeg s = x;
  if (s.inA) s.inA += row0;
  if (s.outA) s.outA += row0;
  if (s.inB) s.inB += row0;
  if (s.outB) s.outB += row0;
  return s;
}

// Leaf solve with the stored leaf factorization (FP64 or the FP32 copy),
// one launch per run of equal leaves.
int leaf_solve(hodlr_plan p, const ColSeg& x, int ncols, bool trans, hipStream_t st) {
  for (const hodlr::LeafRun& R : p->runs) {
    if (R.m == 0) continue;
    const ColSeg xr = shift_rows(x, R.row0);
    const int* perm = p->leaf_perm + R.row0;
    if (p->mixed) {
      HIP_TRY(hodlr::launch_getrs_f32(R.m, p->LU32 + R.lu_off, R.ldl, R.lbs, (int)R.count, perm,
                                      p->leaf_dinv32 + R.dinv_off, xr, R.m, ncols, trans, st));
    } else if (trans) {
      HIP_TRY(hodlr::launch_getrs_t(R.m, p->LU + R.lu_off, R.ldl, R.lbs, (int)R.count, perm,
                                    p->leaf_dinv + R.dinv_off, xr, R.m, ncols, st));
    } else {
      HIP_TRY(hodlr::launch_getrs(R.m, p->LU + R.lu_off, R.ldl, R.lbs, (int)R.count, perm,
                                  p->leaf_dinv + R.dinv_off, xr, R.m, ncols, st));
    }
  }
  return HODLR_OK;
}

// sum over leaves of m^k
double leaf_pow_sum(hodlr_plan p, int k) {
  double s = 0.0;
  for (const hodlr::LeafRun& R : p->runs) s += (double)R.count * std::pow((double)R.m, k);
  return s;
}

// first heap index of level t (level L: the leaves)
int64_t heap0(int t) { return ((int64_t)1 << t) - 1; }

CoreBufs core_level(hodlr_plan p, int t) {
  CoreBufs c;
  c.S = p->coreS + p->s_off[t];
  c.R = nullptr;  // set by the caller: the level's slot of a heap-ordered buffer
  c.ipiv = p->core_ipiv + p->p_off[t];
  c.perm = p->core_perm + p->p_off[t];
  c.info = p->core_info + p->i_off[t];
  c.dinv = p->core_dinv + p->d_off[t];
  c.general = p->core_general + heap0(t);
  c.tmask = p->core_tmask;
  c.schur = p->schur.C ? &p->schur : nullptr;
  return c;
}

int check_ld(int64_t ld, int64_t n) { return ld >= n ? HODLR_OK : HODLR_ERR_ARG; }

// Leaf phase of factor (+ solve): LU of every leaf (out of place, K is
// only read), then [Y_{L-1} | B] = K^-1 [U | B].
int leaf_phase(hodlr_plan p, const double* K, int64_t ldk, const double* U, int64_t ldu,
               double* B, int64_t ldb, int nrhs, hipStream_t st) {
  const double m2 = leaf_pow_sum(p, 2);
  const double kb = (p->mixed ? 4.0 : 8.0) * m2;
  for (const hodlr::LeafRun& R : p->runs) {
    if (R.m == 0) continue;
    const double rm2 = (double)R.count * R.m * R.m;
    ProfScope ps(HODLR_PROF_GETRF_LEAF, st, R.count * lu_flops(R.m),
                 (8.0 + (p->mixed ? 4.0 : 8.0)) * rm2);
    if (p->mixed)
      HIP_TRY(hodlr::launch_getrf_leaf_f32(R.m, K + R.row0, ldk, R.m, p->LU32 + R.lu_off, R.ldl,
                                           R.lbs, (int)R.count, p->leaf_dinv32 + R.dinv_off,
                                           p->leaf_ipiv + R.row0, p->leaf_perm + R.row0,
                                           p->leaf_info + R.first, p->leaf_retry + R.first, st));
    else
      HIP_TRY(hodlr::launch_getrf_leaf(R.m, K + R.row0, ldk, R.m, p->LU + R.lu_off, R.ldl, R.lbs,
                                       (int)R.count, p->leaf_dinv + R.dinv_off,
                                       p->leaf_ipiv + R.row0, p->leaf_perm + R.row0,
                                       p->leaf_info + R.first, p->leaf_retry + R.first, st));
  }
  ColSeg x;
  int ncols;
  if (p->L > 0) {
    x.inA = U; x.outA = p->Y; x.ldiA = ldu; x.ldoA = p->N; x.nA = p->r;
    x.inB = B; x.outB = B; x.ldiB = ldb; x.ldoB = ldb;
    ncols = p->r + nrhs;
  } else {
    x = hodlr::seg_single(B, ldb);
    ncols = nrhs;
  }
  if (ncols > 0) {
    ProfScope ps(HODLR_PROF_GETRS_LEAF, st, 2.0 * m2 * ncols, kb + 16.0 * (double)p->N * ncols);
    return leaf_solve(p, x, ncols, false, st);
  }
  return HODLR_OK;
}

// Per-leaf projections part_i = V_i^T X_i (r x nc, ld r, stride r*nc).
int leaf_projections(hodlr_plan p, const double* Vm, int64_t ldv, const ColSeg& x, int nc,
                     double* part, hipStream_t st) {
  ProfScope ps(HODLR_PROF_REDUCE, st, 2.0 * (double)p->N * p->r * nc,
               8.0 * (double)p->N * (p->r + nc));
  const int64_t pe = (int64_t)p->r * nc;
  for (const hodlr::LeafRun& R : p->runs) {
    if (R.m == 0) {  // empty leaf slots project to zero
      HIP_TRY(hodlr::launch_fill_scale(R.count * pe, 1, part + R.first * pe, R.count * pe, 0, 0.0,
                                       st));
      continue;
    }
    int cpc = 0;
    HIP_TRY(hodlr::launch_reduce(Vm + R.row0, ldv, p->r, shift_rows(x, R.row0), nc, R.m,
                                 (int)R.count, R.m, part + R.first * pe, &cpc, st));
  }
  return HODLR_OK;
}

// B_i -= Y_i G_i for every leaf; G_c for leaf c at Gm + (c>>1)*gs2 +
// ((c&1)^gflip)*gs1 (r x nc, ld ldg): per-leaf blocks (ldg = r, gs1 = r*nc,
// gs2 = 2*gs1, gflip = 0) or the halves of the lowest level's core solutions.
int final_update(hodlr_plan p, const double* Ym, int64_t ldy, const double* Gm, int64_t ldg,
                 int64_t gs2, int64_t gs1, int gflip, int nc, double* B, int64_t ldb,
                 hipStream_t st) {
  ProfScope ps(HODLR_PROF_UPDATE, st, 2.0 * (double)p->N * p->r * nc,
               8.0 * (double)p->N * (p->r + 2.0 * nc));
  for (const hodlr::LeafRun& R : p->runs) {
    if (R.m == 0) continue;
    hodlr::BGemmArgs g;
    g.batch = (int)R.count; g.M = R.m; g.K = p->r; g.Nc = nc; g.csplit = nc;
    g.A = Ym + R.row0; g.lda = ldy; g.sa = R.m;
    g.B = Gm; g.ldb = ldg; g.bshift = 0; g.bs1 = gs1; g.bs2 = gs2; g.bflip = gflip;
    g.ioff = (int)R.first;
    g.alpha1 = -1.0; g.beta1 = 1.0;
    g.D1 = B + R.row0; g.ldd1 = ldb; g.sd1 = R.m;
    g.C1 = B + R.row0; g.ldc1 = ldb; g.sc1 = R.m;
    HIP_TRY(hodlr::launch_bgemm(g, st));
  }
  return HODLR_OK;
}

// B_i -= Y_i G_i with G_i in per-leaf blocks (r x nc, ld r).
int final_update_leaf(hodlr_plan p, const double* Gm, int nc, double* B, int64_t ldb,
                      hipStream_t st) {
  const int64_t gs1 = (int64_t)p->r * nc;
  return final_update(p, p->Y, p->N, Gm, p->r, 2 * gs1, gs1, 0, nc, B, ldb, st);
}

// Top-down pass: G_i for every non-empty leaf from the core solutions
// (E^Y at EY, E^B at EB, heap order, node strides eys / ebs).
int path_down(hodlr_plan p, const double* EY, int64_t eys, const double* EB, int64_t ebs, int nc,
              hipStream_t st) {
  const int r = p->r;
  ProfScope ps(HODLR_PROF_LEAF_LEVEL, st, (double)p->P * p->L * 2.0 * r * r * nc,
               8.0 * (double)p->P * p->L * r * (r + nc));
  for (const hodlr::LeafRun& R : p->runs) {
    if (R.m == 0) continue;
    HIP_TRY(hodlr::launch_path_down(EY, eys, EB, ebs, p->L, r, nc, R.first, R.count, p->G, st));
  }
  return HODLR_OK;
}

// Level loop of factor (+ solve, nrhs > 0) in node-sum form.  Input:
// [Y | B] after the leaf solve.  Output: every node's core factorization
// and core solution E_n (heap order in p->E), and B overwritten by the
// solution.
int level_phase(hodlr_plan p, double* B, int64_t ldb, int nrhs, hipStream_t st) {
  const int r = p->r, L = p->L, n2 = 2 * r;
  const int nc = r + nrhs;
  if (L == 0) return HODLR_OK;
  ColSeg x;
  x.inA = p->Y; x.outA = p->Y; x.ldiA = p->N; x.ldoA = p->N; x.nA = r;
  x.inB = B; x.outB = B; x.ldiB = ldb; x.ldoB = ldb;
  int rc = leaf_projections(p, p->V, p->ldv, x, nc, p->Pb, st);
  if (rc) return rc;
  p->nq = nc;
  const int64_t es = (int64_t)n2 * nc;  // node stride of E
  bool assembled = false;  // this level's S and R formed by the solve below
  for (int t = L - 1; t >= 0; --t) {
    const int nnode = 1 << t;
    CoreBufs c = core_level(p, t);
    c.R = p->E + heap0(t) * es;
    if (t > 0) {
      c.PS = p->coreS + p->s_off[t - 1];
      c.PR = p->E + heap0(t - 1) * es;
    }
    bool up = false;
    c.assembled = &up;
    if (!assembled) {
      ProfScope ps(HODLR_PROF_ASSEMBLE, st, 0.0,
                   8.0 * (double)nnode * (n2 * n2 + 2.0 * n2 * nc));
      if (t == L - 1)  // the children are leaves: their projections
        HIP_TRY(hodlr::launch_assemble(p->Pb, 1, r, nc, 0, 0, nc, nnode, c.S, c.R, st));
      else
        HIP_TRY(hodlr::launch_assemble_pair(p->E + heap0(t + 1) * es, es, r, 0, 0, nc, nnode, c.S,
                                            c.R, es, st));
    }
    rc = core_factor_solve(nnode, r, c, nc, st);
    if (rc) return rc;
    assembled = up;
  }
  if (nrhs == 0) return HODLR_OK;
  rc = path_down(p, p->E, es, p->E + (int64_t)r * n2, es, nrhs, st);
  return rc ? rc : final_update_leaf(p, p->G, nrhs, B, ldb, st);
}

}  // namespace

extern "C" {

const char* hodlr_version(void) { return "hodlr-amd 0.1 (gfx950, FP64 MFMA)"; }

const char* hodlr_strerror(int status) {
  switch (status) {
    case HODLR_OK: return "success";
    case HODLR_ERR_ARG: return "invalid argument";
    case HODLR_ERR_SHAPE: return "inconsistent HODLR shape";
    case HODLR_ERR_UNSUPPORTED: return "size not supported by the kernels";
    case HODLR_ERR_HIP: return "HIP runtime error";
    case HODLR_ERR_NOMEM: return "device allocation failed";
    case HODLR_ERR_STATE: return "plan not factored";
    case HODLR_SINGULAR_LEAF: return "singular leaf block (zero pivot)";
    case HODLR_SINGULAR_CORE: return "singular core block (zero pivot)";
    default: return "unknown status";
  }
}

const char* hodlr_last_error_detail(void) { return g_detail; }

int hodlr_workspace_release(hipStream_t st) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_ws.mu);
  auto it = g_ws.buf.find({dev, (uintptr_t)st});
  if (it == g_ws.buf.end()) return HODLR_OK;
  void* ptr = it->second.first;
  g_ws.buf.erase(it);
  if (ptr) HIP_TRY(hipFreeAsync(ptr, st));
  return HODLR_OK;
}

int hodlr_profile_mask(unsigned mask) {
  std::lock_guard<std::mutex> lk(g_prof.mu);
  g_prof.mask = mask;
  return HODLR_OK;
}

int hodlr_profile_enable(int on) {
  std::lock_guard<std::mutex> lk(g_prof.mu);
  g_prof.on = (on != 0);
  return HODLR_OK;
}

int hodlr_profile_reset(void) {
  std::lock_guard<std::mutex> lk(g_prof.mu);
  for (auto& r : g_prof.recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  g_prof.recs.clear();
  return HODLR_OK;
}

int hodlr_profile_read(int kind, int64_t* count, double* total_ms, double* flops,
                       double* bytes) {
  if (kind < 0 || kind >= HODLR_PROF_NKINDS) return HODLR_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_prof.mu);
  int64_t c = 0;
  double ms = 0.0, f = 0.0, b = 0.0;
  for (auto& r : g_prof.recs) {
    if (r.kind != kind) continue;
    HIP_TRY(hipEventSynchronize(r.b));
    float e = 0.0f;
    HIP_TRY(hipEventElapsedTime(&e, r.a, r.b));
    ++c; ms += e; f += r.flops; b += r.bytes;
  }
  if (count) *count = c;
  if (total_ms) *total_ms = ms;
  if (flops) *flops = f;
  if (bytes) *bytes = b;
  return HODLR_OK;
}

// Plan of a balanced tree of 2^levels leaf slots with the given leaf sizes
// (validated by the callers); d.n is their sum and d.leaf the largest.
static int plan_create_sizes(hodlr_plan* plan, const hodlr_desc& d, const std::vector<int>& sizes) {
  DeviceGuard g(d.device);
  if (!g.ok) return HODLR_ERR_ARG;
  hodlr_plan p = new hodlr_plan_s();
  p->d = d;
  p->N = d.n; p->m = d.leaf; p->L = d.levels; p->r = d.rank; p->smax = d.nrhs_max;
  p->P = (int64_t)1 << d.levels;
  p->mixed = d.precision == HODLR_MIXED_IR;
  const int r = p->r, n2 = 2 * r;
  // runs of equal leaves; leaf factors in contiguous blocks, the column
  // stride padded off a power of two (a leaf's columns would otherwise
  // alias in the HBM channel interleave)
  int64_t row0 = 0, lu = 0, dv = 0;
  for (int64_t i = 0; i < p->P; ++i) {
    const int m = sizes[(size_t)i];
    if (p->runs.empty() || p->runs.back().m != m) {
      hodlr::LeafRun R;
      R.first = i; R.count = 0; R.m = m; R.row0 = row0; R.lu_off = lu; R.dinv_off = dv;
      R.ldl = m > 0 ? m + (p->mixed ? 16 : 8) : 0;
      R.lbs = R.ldl * m;
      p->runs.push_back(R);
    }
    hodlr::LeafRun& R = p->runs.back();
    R.count += 1;
    row0 += m;
    lu += R.lbs;
    dv += m > 0 ? hodlr::dinv_elems(m) : 0;
  }
  if (p->runs.size() == 1) {
    p->ldl = p->runs[0].ldl;
    p->lbs = p->runs[0].lbs;
  }
  int rc = HODLR_OK;
  rc = rc ? rc : palloc(p, &p->leaf_ipiv, std::max<int64_t>(p->N, 1));
  rc = rc ? rc : palloc(p, &p->leaf_perm, std::max<int64_t>(p->N, 1));
  rc = rc ? rc : palloc(p, &p->leaf_info, p->P);
  rc = rc ? rc : palloc(p, &p->leaf_retry, p->P);
  if (p->mixed) {
    rc = rc ? rc : palloc(p, &p->LU32, lu);
    rc = rc ? rc : palloc(p, &p->leaf_dinv32, dv);
  } else {
    rc = rc ? rc : palloc(p, &p->LU, lu);
    rc = rc ? rc : palloc(p, &p->leaf_dinv, dv);
  }
  // empty leaf slots never take a factorization: their INFO stays 0
  if (!rc && p->runs.size() > 1) {
    std::vector<int> zero((size_t)p->P, 0);
    if (hipMemcpy(p->leaf_info, zero.data(), sizeof(int) * zero.size(), hipMemcpyHostToDevice) !=
        hipSuccess)
      rc = HODLR_ERR_HIP;
  }
  rc = rc ? rc : palloc(p, &p->Y, p->L > 0 ? p->N * r : 0);
  int64_t so = 0, po = 0, dof = 0, io = 0;
  for (int t = 0; t < p->L; ++t) {
    const int64_t nn = (int64_t)1 << t;
    p->s_off.push_back(so); p->p_off.push_back(po); p->d_off.push_back(dof); p->i_off.push_back(io);
    so += nn * n2 * n2; po += nn * n2; dof += nn * hodlr::dinv_elems(n2); io += nn;
  }
  rc = rc ? rc : palloc(p, &p->coreS, so);
  rc = rc ? rc : palloc(p, &p->core_ipiv, po);
  rc = rc ? rc : palloc(p, &p->core_perm, po);
  rc = rc ? rc : palloc(p, &p->core_info, io > 0 ? io : 1);
  rc = rc ? rc : palloc(p, &p->core_dinv, dof);
  rc = rc ? rc : palloc(p, &p->core_general, p->L > 0 ? p->P - 1 : 0);
  rc = rc ? rc : palloc(p, &p->core_tmask, p->L > 0 ? ((int64_t)1 << (p->L - 1)) : 0);
  if (p->L > 0 && r > 64 && r <= 128) {
    const int64_t nn = (int64_t)1 << (p->L - 1);
    rc = rc ? rc : palloc(p, &p->schur.C, nn * hodlr::kCoreSchurElems);
    rc = rc ? rc : palloc(p, &p->schur.LU, nn * hodlr::kCoreSchurElems);
    rc = rc ? rc : palloc(p, &p->schur.dinv, nn * hodlr::dinv_elems(128));
    rc = rc ? rc : palloc(p, &p->schur.ipiv, nn * 128);
    rc = rc ? rc : palloc(p, &p->schur.perm, nn * 128);
    rc = rc ? rc : palloc(p, &p->schur.info, nn);
    rc = rc ? rc : palloc(p, &p->schur.retry, nn);
  }
  // per-leaf projections / accumulators, core solutions and per-node sums
  // for the widest operand: [Y | B] with nrhs_max columns, or a solve chunk
  // of that width
  p->ncmax = r + (p->smax > 0 ? p->smax : 1);
  const int64_t pe = p->L > 0 ? p->P * r * p->ncmax : 0;
  const int64_t ee = p->L > 0 ? (p->P - 1) * n2 * p->ncmax : 0;
  const int64_t he = p->L > 0 ? (2 * p->P - 1) * r * p->ncmax : 0;
  rc = rc ? rc : palloc(p, &p->Pb, pe);
  rc = rc ? rc : palloc(p, &p->G, pe);
  rc = rc ? rc : palloc(p, &p->E, ee);
  rc = rc ? rc : palloc(p, &p->EB, ee);
  rc = rc ? rc : palloc(p, &p->Hb, he);
  if (rc) {
    plan_free(p);
    delete p;
    return rc;
  }
  *plan = p;
  return HODLR_OK;
}

static int check_common(const hodlr_desc& d) {
  if (d.rank <= 0 || d.levels < 0 || d.levels > 24 || d.nrhs_max < 0) return HODLR_ERR_ARG;
  if (d.precision != HODLR_FP64 && d.precision != HODLR_MIXED_IR) return HODLR_ERR_ARG;
  if (2 * d.rank > hodlr::kMaxBlock) return HODLR_ERR_UNSUPPORTED;
  return HODLR_OK;
}

// One leaf size: 0 (empty slot) or above the rank (matrix.cc:225) and
// within what the leaf kernels handle.
static int check_leaf(const hodlr_desc& d, int m) {
  if (m < 0) return HODLR_ERR_ARG;
  if (m == 0) return HODLR_OK;
  if (d.levels > 0 && m <= d.rank) return HODLR_ERR_SHAPE;
  if (m > hodlr::kMaxBlock || (d.precision == HODLR_MIXED_IR && m > 512))
    return HODLR_ERR_UNSUPPORTED;
  return HODLR_OK;
}

int hodlr_plan_create(hodlr_plan* plan, const hodlr_desc* desc) {
  if (!plan || !desc) return HODLR_ERR_ARG;
  *plan = nullptr;
  const hodlr_desc& d = *desc;
  if (d.leaf <= 0 || (d.flags & ~HODLR_PLAN_TRAILING_EMPTY)) return HODLR_ERR_ARG;
  int rc = check_common(d);
  if (rc) return rc;
  const int64_t slots = (int64_t)1 << d.levels;
  int64_t nleaf = slots;
  if (d.flags & HODLR_PLAN_TRAILING_EMPTY) {
    if (d.n <= 0 || d.n % d.leaf || d.n / d.leaf > slots) return HODLR_ERR_SHAPE;
    nleaf = d.n / d.leaf;
  } else if (d.n != ((int64_t)d.leaf << d.levels)) {
    return HODLR_ERR_SHAPE;
  }
  if ((rc = check_leaf(d, d.leaf))) return rc;
  std::vector<int> sizes((size_t)slots, 0);
  std::fill(sizes.begin(), sizes.begin() + nleaf, d.leaf);
  return plan_create_sizes(plan, d, sizes);
}

int hodlr_plan_create_uneven(hodlr_plan* plan, const hodlr_desc* desc, const int* leaf_sizes) {
  if (!plan || !desc || !leaf_sizes) return HODLR_ERR_ARG;
  *plan = nullptr;
  hodlr_desc d = *desc;
  if (d.flags) return HODLR_ERR_ARG;
  int rc = check_common(d);
  if (rc) return rc;
  const int64_t slots = (int64_t)1 << d.levels;
  std::vector<int> sizes(leaf_sizes, leaf_sizes + slots);
  int64_t n = 0;
  int mmax = 0;
  for (int m : sizes) {
    if ((rc = check_leaf(d, m))) return rc;
    n += m;
    mmax = std::max(mmax, m);
  }
  if (mmax == 0 || d.n != n) return HODLR_ERR_SHAPE;
  d.leaf = mmax;
  return plan_create_sizes(plan, d, sizes);
}

int hodlr_plan_destroy(hodlr_plan plan) {
  if (!plan) return HODLR_ERR_ARG;
  DeviceGuard g(plan->d.device);
  (void)hipDeviceSynchronize();
  plan_free(plan);
  delete plan;
  return HODLR_OK;
}

int64_t hodlr_plan_device_bytes(hodlr_plan plan) { return plan ? plan->bytes : -1; }

int hodlr_factor(hodlr_plan p, const double* K, int64_t ldk, const double* U, int64_t ldu,
                 const double* V, int64_t ldv, hipStream_t st) {
  if (!p || !K || !U || !V) return HODLR_ERR_ARG;
  if (check_ld(ldk, p->N) || check_ld(ldu, p->N) || check_ld(ldv, p->N)) return HODLR_ERR_ARG;
  DeviceGuard g(p->d.device);
  p->K = K; p->U = U; p->V = V; p->ldk = ldk; p->ldu = ldu; p->ldv = ldv;
  p->factored = false;
  drop_graph(p);
  int rc = leaf_phase(p, K, ldk, U, ldu, nullptr, 0, 0, st);
  if (!rc) rc = level_phase(p, nullptr, 0, 0, st);
  if (!rc) p->factored = true;
  return rc;
}

int hodlr_factor_solve(hodlr_plan p, const double* K, int64_t ldk, const double* U, int64_t ldu,
                       const double* V, int64_t ldv, double* B, int64_t ldb, int nrhs,
                       hipStream_t st) {
  if (!p || !K || !U || !V || (nrhs > 0 && !B) || nrhs < 0) return HODLR_ERR_ARG;
  if (check_ld(ldk, p->N) || check_ld(ldu, p->N) || check_ld(ldv, p->N)) return HODLR_ERR_ARG;
  if (nrhs > 0 && check_ld(ldb, p->N)) return HODLR_ERR_ARG;
  if (nrhs > p->smax) return HODLR_ERR_ARG;
  DeviceGuard g(p->d.device);
  p->K = K; p->U = U; p->V = V; p->ldk = ldk; p->ldu = ldu; p->ldv = ldv;
  p->factored = false;
  drop_graph(p);
  int rc = leaf_phase(p, K, ldk, U, ldu, B, ldb, nrhs, st);
  if (!rc) rc = level_phase(p, B, ldb, nrhs, st);
  if (!rc) p->factored = true;
  return rc;
}

int hodlr_solve(hodlr_plan p, double* B, int64_t ldb, int nrhs, hipStream_t st) {
  if (!p || !B || nrhs < 0) return HODLR_ERR_ARG;
  if (!p->factored) return HODLR_ERR_STATE;
  if (check_ld(ldb, p->N)) return HODLR_ERR_ARG;
  DeviceGuard g(p->d.device);
  const int r = p->r, L = p->L;
  for (int c0 = 0; c0 < nrhs; c0 += p->ncmax) {
    const int nc = std::min(p->ncmax, nrhs - c0);
    double* Bc = B + (int64_t)c0 * ldb;
    ColSeg x = hodlr::seg_single(Bc, ldb);
    int rc0 = leaf_solve(p, x, nc, false, st);
    if (rc0) return rc0;
    if (L == 0) continue;
    // p_i = V_i^T B_i; each level's right-hand sides from the level below's
    // core solutions (node-sum form), then G_i root first
    int rc = leaf_projections(p, p->V, p->ldv, x, nc, p->Pb, st);
    if (rc) return rc;
    const int n2 = 2 * r;
    const int64_t bs = (int64_t)n2 * nc;
    bool assembled = false;  // this level's right-hand sides formed by the solve below
    for (int t = L - 1; t >= 0; --t) {
      const int nnode = 1 << t;
      CoreBufs c = core_level(p, t);
      double* Rt = p->EB + heap0(t) * bs;
      if (t > 0) c.PR = p->EB + heap0(t - 1) * bs;
      bool up = false;
      c.assembled = &up;
      if (t == L - 1)
        HIP_TRY(hodlr::launch_assemble(p->Pb, 1, r, nc, 0, 0, nc, nnode, nullptr, Rt, st));
      else if (!assembled)
        HIP_TRY(hodlr::launch_assemble_pair(p->EB + heap0(t + 1) * bs, bs, r, 0, 0, nc, nnode,
                                            nullptr, Rt, bs, st));
      rc = core_solve(nnode, n2, c, Rt, bs, nc, st);
      if (rc) return rc;
      assembled = up;
    }
    rc = path_down(p, p->E, (int64_t)n2 * p->nq, p->EB, bs, nc, st);
    if (rc) return rc;
    rc = final_update_leaf(p, p->G, nc, Bc, ldb, st);
    if (rc) return rc;
  }
  return HODLR_OK;
}

// Transpose of hodlr_solve's operator (node-sum form), then K^-T.  With
// Hb_i = Y_i^T g_i at the leaves: bottom-up Hb_n = sum_c (I - E^Y_{n,c})^T
// Hb_c (the transpose of the top-down pass); top-down each node's core
// takes [Hb_c0 + Tb_n; Hb_c1 + Tb_n] (Tb_n: its half of the parent's
// transposed core solve) through S_n^-T; the leaves' halves Tb_i give
// g_i -= V_i Tb_i.
int hodlr_solve_adjoint(hodlr_plan p, double* G, int64_t ldg, int nrhs, hipStream_t st) {
  if (!p || !G || nrhs < 0) return HODLR_ERR_ARG;
  if (!p->factored) return HODLR_ERR_STATE;
  if (check_ld(ldg, p->N)) return HODLR_ERR_ARG;
  DeviceGuard g(p->d.device);
  const int r = p->r, L = p->L, n2 = 2 * r;
  const int64_t es = (int64_t)n2 * p->nq;
  for (int c0 = 0; c0 < nrhs; c0 += p->ncmax) {
    const int nc = std::min(p->ncmax, nrhs - c0);
    double* Gc = G + (int64_t)c0 * ldg;
    ColSeg x = hodlr::seg_single(Gc, ldg);
    if (L > 0) {
      const int64_t hs = (int64_t)r * nc, bs = (int64_t)n2 * nc;
      int rc = leaf_projections(p, p->Y, p->N, x, nc, p->Hb + heap0(L) * hs, st);
      if (rc) return rc;
      for (int t = L - 1; t >= 1; --t)
        HIP_TRY(hodlr::launch_adj_up(p->E + heap0(t) * es, es, p->Hb + heap0(t + 1) * hs, r, nc,
                                     1 << t, p->Hb + heap0(t) * hs, st));
      for (int t = 0; t < L; ++t) {
        const int nnode = 1 << t;
        CoreBufs c = core_level(p, t);
        double* Rt = p->EB + heap0(t) * bs;
        HIP_TRY(hodlr::launch_adj_assemble(p->Hb + heap0(t + 1) * hs,
                                           t > 0 ? p->EB + heap0(t - 1) * bs : nullptr, bs, r, nc,
                                           nnode, Rt, bs, st));
        ColSeg rx = hodlr::seg_single(Rt, n2);
        HIP_TRY(hodlr::launch_getrs_t(n2, c.S, n2, (int64_t)n2 * n2, nnode, c.perm, c.dinv, rx,
                                      bs, nc, st));
      }
      // leaf c's Tb: rows [r, 2r) of its parent's solution for an even c,
      // [0, r) for an odd one
      rc = final_update(p, p->V, p->ldv, p->EB + heap0(L - 1) * bs, n2, bs, r, 1, nc, Gc, ldg, st);
      if (rc) return rc;
    }
    int rc1 = leaf_solve(p, x, nc, true, st);
    if (rc1) return rc1;
  }
  return HODLR_OK;
}

// R <- A^-1 R for the refinement's residual block: replays a captured HIP
// graph of hodlr_solve when the stream can be captured (not the legacy null
// stream, profiling off), else runs it eagerly.
static int correction_solve(hodlr_plan p, double* R, int64_t ldr, int nrhs, hipStream_t st) {
  if (p->ir_exec && (p->ir_exec_R != R || p->ir_exec_nrhs != nrhs)) drop_graph(p);
  if (!p->ir_exec && st != nullptr && !g_prof.on) {
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
      const int rc = hodlr_solve(p, R, ldr, nrhs, st);
      hipGraph_t graph = nullptr;
      const hipError_t e = hipStreamEndCapture(st, &graph);
      if (rc == HODLR_OK && e == hipSuccess && graph) {
        if (hipGraphInstantiate(&p->ir_exec, graph, nullptr, nullptr, 0) != hipSuccess)
          p->ir_exec = nullptr;
      }
      if (graph) (void)hipGraphDestroy(graph);
      (void)hipGetLastError();
      if (p->ir_exec) {
        p->ir_exec_R = R;
        p->ir_exec_nrhs = nrhs;
      }
    } else {
      (void)hipGetLastError();
    }
  }
  if (p->ir_exec) {
    HIP_TRY(hipGraphLaunch(p->ir_exec, st));
    return HODLR_OK;
  }
  return hodlr_solve(p, R, ldr, nrhs, st);
}

int hodlr_solve_ir(hodlr_plan p, const double* K, int64_t ldk, const double* U, int64_t ldu,
                   const double* V, int64_t ldv, const double* B, int64_t ldb, double* X,
                   int64_t ldx, int nrhs, double tol, int maxit, int flags,
                   hodlr_ir_stats* stats, hipStream_t st) {
  if (!p || !K || !U || !V || !B || !X || nrhs < 0 || maxit < 0 || maxit > 32 || !(tol >= 0.0))
    return HODLR_ERR_ARG;
  if (flags & ~HODLR_IR_X_GIVEN) return HODLR_ERR_ARG;
  if (!p->factored) return HODLR_ERR_STATE;
  const int64_t N = p->N;
  if (check_ld(ldk, N) || check_ld(ldu, N) || check_ld(ldv, N) || check_ld(ldb, N) ||
      check_ld(ldx, N))
    return HODLR_ERR_ARG;
  DeviceGuard g(p->d.device);
  if (stats) memset(stats, 0, sizeof(*stats));
  if (nrhs == 0) {
    if (stats) stats->converged = 1;
    return HODLR_OK;
  }
  // residual block (N x nrhs), matvec workspace, norms of B and R
  const int64_t mvw = hodlr::matvec_work_elems(p->P, 1, p->r, nrhs);
  const int64_t need = N * nrhs + mvw + 2 * (int64_t)nrhs + (int64_t)hodlr::kColmaxParts * nrhs;
  if (p->ir_elems < need) {
    HIP_TRY(hipStreamSynchronize(st));
    if (p->ir_buf) HIP_TRY(hipFree(p->ir_buf));
    p->ir_buf = nullptr;
    p->ir_elems = 0;
    if (hipMalloc(&p->ir_buf, sizeof(double) * (size_t)need) != hipSuccess) {
      snprintf(g_detail, sizeof g_detail, "refinement buffer of %lld doubles", (long long)need);
      return HODLR_ERR_NOMEM;
    }
    p->ir_elems = need;
  }
  double* R = p->ir_buf;
  double* work = R + N * nrhs;
  double* nb = work + mvw;
  double* nr = nb + nrhs;
  double* cpart = nr + nrhs;  // partial column maxima
  std::vector<double> hb(nrhs), hr(nrhs);
  HIP_TRY(hodlr::launch_colmax(N, nrhs, B, ldb, nb, cpart, st));
  HIP_TRY(hipMemcpyAsync(hb.data(), nb, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
  int rc = HODLR_OK;
  if (!(flags & HODLR_IR_X_GIVEN)) {  // X = solve(B)
    HIP_TRY(hodlr::launch_copy_cols(N, nrhs, B, ldb, X, ldx, st));
    rc = hodlr_solve(p, X, ldx, nrhs, st);
    if (rc) return rc;
  }
  for (int it = 0;; ++it) {
    // R = B - A X
    HIP_TRY(hodlr::launch_copy_cols(N, nrhs, B, ldb, R, N, st));
    HIP_TRY(hodlr::launch_matvec_runs(p->runs.data(), (int)p->runs.size(), p->P, p->r, nrhs, K,
                                      ldk, U, ldu, V, ldv, X, ldx, -1.0, 1.0, R, N, work, st));
    HIP_TRY(hodlr::launch_colmax(N, nrhs, R, N, nr, cpart, st));
    HIP_TRY(hipMemcpyAsync(hr.data(), nr, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    double res = 0.0;
    for (int j = 0; j < nrhs; ++j) {
      const double q = hb[j] > 0.0 ? hr[j] / hb[j] : hr[j];
      res = (q > res || q != q) ? q : res;
    }
    if (stats) {
      stats->history[it] = res;
      stats->residual = res;
      stats->iterations = it;
    }
    if (res <= tol) {
      if (stats) stats->converged = 1;
      break;
    }
    if (it == maxit || res != res) break;
    rc = correction_solve(p, R, N, nrhs, st);
    if (rc) return rc;
    HIP_TRY(hodlr::launch_add_cols(N, nrhs, R, N, X, ldx, st));
  }
  return HODLR_OK;
}

int hodlr_grad(int64_t n, int leaf, int rank, int nrhs, const double* U, int64_t ldu,
               const double* V, int64_t ldv, const double* X, int64_t ldx, const double* Lam,
               int64_t ldl, double* dK, int64_t lddk, double* dU, int64_t lddu, double* dV,
               int64_t lddv, hipStream_t st) {
  if (n <= 0 || leaf <= 0 || rank <= 0 || nrhs < 0 || n % leaf) return HODLR_ERR_ARG;
  if (!X || !Lam || ldx < n || ldl < n) return HODLR_ERR_ARG;
  if ((dU || dV) && (!U || !V || ldu < n || ldv < n)) return HODLR_ERR_ARG;
  if ((dK && lddk < n) || (dU && lddu < n) || (dV && lddv < n)) return HODLR_ERR_ARG;
  if (nrhs == 0) {
    if (dK) HIP_TRY(hodlr::launch_fill_scale(n, leaf, dK, lddk, 0, 0.0, st));
    if (dU) HIP_TRY(hodlr::launch_fill_scale(n, rank, dU, lddu, 0, 0.0, st));
    if (dV) HIP_TRY(hodlr::launch_fill_scale(n, rank, dV, lddv, 0, 0.0, st));
    return HODLR_OK;
  }
  if (dK)
    HIP_TRY(hodlr::launch_leaf_outer(n, leaf, leaf, nrhs, -1.0, Lam, ldl, X, ldx, nullptr, nullptr,
                                     dK, lddk, st));
  if (!dU && !dV) return HODLR_OK;
  const int64_t pe = (n / leaf + 1) * (int64_t)rank * nrhs;
  void* w = nullptr;
  int rc = workspace(sizeof(double) * (size_t)(2 * pe), st, &w);
  if (rc) return rc;
  double* pw = static_cast<double*>(w);
  double* vx_part = pw;
  double* vx_tot = pw + (n / leaf) * (int64_t)rank * nrhs;
  double* ul_part = pw + pe;
  double* ul_tot = ul_part + (n / leaf) * (int64_t)rank * nrhs;
  if (dU) {
    HIP_TRY(hodlr::launch_leaf_projections(n, leaf, rank, nrhs, V, ldv, X, ldx, vx_part, vx_tot,
                                           st));
    HIP_TRY(hodlr::launch_leaf_outer(n, leaf, rank, nrhs, -1.0, Lam, ldl, nullptr, 0, vx_part,
                                     vx_tot, dU, lddu, st));
  }
  if (dV) {
    HIP_TRY(hodlr::launch_leaf_projections(n, leaf, rank, nrhs, U, ldu, Lam, ldl, ul_part, ul_tot,
                                           st));
    HIP_TRY(hodlr::launch_leaf_outer(n, leaf, rank, nrhs, -1.0, X, ldx, nullptr, 0, ul_part,
                                     ul_tot, dV, lddv, st));
  }
  return HODLR_OK;
}

int hodlr_plan_status(hodlr_plan p, hipStream_t st, int* where, int* level) {
  if (!p) return HODLR_ERR_ARG;
  DeviceGuard g(p->d.device);
  HIP_TRY(hipStreamSynchronize(st));
  if (where) *where = -1;
  if (level) *level = -1;
  std::vector<int> li(p->P), ci;
  HIP_TRY(hipMemcpy(li.data(), p->leaf_info, sizeof(int) * p->P, hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < p->P; ++i)
    if (li[i]) {
      if (where) *where = (int)i;
      if (level) *level = p->L;
      return HODLR_SINGULAR_LEAF;
    }
  for (int t = 0; t < p->L; ++t) {
    const int nn = 1 << t;
    ci.resize(nn);
    HIP_TRY(hipMemcpy(ci.data(), p->core_info + p->i_off[t], sizeof(int) * nn,
                      hipMemcpyDeviceToHost));
    for (int k = 0; k < nn; ++k)
      if (ci[k]) {
        if (where) *where = k;
        if (level) *level = t;
        return HODLR_SINGULAR_CORE;
      }
  }
  return HODLR_OK;
}

static int matvec_impl(bool trans, int64_t n, int leaf, int rank, int nrhs, const double* K,
                       int64_t ldk, const double* U, int64_t ldu, const double* V, int64_t ldv,
                       const double* X, int64_t ldx, double alpha, double beta, double* Y,
                       int64_t ldy, hipStream_t st) {
  if (n <= 0 || leaf <= 0 || rank <= 0 || nrhs < 0 || n % leaf) return HODLR_ERR_ARG;
  if (!K || !U || !V || !X || !Y) return HODLR_ERR_ARG;
  if (ldk < n || ldu < n || ldv < n || ldx < n || ldy < n) return HODLR_ERR_ARG;
  if (nrhs == 0) return HODLR_OK;
  void* w = nullptr;
  int rc = workspace(sizeof(double) * (size_t)hodlr::matvec_work_elems(n, leaf, rank, nrhs), st, &w);
  if (rc) return rc;
  HIP_TRY(hodlr::launch_matvec(n, leaf, rank, nrhs, K, ldk, U, ldu, V, ldv, X, ldx, alpha, beta, Y,
                               ldy, static_cast<double*>(w), st, trans));
  return HODLR_OK;
}

int hodlr_plan_matvec(hodlr_plan p, int trans, int nrhs, const double* K, int64_t ldk,
                      const double* U, int64_t ldu, const double* V, int64_t ldv, const double* X,
                      int64_t ldx, double alpha, double beta, double* Y, int64_t ldy,
                      hipStream_t st) {
  if (!p || nrhs < 0 || !K || !U || !V || !X || !Y) return HODLR_ERR_ARG;
  const int64_t n = p->N;
  if (ldk < n || ldu < n || ldv < n || ldx < n || ldy < n) return HODLR_ERR_ARG;
  if (nrhs == 0) return HODLR_OK;
  DeviceGuard g(p->d.device);
  void* w = nullptr;
  int rc = workspace(sizeof(double) * (size_t)hodlr::matvec_work_elems(p->P, 1, p->r, nrhs), st,
                     &w);
  if (rc) return rc;
  HIP_TRY(hodlr::launch_matvec_runs(p->runs.data(), (int)p->runs.size(), p->P, p->r, nrhs, K, ldk,
                                    U, ldu, V, ldv, X, ldx, alpha, beta, Y, ldy,
                                    static_cast<double*>(w), st, trans != 0));
  return HODLR_OK;
}

int hodlr_matvec(int64_t n, int leaf, int rank, int nrhs, const double* K, int64_t ldk,
                 const double* U, int64_t ldu, const double* V, int64_t ldv, const double* X,
                 int64_t ldx, double alpha, double beta, double* Y, int64_t ldy, hipStream_t st) {
  return matvec_impl(false, n, leaf, rank, nrhs, K, ldk, U, ldu, V, ldv, X, ldx, alpha, beta, Y,
                     ldy, st);
}

int hodlr_matvec_noalloc(int64_t n, int leaf, int rank, int nrhs, const double* K, int64_t ldk,
                         const double* U, int64_t ldu, const double* V, int64_t ldv,
                         const double* X, int64_t ldx, double alpha, double beta, double* Y,
                         int64_t ldy, hipStream_t st) {
  if (n <= 0 || leaf <= 0 || rank <= 0 || nrhs < 0 || n % leaf) return HODLR_ERR_ARG;
  if (!K || !U || !V || !X || !Y) return HODLR_ERR_ARG;
  if (ldk < n || ldu < n || ldv < n || ldx < n || ldy < n) return HODLR_ERR_ARG;
  const int64_t need = hodlr::matvec_work_elems(n, leaf, rank, nrhs);
  if (need > hodlr::kNoAllocScratchElems) {
    snprintf(g_detail, sizeof g_detail,
             "hodlr_matvec_noalloc needs (N/leaf + 1) * rank * nrhs = %lld doubles of scratch, "
             "more than its fixed %lld",
             (long long)need, (long long)hodlr::kNoAllocScratchElems);
    return HODLR_ERR_UNSUPPORTED;
  }
  if (nrhs == 0) return HODLR_OK;
  double* work = nullptr;
  HIP_TRY(hodlr::noalloc_scratch(&work));
  HIP_TRY(hodlr::launch_matvec(n, leaf, rank, nrhs, K, ldk, U, ldu, V, ldv, X, ldx, alpha, beta,
                               Y, ldy, work, st, false));
  return HODLR_OK;
}

int hodlr_matvec_t(int64_t n, int leaf, int rank, int nrhs, const double* K, int64_t ldk,
                   const double* U, int64_t ldu, const double* V, int64_t ldv, const double* X,
                   int64_t ldx, double alpha, double beta, double* Y, int64_t ldy,
                   hipStream_t st) {
  return matvec_impl(true, n, leaf, rank, nrhs, K, ldk, U, ldu, V, ldv, X, ldx, alpha, beta, Y,
                     ldy, st);
}

// ---------------------------------------------------------------------
// Reference-layout hsolve on the device (src/tasks/leaf_solve.cc:73-154).
// ---------------------------------------------------------------------
int hodlr_hsolve_ref(int nrow, int nrhs, int rank, int nPart, int LD, double* K, double* X,
                     const double* V, hipStream_t st) {
  if (nrow <= 0 || nrhs < 0 || rank <= 0 || nPart <= 0 || LD < nrow) return HODLR_ERR_ARG;
  if (!K || !X || !V) return HODLR_ERR_ARG;
  if (!is_pow2(nPart) || nrow % nPart) return HODLR_ERR_SHAPE;  // leaf_solve.cc:64,95-96
  const int Ls = ilog2(nPart);
  const int m = nrow / nPart;
  const int r = rank, n2 = 2 * r;
  if (m > hodlr::kMaxBlock || n2 > hodlr::kMaxBlock) return HODLR_ERR_UNSUPPORTED;
  const int wtot = nrhs + r * Ls;
  // workspace: leaf ipiv/perm/info/dinv, core S/ipiv/perm/info/dinv, R, part
  int64_t pe = 0, re = 0, se = 0, ce = 0;
  for (int t = 0; t < Ls; ++t) {
    const int64_t n_child = (int64_t)nrow >> (t + 1);
    const int w = nrhs + r * t;
    const int64_t chunks = ((int64_t)2 << t) * hodlr::reduce_chunks_per_child(n_child, reduce_rc(n_child));
    pe = std::max<int64_t>(pe, chunks * r * (w + r));
    re = std::max<int64_t>(re, ((int64_t)1 << t) * n2 * std::max(w, 1));
    se = std::max<int64_t>(se, ((int64_t)1 << t) * n2 * n2);
    ce = std::max<int64_t>(ce, ((int64_t)1 << t));
  }
  const int64_t dl = (int64_t)nPart * hodlr::dinv_elems(m);
  const int64_t dc = ce * hodlr::dinv_elems(n2);
  const int64_t ndbl = dl + pe + re + se + dc;
  // core INFO slots for every level (level t at offset 2^t - 1)
  const int64_t nint = (int64_t)nPart * m * 2 + nPart + ce * n2 * 2 + nPart + 64;
  void* w = nullptr;
  int rc = workspace(sizeof(double) * (size_t)ndbl + sizeof(int) * (size_t)nint, st, &w);
  if (rc) return rc;
  double* dp = static_cast<double*>(w);
  double* leaf_dinv = dp; dp += dl;
  double* part = dp; dp += pe;
  double* R = dp; dp += re;
  double* S = dp; dp += se;
  double* cdinv = dp; dp += dc;
  int* ip = reinterpret_cast<int*>(dp);
  int* leaf_ipiv = ip; ip += (int64_t)nPart * m;
  int* leaf_perm = ip; ip += (int64_t)nPart * m;
  int* leaf_info = ip; ip += nPart;
  int* cipiv = ip; ip += ce * n2;
  int* cperm = ip; ip += ce * n2;
  int* cinfo_all = ip; ip += nPart;
  // leaf: dgesv on all columns (leaf_solve.cc:80-92)
  HIP_TRY(hodlr::launch_getrf(m, K, LD, m, nPart, leaf_ipiv, leaf_perm, leaf_info, st));
  HIP_TRY(hodlr::launch_trtri_diag(m, K, LD, m, nPart, leaf_dinv, st));
  ColSeg xall = hodlr::seg_single(X, LD);
  if (wtot > 0)
    HIP_TRY(hodlr::launch_getrs(m, K, LD, m, nPart, leaf_perm, leaf_dinv, xall, m, wtot, st));
  // levels, deepest first (the recursion's post-order, level-synchronous)
  for (int t = Ls - 1; t >= 0; --t) {
    const int nchild = 1 << (t + 1), nnode = 1 << t;
    const int64_t n_child = (int64_t)nrow >> (t + 1);
    const int wd = nrhs + r * t;  // d = cols [0, wd), u = cols [wd, wd + r)
    int cpc = 0;
    HIP_TRY(hodlr::launch_reduce(V, LD, r, xall, wd + r, n_child, nchild, reduce_rc(n_child), part,
                                 &cpc, st));
    HIP_TRY(hodlr::launch_assemble(part, cpc, r, wd + r, wd, 0, wd, nnode, S, R, st));
    CoreBufs c{S, R, cipiv, cperm, cinfo_all + (nnode - 1), cdinv};
    rc = core_factor_solve(nnode, r, c, wd, st);
    if (rc) return rc;
    if (wd > 0)
      HIP_TRY(hodlr::launch_update(X + (int64_t)wd * LD, LD, r, R, n2, (int64_t)n2 * wd, 1, n_child,
                                   nchild, -1.0, xall, wd, st));
  }
  HIP_TRY(hipStreamSynchronize(st));
  std::vector<int> li(nPart);
  HIP_TRY(hipMemcpy(li.data(), leaf_info, sizeof(int) * nPart, hipMemcpyDeviceToHost));
  for (int i = 0; i < nPart; ++i)
    if (li[i]) return HODLR_SINGULAR_LEAF;
  if (nPart > 1) {
    std::vector<int> ci(nPart - 1);
    HIP_TRY(hipMemcpy(ci.data(), cinfo_all, sizeof(int) * (nPart - 1), hipMemcpyDeviceToHost));
    for (int k = 0; k < nPart - 1; ++k)
      if (ci[k]) return HODLR_SINGULAR_CORE;
  }
  return HODLR_OK;
}

// ---------------------------------------------------------------------
// LMatrix operators
// ---------------------------------------------------------------------
int hodlr_gemm_red(int64_t rows, int nchunk, int nslot, double alpha, const double* A, int64_t lda,
                   int ka, const double* B, int64_t ldb, int kb, double* C, int64_t ldc,
                   hipStream_t st) {
  if (rows <= 0 || nchunk <= 0 || nslot <= 0 || ka <= 0 || kb <= 0) return HODLR_ERR_ARG;
  if (rows % nchunk || nchunk % nslot) return HODLR_ERR_SHAPE;  // lmatrix.cc:725-726
  if (!A || !B || !C || lda < rows || ldb < rows || ldc < (int64_t)nslot * ka) return HODLR_ERR_ARG;
  const int64_t blk = rows / nchunk;
  const int64_t rc = blk < 512 ? blk : 512;
  const int64_t cpc = hodlr::reduce_chunks_per_child(blk, rc);
  void* w = nullptr;
  int rc2 = workspace(sizeof(double) * (size_t)(nchunk * cpc * (int64_t)ka * kb), st, &w);
  if (rc2) return rc2;
  double* part = static_cast<double*>(w);
  ColSeg x = hodlr::seg_single(const_cast<double*>(B), ldb);
  int cpc_out = 0;
  HIP_TRY(hodlr::launch_reduce(A, lda, ka, x, kb, blk, nchunk, rc, part, &cpc_out, st));
  HIP_TRY(hodlr::launch_sum_slots(part, (nchunk / nslot) * cpc_out, nslot, ka, kb, alpha, C, ldc,
                                  st));
  return HODLR_OK;
}

int hodlr_gemm_bro(int64_t rows, int nchunk, int nslot, double alpha, const double* A, int64_t lda,
                   int k, const double* E, int64_t lde, double* D, int64_t ldd, int nd,
                   hipStream_t st) {
  if (rows <= 0 || nchunk <= 0 || nslot <= 0 || k <= 0 || nd < 0) return HODLR_ERR_ARG;
  if (rows % nchunk || nchunk % nslot) return HODLR_ERR_SHAPE;  // lmatrix.cc:824-825
  if (!A || !E || !D || lda < rows || ldd < rows || lde < (int64_t)nslot * k) return HODLR_ERR_ARG;
  if (nd == 0) return HODLR_OK;
  const int64_t blk = rows / nchunk;
  ColSeg x = hodlr::seg_single(D, ldd);
  // slot s rows [s*k, s*k+k) of E  ==  (s/2)*(2k) + (s%2)*k
  HIP_TRY(hodlr::launch_update(A, lda, k, E, lde, (int64_t)2 * k, nchunk / nslot, blk, nchunk,
                               alpha, x, nd, st));
  return HODLR_OK;
}

int hodlr_node_solve(int nnode, int r, const double* VTu, int64_t ldu, double* VTd, int64_t ldd,
                     int nrhs, int* info, hipStream_t st) {
  if (nnode <= 0 || r <= 0 || nrhs < 0 || !VTu || !VTd || !info) return HODLR_ERR_ARG;
  if (ldu < (int64_t)2 * r * nnode || ldd < (int64_t)2 * r * nnode) return HODLR_ERR_ARG;
  const int n2 = 2 * r;
  if (n2 > hodlr::kMaxBlock) return HODLR_ERR_UNSUPPORTED;
  const int64_t nS = (int64_t)nnode * n2 * n2, nR = (int64_t)nnode * n2 * std::max(nrhs, 1);
  const int64_t nD = (int64_t)nnode * hodlr::dinv_elems(n2);
  void* w = nullptr;
  int rc = workspace(sizeof(double) * (size_t)(nS + nR + nD) + sizeof(int) * (size_t)(nnode * 2 * n2 + 16), st, &w);
  if (rc) return rc;
  double* S = static_cast<double*>(w);
  double* R = S + nS;
  double* dinv = R + nR;
  int* ipiv = reinterpret_cast<int*>(dinv + nD);
  int* perm = ipiv + (int64_t)nnode * n2;
  HIP_TRY(hodlr::launch_node_assemble(nnode, r, VTu, ldu, VTd, ldd, nrhs, S, R, st));
  CoreBufs c{S, R, ipiv, perm, info, dinv};
  rc = core_factor_solve(nnode, r, c, nrhs, st);
  if (rc) return rc;
  // eta (R, node-major 2r x nrhs) back into the VTd slots
  for (int k = 0; k < nnode; ++k)
    HIP_TRY(hodlr::launch_copy_cols(n2, nrhs, R + (int64_t)k * n2 * nrhs, n2,
                                    VTd + (int64_t)k * n2, ldd, st));
  return HODLR_OK;
}

int hodlr_node_solve_region(int nnode, int r0, int r1, int nrhs, const double* VTu0, int64_t ldu0,
                            int64_t su0, const double* VTu1, int64_t ldu1, int64_t su1,
                            double* VTd0, int64_t ldd0, int64_t sd0, double* VTd1, int64_t ldd1,
                            int64_t sd1, int* info, hipStream_t st) {
  if (nnode <= 0 || r0 <= 0 || r1 <= 0 || nrhs < 0 || !VTu0 || !VTu1 || !info ||
      (nrhs > 0 && (!VTd0 || !VTd1)))
    return HODLR_ERR_ARG;
  if (ldu0 < r1 || ldu1 < r0 || (nrhs > 0 && (ldd0 < r1 || ldd1 < r0))) return HODLR_ERR_ARG;
  if (nnode > 1 && (su0 < 0 || su1 < 0 || sd0 < 0 || sd1 < 0)) return HODLR_ERR_ARG;
  const int n = r0 + r1;
  if (n > hodlr::kMaxBlock) return HODLR_ERR_UNSUPPORTED;
  const int64_t nS = (int64_t)nnode * n * n, nR = (int64_t)nnode * n * std::max(nrhs, 1);
  const int64_t nD = (int64_t)nnode * hodlr::dinv_elems(n);
  void* w = nullptr;
  int rc = workspace(sizeof(double) * (size_t)(nS + nR + nD) +
                         sizeof(int) * (size_t)((int64_t)nnode * 2 * n + 16), st, &w);
  if (rc) return rc;
  double* S = static_cast<double*>(w);
  double* R = S + nS;
  double* dinv = R + nR;
  int* ipiv = reinterpret_cast<int*>(dinv + nD);
  int* perm = ipiv + (int64_t)nnode * n;
  HIP_TRY(hodlr::launch_region_assemble(nnode, r0, r1, nrhs, VTu0, ldu0, su0, VTu1, ldu1, su1,
                                        VTd0, ldd0, sd0, VTd1, ldd1, sd1, S, R, st));
  CoreBufs c{S, R, ipiv, perm, info, dinv};
  rc = core_factor_solve_n(nnode, n, c, nrhs, st);
  if (rc) return rc;
  HIP_TRY(hodlr::launch_region_scatter(nnode, r0, r1, nrhs, R, VTd0, ldd0, sd0, VTd1, ldd1, sd1,
                                       st));
  return HODLR_OK;
}

// ---------------------------------------------------------------------
// batched LAPACK
// ---------------------------------------------------------------------
int hodlr_dgetrf_batched(int n, double* A, int64_t lda, int64_t bstride, int batch, int* ipiv,
                         int* info, hipStream_t st) {
  if (n <= 0 || batch <= 0 || !A || !ipiv || !info || lda < n) return HODLR_ERR_ARG;
  if (n > hodlr::kMaxBlock) return HODLR_ERR_UNSUPPORTED;
  void* w = nullptr;
  int rc = workspace(sizeof(int) * (size_t)batch * n, st, &w);
  if (rc) return rc;
  HIP_TRY(hodlr::launch_getrf(n, A, lda, bstride, batch, ipiv, static_cast<int*>(w), info, st));
  return HODLR_OK;
}

int hodlr_dgetrf_batched_oop(int n, const double* K, int64_t ldk, int64_t kbs, double* LU,
                             int64_t ldl, int64_t lbs, int batch, int* ipiv, int* info,
                             hipStream_t st) {
  if (n <= 0 || batch <= 0 || !K || !LU || !ipiv || !info || ldk < n || ldl < n) return HODLR_ERR_ARG;
  if (n > hodlr::kMaxBlock) return HODLR_ERR_UNSUPPORTED;
  const int64_t nD = (int64_t)batch * hodlr::dinv_elems(n);
  void* w = nullptr;
  int rc = workspace(sizeof(double) * (size_t)nD + sizeof(int) * (size_t)batch * (n + 1) + 64, st, &w);
  if (rc) return rc;
  double* dinv = static_cast<double*>(w);
  int* perm = reinterpret_cast<int*>(dinv + nD);
  int* retry = perm + (int64_t)batch * n;
  HIP_TRY(hodlr::launch_getrf_leaf(n, K, ldk, kbs, LU, ldl, lbs, batch, dinv, ipiv, perm, info,
                                   retry, st));
  return HODLR_OK;
}

int hodlr_dgesv_batched(int n, int nrhs, double* A, int64_t lda, int64_t bstride, int batch,
                        int* ipiv, int* info, double* B, int64_t ldb, int64_t bbstride,
                        hipStream_t st) {
  if (n <= 0 || batch <= 0 || nrhs < 0 || !A || !ipiv || !info || lda < n ||
      (nrhs && (!B || ldb < n)))
    return HODLR_ERR_ARG;
  if (n > hodlr::kMaxBlock) return HODLR_ERR_UNSUPPORTED;
  const int64_t nD = (int64_t)batch * hodlr::dinv_elems(n);
  void* w = nullptr;
  int rc = workspace(sizeof(double) * (size_t)nD + sizeof(int) * (size_t)batch * n + 64, st, &w);
  if (rc) return rc;
  double* dinv = static_cast<double*>(w);
  int* perm = reinterpret_cast<int*>(dinv + nD);
  HIP_TRY(hodlr::launch_getrf(n, A, lda, bstride, batch, ipiv, perm, info, st));
  HIP_TRY(hodlr::launch_trtri_diag(n, A, lda, bstride, batch, dinv, st));
  if (nrhs > 0) {
    ColSeg x = hodlr::seg_single(B, ldb);
    HIP_TRY(hodlr::launch_getrs(n, A, lda, bstride, batch, perm, dinv, x, bbstride, nrhs, st));
  }
  return HODLR_OK;
}

// ---------------------------------------------------------------------
// generators
// ---------------------------------------------------------------------
int hodlr_generate(uint64_t seed, uint32_t stream_id, uint32_t system, int kind, int64_t rows,
                   int cols, double scale, double* out, int64_t ld, hipStream_t st) {
  if (rows < 0 || cols < 0 || (kind != 0 && kind != 1) || (rows * (int64_t)cols > 0 && !out) ||
      ld < rows)
    return HODLR_ERR_ARG;
  HIP_TRY(hodlr::launch_generate(seed, stream_id, system, kind, rows, cols, scale, out, ld, st));
  return HODLR_OK;
}

int hodlr_generate_ref(long master, int npart, int64_t rows_per_part, int cols, double offset,
                       double* out, int64_t ld, hipStream_t st) {
  if (npart <= 0 || rows_per_part <= 0 || cols <= 0 || !out) return HODLR_ERR_ARG;
  if (npart > 65535 || ld < (int64_t)npart * rows_per_part) return HODLR_ERR_ARG;
  HIP_TRY(hodlr::launch_ref_fill((uint64_t)master, npart, rows_per_part, cols, offset, out, ld,
                                 st));
  return HODLR_OK;
}

int hodlr_selftest_mfma(double* out, hipStream_t st) {
  if (!out) return HODLR_ERR_ARG;
  HIP_TRY(hodlr::launch_selftest_mfma(out, st));
  return HODLR_OK;
}

int hodlr_selftest_mfma_f32(double* out, hipStream_t st) {
  if (!out) return HODLR_ERR_ARG;
  HIP_TRY(hodlr::launch_selftest_mfma_f32(out, st));
  return HODLR_OK;
}

int hodlr_dense_blocks(int64_t n, int leaf, int rank, const double* U, int64_t ldu,
                       const double* V, int64_t ldv, const double* D, double offset, double* K,
                       int64_t ldk, hipStream_t st) {
  if (n <= 0 || leaf <= 0 || rank < 0 || n % leaf || !K || !D || ldk < n) return HODLR_ERR_ARG;
  if (rank > 0 && (!U || !V || ldu < n || ldv < n)) return HODLR_ERR_ARG;
  HIP_TRY(hodlr::launch_dense_blocks(n, leaf, rank, U, ldu, V, ldv, D, offset, K, ldk, st));
  return HODLR_OK;
}

int hodlr_shift_leaf_diagonal(int64_t n, int leaf, double shift, double* K, int64_t ldk,
                              hipStream_t st) {
  if (n <= 0 || leaf <= 0 || n % leaf || !K || ldk < n) return HODLR_ERR_ARG;
  HIP_TRY(hodlr::launch_shift_diag(n, leaf, shift, K, ldk, st));
  return HODLR_OK;
}

}  // extern "C"
