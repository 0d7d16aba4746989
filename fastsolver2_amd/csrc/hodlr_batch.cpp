// hodlr_batch.cpp -- many independent HODLR systems of uneven size, leaf
// size and rank, factored and solved together (include/hodlr.h, "batches").
//
// Each system keeps its own plan (the level-synchronous factor + solve of
// hodlr_capi.cpp).  hodlr_batch_factor_solve captures the whole batch once
// as ONE HIP graph: the caller's stream forks into `lanes` internal streams
// (default priority), systems are dealt to the lanes longest-first onto the
// least-loaded lane (estimated work m^2 N + r^2 N per system), and the lanes
// join back.  Later calls with the same operand pointers replay the graph, so
// the batch costs one graph launch instead of ~50 kernel launches per
// system, and independent systems overlap on the GPU.  hodlr_batch_tune is
// the autotuner: it times the replay for every (lanes, leaf-solve column
// tile) pair and keeps the fastest.
#include "../../include/hodlr.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "hodlr_internal.hpp"

struct hodlr_batch_s {
  int device = 0;
  std::vector<hodlr_desc> desc;
  std::vector<hodlr_plan> plans;
  int lanes = 8;
  int tile = 0;  // leaf-solve column tile (0 = by width)
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;  // [0] fork, [1 + l] join of lane l
  hipGraphExec_t exec = nullptr;
  std::vector<uint64_t> key;       // operands and knobs baked into exec
};

namespace {

constexpr int kMaxLanes = 16;

void drop_exec(hodlr_batch b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  b->exec = nullptr;
  b->key.clear();
}

void release(hodlr_batch b) {
  drop_exec(b);
  for (hipStream_t s : b->streams) (void)hipStreamDestroy(s);
  for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
  for (hodlr_plan p : b->plans)
    if (p) (void)hodlr_plan_destroy(p);
  b->streams.clear();
  b->events.clear();
  b->plans.clear();
}

struct Ops {
  double* const* K; const int64_t* ldk;
  const double* const* U; const int64_t* ldu;
  const double* const* V; const int64_t* ldv;
  double* const* B; const int64_t* ldb;
  const int* nrhs;
};

// systems longest-first onto the least-loaded lane
std::vector<int> assign_lanes(const std::vector<hodlr_desc>& d, int lanes) {
  const int n = (int)d.size();
  std::vector<double> w(n);
  for (int i = 0; i < n; ++i)
    w[i] = (double)d[i].n * ((double)d[i].leaf * d[i].leaf + (double)d[i].rank * d[i].rank);
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return w[a] > w[c]; });
  std::vector<double> load(lanes, 0.0);
  std::vector<int> lane(n, 0);
  for (int i : order) {
    const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    lane[i] = l;
    load[l] += w[i];
  }
  return lane;
}

// Enqueue every system's factor + solve, forked over the lanes and joined
// back onto `st` (captured or not).
int enqueue(hodlr_batch b, const Ops& o, hipStream_t st) {
  const int n = (int)b->plans.size();
  const int lanes = std::max(1, std::min(b->lanes, kMaxLanes));
  const std::vector<int> lane = assign_lanes(b->desc, lanes);
  if (hipEventRecord(b->events[0], st) != hipSuccess) return HODLR_ERR_HIP;
  for (int l = 0; l < lanes; ++l)
    if (hipStreamWaitEvent(b->streams[l], b->events[0], 0) != hipSuccess) return HODLR_ERR_HIP;
  const int saved_tile = hodlr::getrs_tile();
  hodlr::set_getrs_tile(b->tile);
  int rc = HODLR_OK;
  for (int i = 0; i < n && rc == HODLR_OK; ++i)
    rc = hodlr_factor_solve(b->plans[i], o.K[i], o.ldk[i], o.U[i], o.ldu[i], o.V[i], o.ldv[i],
                            o.B[i], o.ldb[i], o.nrhs[i], b->streams[lane[i]]);
  hodlr::set_getrs_tile(saved_tile);
  for (int l = 0; l < lanes; ++l) {
    if (hipEventRecord(b->events[1 + l], b->streams[l]) != hipSuccess) return HODLR_ERR_HIP;
    if (hipStreamWaitEvent(st, b->events[1 + l], 0) != hipSuccess) return HODLR_ERR_HIP;
  }
  return rc;
}

std::vector<uint64_t> make_key(hodlr_batch b, const Ops& o) {
  std::vector<uint64_t> k;
  const int n = (int)b->plans.size();
  k.reserve(9 * n + 2);
  k.push_back((uint64_t)b->lanes);
  k.push_back((uint64_t)(int64_t)b->tile);
  for (int i = 0; i < n; ++i) {
    k.push_back((uint64_t)(uintptr_t)o.K[i]); k.push_back((uint64_t)o.ldk[i]);
    k.push_back((uint64_t)(uintptr_t)o.U[i]); k.push_back((uint64_t)o.ldu[i]);
    k.push_back((uint64_t)(uintptr_t)o.V[i]); k.push_back((uint64_t)o.ldv[i]);
    k.push_back((uint64_t)(uintptr_t)o.B[i]); k.push_back((uint64_t)o.ldb[i]);
    k.push_back((uint64_t)(int64_t)o.nrhs[i]);
  }
  return k;
}

void unmark(hodlr_batch b) {
  for (hodlr_plan p : b->plans) hodlr::plan_mark_unfactored(p);
}
// a replayed graph holds each plan's complete factor + solve
void mark(hodlr_plan p) { hodlr::plan_mark_factored(p); }

struct DeviceScope {
  int prev = -1;
  explicit DeviceScope(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceScope() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

}  // namespace

extern "C" {

int hodlr_batch_create(hodlr_batch* out, int nsys, const hodlr_desc* descs) {
  if (!out || nsys <= 0 || !descs) return HODLR_ERR_ARG;
  *out = nullptr;
  for (int i = 1; i < nsys; ++i)
    if (descs[i].device != descs[0].device) return HODLR_ERR_ARG;
  DeviceScope ds(descs[0].device);
  hodlr_batch b = new hodlr_batch_s();
  b->device = descs[0].device;
  b->desc.assign(descs, descs + nsys);
  b->plans.assign(nsys, nullptr);
  int rc = HODLR_OK;
  for (int i = 0; i < nsys && rc == HODLR_OK; ++i) rc = hodlr_plan_create(&b->plans[i], &descs[i]);
  for (int l = 0; l < kMaxLanes && rc == HODLR_OK; ++l) {
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) rc = HODLR_ERR_HIP;
    else b->streams.push_back(s);
  }
  for (int e = 0; e <= kMaxLanes && rc == HODLR_OK; ++e) {
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) rc = HODLR_ERR_HIP;
    else b->events.push_back(ev);
  }
  if (rc) {
    release(b);
    delete b;
    return rc;
  }
  *out = b;
  return HODLR_OK;
}

int hodlr_batch_destroy(hodlr_batch b) {
  if (!b) return HODLR_ERR_ARG;
  DeviceScope ds(b->device);
  (void)hipDeviceSynchronize();
  release(b);
  delete b;
  return HODLR_OK;
}

int hodlr_batch_size(hodlr_batch b) { return b ? (int)b->plans.size() : -1; }

int hodlr_batch_factor_solve(hodlr_batch b, double* const* K, const int64_t* ldk,
                             const double* const* U, const int64_t* ldu, const double* const* V,
                             const int64_t* ldv, double* const* B, const int64_t* ldb,
                             const int* nrhs, hipStream_t stream) {
  if (!b || !K || !ldk || !U || !ldu || !V || !ldv || !B || !ldb || !nrhs) return HODLR_ERR_ARG;
  const int n = (int)b->plans.size();
  for (int i = 0; i < n; ++i) {
    if (!K[i] || !U[i] || !V[i] || nrhs[i] < 0 || (nrhs[i] > 0 && !B[i])) return HODLR_ERR_ARG;
    if (nrhs[i] > b->desc[i].nrhs_max) return HODLR_ERR_ARG;
  }
  DeviceScope ds(b->device);
  const Ops o{K, ldk, U, ldu, V, ldv, B, ldb, nrhs};
  // eager when the stream cannot be captured (legacy null stream) or the
  // per-launch profiler is recording events
  if (stream == nullptr || hodlr::profiling_enabled()) {
    drop_exec(b);
    return enqueue(b, o, stream);
  }
  std::vector<uint64_t> key = make_key(b, o);
  if (!b->exec || key != b->key) {
    drop_exec(b);
    if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipGetLastError();
      return enqueue(b, o, stream);
    }
    const int rc = enqueue(b, o, stream);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(stream, &graph);
    // the capture marked every plan factored, but nothing has run yet: until
    // the graph launches successfully no plan holds a factorization
    unmark(b);
    if (rc != HODLR_OK || e != hipSuccess || !graph) {
      if (graph) (void)hipGraphDestroy(graph);
      (void)hipGetLastError();
      return rc != HODLR_OK ? rc : HODLR_ERR_HIP;
    }
    const hipError_t ie = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
      b->exec = nullptr;
      (void)hipGetLastError();
      return HODLR_ERR_HIP;
    }
    b->key = std::move(key);
  }
  unmark(b);
  if (hipGraphLaunch(b->exec, stream) != hipSuccess) return HODLR_ERR_HIP;
  for (hodlr_plan p : b->plans) mark(p);
  return HODLR_OK;
}

int hodlr_batch_status(hodlr_batch b, hipStream_t stream, int* sys, int* where, int* level) {
  if (!b) return HODLR_ERR_ARG;
  DeviceScope ds(b->device);
  if (hipStreamSynchronize(stream) != hipSuccess) return HODLR_ERR_HIP;
  for (size_t i = 0; i < b->plans.size(); ++i) {
    int w = -1, l = -1;
    const int st = hodlr_plan_status(b->plans[i], b->streams[0], &w, &l);
    if (st != HODLR_OK) {
      if (sys) *sys = (int)i;
      if (where) *where = w;
      if (level) *level = l;
      return st;
    }
  }
  if (sys) *sys = -1;
  if (where) *where = -1;
  if (level) *level = -1;
  return HODLR_OK;
}

hodlr_plan hodlr_batch_plan(hodlr_batch b, int i) {
  if (!b || i < 0 || i >= (int)b->plans.size()) return nullptr;
  return b->plans[i];
}

int hodlr_batch_set_tuning(hodlr_batch b, int lanes, int tile) {
  if (!b || lanes < 1 || lanes > kMaxLanes) return HODLR_ERR_ARG;
  if (tile != 0 && tile != 1 && tile != 2 && tile != 5) return HODLR_ERR_ARG;
  b->lanes = lanes;
  b->tile = tile;
  return HODLR_OK;
}

int hodlr_batch_get_tuning(hodlr_batch b, int* lanes, int* tile) {
  if (!b) return HODLR_ERR_ARG;
  if (lanes) *lanes = b->lanes;
  if (tile) *tile = b->tile;
  return HODLR_OK;
}

int hodlr_batch_tune(hodlr_batch b, double* const* K, const int64_t* ldk, const double* const* U,
                     const int64_t* ldu, const double* const* V, const int64_t* ldv,
                     double* const* B, const int64_t* ldb, const int* nrhs, hipStream_t stream,
                     double* best_ms) {
  if (!b || stream == nullptr) return HODLR_ERR_ARG;
  DeviceScope ds(b->device);
  static const int kLanes[] = {1, 2, 4, 8, 16};
  static const int kTiles[] = {0, 1, 2, 5};
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess) return HODLR_ERR_HIP;
  double best = 1e300;
  int bl = b->lanes, bt = b->tile;
  int rc = HODLR_OK;
  for (int lanes : kLanes) {
    for (int tile : kTiles) {
      b->lanes = lanes;
      b->tile = tile;
      // the first call captures (and runs), the second is an untimed warm
      // replay (a graph's first launch uploads it), then the fastest of
      // two timed replays
      for (int w = 0; w < 2 && !rc; ++w)
        rc = hodlr_batch_factor_solve(b, K, ldk, U, ldu, V, ldv, B, ldb, nrhs, stream);
      if (rc) break;
      float ms = 1e30f;
      for (int rep = 0; rep < 2 && !rc; ++rep) {
        // from an idle stream, as a caller that waits for each batch sees
        // it (replays queued behind a running one overlap its tail and
        // favour fewer lanes)
        if (hipStreamSynchronize(stream) != hipSuccess) { rc = HODLR_ERR_HIP; break; }
        (void)hipEventRecord(t0, stream);
        rc = hodlr_batch_factor_solve(b, K, ldk, U, ldu, V, ldv, B, ldb, nrhs, stream);
        (void)hipEventRecord(t1, stream);
        if (rc) break;
        if (hipEventSynchronize(t1) != hipSuccess) { rc = HODLR_ERR_HIP; break; }
        float m = 0.f;
        (void)hipEventElapsedTime(&m, t0, t1);
        ms = m < ms ? m : ms;
      }
      if (rc) break;
      if (ms < best) { best = ms; bl = lanes; bt = tile; }
    }
    if (rc) break;
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  b->lanes = bl;
  b->tile = bt;
  drop_exec(b);
  if (best_ms) *best_ms = best;
  return rc;
}

}  // extern "C"
