// hodlr_kernels.hip -- gfx950 (MI355X) kernels of the HODLR factor-and-solve
// hot path.  FP64 throughout; matrix products use the CDNA4 FP64 matrix core
// instruction v_mfma_f64_16x16x4_f64 on 64-wide wavefronts.
//
// MFMA 16x16x4 f64 operand layout (lane l of a 64-lane wave):
//   A (16x4):  A[l % 16][l / 16]          one double per lane
//   B (4x16):  B[l / 16][l % 16]          one double per lane
//   C/D:       C[(l/16) + 4*v][l % 16]    v = 0..3, a 4-double vector
// (the f64 16x16 variant interleaves rows across the four lane groups;
// crow() below is the one place that encodes it, and
// hodlr_selftest_mfma checks it on the device).
//
// Reference algorithm: Charles-Chao-Chen/fastSolver2 hsolve
// (src/tasks/leaf_solve.cc:73-154) and the level loop gemmRed -> node_solve
// -> gemmBro (spmd_driver/solver_irregular_part.cc:117-140).  The kernels
// here are the batched, level-synchronous building blocks of that loop.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <mutex>
#include <utility>
#include <vector>

#include <cfloat>
#include <climits>
#include <cstdint>

#include "hodlr_internal.hpp"

#ifndef GETRF_NB
#define GETRF_NB 32
#endif

namespace hodlr {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
// One DPP row operation on a 64-bit value (two 32-bit halves).
template <int CTRL>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long v) {
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)v, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(v >> 32), CTRL, 0xF, 0xF, false);
  return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}
__device__ __forceinline__ unsigned long long umax64(unsigned long long a, unsigned long long b) {
  return a > b ? a : b;
}
__device__ __forceinline__ unsigned long long readlane_u64(unsigned long long v, int l) {
  const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)v, l);
  const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(v >> 32), l);
  return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ double readlane_f64(double v, int l) {
  return __longlong_as_double((long long)readlane_u64((unsigned long long)__double_as_longlong(v), l));
}
// Maximum over the 64 lanes: quad butterflies (xor 1, xor 2), half-row and
// row mirrors leave every lane of a 16-lane row with the row maximum; the
// four rows are combined through readlane.
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
  v = umax64(v, dpp_u64<0xB1>(v));   // quad_perm [1,0,3,2]
  v = umax64(v, dpp_u64<0x4E>(v));   // quad_perm [2,3,0,1]
  v = umax64(v, dpp_u64<0x141>(v));  // row_half_mirror
  v = umax64(v, dpp_u64<0x140>(v));  // row_mirror
  return umax64(umax64(readlane_u64(v, 0), readlane_u64(v, 16)),
                umax64(readlane_u64(v, 32), readlane_u64(v, 48)));
}

// 32-bit maximum over the wave, same butterfly; 0 is the identity, so the
// DPP moves fold into v_max_u32.
template <int CTRL> __device__ __forceinline__ unsigned dpp_u32(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
  v = max(v, dpp_u32<0xB1>(v));
  v = max(v, dpp_u32<0x4E>(v));
  v = max(v, dpp_u32<0x141>(v));
  v = max(v, dpp_u32<0x140>(v));
  return max(max((unsigned)__builtin_amdgcn_readlane((int)v, 0), (unsigned)__builtin_amdgcn_readlane((int)v, 16)),
             max((unsigned)__builtin_amdgcn_readlane((int)v, 32), (unsigned)__builtin_amdgcn_readlane((int)v, 48)));
}
// Lane of the largest 64-bit key, the smallest `q` among equal keys (0 if
// every key is 0); `key` gives the maximum.  Stage one reduces a 32-bit
// summary `k32` that is non-decreasing in the key (exact for keys below
// 2^32): a single lane at its maximum is the answer; otherwise the lanes at
// that maximum go through the full 64-bit reduction and the position rule.
__device__ __forceinline__ int wave_argmax_key(unsigned long long k, unsigned k32, int q,
                                               unsigned long long& key) {
  if (__ballot(k != 0ULL) == 0ULL) {  // no candidate in the wave
    key = 0ULL;
    return 0;
  }
  const unsigned h = wave_max_u32(k32);
  const unsigned long long bal = __ballot(k32 == h);
  if (__popcll(bal) == 1) {
    const int l = __ffsll((long long)bal) - 1;
    key = readlane_u64(k, l);
    return l;
  }
  const unsigned long long kk = (k32 == h) ? k : 0ULL;
  const unsigned long long m = wave_max_u64(kk);
  key = m;
  if (m == 0ULL) return 0;
  const unsigned long long b2 = __ballot(kk == m);
  if (__popcll(b2) == 1) return __ffsll((long long)b2) - 1;
  int qq = (kk == m) ? q : INT_MAX;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) qq = min(qq, __shfl_xor(qq, off, 64));
  return __ffsll((long long)__ballot(kk == m && q == qq)) - 1;
}

// The interchanges of a panel on the columns outside it.  dlaswp's
// sequence of nb swaps is one permutation of the rows from j0 on that moves
// at most 2*nb <= 64 of them: mv[k] = src | dst << 16 sends the row at
// position j0 + src to j0 + dst (k < M).  A wave takes whole columns, lane k
// moving row k: a column's load (all its moving rows) completes before its
// store, and U columns are in flight at once, so a column costs a fraction
// of one memory round trip instead of nb dependent ones.
template <int U, typename T>
__device__ __forceinline__ void move_rows(T* A, int64_t lda, int n, int j0, int nb, const int* mv,
                                          int M, int wid, int nw, int lane) {
  if (M == 0) return;
  const int m = lane < M ? mv[lane] : 0;
  const bool act = lane < M;
  const int64_t src = j0 + (m & 0xFFFF), dst = j0 + (m >> 16);
  const int ncol = n - nb;
  for (int c0 = wid; c0 < ncol; c0 += U * nw) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int c = c0 + u * nw;
      const int64_t col = c < j0 ? c : c + nb;
      v[u] = (act && c < ncol) ? A[col * lda + src] : T(0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int c = c0 + u * nw;
      const int64_t col = c < j0 ? c : c + nb;
      if (act && c < ncol) A[col * lda + dst] = v[u];
    }
  }
}

// row of accumulator element v held by `lane`
__device__ __forceinline__ int crow(int lane, int v) { return (lane >> 4) + 4 * v; }

__device__ __forceinline__ d4 d4zero() {
  d4 z = {0.0, 0.0, 0.0, 0.0};
  return z;
}

// Element-type traits of the LU kernels: FP64 (the reference's precision)
// and FP32 (the mixed-precision leaf factorization).  The FP32 matrix core
// (v_mfma_f32_16x16x4f32) lays its accumulator out as row 4*(lane/16) + v,
// the FP64 one as row lane/16 + 4v.
typedef float f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ f4 mfma(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
template <typename T> struct Elem;
template <> struct Elem<double> {
  typedef d4 V4;
  static constexpr int DV = 4;  // crow(lane, v + 1) - crow(lane, v)
  __device__ static int crow(int lane, int v) { return (lane >> 4) + 4 * v; }
  __device__ static double sfmin() { return DBL_MIN; }
  // bits(|a|) + 1 orders non-negative values; NaN -> 1 (counts as zero)
  __device__ static unsigned long long key(double a) {
    return (a == a) ? (unsigned long long)__double_as_longlong(fabs(a)) + 1ULL : 1ULL;
  }
  __device__ static unsigned key32(unsigned long long k) { return (unsigned)(k >> 32); }
  __device__ static double lane(double a, int l) { return readlane_f64(a, l); }
};
template <> struct Elem<float> {
  typedef f4 V4;
  static constexpr int DV = 1;
  __device__ static int crow(int lane, int v) { return 4 * (lane >> 4) + v; }
  __device__ static float sfmin() { return FLT_MIN; }
  __device__ static unsigned long long key(float a) {
    return (a == a) ? (unsigned long long)__float_as_uint(fabsf(a)) + 1ULL : 1ULL;
  }
  __device__ static unsigned key32(unsigned long long k) { return (unsigned)k; }  // k < 2^32
  __device__ static float lane(float a, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a), l)); }
};
template <typename T>
__device__ __forceinline__ typename Elem<T>::V4 zero4() {
  typename Elem<T>::V4 z = {T(0), T(0), T(0), T(0)};
  return z;
}

// Raw buffer access: a wave-uniform base in SGPRs and 32-bit byte offsets
// (lane part in a VGPR, wave part in an SGPR), so a tile of loads costs one
// offset register per lane instead of a 64-bit address per element.  The
// base must be wave-uniform; the range check is off (num_records = max).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)0xFFFFFFFFu, 0x00020000);
}
template <typename T> __device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so);
template <> __device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
}
template <> __device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
}
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
// An opaque copy: the compiler cannot see through it, so address arithmetic
// built on it stays in the loop body where the constant parts fold into the
// instructions' offset fields (instead of one hoisted register per offset
// held across the whole kernel).  Emits no instruction.
template <typename V> __device__ __forceinline__ V opaque(V x) {
  asm volatile("" : "+v"(x));
  return x;
}
// The same for a wave-uniform value kept in a scalar register (a buffer
// access's scalar offset is built on it inside the loop that uses it).
template <typename V> __device__ __forceinline__ V opaque_s(V x) {
  asm volatile("" : "+s"(x));
  return x;
}
// Workgroup barrier that orders LDS only: it waits for this wave's LDS
// operations, not for its global stores (which the workgroup fence of
// __syncthreads also covers).  For exchanges through LDS where no wave reads
// another's global writes before the next full barrier.  Neither barrier
// waits for global_load_lds copies into LDS: the issuing wave waits for
// those itself (s_waitcnt vmcnt) before the barrier that publishes them.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
__device__ __forceinline__ void buf_store(double x, __amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, vo, so, 0);
}
__device__ __forceinline__ void buf_store(float x, __amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, x), r, vo, so, 0);
}

__device__ __forceinline__ const double* seg_in(const ColSeg& s, int j) {
  return j < s.nA ? s.inA + (int64_t)j * s.ldiA : s.inB + (int64_t)(j - s.nA) * s.ldiB;
}
__device__ __forceinline__ double* seg_out(const ColSeg& s, int j) {
  return j < s.nA ? s.outA + (int64_t)j * s.ldoA : s.outB + (int64_t)(j - s.nA) * s.ldoB;
}

// =====================================================================
// Batched LU with partial pivoting, one workgroup (4 waves) per block.
// Right-looking, NB-column panels held in LDS; the panel is factored
// column by column (LAPACK dgetf2/dgetrf2 pivot rule: first row of largest
// |a|), interchanges are applied to the whole rows, U12 = L11^-1 A12 is an
// MFMA product with the inverse of L11, and the trailing update
// A22 -= L21 U12 runs on MFMA with L21 read from the LDS panel.
// =====================================================================
template <int NB>
__global__ void __launch_bounds__(256)
k_getrf(double* __restrict__ A0, int64_t lda, int64_t bstride, int n,
        int* __restrict__ ipiv0, int* __restrict__ perm0, int* __restrict__ info0) {
  extern __shared__ double lds[];
  double* pan = lds;                 // [NB][n], column c at pan + c*n
  double* prow = pan + NB * n;       // [NB] pivot row
  double* orow = prow + NB;          // [NB] row displaced by the pivot
  double* linv = orow + NB;          // [NB][NB+1] inverse of L11
  double* redv = linv + NB * (NB + 1);  // [4]
  int* redi = reinterpret_cast<int*>(redv + 4);  // [4]
  int* pivs = redi + 4;              // [NB] panel-relative pivot rows
  int* perm = pivs + NB;             // [n]

  const int b = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  double* A = A0 + (int64_t)b * bstride;
  int info = 0;  // meaningful in thread 0 only

  for (int i = tid; i < n; i += 256) perm[i] = i;

  for (int j0 = 0; j0 < n; j0 += NB) {
    const int nb = min(NB, n - j0);
    const int rows = n - j0;
    // 1. load the panel rows [j0, n) x cols [j0, j0+nb)
    for (int idx = tid; idx < nb * rows; idx += 256) {
      const int c = idx / rows, i = idx - c * rows;
      pan[c * n + i] = A[(int64_t)(j0 + i) + (int64_t)(j0 + c) * lda];
    }
    __syncthreads();

    // 2. unblocked factorization of the panel
    for (int jj = 0; jj < nb; ++jj) {
      double best = -1.0;
      int bi = INT_MAX;
      for (int i = jj + tid; i < rows; i += 256) {
        const double v = fabs(pan[jj * n + i]);
        if (v > best) { best = v; bi = i; }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(best, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
      }
      if (lane == 0) { redv[wid] = best; redi[wid] = bi; }
      __syncthreads();
      best = redv[0];
      bi = redi[0];
#pragma unroll
      for (int w = 1; w < 4; ++w) {
        const double ov = redv[w];
        const int oi = redi[w];
        if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
      }
      const int p = (bi < rows) ? bi : jj;
      if (tid < nb) {
        prow[tid] = pan[tid * n + p];
        orow[tid] = pan[tid * n + jj];
      }
      if (tid == 0) pivs[jj] = p;
      __syncthreads();
      const double pv = prow[jj];
      if (pv == 0.0) {
        // whole column is zero: p == jj, nothing to swap or eliminate
        if (tid == 0 && info == 0) info = j0 + jj + 1;
        continue;
      }
      const bool use_rcp = fabs(pv) >= DBL_MIN;
      const double rcp = 1.0 / pv;
      for (int i = jj + tid; i < rows; i += 256) {
        if (i == jj) {
          for (int c = 0; c < nb; ++c) pan[c * n + jj] = prow[c];
        } else {
          const bool isp = (i == p);
          double l = isp ? orow[jj] : pan[jj * n + i];
          l = use_rcp ? l * rcp : l / pv;
          pan[jj * n + i] = l;
          if (isp)
            for (int c = 0; c < jj; ++c) pan[c * n + i] = orow[c];
          for (int c = jj + 1; c < nb; ++c) {
            const double a = isp ? orow[c] : pan[c * n + i];
            pan[c * n + i] = a - l * prow[c];
          }
        }
      }
      __syncthreads();
    }

    // 3. write the panel back, record pivots, update the permutation
    for (int idx = tid; idx < nb * rows; idx += 256) {
      const int c = idx / rows, i = idx - c * rows;
      A[(int64_t)(j0 + i) + (int64_t)(j0 + c) * lda] = pan[c * n + i];
    }
    if (tid < nb) ipiv0[(int64_t)b * n + j0 + tid] = j0 + pivs[tid] + 1;  // 1-based
    if (tid == 0) {
      for (int jj = 0; jj < nb; ++jj) {
        const int p = pivs[jj];
        if (p != jj) {
          const int t = perm[j0 + jj];
          perm[j0 + jj] = perm[j0 + p];
          perm[j0 + p] = t;
        }
      }
    }

    // 4. interchanges on the other columns (sequential in jj, as dlaswp)
    for (int c = tid; c < n - nb; c += 256) {
      const int col = c < j0 ? c : c + nb;
      double* Ac = A + (int64_t)col * lda;
      for (int jj = 0; jj < nb; ++jj) {
        const int p = pivs[jj];
        if (p != jj) {
          const double t = Ac[j0 + jj];
          Ac[j0 + jj] = Ac[j0 + p];
          Ac[j0 + p] = t;
        }
      }
    }
    // 5a. linv = L11^-1 (unit lower, padded with the identity past nb):
    // lane c < NB solves L11 x = e_c.
    const int m2 = n - j0 - nb;
    if (m2 > 0 && tid < NB) {
      const int c = tid;
      double* xc = linv + c * (NB + 1);  // column c of the inverse
      for (int i = 0; i < NB; ++i) {
        double acc = (i == c) ? 1.0 : 0.0;
        if (i > c && i < nb)
          for (int k = c; k < i; ++k) acc -= pan[k * n + i] * xc[k];
        xc[i] = (i < c) ? 0.0 : acc;
      }
    }
    __syncthreads();
    // 5b. U12 = linv * A12 on MFMA; each wave owns whole 16-column tiles, so
    // it reads a tile's 32 source rows before overwriting them.
    if (m2 > 0) {
      const int o = j0 + nb;
      const int nct = (m2 + 15) >> 4;
      for (int ct = wid; ct < nct; ct += 4) {
        const int col = ct * 16 + (lane & 15);
        d4 u0 = d4zero(), u1 = d4zero();
#pragma unroll
        for (int kk = 0; kk < NB / 4; ++kk) {
          const int k = kk * 4 + (lane >> 4);
          const double bf = (k < nb && col < m2) ? A[(int64_t)(j0 + k) + (int64_t)(o + col) * lda] : 0.0;
          const double a0 = linv[(lane & 15) + k * (NB + 1)];
          u0 = mfma(a0, bf, u0);
          if (NB > 16) {
            const double a1 = linv[16 + (lane & 15) + k * (NB + 1)];
            u1 = mfma(a1, bf, u1);
          }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r0 = crow(lane, v);
          if (col < m2 && r0 < nb) A[(int64_t)(j0 + r0) + (int64_t)(o + col) * lda] = u0[v];
          if (NB > 16 && col < m2 && 16 + r0 < nb)
            A[(int64_t)(j0 + 16 + r0) + (int64_t)(o + col) * lda] = u1[v];
        }
      }
    }
    __syncthreads();

    // 6. trailing update A22 -= L21 * U12 (MFMA), 32x32 blocks per wave
    if (m2 > 0) {
      const int t32 = (m2 + 31) >> 5;
      const int o = j0 + nb;
      for (int blk = wid; blk < t32 * t32; blk += 4) {
        const int r0 = (blk % t32) * 32, c0 = (blk / t32) * 32;
        d4 acc[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int bb = 0; bb < 2; ++bb) {
            const int col = c0 + bb * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int row = r0 + a * 16 + crow(lane, v);
              acc[a][bb][v] = (row < m2 && col < m2)
                                  ? A[(int64_t)(o + row) + (int64_t)(o + col) * lda]
                                  : 0.0;
            }
          }
#pragma unroll
        for (int kk = 0; kk < NB / 4; ++kk) {
          const int k = kk * 4 + (lane >> 4);
          double af[2], bf[2];
#pragma unroll
          for (int a = 0; a < 2; ++a) {
            const int row = r0 + a * 16 + (lane & 15);
            af[a] = (k < nb && row < m2) ? -pan[k * n + nb + row] : 0.0;
          }
#pragma unroll
          for (int bb = 0; bb < 2; ++bb) {
            const int col = c0 + bb * 16 + (lane & 15);
            bf[bb] = (k < nb && col < m2) ? A[(int64_t)(j0 + k) + (int64_t)(o + col) * lda] : 0.0;
          }
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int bb = 0; bb < 2; ++bb) acc[a][bb] = mfma(af[a], bf[bb], acc[a][bb]);
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int bb = 0; bb < 2; ++bb) {
            const int col = c0 + bb * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int row = r0 + a * 16 + crow(lane, v);
              if (row < m2 && col < m2)
                A[(int64_t)(o + row) + (int64_t)(o + col) * lda] = acc[a][bb][v];
            }
          }
      }
    }
    __syncthreads();
  }

  for (int i = tid; i < n; i += 256) perm0[(int64_t)b * n + i] = perm[i];
  if (tid == 0) info0[b] = info;
}

// =====================================================================
// Batched LU for n <= 512: one workgroup (NT threads) per block, thread t
// owning row j0 + t of the current NB-column panel in registers.  Pivoting
// is implicit: rows stay in their threads and carry their LAPACK position;
// per pivot column one wave argmax (DPP, first row of largest |a|, as
// idamax), one LDS exchange of the waves' candidates and one barrier.
// Panels are grouped in outer blocks of S: U = L11^-1 A12 and the rank-NB
// update run on MFMA only inside the outer block; once its S panels are
// factored, a block forward substitution forms U12 for the columns to the
// right and the trailing matrix takes ONE update of inner dimension S*NB,
// so the trailing matrix crosses HBM once per S panels.
// =====================================================================
template <int NB, int NT, int S, typename T>
__global__ void __launch_bounds__(NT, 2)  // two waves per SIMD: 256 registers
k_getrf_reg(T* __restrict__ A0, int64_t lda, int64_t bstride, int n,
            int* __restrict__ ipiv0, int* __restrict__ perm0, int* __restrict__ info0,
            const int* __restrict__ only, const double* __restrict__ src, int64_t lds,
            int64_t sbstride) {
  // `only` (optional): factor block b only where only[b] != 0; `src`
  // (optional): block b is first copied from src + b*sbstride (ld lds),
  // converted to T, and factored in A
  // per-column candidates, T-buffered by column parity: each wave
  // publishes its best row (key, index and the whole panel row) so the
  // winner's row is readable right after the single barrier of the column
  constexpr int NW = NT / 64;
  static_assert(NW >= 2 && (NW & (NW - 1)) == 0, "a power-of-two count of waves");
  // outer blocks of two panels are the only size checked against the
  // oracle (a four-panel FP32 build gave a wrong factor)
  static_assert(S == 2, "outer blocks of two panels");
  typedef typename Elem<T>::V4 V4;
  __shared__ unsigned long long ckey[2][NW];
  __shared__ int cidx[2][NW];
  __shared__ T crowv[2][NW][NB];
  __shared__ int pivs[NB];
  __shared__ int mv[2 * NB];  // the panel's moving rows (move_rows)
  static_assert(2 * NB <= 64, "a moving row per lane");
  __shared__ int mcount;
  // LDS pool, in two lives.  Panel phase: L11 of each panel of the outer
  // block ([i*(NB+1) + k]), their inverses (same layout) and the doubling
  // scratch.  Trailing update: the negated L panel of a row band (k-major,
  // [k*(BR+UPAD) + row]) and the U panel of a column tile ([k*(BC+UPAD) +
  // col]).  The trailing update runs between two barriers after the last
  // reader of the panel-phase arrays of its outer block.
  constexpr int KC = S * NB;              // inner dimension of the trailing update
  constexpr int TC = (NW >= 8) ? 4 : 2;   // 16-column tiles per wave
  constexpr int BR = 32 * (NW / 2);       // macro tile: NW/2 x 2 waves of 32 x 16*TC
  constexpr int BC = 2 * 16 * TC;
  constexpr int UPAD = 2;
  constexpr int LU_ELEMS = (2 * S + 1) * NB * (NB + 1);
  constexpr int UP_ELEMS = KC * (BR + UPAD) + KC * (BC + UPAD);
  __shared__ T pool[LU_ELEMS > UP_ELEMS ? LU_ELEMS : UP_ELEMS];
  T* const L11s = pool;
  T* const Xis = pool + S * NB * (NB + 1);
  T* const Tt = pool + 2 * S * NB * (NB + 1);
  T* const As = pool;
  T* const Bs = pool + KC * (BR + UPAD);
  __shared__ int perm[NT];
  const int b = blockIdx.x;
  if (only && !only[b]) return;
  const int t = threadIdx.x;
  T* A = A0 + (int64_t)b * bstride;
  if (src) {
    const double* S0 = src + (int64_t)b * sbstride;
    for (int c = 0; c < n; ++c)
      if (t < n) A[(int64_t)t + (int64_t)c * lda] = (T)S0[(int64_t)t + (int64_t)c * lds];
    __syncthreads();
  }
  int info = 0;  // thread 0
  if (t < n) perm[t] = t;

  for (int j0 = 0; j0 < n; j0 += NB) {
    // the lane indices again, opaque: lane-dependent addresses are formed
    // per panel instead of being hoisted into registers held for the whole
    // kernel (the register file is the limit of this kernel)
    const int t = opaque((int)threadIdx.x), lane = t & 63;
    const int wid = __builtin_amdgcn_readfirstlane(t >> 6);
    const int nb = min(NB, n - j0);
    const int rows = n - j0;
    // outer block [J0, cend) of S panels; this is its panel sp
    const int sp = (j0 / NB) % S;
    const int J0 = j0 - sp * NB;
    const int cend = min(J0 + S * NB, n);
    T* L11 = L11s + sp * NB * (NB + 1);
    T* Xi = Xis + sp * NB * (NB + 1);
    const bool own = t < rows;
    if (t == 0) mcount = 0;  // read after the panel's first barrier
    T p[NB];
    if (own) {
      const T* Arow = A + (int64_t)(j0 + t) + (int64_t)j0 * lda;
#pragma unroll
      for (int c = 0; c < NB; ++c) p[c] = (c < nb) ? Arow[(int64_t)c * lda] : 0.0;
    } else {
#pragma unroll
      for (int c = 0; c < NB; ++c) p[c] = 0.0;
    }
    // Implicit pivoting: thread t keeps its row; `pos` is the row's current
    // position in LAPACK's interchange order (ties go to the smallest
    // position, as idamax), `live` until the row becomes a pivot.
    int pos = t;
    bool live = own;
    const bool wact = __ballot(own) != 0ULL;  // wave-uniform

    // a wave with no row in the panel publishes "no candidate" and keeps
    // the barrier count, in a loop of its own (no merge with the register-
    // resident panel below, so no register copies per column)
    if (!wact) {
      for (int jj = 0; jj < nb; ++jj) {
        if (lane == 0) {
          ckey[jj & 1][wid] = 0ULL;
          cidx[jj & 1][wid] = INT_MAX;
        }
        lds_barrier();
      }
    } else
#pragma unroll
    for (int jj = 0; jj < NB; ++jj) {
      if (jj < nb) {
        const int par = jj & 1;
        // key = bits(|a|) + 1 orders non-negative doubles; a NaN counts as
        // zero; 0 = not a candidate (retired rows, rows past n)
        const T aj = p[jj];
        const unsigned long long mykey = live ? Elem<T>::key(aj) : 0ULL;
        const int cpos = mykey ? pos : INT_MAX;
        // this wave's candidate: largest key, then smallest position
        unsigned long long wmax;
        const int wl = wave_argmax_key(mykey, Elem<T>::key32(mykey), cpos, wmax);
        if (lane == wl) {
#pragma unroll
          for (int c = 0; c < NB; ++c)
            if (c >= jj) crowv[par][wid][c] = p[c];
          ckey[par][wid] = wmax;
          cidx[par][wid] = wmax ? cpos : INT_MAX;
        }
        lds_barrier();
        // winner across the waves, lane-parallel: lane w < NW holds wave w's
        // candidate (key, position and pivot value)
        // (every lane reads, so the three reads are in flight together)
        const int lw = lane & (NW - 1);
        const unsigned long long kw0 = ckey[par][lw];
        const int qw0 = cidx[par][lw];
        const T pvw0 = crowv[par][lw][jj];
        const unsigned long long kw = (lane < NW) ? kw0 : 0ULL;
        const int qw = (lane < NW) ? qw0 : INT_MAX;
        const T pvw = (lane < NW) ? pvw0 : T(0);
        unsigned long long kbest;
        const int bw = wave_argmax_key(kw, Elem<T>::key32(kw), qw, kbest);
        const int bpos = __builtin_amdgcn_readlane(qw, bw);
        const bool winner = (lane == wl) && (wmax != 0ULL);
        // the pivot row, every read issued before the first use (under the
        // reciprocal), not one LDS round trip per pair of columns
        const T* prow = crowv[par][bw];
        T w[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c)
          if (c > jj) w[c] = prow[c];
        if (t == 0) pivs[jj] = bpos;
        // the global winner retires to position jj; the row at jj takes the
        // winner's old position
        if (live && winner && wid == bw) {
          live = false;
          pos = jj;
        } else if (pos == jj) {
          pos = bpos;
        }
        const T pv = Elem<T>::lane(pvw, bw);  // wave-uniform
        if (pv != 0.0) {
          if (live) {
            T l;
            if (fabs(pv) >= Elem<T>::sfmin())
              l = p[jj] * (T(1) / pv);
            else
              l = p[jj] / pv;
            p[jj] = l;
#pragma unroll
            for (int c = 0; c < NB; ++c)
              if (c > jj) p[c] -= l * w[c];
          }
        } else if (t == 0 && info == 0) {
          info = j0 + jj + 1;  // the live part of the column is zero
        }
      }
    }
    lds_barrier();  // pivs

    // write the panel rows to their final positions; L11 to LDS
    if (own) {
      T* Arow = A + (int64_t)(j0 + pos) + (int64_t)j0 * lda;
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c < nb) Arow[(int64_t)c * lda] = p[c];
    }
    if (own && pos < NB) {
#pragma unroll
      for (int c = 0; c < NB; ++c) L11[pos * (NB + 1) + c] = p[c];
    }
    if (t >= nb && t < NB) {
      // rows past nb are identity rows of the padded L11
#pragma unroll
      for (int c = 0; c < NB; ++c) L11[t * (NB + 1) + c] = 0.0;
    }
    if (t < nb) ipiv0[(int64_t)b * n + j0 + t] = j0 + pivs[t] + 1;  // 1-based
    // the panel's interchanges on the row permutation: the row that started
    // at j0 + t ends at j0 + pos, so perm moves the same way (one gather
    // before the barrier, one scatter after it)
    const int pold = own ? perm[j0 + t] : 0;
    if (own && pos != t) mv[atomicAdd(&mcount, 1)] = t | (pos << 16);
    lds_barrier();  // L11, perm, mv; the row stores above are read after a full barrier
    if (own && pos != t) perm[j0 + pos] = pold;

    // the panel's interchanges on the other columns
    const int m2 = n - j0 - nb;
    move_rows<32>(A, lda, n, j0, nb, mv, mcount, wid, NW, lane);
    // Xi = L11^-1 by doubling: after the step of size h the diagonal blocks
    // of size 2h of Xi invert those of L11; the off-diagonal h x h block of
    // a 2h block is -B^-1 (C A^-1), C the matching block of L11.
    if (m2 > 0) {
      for (int e = t; e < NB * NB; e += NT) {
        const int i = e / NB, j = e % NB;
        Xi[i * (NB + 1) + j] = (i == j) ? 1.0 : 0.0;
      }
      lds_barrier();
      for (int h = 1; h < NB; h <<= 1) {
        for (int e = t; e < NB * h / 2; e += NT) {
          const int q = e / (h * h), r = e % (h * h);
          const int s0 = q * 2 * h, i = s0 + h + r / h, j = s0 + r % h;
          T acc = 0.0;
          for (int k = j; k < s0 + h; ++k) acc += L11[i * (NB + 1) + k] * Xi[k * (NB + 1) + j];
          Tt[i * (NB + 1) + j] = acc;
        }
        lds_barrier();
        for (int e = t; e < NB * h / 2; e += NT) {
          const int q = e / (h * h), r = e % (h * h);
          const int s0 = q * 2 * h, i = s0 + h + r / h, j = s0 + r % h;
          T acc = 0.0;
          for (int k = s0 + h; k <= i; ++k) acc += Xi[i * (NB + 1) + k] * Tt[k * (NB + 1) + j];
          Xi[i * (NB + 1) + j] = -acc;
        }
        lds_barrier();
      }
    }
    __syncthreads();  // the interchanged columns are read by other waves from here on

    // A(rb.., cb..) -= A(rb.., kb..kb+kc) * A(kb..kb+kc, cb..) over nr x ncl,
    // 32 x 32 blocks per wave, the inner dimension in 16-wide chunks.
    auto update = [&](int rb, int nr, int cb, int ncl, int kb, int kc) {
      const int tr = (nr + 31) >> 5, tc = (ncl + 31) >> 5;
      for (int blk = wid; blk < tr * tc; blk += NW) {
        const int r0 = (blk % tr) * 32, c0 = (blk / tr) * 32;
        const bool full = (r0 + 32 <= nr) && (c0 + 32 <= ncl);
        V4 c[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int bb = 0; bb < 2; ++bb) {
            const int col = c0 + bb * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int row = r0 + a * 16 + Elem<T>::crow(lane, v);
              c[a][bb][v] = (full || (row < nr && col < ncl))
                                ? A[(int64_t)(rb + row) + (int64_t)(cb + col) * lda] : 0.0;
            }
          }
        for (int k0 = 0; k0 < kc; k0 += 16) {
          T af[4][2], bf[4][2];
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) {
            const int k = k0 + kk * 4 + (lane >> 4);
#pragma unroll
            for (int a = 0; a < 2; ++a) {
              const int row = r0 + a * 16 + (lane & 15);
              af[kk][a] = (k < kc && (full || row < nr)) ? A[(int64_t)(rb + row) + (int64_t)(kb + k) * lda] : 0.0;
            }
#pragma unroll
            for (int bb = 0; bb < 2; ++bb) {
              const int col = c0 + bb * 16 + (lane & 15);
              bf[kk][bb] = (k < kc && (full || col < ncl)) ? A[(int64_t)(kb + k) + (int64_t)(cb + col) * lda] : 0.0;
            }
          }
#pragma unroll
          for (int kk = 0; kk < 4; ++kk)
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
              for (int bb = 0; bb < 2; ++bb) c[a][bb] = mfma(-af[kk][a], bf[kk][bb], c[a][bb]);
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int bb = 0; bb < 2; ++bb) {
            const int col = c0 + bb * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int row = r0 + a * 16 + Elem<T>::crow(lane, v);
              if (full || (row < nr && col < ncl))
                A[(int64_t)(rb + row) + (int64_t)(cb + col) * lda] = c[a][bb][v];
            }
          }
      }
    };

    // The trailing update of a complete outer block, both panels staged in
    // LDS: C(rb.., cb..) -= L(rb.., kb..kb+KC) U(kb..kb+KC, cb..) over
    // nr x nc.  Macro tiles of BR x BC in row-band order; the L panel of a
    // band is staged once (negated), the U panel once per tile, fetched into
    // registers under the previous tile's MFMAs.  Accumulators
    // start at zero and C, loaded under the MFMAs, is added at the end, so
    // every element of C crosses HBM once each way per outer block.
    static_assert(NT % BR == 0 && NT % KC == 0 && KC % 64 == 0 && (KC * BC) % NT == 0, "staging split");
    // 32-bit byte offsets need 127 * lda * sizeof(T) < 2^32; the caller
    // takes the plain update beyond lda = 2^22.
    auto update_lds = [&](int rb, int nr, int cb, int nc, int kb) {
      constexpr int PA = BR * KC / NT, PB = KC * BC / NT;
      constexpr uint32_t SZ = sizeof(T);
      const int nbc = (nc + BC - 1) / BC;
      const int ntile = ((nr + BR - 1) / BR) * nbc;
      const int wr = wid >> 1, wc = wid & 1;
      const uint32_t ldb = (uint32_t)lda * SZ;  // column stride in bytes
      // thread t stages row t % BR of the L panel (k = t / BR + i * NT / BR)
      // and row t % KC of the U panel (column t / KC + i * NT / KC, the
      // same for the whole wave since KC is a multiple of 64)
      const int ra = t % BR, ka = t / BR;
      const int kr = t % KC, cbt = __builtin_amdgcn_readfirstlane(t / KC);
      const uint32_t voa = ra * SZ + ka * ldb;
      const uint32_t vob = kr * SZ + cbt * ldb;
      // The MFMAs form the transposed tile (U^T L^T, the U panel as the A
      // operand): accumulator element (i, j) is C(row j, column i), so a
      // lane holds rows lane & 15 of columns crow(lane, v) and each 16-lane
      // group of a load or store covers 16 consecutive rows of a column
      // (128-byte lines).  Lane offset of (lane & 15, crow(lane, 0)):
      const int lc = Elem<T>::crow(lane, 0);
      const uint32_t voc = (lane & 15) * SZ + lc * ldb;
      T pa[PA], pb[PB];
      auto fetch_u = [&](int q) {
        const int C0 = (q % nbc) * BC;
        const __amdgpu_buffer_rsrc_t rs = buf_rsrc(A + kb + (int64_t)(cb + C0) * lda);
        const uint32_t vo = opaque(vob);
#pragma unroll
        for (int i = 0; i < PB; ++i)  // wave-uniform column test
          pb[i] = (C0 + cbt + i * (NT / KC) < nc) ? buf_load<T>(rs, vo, i * (NT / KC) * ldb) : T(0);
      };
      fetch_u(0);
      for (int q = 0; q < ntile; ++q) {
        const int R0 = (q / nbc) * BR, C0 = (q % nbc) * BC;
        if (q % nbc == 0) {
          // a new row band: its L panel is read here, under the wait for
          // the other waves' last MFMAs of the previous band
          const __amdgpu_buffer_rsrc_t rs = buf_rsrc(A + (rb + R0) + (int64_t)kb * lda);
          const uint32_t vo = opaque(voa);
          if (R0 + ra < nr) {
#pragma unroll
            for (int i = 0; i < PA; ++i) pa[i] = buf_load<T>(rs, vo, i * (NT / BR) * ldb);
          } else {
#pragma unroll
            for (int i = 0; i < PA; ++i) pa[i] = T(0);
          }
        }
        lds_barrier();  // the previous tile's MFMAs are done with As, Bs
        if (q % nbc == 0) {
          T* const dst = As + opaque(ka * (BR + UPAD) + ra);
#pragma unroll
          for (int i = 0; i < PA; ++i) dst[i * (NT / BR) * (BR + UPAD)] = -pa[i];
        }
        {
          T* const dst = Bs + opaque(kr * (BC + UPAD) + cbt);
#pragma unroll
          for (int i = 0; i < PB; ++i) dst[i * (NT / KC)] = pb[i];
        }
        lds_barrier();
        const int r0 = R0 + wr * 32, c0 = C0 + wc * 16 * TC;
        const bool act = r0 < nr && c0 < nc;  // wave-uniform
        const bool full = r0 + 32 <= nr && c0 + 16 * TC <= nc;
        const __amdgpu_buffer_rsrc_t rc = buf_rsrc(A + (rb + r0) + (int64_t)(cb + c0) * lda);
        const uint32_t vc = opaque(voc);
        auto in_c = [&](int a, int bb, int v) {
          return r0 + a * 16 + (lane & 15) < nr && c0 + bb * 16 + lc + v * Elem<T>::DV < nc;
        };
        T cv[2][TC][4];
        if (full) {
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int bb = 0; bb < TC; ++bb)
#pragma unroll
              for (int v = 0; v < 4; ++v)
                cv[a][bb][v] = buf_load<T>(rc, vc, (bb * 16 + v * Elem<T>::DV) * ldb + a * 16 * SZ);
        } else if (act) {
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int bb = 0; bb < TC; ++bb)
#pragma unroll
              for (int v = 0; v < 4; ++v)
                cv[a][bb][v] = in_c(a, bb, v)
                                   ? buf_load<T>(rc, vc, (bb * 16 + v * Elem<T>::DV) * ldb + a * 16 * SZ)
                                   : T(0);
        }
        if (q + 1 < ntile) fetch_u(q + 1);
        if (act) {
          V4 acc[2][TC];
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int bb = 0; bb < TC; ++bb) acc[a][bb] = zero4<T>();
          const T* const sa = As + opaque((lane >> 4) * (BR + UPAD) + wr * 32 + (lane & 15));
          const T* const sb = Bs + opaque((lane >> 4) * (BC + UPAD) + wc * 16 * TC + (lane & 15));
#pragma unroll
          for (int ks = 0; ks < KC / 4; ++ks) {
            T af[2], bf[TC];
#pragma unroll
            for (int a = 0; a < 2; ++a) af[a] = sa[ks * 4 * (BR + UPAD) + a * 16];
#pragma unroll
            for (int bb = 0; bb < TC; ++bb) bf[bb] = sb[ks * 4 * (BC + UPAD) + bb * 16];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
              for (int bb = 0; bb < TC; ++bb) acc[a][bb] = mfma(bf[bb], af[a], acc[a][bb]);
          }
          if (full) {
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
              for (int bb = 0; bb < TC; ++bb)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                  buf_store(acc[a][bb][v] + cv[a][bb][v], rc, vc, (bb * 16 + v * Elem<T>::DV) * ldb + a * 16 * SZ);
          } else {
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
              for (int bb = 0; bb < TC; ++bb)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                  if (in_c(a, bb, v))
                    buf_store(acc[a][bb][v] + cv[a][bb][v], rc, vc, (bb * 16 + v * Elem<T>::DV) * ldb + a * 16 * SZ);
          }
        }
      }
    };

    const int o = j0 + nb;
    // inside the outer block: U = L11^-1 A(j0.., o..cend), then the rows
    // below take the rank-nb update on those columns only.  U also goes to
    // the part of the LDS pool the panel phase does not use, so the update
    // reads it from there; each wave's 16-row tiles load their L fragments
    // and C values together.
    if (o < cend) {
      const int m2b = cend - o;  // <= NB: the outer block's remaining panels
      static_assert(LU_ELEMS + NB * (NB + 1) <= (LU_ELEMS > UP_ELEMS ? LU_ELEMS : UP_ELEMS),
                    "U staging fits the pool");
      T* const Us = pool + LU_ELEMS;  // [k*(NB+1) + col]
      const int nct = (m2b + 15) >> 4;
      for (int ct = wid; ct < nct; ct += NW) {
        const int col = ct * 16 + (lane & 15);
        V4 u[NB / 16];
#pragma unroll
        for (int q = 0; q < NB / 16; ++q) u[q] = zero4<T>();
#pragma unroll
        for (int kk = 0; kk < NB / 4; ++kk) {
          const int k = kk * 4 + (lane >> 4);
          const T bf = (k < nb && col < m2b) ? A[(int64_t)(j0 + k) + (int64_t)(o + col) * lda] : 0.0;
#pragma unroll
          for (int q = 0; q < NB / 16; ++q)
            u[q] = mfma(Xi[(q * 16 + (lane & 15)) * (NB + 1) + k], bf, u[q]);
        }
#pragma unroll
        for (int q = 0; q < NB / 16; ++q)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = q * 16 + Elem<T>::crow(lane, v);
            // rows past nb and columns past m2b are exact zeros
            Us[r * (NB + 1) + col] = u[q][v];
            if (col < m2b && r < nb) A[(int64_t)(j0 + r) + (int64_t)(o + col) * lda] = u[q][v];
          }
      }
      lds_barrier();  // Us; the L panel was written before the last full barrier
      const int nr = n - o;
      const int nrt = (nr + 15) >> 4;
      constexpr int TN = NB / 16;  // column tiles of the outer block's remainder
      for (int rt = wid; rt < nrt; rt += NW) {
        const bool full = rt * 16 + 16 <= nr && m2b == NB;
        T af[NB / 4];
        V4 c[TN];
        const int arow = rt * 16 + (lane & 15);
#pragma unroll
        for (int kk = 0; kk < NB / 4; ++kk) {
          const int k = kk * 4 + (lane >> 4);
          af[kk] = (full || (arow < nr && k < nb)) ? A[(int64_t)(o + arow) + (int64_t)(j0 + k) * lda] : T(0);
        }
        // transposed tiles (U as the A operand): a lane holds row lane & 15
        // of columns crow(lane, v), so C moves in whole 128-byte lines
#pragma unroll
        for (int tn = 0; tn < TN; ++tn)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = rt * 16 + (lane & 15), col = tn * 16 + Elem<T>::crow(lane, v);
            c[tn][v] = (full || (row < nr && col < m2b)) ? A[(int64_t)(o + row) + (int64_t)(o + col) * lda] : T(0);
          }
#pragma unroll
        for (int kk = 0; kk < NB / 4; ++kk) {
          const int k = kk * 4 + (lane >> 4);
#pragma unroll
          for (int tn = 0; tn < TN; ++tn)
            c[tn] = mfma(Us[k * (NB + 1) + tn * 16 + (lane & 15)], -af[kk], c[tn]);
        }
#pragma unroll
        for (int tn = 0; tn < TN; ++tn)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = rt * 16 + (lane & 15), col = tn * 16 + Elem<T>::crow(lane, v);
            if (full || (row < nr && col < m2b)) A[(int64_t)(o + row) + (int64_t)(o + col) * lda] = c[tn][v];
          }
      }
    } else if (cend < n) {
      // outer block complete (all its panels full): block forward
      // substitution U_s = L_ss^-1 (A_s - sum_{q<s} L_sq U_q) on the columns
      // right of it, then the trailing update with inner dimension cend - J0
      // (cend = J0 + S*NB here, so all S panels are full).  The L_sq
      // blocks (q < s) are the same for every column tile: staged once in
      // the part of the LDS pool the panel phase does not use; a column
      // tile's right-hand sides of all S panels are fetched together.
      const int ntr = n - cend;
      const int nct = (ntr + 15) >> 4;
      static_assert(LU_ELEMS + S * (S - 1) / 2 * NB * (NB + 1) <= (LU_ELEMS > UP_ELEMS ? LU_ELEMS : UP_ELEMS),
                    "L_sq staging fits the pool");
      // block (s2, q), q < s2, at (s2*(s2-1)/2 + q)*NB*(NB+1), [row*(NB+1) + k]
      T* const Lq = pool + LU_ELEMS;
#pragma unroll
      for (int s2 = 0; s2 < S; ++s2)
#pragma unroll
        for (int q = 0; q < S; ++q)
          if (q < s2) {
            for (int e = t; e < NB * NB; e += NT) {
              const int row = e % NB, k = e / NB;
              Lq[(s2 * (s2 - 1) / 2 + q) * NB * (NB + 1) + row * (NB + 1) + k] =
                  A[(int64_t)(J0 + s2 * NB + row) + (int64_t)(J0 + q * NB + k) * lda];
            }
          }
      lds_barrier();
      // element (s2, a, v) of a tile: row J0 + s2*NB + a*16 + crow(lane, v),
      // column cend + ct*16 + (lane & 15)
      const int64_t lo = (int64_t)Elem<T>::crow(lane, 0) + (int64_t)(lane & 15) * lda;
      T cur[S][NB / 16][4];
      auto fetch = [&](int ct, T (&d)[S][NB / 16][4]) {
        const T* bp = A + J0 + (int64_t)(cend + ct * 16) * lda + lo;
        const bool cok = ct * 16 + (lane & 15) < ntr;
#pragma unroll
        for (int s2 = 0; s2 < S; ++s2)
#pragma unroll
          for (int a = 0; a < NB / 16; ++a)
#pragma unroll
            for (int v = 0; v < 4; ++v) d[s2][a][v] = cok ? bp[s2 * NB + a * 16 + v * Elem<T>::DV] : T(0);
      };
      for (int ct = wid; ct < nct; ct += NW) {
        fetch(ct, cur);
        const bool cok = ct * 16 + (lane & 15) < ntr;
        T* const bp = A + J0 + (int64_t)(cend + ct * 16) * lda + lo;
        V4 u[S][NB / 16];
#pragma unroll
        for (int s2 = 0; s2 < S; ++s2) {
          V4 acc[NB / 16];
#pragma unroll
          for (int a = 0; a < NB / 16; ++a)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[a][v] = cur[s2][a][v];
#pragma unroll
          for (int q = 0; q < S; ++q) {
            if (q < s2) {
              // k-step (a2, v): rows a2*16 + crow(lane, v) of U_q are
              // exactly the accumulator element v of tile a2
#pragma unroll
              for (int a2 = 0; a2 < NB / 16; ++a2)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                  const T bf = u[q][a2][v];
                  const int k = a2 * 16 + Elem<T>::crow(lane, v);
#pragma unroll
                  for (int a = 0; a < NB / 16; ++a)
                    acc[a] = mfma(-Lq[(s2 * (s2 - 1) / 2 + q) * NB * (NB + 1) + (a * 16 + (lane & 15)) * (NB + 1) + k], bf, acc[a]);
                }
            }
          }
#pragma unroll
          for (int a = 0; a < NB / 16; ++a) u[s2][a] = zero4<T>();
#pragma unroll
          for (int a2 = 0; a2 < NB / 16; ++a2)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int k = a2 * 16 + Elem<T>::crow(lane, v);
              const T bf = acc[a2][v];
#pragma unroll
              for (int a = 0; a < NB / 16; ++a)
                u[s2][a] = mfma(Xis[s2 * NB * (NB + 1) + (a * 16 + (lane & 15)) * (NB + 1) + k], bf, u[s2][a]);
            }
#pragma unroll
          for (int a = 0; a < NB / 16; ++a)
#pragma unroll
            for (int v = 0; v < 4; ++v)
              if (cok) bp[s2 * NB + a * 16 + v * Elem<T>::DV] = u[s2][a][v];
        }
      }
      __syncthreads();
      if (lda <= (1 << 22))
        update_lds(cend, ntr, cend, ntr, J0);
      else
        update(cend, ntr, cend, ntr, J0, cend - J0);
    }
    __syncthreads();
  }
  if (t < n) perm0[(int64_t)b * n + t] = perm[t];
  if (t == 0) info0[b] = info;
}

// =====================================================================
// Batched leaf LU without row interchanges, verified against the partial
// pivoting rule, out of place (K -> LU) -- the fast path of the leaf
// factorization.  n = 128*RT; one workgroup of 8 waves per block.
//
// Left-looking over 64-column slabs.  A slab (all n rows) lives in MFMA
// accumulators as transposed 16x16 tiles -- lane l holds rows l & 15 of
// columns crow(l, v), so each 16-lane group moves 16 consecutive rows of a
// column (a 128-byte line) -- wave w owning the 16-row tiles w, w+8, ...
// (interleaved, so every wave keeps work as the factored rows move down).
// For each 32-column chunk c left of the slab, in order:
//     U_c = L_cc^-1 X_c      (X_c: the slab's rows of chunk c; MFMA with
//                            the stored inverse of the unit-lower L_cc)
//     X_r -= L_rc U_c        (every row tile r below the chunk; L_rc read
//                            once per slab from the factors already written)
// then the slab's two 32-column panels are factored: the 32x32 diagonal
// block by one wave (dgetf2 order, no interchanges), its L and U inverted
// (the stored 32x32 inverses the solves use), L21 = A21 U11^-1 and, in the
// first panel, A22 -= L21 (L11^-1 A12) on MFMA.
//
// Partial pivoting (dgetrf2: the first row of largest |a| in the column)
// keeps the diagonal at step j exactly when every multiplier of that step
// has |l_ij| <= 1.  Every multiplier is checked against 1 - margin (the
// margin absorbs rounding differences to the netlib operation order); a
// block where any check fails -- or a zero or non-finite pivot -- stops and
// sets retry[b], and the caller refactors it from K with the pivoted
// kernel.  Accepted blocks are exactly the factorization dgetrf computes
// with IPIV = identity, up to rounding.
// =====================================================================
template <typename T> struct PivotMargin;
template <> struct PivotMargin<double> {
  __device__ static double thr() { return 1.0 - 1.0 / 1048576.0; }  // 1 - 2^-20
};
template <> struct PivotMargin<float> {
  __device__ static float thr() { return 1.0f - 1.0f / 4096.0f; }   // 1 - 2^-12
};

// Reciprocal of a pivot: the hardware estimate and one Newton-Raphson
// step, r = r0 + r0 (1 - x r0), within an ulp of 1/x (exact for powers of
// two, inf for x = 0, NaN for NaN) -- a short dependent chain where the
// IEEE divide is a long one.
__device__ __forceinline__ double recip_nr(double x) {
  const double r0 = __builtin_amdgcn_rcp(x);
  return fma(r0, fma(-x, r0, 1.0), r0);
}
__device__ __forceinline__ float recip_nr(float x) {
  const float r0 = __builtin_amdgcn_rcpf(x);
  return fmaf(r0, fmaf(-x, r0, 1.0f), r0);
}

// Orders one wave's LDS accesses at this point: in-order LDS execution
// makes the values written before it visible to the wave's reads after it;
// the empty asm with a memory clobber also keeps the compiler from moving
// memory operations across it (and so from overlapping the phases of the
// diagonal step, which would need more registers than there are).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_wave_barrier();
  asm volatile("" ::: "memory");
}

// Unblocked LU without interchanges of a 16x16 block held by lanes 0..15
// (lane i: row i in a[]), dgetf2's operation order; every multiplier is
// checked against thr.  Row j, final after step j - 1, reaches every lane
// from lane j's registers by readlane (register-to-scalar moves: no LDS
// round trip and no wave synchronization on the serial chain; half the
// time of publishing the row through LDS, measured in the leaf and core
// LUs once the leaf LU's scalar offsets stopped occupying the scalar
// registers across the kernel).
template <typename T>
__device__ __forceinline__ void ll_lu16(T (&a)[16], int lane, T thr, bool& fail) {
  const bool own = lane < 16;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    T u[16];
#pragma unroll
    for (int k = j; k < 16; ++k) u[k] = Elem<T>::lane(a[k], j);
    const T r = recip_nr(u[j]);
    if (own && lane > j) {
      const T l = a[j] * r;
      fail |= !(fabs(l) <= thr);  // NaN, inf (zero pivot) included
      a[j] = l;
#pragma unroll
      for (int k = j + 1; k < 16; ++k) a[k] -= l * u[k];
    }
  }
}

// Column stride of the 16x16 inversion tiles: 16 elements and 16 bytes of
// padding, so every column starts 16-byte aligned and a step's reads are
// wide LDS reads off one base address.
template <typename T> struct InvTile {
  static constexpr int S = 16 + 16 / (int)sizeof(T);
};

// Inverses of the unit-lower L and upper U of a 16x16 LU held in the
// inversion tiles Mt[0] (L with its unit diagonal, [col][row]) and Mt[1]
// (U index-reversed, also lower triangular): lanes 0..15 solve for the
// columns of L^-1, lanes 16..31 for those of U^-1 (reversed), by
// column-oriented substitution; written to Li / Ui at offset (o, o).
template <typename T>
__device__ __forceinline__ void ll_inv16(T (*Mt)[16][InvTile<T>::S], T (*Li)[33], T (*Ui)[33],
                                         int o, int lane) {
  const int h = (lane >> 4) & 1, c = lane & 15;
  const int c0 = h ? 15 - c : c;
  T x[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) x[q] = (q == c0) ? T(1) : T(0);
  const T* const tile = &Mt[h][0][0];
  // The diagonal's reciprocals first (16 independent load + reciprocal
  // chains, issued together), then the substitution with each step's column
  // read one step ahead: the scheduling barriers keep the reads of step
  // k + 1 above the arithmetic of step k, so a step's dependent chain is
  // its multiply and the next element's update, not an LDS round trip and
  // a reciprocal (left to the scheduler, each step waited for its own
  // column and its reciprocal).  The arithmetic is unchanged.
  T rd[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) rd[k] = recip_nr(tile[k * InvTile<T>::S + k]);
  T cn[16];
#pragma unroll
  for (int q = 1; q < 16; ++q) cn[q] = tile[q];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    T cc[16];
#pragma unroll
    for (int q = k + 1; q < 16; ++q) cc[q] = cn[q];
    if (k + 1 < 16) {
#pragma unroll
      for (int q = k + 2; q < 16; ++q) cn[q] = tile[(k + 1) * InvTile<T>::S + q];
    }
    __builtin_amdgcn_sched_barrier(0);
    const T d = x[k] * rd[k];
    const T xk = (k < c0) ? T(0) : d;
    x[k] = xk;
#pragma unroll
    for (int q = k + 1; q < 16; ++q) x[q] -= cc[q] * xk;
    __builtin_amdgcn_sched_barrier(0);
  }
  if (lane < 16) {
#pragma unroll
    for (int q = 0; q < 16; ++q) Li[o + q][o + c] = x[q];
  } else if (lane < 32) {
#pragma unroll
    for (int q = 0; q < 16; ++q) Ui[o + q][o + c] = x[15 - q];
  }
}

// The diagonal step of a panel, one wave: factor the 32x32 block Ds in
// place without interchanges, check every multiplier, invert L11 (unit
// lower) and U11 into Li / Ui ([row][col]) and the stored inverse `out`
// ([Linv | Uinv], column-major), and with A12 given form U12 = L11^-1 A12
// in place.  Blocked 2 x 2 on 16x16 blocks, so the serial chains are 16
// steps long and the couplings are single-tile MFMA products:
//   A11 = L1 U1;  L21 = A21 U1^-1;  U12 = L1^-1 A12;  A22 - L21 U12 = L2 U2;
//   L^-1 = [L1^-1, 0; -L2^-1 (L21 L1^-1), L2^-1],
//   U^-1 = [U1^-1, -(U1^-1 U12) U2^-1; 0, U2^-1].
template <typename T>
__device__ __forceinline__ void ll_diag_step(T (*Ds)[33], T (*As12)[33], T (*Li)[33],
                                             T (*Ui)[33], T (*Ms)[32][33],
                                             T* __restrict__ out, bool with_a12, T thr,
                                             bool& fail, int lane_in) {
  typedef typename Elem<T>::V4 V4;
  // opaque: the lane-dependent addresses below are formed here, not
  // hoisted out of the caller's slab loop into registers held throughout
  const int lane = opaque(lane_in);
  const int lr = lane & 15, lk = lane >> 4;
  // the inversion tiles and a product scratch, in the space of Ms
  T (*const Mt)[16][InvTile<T>::S] = reinterpret_cast<T (*)[16][InvTile<T>::S]>(&Ms[0][0][0]);
  T (*const Pt)[16][17] = reinterpret_cast<T (*)[16][17]>(&Ms[0][0][0] + 2 * 16 * InvTile<T>::S);
  const int i = lane & 15;
  auto factor16 = [&](int o) {
    // LU of the diagonal 16x16 block at (o, o), then its inversion tiles
    T a[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = Ds[o + i][o + k];
    ll_lu16<T>(a, lane, thr, fail);
    if (lane < 16) {
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        Ds[o + i][o + k] = a[k];
        Mt[0][k][i] = (i > k) ? a[k] : ((i == k) ? T(1) : T(0));
        Mt[1][15 - k][15 - i] = (i <= k) ? a[k] : T(0);
      }
    }
    wave_sync();
    ll_inv16<T>(Mt, Li, Ui, o, lane);
    wave_sync();
  };
  factor16(0);
  {
    // L21 = A21 U1^-1 (checked) and U12 = L1^-1 A12
    V4 d = zero4<T>(), e = zero4<T>();
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      d = mfma(Ds[16 + lr][4 * s + lk], Ui[4 * s + lk][lr], d);
      e = mfma(Li[lr][4 * s + lk], Ds[4 * s + lk][16 + lr], e);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) fail |= !(fabs(d[v]) <= thr);
    wave_sync();
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      Ds[16 + Elem<T>::crow(lane, v)][lr] = d[v];
      Ds[Elem<T>::crow(lane, v)][16 + lr] = e[v];
    }
    wave_sync();
    // A22 -= L21 U12
    V4 c;
#pragma unroll
    for (int v = 0; v < 4; ++v) c[v] = Ds[16 + Elem<T>::crow(lane, v)][16 + lr];
#pragma unroll
    for (int s = 0; s < 4; ++s) c = mfma(-Ds[16 + lr][4 * s + lk], Ds[4 * s + lk][16 + lr], c);
    wave_sync();
#pragma unroll
    for (int v = 0; v < 4; ++v) Ds[16 + Elem<T>::crow(lane, v)][16 + lr] = c[v];
    wave_sync();
  }
  factor16(16);
  {
    // off-diagonal blocks of the inverses: P0 = L21 L1^-1, P1 = U1^-1 U12,
    // then Linv21 = -L2^-1 P0 and Uinv12 = -P1 U2^-1
    V4 p0 = zero4<T>(), p1 = zero4<T>();
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      p0 = mfma(Ds[16 + lr][4 * s + lk], Li[4 * s + lk][lr], p0);
      p1 = mfma(Ui[lr][4 * s + lk], Ds[4 * s + lk][16 + lr], p1);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      Pt[0][Elem<T>::crow(lane, v)][lr] = p0[v];
      Pt[1][Elem<T>::crow(lane, v)][lr] = p1[v];
    }
    wave_sync();
    V4 l21 = zero4<T>(), u12 = zero4<T>();
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      l21 = mfma(Li[16 + lr][16 + 4 * s + lk], Pt[0][4 * s + lk][lr], l21);
      u12 = mfma(Pt[1][lr][4 * s + lk], Ui[16 + 4 * s + lk][16 + lr], u12);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      Li[16 + Elem<T>::crow(lane, v)][lr] = -l21[v];
      Ui[Elem<T>::crow(lane, v)][16 + lr] = -u12[v];
      Li[Elem<T>::crow(lane, v)][16 + lr] = T(0);
      Ui[16 + Elem<T>::crow(lane, v)][lr] = T(0);
    }
  }
  wave_sync();
  // the stored inverses, column-major, 512 consecutive bytes per store
  // (none when out is null)
  if (out) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int idx = lane + 64 * q;
      out[idx] = Li[idx % 32][idx / 32];
      out[1024 + idx] = Ui[idx % 32][idx / 32];
    }
  }
  if (with_a12) {
    // U12 of the slab = L11^-1 A12: D(i = row, j = col), A(i, k) =
    // Linv(row, k), B(k, j) = A12(k, col); Linv is lower triangular
    V4 d[2][2];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int ct = 0; ct < 2; ++ct) {
        d[rt][ct] = zero4<T>();
#pragma unroll
        for (int s = 0; s < 4 * (rt + 1); ++s)
          d[rt][ct] = mfma(Li[16 * rt + lr][4 * s + lk], As12[4 * s + lk][16 * ct + lr], d[rt][ct]);
      }
    wave_sync();
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int v = 0; v < 4; ++v) As12[16 * rt + Elem<T>::crow(lane, v)][16 * ct + lr] = d[rt][ct][v];
  }
}

template <int RT, typename T>
__global__ void __launch_bounds__(512, 1)
k_getrf_ll(const double* __restrict__ K0, int64_t ldk, int64_t kbs, T* __restrict__ A0,
           int64_t lda, int64_t abs_, T* __restrict__ dinv0, int* __restrict__ ipiv0,
           int* __restrict__ perm0, int* __restrict__ info0, int* __restrict__ retry0) {
  constexpr int n = 128 * RT;
  constexpr int NS = n / 64;  // slabs
  constexpr int DV = Elem<T>::DV;
  typedef typename Elem<T>::V4 V4;
  // X_c (the slab's rows of chunk c) and U_c = L_cc^-1 X_c in the chunk
  // loop; in the panel phase the same space parks wave 0's accumulators
  // while it runs the diagonal step
  __shared__ T pool[2 * 32 * 65];
  static_assert(RT * 16 * 64 <= 2 * 32 * 65, "accumulator parking fits");
  T (*const Xs)[65] = reinterpret_cast<T (*)[65]>(pool);
  T (*const Us)[65] = reinterpret_cast<T (*)[65]>(pool + 32 * 65);
  __shared__ T Ds[32][33];     // diagonal block: A11 in, packed LU out
  __shared__ T As12[32][33];   // A12 in, U12 out
  __shared__ T Li[32][33];     // L11^-1 [row][col]
  __shared__ T Ui[32][33];     // U11^-1 [row][col]
  __shared__ __align__(16) T Ms[2][32][33];  // inversion tiles
  __shared__ int bad;
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
  const double* K = K0 + (int64_t)b * kbs;
  T* A = A0 + (int64_t)b * abs_;
  T* dinv = dinv0 + (int64_t)b * dinv_elems(n);
  const T thr = PivotMargin<T>::thr();
  // byte strides of a column (32-bit offsets: the launcher checks
  // 64 * ldk * 8 < 2^32 and n * lda * sizeof(T) < 2^32)
  const uint32_t ldkb = (uint32_t)ldk * 8u;
  const uint32_t ldab = (uint32_t)lda * (uint32_t)sizeof(T);
  bool fail = false;
  if (tid == 0) bad = 0;
  V4 acc[RT][4];

  for (int J = 0; J < NS; ++J) {
    const int j0 = 64 * J;
    // the column strides again, opaque per slab: the scalar offsets built
    // on them (64 for the slab's loads, 32 for the L fragments) are formed
    // here instead of being hoisted and held across the kernel, which had
    // spilled 307 scalar registers (19 vector registers to scratch) and
    // now spills 145 (13)
    const uint32_t ldkb_j = opaque_s(ldkb), ldab_j = opaque_s(ldab);
    // Barriers.  The chunk loop exchanges only LDS data (X_c, U_c), so its
    // barriers order LDS alone and the slab's loads of K stay in flight
    // under the first chunks.  Factors a chunk reads from global memory
    // were stored by earlier slabs; those of the previous slab (chunks
    // 2J-2, 2J-1, prefetched from the end of chunk 2J-3) become visible at
    // one full barrier: here for J = 1, at chunk 2J-3 otherwise.
    if (J == 1) __syncthreads();
    // the lane indices again, opaque, so the lane parts of the addresses
    // are formed here instead of being hoisted into registers for the
    // whole kernel
    const int lane = opaque((int)threadIdx.x) & 63;
    const int lr = lane & 15, lk = lane >> 4;
    const int cr0 = Elem<T>::crow(lane, 0);  // column of element v: cr0 + v*DV
    {
      // the slab, all n rows, converted to T: element (q, ct, v) is row
      // 16 (wid + 8q) + lr, column j0 + 16 ct + cr0 + v DV
      const __amdgpu_buffer_rsrc_t rk = buf_rsrc(K + (int64_t)j0 * ldk);
      const uint32_t vo = lr * 8u + (uint32_t)cr0 * ldkb_j;
#pragma unroll
      for (int q = 0; q < RT; ++q)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
          for (int v = 0; v < 4; ++v)
            acc[q][ct][v] = (T)buf_load<double>(rk, vo, 128u * (wid + 8 * q) + (16 * ct + v * DV) * ldkb_j);
    }
    if (J > 0 && wid < 2) {
      // X_0 = rows 0..31 (tiles 0, 1: waves 0, 1, q = 0)
#pragma unroll
      for (int ct = 0; ct < 4; ++ct)
#pragma unroll
        for (int v = 0; v < 4; ++v) Xs[16 * wid + lr][16 * ct + cr0 + v * DV] = acc[0][ct][v];
    }
    // Prefetches, issued where their registers are free so that the loads
    // run under the next barrier: the stored L_cc^-1 fragments of a chunk's
    // U step (B(k, j = row) = Linv(16 rl + row, k)), and the chunk's L
    // fragments (B(k, j = row) = L(row, 32c + k)), four k-steps in flight:
    // step s2 + 4 is fetched into the registers step s2 has just released.
    const int rl = wid & 1, uct = wid >> 1;  // this wave's U tile
    T lcv[8];
    T lf[RT][8];
    const uint32_t vl = opaque(lr * (uint32_t)sizeof(T) + lk * ldab_j);
    auto fetch_lc = [&](int c) {
      const T* Lc = dinv + (int64_t)c * 2 * kTri * kTri + 16 * rl + lr;
#pragma unroll
      for (int s2 = 0; s2 < 8; ++s2) lcv[s2] = Lc[(4 * s2 + lk) * 32];
    };
    auto fetch_l = [&](int c, int s2) {
      const __amdgpu_buffer_rsrc_t ra = buf_rsrc(A + (int64_t)(32 * c) * lda);
#pragma unroll
      for (int q = 0; q < RT; ++q) {
        const int rt = wid + 8 * q;
        if (rt >= 2 * c + 2) lf[q][s2] = buf_load<T>(ra, vl, 16u * sizeof(T) * rt + 4 * s2 * ldab_j);
      }
    };
    if (J > 0) {
      fetch_lc(0);
#pragma unroll
      for (int s2 = 0; s2 < 4; ++s2) fetch_l(0, s2);
    }
    lds_barrier();  // Xs
    for (int c = 0; c < 2 * J; ++c) {
      {
        // U_c tile (rl, uct) = sum_k Linv(16 rl + row, k) X(k, 16 uct + col),
        // transposed: A(i = col, k) = X(k, col), B(k, j = row) = Linv(row, k)
        V4 u = zero4<T>();
#pragma unroll
        for (int s2 = 0; s2 < 8; ++s2)
          if (s2 < 4 * (rl + 1)) u = mfma(Xs[4 * s2 + lk][16 * uct + lr], lcv[s2], u);
#pragma unroll
        for (int v = 0; v < 4; ++v) Us[16 * rl + lr][16 * uct + cr0 + v * DV] = u[v];
      }
      // Us ready, Xs free; at chunk 2J-3 also the previous slab's factors
      if (c == 2 * J - 3)
        __syncthreads();
      else
        lds_barrier();
#pragma unroll
      for (int q = 0; q < RT; ++q) {
        const int rt = wid + 8 * q;
        if ((rt >> 1) == c) {
          // this tile's rows are final U rows
#pragma unroll
          for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[q][ct][v] = Us[16 * (rt & 1) + lr][16 * ct + cr0 + v * DV];
        }
      }
      // X_r -= L_rc U_c: A(i = col, k) = U(k, col), B(k, j = row) = L(row, k);
      // k-step outermost, so only the step's four U operands are live
#pragma unroll
      for (int s2 = 0; s2 < 8; ++s2) {
        T ua[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) ua[ct] = -Us[4 * s2 + lk][16 * ct + lr];
#pragma unroll
        for (int q = 0; q < RT; ++q) {
          const int rt = wid + 8 * q;
          if (rt >= 2 * c + 2) {
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) acc[q][ct] = mfma(ua[ct], lf[q][s2], acc[q][ct]);
          }
        }
        if (s2 + 4 < 8) fetch_l(c, s2 + 4);
      }
      if (c + 1 < 2 * J) {
        // the next chunk's first loads, under the barrier and its U step
        fetch_lc(c + 1);
#pragma unroll
        for (int s2 = 0; s2 < 4; ++s2) fetch_l(c + 1, s2);
      }
#pragma unroll
      for (int q = 0; q < RT; ++q) {
        const int rt = wid + 8 * q;
        if ((rt >> 1) == c + 1 && c + 1 < 2 * J) {
          // the next chunk's X
#pragma unroll
          for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v) Xs[16 * (rt & 1) + lr][16 * ct + cr0 + v * DV] = acc[q][ct][v];
        }
      }
      lds_barrier();  // Xs (next chunk) ready, Us free
    }
    // the slab's two panels
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int ib = 2 * J + p;  // 32-row diagonal block
      const int cp = 2 * p;      // its first column tile in the slab
#pragma unroll
      for (int q = 0; q < RT; ++q) {
        const int rt = wid + 8 * q;
        if ((rt >> 1) == ib) {
#pragma unroll
          for (int cc = 0; cc < 2; ++cc)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              Ds[16 * (rt & 1) + lr][16 * cc + cr0 + v * DV] = acc[q][cp + cc][v];
              if (p == 0) As12[16 * (rt & 1) + lr][16 * cc + cr0 + v * DV] = acc[q][2 + cc][v];
            }
        }
      }
      __syncthreads();
      if (wid == 0) {
        // the accumulators wait in LDS, so the diagonal step has the registers
        T* const park = pool + lane;
#pragma unroll
        for (int q = 0; q < RT; ++q)
#pragma unroll
          for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v) park[((q * 4 + ct) * 4 + v) * 64] = acc[q][ct][v];
        // compiler barrier: the parking stores stay here (sunk into the
        // step they would keep the accumulators live)
        asm volatile("" ::: "memory");
        // (the stored inverses go out from all waves after the barrier)
        ll_diag_step<T>(Ds, As12, Li, Ui, Ms, nullptr, p == 0, thr, fail, lane);
        asm volatile("" ::: "memory");
#pragma unroll
        for (int q = 0; q < RT; ++q)
#pragma unroll
          for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[q][ct][v] = park[((q * 4 + ct) * 4 + v) * 64];
      }
      __syncthreads();
      {
        // the diagonal block's stored inverses [Linv | Uinv], column-major,
        // from every wave (off the diagonal step's serial path; Li and Ui
        // are next written by a diagonal step several barriers later), one
        // element in flight per lane at a time: the slab's accumulators
        // are live here
        T* const out = dinv + (int64_t)ib * 2 * kTri * kTri;
        const int t0 = opaque(tid);
#pragma unroll 1
        for (int q = 0; q < 4; ++q) {
          const int idx = t0 + 512 * q;
          const int e = idx & 1023;
          out[idx] = (q < 2 ? &Li[0][0] : &Ui[0][0])[(e % 32) * 33 + e / 32];
        }
      }
#pragma unroll
      for (int q = 0; q < RT; ++q) {
        const int rt = wid + 8 * q;
        if ((rt >> 1) == ib) {
#pragma unroll
          for (int cc = 0; cc < 2; ++cc)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              acc[q][cp + cc][v] = Ds[16 * (rt & 1) + lr][16 * cc + cr0 + v * DV];
              if (p == 0) acc[q][2 + cc][v] = As12[16 * (rt & 1) + lr][16 * cc + cr0 + v * DV];
            }
        } else if (rt >= 2 * ib + 2) {
          // L21 = A21 U11^-1: D(i = col, j = row) = sum_k Uinv(k, col) A21(row, k);
          // k-step (cc, v) takes B from accumulator element v (k = 16 cc + crow)
          V4 l0 = zero4<T>(), l1 = zero4<T>();
#pragma unroll
          for (int cc = 0; cc < 2; ++cc)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int k = 16 * cc + cr0 + v * DV;
              const T bb = acc[q][cp + cc][v];
              if (cc == 0) l0 = mfma(Ui[k][lr], bb, l0);
              l1 = mfma(Ui[k][16 + lr], bb, l1);
            }
#pragma unroll
          for (int v = 0; v < 4; ++v) fail |= !(fabs(l0[v]) <= thr) || !(fabs(l1[v]) <= thr);
          acc[q][cp] = l0;
          acc[q][cp + 1] = l1;
          if (p == 0) {
            // A22 -= L21 U12: D(i = col, j = row) -= sum_k U12(k, col) L21(row, k)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc)
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                const int k = 16 * cc + cr0 + v * DV;
                const T bb = acc[q][cc][v];
#pragma unroll
                for (int oc = 0; oc < 2; ++oc) acc[q][2 + oc] = mfma(-As12[k][16 * oc + lr], bb, acc[q][2 + oc]);
              }
          }
        }
      }
      if (p == 0) __syncthreads();  // Ds, As12 reused by the second panel
    }
    if (fail) bad = 1;
    __syncthreads();
    if (bad) {
      if (tid == 0) retry0[b] = 1;
      return;
    }
    {
      // the slab's factors
      const __amdgpu_buffer_rsrc_t ra = buf_rsrc(A + (int64_t)j0 * lda);
      const uint32_t vo = lr * (uint32_t)sizeof(T) + (uint32_t)cr0 * ldab_j;
#pragma unroll
      for (int q = 0; q < RT; ++q)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
          for (int v = 0; v < 4; ++v)
            buf_store(acc[q][ct][v], ra, vo, 16u * sizeof(T) * (wid + 8 * q) + (16 * ct + v * DV) * ldab_j);
    }
  }
  for (int t = tid; t < n; t += 512) {
    ipiv0[(int64_t)b * n + t] = t + 1;
    perm0[(int64_t)b * n + t] = t;
  }
  if (tid == 0) {
    info0[b] = 0;
    retry0[b] = 0;
  }
}

// =====================================================================
// Batched LU for n <= 128 (the 2r x 2r cores at rank <= 64): ONE wave per
// block, each lane holding panel rows lane and lane + 64 in registers.  The
// pivot search is a DPP reduction inside the wave and the pivot row is
// broadcast through LDS, so a pivot column costs no workgroup barrier.
// Pivot rule, implicit row tracking, U12 and the trailing update are those
// of k_getrf_reg.
// =====================================================================
template <typename T>
__device__ __forceinline__ void trtri_block(const T* __restrict__ LU, int64_t lda, int n, int ib,
                                            T* __restrict__ out, T (*M0)[kTri + 1],
                                            T (*M1)[kTri + 1], int lane);

// Orders the accesses of a workgroup's one running wave (LDS and global):
// the fences of a workgroup barrier without the barrier, so the LU below
// may also run in one wave of a wider workgroup whose other waves have left.
__device__ __forceinline__ void w64_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// LDS of the one-wave pivoted LU (w64_factor).
struct W64Lds {
  double* prow;  // NB
  int* pivs;     // NB
  double* L11;   // NB (NB + 1), also an inversion tile
  double* Xi;    // NB (NB + 1), also an inversion tile
  double* Tt;    // NB (NB + 1)
  int* perm;     // 128
  int* mv;       // 2 NB: the panel's moving rows (move_rows)
  int* mcount;
};

// Pivoted LU of one block of n <= 128 rows by ONE wave (the body of
// k_getrf_w64): rows lane and lane + 64 in registers, per pivot column a
// DPP argmax and an LDS broadcast of the pivot row; ipiv (1-based), permo,
// *infoo; with dinv non-null the inverses of the factor's 32x32 diagonal
// blocks as well.
template <int NB>
__device__ __forceinline__ void w64_factor(double* __restrict__ A, int64_t lda, int n,
                                           int* __restrict__ ipiv, int* __restrict__ permo,
                                           int* __restrict__ infoo, double* __restrict__ dinv,
                                           const W64Lds& sm, int lane) {
  static_assert(2 * NB <= 64, "a moving row per lane");
  constexpr int R = 2;  // rows per lane
  double* const prow = sm.prow;
  int* const pivs = sm.pivs;
  double* const L11 = sm.L11;
  double* const Xi = sm.Xi;
  double* const Tt = sm.Tt;
  int* const perm = sm.perm;
  int* const mv = sm.mv;
  int& mcount = *sm.mcount;
  int info = 0;  // lane 0
  for (int q = 0; q < R; ++q)
    if (q * 64 + lane < n) perm[q * 64 + lane] = q * 64 + lane;

  for (int j0 = 0; j0 < n; j0 += NB) {
    const int nb = min(NB, n - j0);
    const int rows = n - j0;
    double p[R][NB];
    int pos[R];
    bool live[R];
    if (lane == 0) mcount = 0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int rr = q * 64 + lane;
      pos[q] = rr;
      live[q] = rr < rows;
      const double* Arow = A + (int64_t)(j0 + rr) + (int64_t)j0 * lda;
#pragma unroll
      for (int c = 0; c < NB; ++c) p[q][c] = (live[q] && c < nb) ? Arow[(int64_t)c * lda] : 0.0;
    }
#pragma unroll
    for (int jj = 0; jj < NB; ++jj) {
      if (jj < nb) {
        // this lane's best row (larger key, then smaller position)
        unsigned long long lk = 0ULL;
        int lq = 0, lp = INT_MAX;
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const double a = p[q][jj];
          const unsigned long long k =
              live[q] ? ((a == a) ? (unsigned long long)__double_as_longlong(fabs(a)) + 1ULL : 1ULL) : 0ULL;
          const int kp = k ? pos[q] : INT_MAX;
          const bool tk = k > lk || (k == lk && kp < lp);
          lk = tk ? k : lk;
          lp = tk ? kp : lp;
          lq = tk ? q : lq;
        }
        const unsigned long long wmax = wave_max_u64(lk);
        const unsigned long long bal = __ballot(lk == wmax);
        int wl;
        if (__popcll(bal) == 1) {
          wl = __ffsll((long long)bal) - 1;
        } else {
          int qq = (lk == wmax) ? lp : INT_MAX;
#pragma unroll
          for (int off = 32; off > 0; off >>= 1) qq = min(qq, __shfl_xor(qq, off, 64));
          wl = __ffsll((long long)__ballot(lk == wmax && lp == qq)) - 1;
        }
        const int wq = __builtin_amdgcn_readlane(lq, wl);
        const int bpos = __builtin_amdgcn_readlane(lp, wl);
        // two statically indexed stores (a select between p[0] and p[1]
        // would become a dynamic index and push p to scratch)
#pragma unroll
        for (int q = 0; q < R; ++q) {
          if (lane == wl && wq == q) {
#pragma unroll
            for (int c = 0; c < NB; ++c)
              if (c >= jj) prow[c] = p[q][c];
          }
        }
        if (lane == 0) pivs[jj] = bpos;
        w64_sync();  // one wave: orders the LDS broadcast
        const double pv = prow[jj];
#pragma unroll
        for (int q = 0; q < R; ++q) {
          if (lane == wl && q == wq) {
            live[q] = false;
            pos[q] = jj;
          } else if (pos[q] == jj) {
            pos[q] = bpos;
          }
        }
        if (pv != 0.0) {
          const bool use_rcp = fabs(pv) >= DBL_MIN;
          const double rcp = 1.0 / pv;
#pragma unroll
          for (int q = 0; q < R; ++q) {
            if (live[q]) {
              const double l = use_rcp ? p[q][jj] * rcp : p[q][jj] / pv;
              p[q][jj] = l;
#pragma unroll
              for (int c = 0; c < NB; ++c)
                if (c > jj) p[q][c] -= l * prow[c];
            }
          }
        } else if (lane == 0 && info == 0) {
          info = j0 + jj + 1;
        }
        w64_sync();  // prow is rewritten by the next column
      }
    }
    // rows to their final positions; L11 (padded) to LDS
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int rr = q * 64 + lane;
      if (rr < rows) {
        double* Arow = A + (int64_t)(j0 + pos[q]) + (int64_t)j0 * lda;
#pragma unroll
        for (int c = 0; c < NB; ++c)
          if (c < nb) Arow[(int64_t)c * lda] = p[q][c];
        if (pos[q] < NB) {
#pragma unroll
          for (int c = 0; c < NB; ++c) L11[pos[q] * (NB + 1) + c] = p[q][c];
        }
      }
    }
    for (int t = nb + lane; t < NB; t += 64)
#pragma unroll
      for (int c = 0; c < NB; ++c) L11[t * (NB + 1) + c] = 0.0;
    if (lane < nb) ipiv[j0 + lane] = j0 + pivs[lane] + 1;
    if (lane == 0) {
      for (int jj = 0; jj < nb; ++jj) {
        const int pp = pivs[jj];
        if (pp != jj) {
          const int tmp = perm[j0 + jj];
          perm[j0 + jj] = perm[j0 + pp];
          perm[j0 + pp] = tmp;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int rr = q * 64 + lane;
      if (rr < rows && pos[q] != rr) mv[atomicAdd(&mcount, 1)] = rr | (pos[q] << 16);
    }
    w64_sync();
    const int m2 = n - j0 - nb;
    move_rows<32>(A, lda, n, j0, nb, mv, mcount, 0, 1, lane);
    if (m2 > 0) {
      for (int e = lane; e < NB * NB; e += 64) {
        const int i = e / NB, j = e % NB;
        Xi[i * (NB + 1) + j] = (i == j) ? 1.0 : 0.0;
      }
      w64_sync();
      for (int h = 1; h < NB; h <<= 1) {
        for (int e = lane; e < NB * h / 2; e += 64) {
          const int q = e / (h * h), r = e % (h * h);
          const int s0 = q * 2 * h, i = s0 + h + r / h, j = s0 + r % h;
          double acc = 0.0;
          for (int k = j; k < s0 + h; ++k) acc += L11[i * (NB + 1) + k] * Xi[k * (NB + 1) + j];
          Tt[i * (NB + 1) + j] = acc;
        }
        w64_sync();
        for (int e = lane; e < NB * h / 2; e += 64) {
          const int q = e / (h * h), r = e % (h * h);
          const int s0 = q * 2 * h, i = s0 + h + r / h, j = s0 + r % h;
          double acc = 0.0;
          for (int k = s0 + h; k <= i; ++k) acc += Xi[i * (NB + 1) + k] * Tt[k * (NB + 1) + j];
          Xi[i * (NB + 1) + j] = -acc;
        }
        w64_sync();
      }
      const int o = j0 + nb;
      const int nct = (m2 + 15) >> 4;
      for (int ct = 0; ct < nct; ++ct) {
        const int col = ct * 16 + (lane & 15);
        d4 u[NB / 16];
#pragma unroll
        for (int q = 0; q < NB / 16; ++q) u[q] = d4zero();
#pragma unroll
        for (int kk = 0; kk < NB / 4; ++kk) {
          const int k = kk * 4 + (lane >> 4);
          const double bf = (k < nb && col < m2) ? A[(int64_t)(j0 + k) + (int64_t)(o + col) * lda] : 0.0;
#pragma unroll
          for (int q = 0; q < NB / 16; ++q)
            u[q] = mfma(Xi[(q * 16 + (lane & 15)) * (NB + 1) + k], bf, u[q]);
        }
#pragma unroll
        for (int q = 0; q < NB / 16; ++q)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = q * 16 + crow(lane, v);
            if (col < m2 && r < nb) A[(int64_t)(j0 + r) + (int64_t)(o + col) * lda] = u[q][v];
          }
      }
      w64_sync();
      // trailing update, 32x32 blocks
      const int t32 = (m2 + 31) >> 5;
      for (int blk = 0; blk < t32 * t32; ++blk) {
        const int r0 = (blk % t32) * 32, c0 = (blk / t32) * 32;
        d4 acc[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int bb = 0; bb < 2; ++bb) {
            const int col = c0 + bb * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int row = r0 + a * 16 + crow(lane, v);
              acc[a][bb][v] = (row < m2 && col < m2) ? A[(int64_t)(o + row) + (int64_t)(o + col) * lda] : 0.0;
            }
          }
        double af[NB / 4][2], bf[NB / 4][2];
#pragma unroll
        for (int kk = 0; kk < NB / 4; ++kk) {
          const int k = kk * 4 + (lane >> 4);
#pragma unroll
          for (int a = 0; a < 2; ++a) {
            const int row = r0 + a * 16 + (lane & 15);
            af[kk][a] = (k < nb && row < m2) ? A[(int64_t)(o + row) + (int64_t)(j0 + k) * lda] : 0.0;
          }
#pragma unroll
          for (int bb = 0; bb < 2; ++bb) {
            const int col = c0 + bb * 16 + (lane & 15);
            bf[kk][bb] = (k < nb && col < m2) ? A[(int64_t)(j0 + k) + (int64_t)(o + col) * lda] : 0.0;
          }
        }
#pragma unroll
        for (int kk = 0; kk < NB / 4; ++kk)
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int bb = 0; bb < 2; ++bb) acc[a][bb] = mfma(-af[kk][a], bf[kk][bb], acc[a][bb]);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int bb = 0; bb < 2; ++bb) {
            const int col = c0 + bb * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int row = r0 + a * 16 + crow(lane, v);
              if (row < m2 && col < m2) A[(int64_t)(o + row) + (int64_t)(o + col) * lda] = acc[a][bb][v];
            }
          }
      }
    }
    w64_sync();
  }
  for (int q = 0; q < R; ++q)
    if (q * 64 + lane < n) permo[q * 64 + lane] = perm[q * 64 + lane];
  if (lane == 0) *infoo = info;
  if (dinv) {
    static_assert(NB >= kTri, "the inversion tiles fit L11 and Xi");
    // the factor is this wave's own stores: the barrier makes every lane's
    // visible to all of them
    w64_sync();
    double (*const M0)[kTri + 1] = reinterpret_cast<double (*)[kTri + 1]>(L11);
    double (*const M1)[kTri + 1] = reinterpret_cast<double (*)[kTri + 1]>(Xi);
    for (int ib = 0; ib < ntri(n); ++ib)
      trtri_block<double>(A, lda, n, ib,
                          dinv + (int64_t)ib * 2 * kTri * kTri, M0,
                          M1, lane);
  }
}

template <int NB>
__global__ void __launch_bounds__(64)
k_getrf_w64(double* __restrict__ A0, int64_t lda, int64_t bstride, int n,
            int* __restrict__ ipiv0, int* __restrict__ perm0, int* __restrict__ info0,
            const int* __restrict__ only, const double* __restrict__ src = nullptr,
            int64_t lds = 0, int64_t sbstride = 0, double* __restrict__ dinv0 = nullptr) {
  // `only` (optional): factor block b only where only[b] != 0; `src`
  // (optional): block b is first copied from src + b*sbstride (ld lds);
  // `dinv0` (optional): the inverses of the factor's 32x32 diagonal blocks
  // are formed too (dinv_elems(n) per block)
  if (only && !only[blockIdx.x]) return;
  if (src) {
    const double* S0 = src + (int64_t)blockIdx.x * sbstride;
    double* D0 = A0 + (int64_t)blockIdx.x * bstride;
    for (int c = 0; c < n; ++c)
      for (int r = threadIdx.x; r < n; r += 64) D0[(int64_t)r + (int64_t)c * lda] = S0[(int64_t)r + (int64_t)c * lds];
    __syncthreads();
  }
  __shared__ double prow[NB];
  __shared__ int pivs[NB];
  __shared__ double L11[NB * (NB + 1)];
  __shared__ double Xi[NB * (NB + 1)];
  __shared__ double Tt[NB * (NB + 1)];
  __shared__ int perm[128];
  __shared__ int mv[2 * NB];
  __shared__ int mcount;
  const int b = blockIdx.x;
  const W64Lds sm{prow, pivs, L11, Xi, Tt, perm, mv, &mcount};
  w64_factor<NB>(A0 + (int64_t)b * bstride, lda, n, ipiv0 + (int64_t)b * n, perm0 + (int64_t)b * n,
                 info0 + b, dinv0 ? dinv0 + (int64_t)b * dinv_elems(n) : nullptr, sm,
                 (int)threadIdx.x);
}

static size_t getrf_lds_bytes(int nb, int n) {
  return (size_t)(nb * n + 2 * nb + nb * (nb + 1) + 4) * sizeof(double) +
         (size_t)(4 + nb + n) * sizeof(int);
}

hipError_t launch_getrf(int n, double* A, int64_t lda, int64_t bstride, int batch, int* ipiv,
                        int* perm, int* info, hipStream_t st) {
  if (batch <= 0 || n <= 0) return hipSuccess;
  if (n > kMaxBlock) return hipErrorInvalidValue;
  if (n <= 128) {
    hipLaunchKernelGGL(k_getrf_w64<32>, dim3(batch), dim3(64), 0, st, A, lda, bstride, n, ipiv,
                       perm, info, (const int*)nullptr);
    return hipGetLastError();
  }
  if (n <= 256) {
    hipLaunchKernelGGL((k_getrf_reg<GETRF_NB, 256, 2, double>), dim3(batch), dim3(256), 0, st, A, lda, bstride,
                       n, ipiv, perm, info, (const int*)nullptr, (const double*)nullptr, (int64_t)0, (int64_t)0);
    return hipGetLastError();
  }
  if (n <= 512) {
    hipLaunchKernelGGL((k_getrf_reg<GETRF_NB, 512, 2, double>), dim3(batch), dim3(512), 0, st, A, lda, bstride,
                       n, ipiv, perm, info, (const int*)nullptr, (const double*)nullptr, (int64_t)0, (int64_t)0);
    return hipGetLastError();
  }
  const size_t lim = 160 * 1024;
  if (getrf_lds_bytes(32, n) <= lim) {
    const size_t bytes = getrf_lds_bytes(32, n);
    static bool attr = false;
    if (!attr) {
      (void)hipFuncSetAttribute((const void*)k_getrf<32>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)lim);
      attr = true;
    }
    hipLaunchKernelGGL(k_getrf<32>, dim3(batch), dim3(256), bytes, st, A, lda, bstride, n, ipiv,
                       perm, info);
  } else {
    const size_t bytes = getrf_lds_bytes(16, n);
    if (bytes > lim) return hipErrorInvalidValue;
    static bool attr = false;
    if (!attr) {
      (void)hipFuncSetAttribute((const void*)k_getrf<16>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)lim);
      attr = true;
    }
    hipLaunchKernelGGL(k_getrf<16>, dim3(batch), dim3(256), bytes, st, A, lda, bstride, n, ipiv,
                       perm, info);
  }
  return hipGetLastError();
}

__global__ void k_fill_int(int* __restrict__ a, int n, int v) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) a[i] = v;
}

// =====================================================================
// LU of the HODLR cores S = [I, A; B, I] (2r x 2r, r <= 64), one workgroup
// of four waves each.  With max|B| <= 1 partial pivoting takes the diagonal
// in each of the first r columns (ties go to the first index, as idamax), so
//   L = [I, 0; P2 B, L2],  U = [I, A; 0, U2],  L2 U2 = P2 (I - B A)
// -- the same factors dgetrf produces (its first 64-column panel is the
// identity block, its trailing update the GEMM I - B A).  The four waves
// stage A and B in LDS and form C = I - B A on MFMA (one 16-row block of C
// each); wave 0 then factors C alone, without barriers (row i of C in lane
// i, 64 columns in registers, the pivot row broadcast by readlane) and
// writes S's LU in LAPACK layout.  Cores with max|B| > 1 or a NaN in B are
// flagged in `general` and left to k_getrf_w64.
// =====================================================================
__global__ void __launch_bounds__(256)
k_getrf_core(double* __restrict__ A0, int64_t lda, int64_t bstride, int n2,
             int* __restrict__ ipiv0, int* __restrict__ perm0, int* __restrict__ info0,
             int* __restrict__ general, double* __restrict__ dinv0, int* __restrict__ tmask) {
  __shared__ double Cs[64 * 65];  // C = I - B A, [row*65 + col]
  __shared__ double Bs[64 * 65];  // B, [row*65 + k]
  __shared__ double As[64 * 65];  // A, [k*65 + col]; then the fast path's packed LU
  // the no-interchange attempt's 32x32 tiles (ll_diag_step)
  __shared__ double Dt[32][33], A12t[32][33], Lit[32][33], Uit[32][33];
  __shared__ __align__(16) double Mst[2][32][33];
  __shared__ unsigned long long wmk[4];
  __shared__ int needinv;  // the diagonal inverses are still to be formed
  const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
  const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 15, lk = lane >> 4;
  const int r = n2 >> 1;
  double* S = A0 + (int64_t)b * bstride;
  // (1) stage B = S(r:2r, 0:r) and A = S(0:r, r:2r) (coalesced columns,
  // wave w takes columns w, w+4, ...) and take max |B|; a NaN makes the key
  // maximal
  unsigned long long mk = 0ULL;
  if (lane < r) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int c = wid + 4 * i;
      if (c < r) {
        const double v = S[(int64_t)(r + lane) + (int64_t)c * lda];
        Bs[lane * 65 + c] = v;
        const unsigned long long k = (v == v) ? (unsigned long long)__double_as_longlong(fabs(v))
                                              : ~0ULL;
        mk = k > mk ? k : mk;
        As[lane * 65 + c] = S[(int64_t)lane + (int64_t)(r + c) * lda];
      }
    }
  }
  mk = wave_max_u64(mk);
  if (lane == 0) wmk[wid] = mk;
  __syncthreads();
  mk = umax64(umax64(wmk[0], wmk[1]), umax64(wmk[2], wmk[3]));
  const bool fallback = mk > (unsigned long long)__double_as_longlong(1.0);
  if (tid == 0) {
    general[b] = fallback ? 1 : 0;
    tmask[b] = 1;  // cleared below once this kernel has formed the inverses
  }
  if (fallback) {  // uniform over the workgroup
    // a general core: the pivoted LU of S and its diagonal inverses by wave
    // 0 alone (the others leave; it synchronizes only itself), in the LDS
    // of the structured path (B and A are read again from S)
    if (wid != 0) return;
    const W64Lds sm{Cs, reinterpret_cast<int*>(Bs), Cs + 64, Cs + 64 + 32 * 33,
                    Cs + 64 + 2 * 32 * 33, reinterpret_cast<int*>(Bs) + 32,
                    reinterpret_cast<int*>(Bs) + 160, reinterpret_cast<int*>(Bs) + 224};
    w64_factor<32>(S, lda, n2, ipiv0 + (int64_t)b * n2, perm0 + (int64_t)b * n2, info0 + b,
                   dinv0 + (int64_t)b * dinv_elems(n2), sm, lane);
    return;
  }
  // (2) C = I - B A on MFMA from LDS: wave w forms 16-row block w of C,
  // its four column tiles side by side (independent MFMA chains)
  const int nt = (r + 15) >> 4;
  if (wid < nt) {
    const int tm = wid;
    const int ar = tm * 16 + lr;
    d4 acc[4];
#pragma unroll
    for (int tn = 0; tn < 4; ++tn) acc[tn] = d4zero();
    for (int k0 = 0; k0 < r; k0 += 4) {
      const int k = k0 + lk;
      const double af = (k < r && ar < r) ? Bs[ar * 65 + k] : 0.0;
#pragma unroll
      for (int tn = 0; tn < 4; ++tn) {
        const int bc = tn * 16 + lr;
        const double bf = (tn < nt && k < r && bc < r) ? As[k * 65 + bc] : 0.0;
        acc[tn] = mfma(af, bf, acc[tn]);
      }
    }
#pragma unroll
    for (int tn = 0; tn < 4; ++tn)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = tm * 16 + crow(lane, v), col = tn * 16 + lr;
        if (row < r && col < r) Cs[row * 65 + col] = (row == col ? 1.0 : 0.0) - acc[tn][v];
      }
  }
  __syncthreads();
  // (3) the LU of C by wave 0; waves 1-3 wait at the barrier before (5)
  if (wid == 0) {
  bool done = false;
  // (3a) No interchanges when partial pivoting keeps every diagonal: C,
  // padded with the identity to rp = 32 or 64 rows, factored in 32x32
  // blocks by the leaf LU's diagonal step (every multiplier checked
  // against 1 - 2^-20, so an accepted factorization is the one dgetrf
  // computes with P2 = I), L21 = C21 U11^-1 (checked) and
  // A22 = C22 - L21 U12 on MFMA.  The packed LU goes to As (A is no longer
  // needed: S keeps it); Cs is left intact for the pivoted path.  With r a
  // multiple of 32 the blocks of S's diagonal inverses are the identity
  // (rows < r) and the diagonal steps' L^-1 / U^-1, written here.
  {
    const int rp = r <= 32 ? 32 : 64;
    const bool inv = (r & 31) == 0;
    double* dv = dinv0 + (int64_t)b * dinv_elems(n2);
    const double thr = PivotMargin<double>::thr();
    bool fail = false;
    double (*const LUp)[65] = reinterpret_cast<double (*)[65]>(As);
    auto cpad = [&](int i, int j) -> double {
      return (i < r && j < r) ? Cs[i * 65 + j] : (i == j ? 1.0 : 0.0);
    };
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int idx = lane + 64 * q, i = idx >> 5, j = idx & 31;
      Dt[i][j] = cpad(i, j);
      if (rp == 64) A12t[i][j] = cpad(i, 32 + j);
    }
    wave_sync();
    ll_diag_step<double>(Dt, A12t, Lit, Uit, Mst, inv ? dv + (int64_t)(r / 32) * 2 * kTri * kTri : nullptr,
                         rp == 64, thr, fail, lane);
    wave_sync();
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int idx = lane + 64 * q, i = idx >> 5, j = idx & 31;
      LUp[i][j] = Dt[i][j];
      if (rp == 64) LUp[i][32 + j] = A12t[i][j];
    }
    if (rp == 64) {
      const int lr = lane & 15, lk = lane >> 4;
      // L21 = C21 U11^-1 (Uit holds U11^-1), checked
#pragma unroll
      for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
          d4 d = d4zero();
#pragma unroll
          for (int s2 = 0; s2 < 8; ++s2)
            d = mfma(cpad(32 + 16 * rt + lr, 4 * s2 + lk), Uit[4 * s2 + lk][16 * ct + lr], d);
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            fail |= !(fabs(d[v]) <= thr);
            LUp[32 + 16 * rt + Elem<double>::crow(lane, v)][16 * ct + lr] = d[v];
          }
        }
      wave_sync();
      // A22 = C22 - L21 U12 into the diagonal tile
#pragma unroll
      for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
          d4 c;
#pragma unroll
          for (int v = 0; v < 4; ++v)
            c[v] = cpad(32 + 16 * rt + Elem<double>::crow(lane, v), 32 + 16 * ct + lr);
#pragma unroll
          for (int s2 = 0; s2 < 8; ++s2)
            c = mfma(-LUp[32 + 16 * rt + lr][4 * s2 + lk], A12t[4 * s2 + lk][16 * ct + lr], c);
#pragma unroll
          for (int v = 0; v < 4; ++v) Dt[16 * rt + Elem<double>::crow(lane, v)][16 * ct + lr] = c[v];
        }
      wave_sync();
      ll_diag_step<double>(Dt, A12t, Lit, Uit, Mst,
                           inv ? dv + (int64_t)(r / 32 + 1) * 2 * kTri * kTri : nullptr, false,
                           thr, fail, lane);
      wave_sync();
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int idx = lane + 64 * q, i = idx >> 5, j = idx & 31;
        LUp[32 + i][32 + j] = Dt[i][j];
      }
    }
    wave_sync();
    if (__ballot(fail) == 0ULL) {
      // S's LU: rows 0..r-1 ([I | A]) and the B block stay as they are;
      // rows r..2r-1, columns r..2r-1 take L2\U2
      if (lane < r) {
        double* row = S + (int64_t)(r + lane);
        for (int c = 0; c < r; ++c) row[(int64_t)(r + c) * lda] = LUp[lane][c];
      }
      int* ipiv = ipiv0 + (int64_t)b * n2;
      int* perm = perm0 + (int64_t)b * n2;
      for (int i = lane; i < n2; i += 64) {
        ipiv[i] = i + 1;
        perm[i] = i;
      }
      if (inv) {
        // identity inverses of the blocks of rows < r
        for (int q = 0; q < r / 32; ++q)
#pragma unroll
          for (int e = 0; e < 16; ++e) {
            const int idx = lane + 64 * e;
            const double one = ((idx & 31) == (idx >> 5)) ? 1.0 : 0.0;
            dv[(int64_t)q * 2 * kTri * kTri + idx] = one;
            dv[(int64_t)q * 2 * kTri * kTri + 1024 + idx] = one;
          }
      }
      if (lane == 0) {
        info0[b] = 0;
        needinv = inv ? 0 : 1;
      }
      done = true;
    }
  }
  if (!done) {
  // (3b) LU of C with implicit pivoting (lane i keeps row i of C; rows
  // never move, pos is a row's LAPACK position), blocked like dgetrf: a
  // 16-column panel is factored with rank-1 updates inside it (the pivot
  // row's panel elements broadcast by readlane), then the panel's pivot
  // rows form U12 = L11^-1 A12 (lane = trailing column) and the live rows
  // take A22 -= L21 U12 (lane = row), both through Cs.  Per column only the
  // panel's <= 16 elements cross lanes, where the unblocked form moved all
  // 64.  Cs ends up holding every row of the LU.
  int pos = lane;
  bool live = lane < r;
  int info = 0;   // lane 0
  int mypiv = 0;  // lane jj: pivot position chosen at step jj
  double* const crow_ = Cs + lane * 65;
  // one panel's code, looped: the unrolled form ran to 150 KB of code
#pragma unroll 1
  for (int k0 = 0; k0 < 64; k0 += 16) {
    if (k0 < r) {
      double p[16];
#pragma unroll
      for (int c = 0; c < 16; ++c) p[c] = (lane < r && k0 + c < r) ? crow_[k0 + c] : 0.0;
      int pl[16];  // pivot lane of each panel step (wave-uniform)
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) {
        pl[jj] = 0;
        const int j = k0 + jj;
        if (j < r) {
          const unsigned long long key = live ? Elem<double>::key(p[jj]) : 0ULL;
          const int cpos = key ? pos : INT_MAX;
          unsigned long long wmax;
          const int wl = wave_argmax_key(key, Elem<double>::key32(key), cpos, wmax);
          pl[jj] = wl;
          const int bpos = __builtin_amdgcn_readlane(pos, wl);
          if (lane == j) mypiv = bpos;
          const double pv = readlane_f64(p[jj], wl);
          if (lane == wl) {
            live = false;
            pos = j;
          } else if (pos == j) {
            pos = bpos;
          }
          if (pv != 0.0) {
            double l;
            if (fabs(pv) >= DBL_MIN)
              l = p[jj] * (1.0 / pv);
            else
              l = p[jj] / pv;
            if (live) {
              p[jj] = l;
#pragma unroll
              for (int c = 0; c < 16; ++c)
                if (c > jj) p[c] -= l * readlane_f64(p[c], wl);
            }
          } else if (lane == 0 && info == 0) {
            info = r + j + 1;
          }
        }
      }
      if (lane < r) {
#pragma unroll
        for (int c = 0; c < 16; ++c)
          if (k0 + c < r) crow_[k0 + c] = p[c];
      }
      if (k0 + 16 < r) {
        wave_sync();
        // U12: lane t takes trailing column k0 + 16 + t through the
        // panel's pivot rows, in pivot order
        const int c = k0 + 16 + lane;
        if (c < r) {
          double u[16];
#pragma unroll
          for (int k = 0; k < 16; ++k) {
            double* prow = Cs + pl[k] * 65;
            double v = prow[c];
#pragma unroll
            for (int m = 0; m < 16; ++m)
              if (m < k) v -= prow[k0 + m] * u[m];
            u[k] = v;
            prow[c] = v;
          }
        }
        wave_sync();
        // A22 -= L21 U12 on the live rows
        if (live) {
          for (int cc = k0 + 16; cc < r; ++cc) {
            double a = crow_[cc];
#pragma unroll
            for (int k = 0; k < 16; ++k) a -= p[k] * Cs[pl[k] * 65 + cc];
            crow_[cc] = a;
          }
        }
        wave_sync();
      }
    }
  }
  wave_sync();
  // (4) S's LU in LAPACK layout: rows r + pos get this lane's L2\U2 row and
  // its B row (the interchanges apply to the first r columns too)
  if (lane < r) {
    double* row = S + (int64_t)(r + pos);
#pragma unroll
    for (int c = 0; c < 64; ++c) {
      if (c < r) {
        row[(int64_t)c * lda] = Bs[lane * 65 + c];
        row[(int64_t)(r + c) * lda] = crow_[c];
      }
    }
  }
  int* ipiv = ipiv0 + (int64_t)b * n2;
  int* perm = perm0 + (int64_t)b * n2;
  if (lane < r) {
    ipiv[lane] = lane + 1;
    perm[lane] = lane;
    ipiv[r + lane] = r + mypiv + 1;
    perm[r + pos] = r + lane;
  }
  if (lane == 0) {
    info0[b] = info;
    needinv = 1;
  }
  }  // !done
  }  // wave 0
  // (5) the inverses of S's 32x32 diagonal blocks where (3) did not form
  // them, one block per wave, from S's LU as (3) stored it; each wave's
  // tiles are LDS the factorization no longer uses
  __syncthreads();
  if (needinv) {
    double(*M0)[33];
    double(*M1)[33];
    if (wid == 0) {
      M0 = Mst[0]; M1 = Mst[1];
    } else if (wid == 1) {
      M0 = Dt; M1 = A12t;
    } else if (wid == 2) {
      M0 = Lit; M1 = Uit;
    } else {
      M0 = reinterpret_cast<double (*)[33]>(Cs);
      M1 = reinterpret_cast<double (*)[33]>(Cs + 32 * 33);
    }
    double* dv = dinv0 + (int64_t)b * dinv_elems(n2);
    for (int ib = wid; ib < ntri(n2); ib += 4)
      trtri_block<double>(S, lda, n2, ib, dv + (int64_t)ib * 2 * kTri * kTri, M0, M1, lane);
  }
  if (tid == 0) tmask[b] = 0;
}

// Cores of rank 64 < r <= 128 (S = [I, A; B, I], n2 = 2r <= 256).  The
// same structure as k_getrf_core, with the Schur complement too large for
// one wave: k_core_schur128 forms C = I - B A per core on MFMA (operands
// read from S; padded with the identity to 128 x 128, ld kC) and flags
// cores with max|B| > 1 or a NaN in B as general; C is then factored by
// the 128-row leaf LU (its fast path, else the pivoted kernel: P2 C =
// L2 U2 with dgetrf's pivots -- the identity padding is never chosen as a
// pivot of the first r columns, where it holds zeros); k_core_compose
// writes S's LU, L = [I, 0; P2 B, L2], U = [I, A; 0, U2], its pivots
// [1..r, r + ipiv2] and INFO.  General cores are factored by the pivoted
// kernel in a masked launch.
constexpr int kC = 136;  // leading dimension of the padded Schur complements
template <typename T>
static hipError_t getrf_leaf_impl(int n, const double* K, int64_t ldk, int64_t kbs, T* LU,
                                  int64_t ldl, int64_t lbs, int batch, T* dinv, int* ipiv,
                                  int* perm, int* info, int* retry, hipStream_t st,
                                  bool fallback_inverses);

__global__ void __launch_bounds__(256)
k_core_schur128(const double* __restrict__ S0, int64_t lda, int64_t bstride, int r,
                double* __restrict__ C0, int* __restrict__ general, int* __restrict__ tmask) {
  // k-chunks of 16: B(:, k0:k0+16) and A(k0:k0+16, :) staged in LDS, the
  // next chunk's loads in registers under the current chunk's MFMAs
  __shared__ double Bk[128][17];   // B(row, k0 + k)
  __shared__ double Ak[16][129];   // A(k0 + k, col)
  __shared__ unsigned long long wmk[4];
  const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const double* S = S0 + (int64_t)b * bstride;
  double* C = C0 + (int64_t)b * kC * 128;
  unsigned long long mk = 0ULL;
  for (int c = wid; c < r; c += 4)
    for (int i = lane; i < r; i += 64) {
      const double v = S[(int64_t)(r + i) + (int64_t)c * lda];
      const unsigned long long k = (v == v) ? (unsigned long long)__double_as_longlong(fabs(v)) : ~0ULL;
      mk = k > mk ? k : mk;
    }
  mk = wave_max_u64(mk);
  if (lane == 0) wmk[wid] = mk;
  __syncthreads();
  mk = umax64(umax64(wmk[0], wmk[1]), umax64(wmk[2], wmk[3]));
  const bool gen = mk > (unsigned long long)__double_as_longlong(1.0);
  if (tid == 0) {
    general[b] = gen ? 1 : 0;
    tmask[b] = 1;  // every wide core's inverses are formed afterwards
  }
  const int nt = (r + 15) >> 4;  // tiles per side (<= 8)
  // wave w owns tiles t = w + 4q (q < 16): tm = t & 7, tn = t >> 3
  d4 acc[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) acc[q] = d4zero();
  // staging: thread e loads B(e & 127, k0 + (e >> 7) + 2j) and
  // A(k0 + (e >> 7) + 2j, e & 127), j < 8
  const int sr = tid & 127, sk = tid >> 7;
  double pb[8], pa[8];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = k0 + sk + 2 * j;
      const bool ok = !gen && k < r && sr < r;
      pb[j] = ok ? S[(int64_t)(r + sr) + (int64_t)k * lda] : 0.0;
      pa[j] = ok ? S[(int64_t)k + (int64_t)(r + sr) * lda] : 0.0;
    }
  };
  if (!gen) fetch(0);
  for (int k0 = 0; k0 < r && !gen; k0 += 16) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      Bk[sr][sk + 2 * j] = pb[j];
      Ak[sk + 2 * j][sr] = pa[j];
    }
    __syncthreads();
    if (k0 + 16 < r) fetch(k0 + 16);
#pragma unroll
    for (int s2 = 0; s2 < 4; ++s2) {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int t = wid + 4 * q, tm = t & 7, tn = t >> 3;
        if (tm < nt && tn < nt)
          acc[q] = mfma(Bk[tm * 16 + lr][4 * s2 + lk], Ak[4 * s2 + lk][tn * 16 + lr], acc[q]);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int t = wid + 4 * q, tm = t & 7, tn = t >> 3;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = tm * 16 + crow(lane, v), col = tn * 16 + lr;
      const double e = (row == col) ? 1.0 : 0.0;
      C[row + (int64_t)col * kC] = (!gen && row < r && col < r) ? e - acc[q][v] : e;
    }
  }
}

__global__ void __launch_bounds__(256)
k_core_compose(double* __restrict__ S0, int64_t lda, int64_t bstride, int r,
               const double* __restrict__ LU0, const int* __restrict__ ipiv2,
               const int* __restrict__ perm2, const int* __restrict__ info2,
               const int* __restrict__ general, int* __restrict__ ipiv0, int* __restrict__ perm0,
               int* __restrict__ info0) {
  __shared__ double Bt[128][65];  // a 64-column chunk of B, [row][col]
  const int b = blockIdx.x, tid = threadIdx.x;
  if (general[b]) return;  // uniform over the workgroup
  const int n2 = 2 * r;
  double* S = S0 + (int64_t)b * bstride;
  const double* LU = LU0 + (int64_t)b * kC * 128;
  const int* pv = perm2 + (int64_t)b * 128;
  // L2 \ U2 into S(r:2r, r:2r)
  for (int e = tid; e < r * r; e += 256) {
    const int i = e % r, c = e / r;
    S[(int64_t)(r + i) + (int64_t)(r + c) * lda] = LU[i + (int64_t)c * kC];
  }
  // P2 B into S(r:2r, 0:r): row i takes B's row perm2[i], staged by chunks
  for (int c0 = 0; c0 < r; c0 += 64) {
    const int w = min(64, r - c0);
    for (int e = tid; e < r * w; e += 256) {
      const int i = e % r, c = e / r;
      Bt[i][c] = S[(int64_t)(r + i) + (int64_t)(c0 + c) * lda];
    }
    __syncthreads();
    for (int e = tid; e < r * w; e += 256) {
      const int i = e % r, c = e / r;
      S[(int64_t)(r + i) + (int64_t)(c0 + c) * lda] = Bt[pv[i]][c];
    }
    __syncthreads();
  }
  int* ipiv = ipiv0 + (int64_t)b * n2;
  int* perm = perm0 + (int64_t)b * n2;
  for (int i = tid; i < r; i += 256) {
    ipiv[i] = i + 1;
    perm[i] = i;
    ipiv[r + i] = r + ipiv2[(int64_t)b * 128 + i];
    perm[r + i] = r + pv[i];
  }
  if (tid == 0) {
    const int f = info2[b];
    info0[b] = (f > 0 && f <= r) ? r + f : 0;
  }
}

hipError_t launch_getrf_core(int n2, double* S, int64_t lda, int64_t bstride, int batch, int* ipiv,
                             int* perm, int* info, int* general, double* dinv, int* tmask,
                             const CoreSchur* cs, hipStream_t st) {
  if (batch <= 0 || n2 <= 0) return hipSuccess;
  if (n2 > 128 || (n2 & 1)) {
    const int r = n2 >> 1;
    if ((n2 & 1) || n2 > 256 || !cs) {
      // every core needs its inverses formed
      hipLaunchKernelGGL(k_fill_int, dim3((batch + 255) / 256), dim3(256), 0, st, tmask, batch, 1);
      return launch_getrf(n2, S, lda, bstride, batch, ipiv, perm, info, st);
    }
    hipLaunchKernelGGL(k_core_schur128, dim3(batch), dim3(256), 0, st, S, lda, bstride, r, cs->C,
                       general, tmask);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // C's diagonal inverses serve the fast kernel's own chunk solves; those
    // of the blocks refactored with interchanges are not needed (S's
    // inverses are formed afterwards)
    e = getrf_leaf_impl<double>(128, cs->C, kC, (int64_t)kC * 128, cs->LU, kC, (int64_t)kC * 128,
                                batch, cs->dinv, cs->ipiv, cs->perm, cs->info, cs->retry, st,
                                false);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_core_compose, dim3(batch), dim3(256), 0, st, S, lda, bstride, r, cs->LU,
                       cs->ipiv, cs->perm, cs->info, general, ipiv, perm, info);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // general cores (max|B| > 1): the pivoted LU of S in place
    hipLaunchKernelGGL((k_getrf_reg<GETRF_NB, 256, 2, double>), dim3(batch), dim3(256), 0, st, S,
                       lda, bstride, n2, ipiv, perm, info, (const int*)general,
                       (const double*)nullptr, (int64_t)0, (int64_t)0);
    return hipGetLastError();
  }
  // (general cores take the pivoted LU inside the kernel)
  hipLaunchKernelGGL(k_getrf_core, dim3(batch), dim3(256), 0, st, S, lda, bstride, n2, ipiv, perm,
                     info, general, dinv, tmask);
  return hipGetLastError();
}


// =====================================================================
// Inverses of the 32x32 diagonal blocks of L (unit lower) and U (upper),
// stored per block as [Linv | Uinv], each 32x32 column-major.  A partial
// last block is padded with the identity.  Lane c < 32 computes column c
// of Linv, lane 32 + c column c of Uinv (substitution on e_c).
// =====================================================================
// One wave: the inverses of diagonal block ib of the LU at LU (ld lda, n
// rows) into out ([Linv | Uinv], 32x32 column-major each), through the two
// LDS tiles M0, M1 ([32][33] each).  Both triangles go in as
// lower-triangular tiles, [col][row]: M0 = L with its unit diagonal, M1 =
// U index-reversed (M1[31-c][31-r] = U[r][c]), so lanes 0..31 (columns of
// Linv) and 32..63 (columns of Uinv, reversed) run the same forward
// substitution side by side.  Only the calling wave touches the tiles:
// its LDS accesses execute in order, so wave syncs order them.
template <typename T>
__device__ __forceinline__ void trtri_block(const T* __restrict__ LU, int64_t lda, int n, int ib,
                                            T* __restrict__ out, T (*M0)[kTri + 1],
                                            T (*M1)[kTri + 1], int lane) {
  const int i0 = ib * kTri;
  const int bs = min(kTri, n - i0);
  // all 16 loads of a lane in flight together
  T v[kTri * kTri / 64];
#pragma unroll
  for (int q = 0; q < kTri * kTri / 64; ++q) {
    const int idx = lane + 64 * q, c = idx / kTri, r = idx % kTri;
    v[q] = (r < bs && c < bs) ? LU[(int64_t)(i0 + r) + (int64_t)(i0 + c) * lda] : ((r == c) ? T(1) : T(0));
  }
#pragma unroll
  for (int q = 0; q < kTri * kTri / 64; ++q) {
    const int idx = lane + 64 * q, c = idx / kTri, r = idx % kTri;
    M0[c][r] = (r > c) ? v[q] : ((r == c) ? T(1) : T(0));
    M1[kTri - 1 - c][kTri - 1 - r] = (r <= c) ? v[q] : T(0);
  }
  wave_sync();
  const int h = lane >> 5, c = lane & 31;
  // column c of Linv solves L x = e_c; column c of Uinv, reversed, solves
  // the reversed U with the unit vector at 31 - c
  const int c0 = h ? kTri - 1 - c : c;
  // column-oriented substitution: once x[k] is final, every later row takes
  // its term at once (independent updates, the LDS reads of a tile column
  // in flight together) -- the dependent chain is 32 steps, not 496
  T x[kTri];
#pragma unroll
  for (int i = 0; i < kTri; ++i) x[i] = (i == c0) ? T(1) : T(0);
  const T* const tile = h ? &M1[0][0] : &M0[0][0];
#pragma unroll
  for (int k = 0; k < kTri; ++k) {
    // an opaque column offset per step keeps the step's reads in the step
    const T* const col = tile + opaque(k * (kTri + 1));
    // opaque: x[k] is formed here (the divide by L's unit diagonal is
    // exact), so the updates below cannot be sunk into later per-row chains
    const T d = opaque(x[k] / col[k]);
    const T xk = opaque((k < c0) ? T(0) : d);
    x[k] = xk;
#pragma unroll
    for (int i = k + 1; i < kTri; ++i) x[i] -= col[i] * xk;
  }
  // the columns go out through an LDS tile (Linv, then Uinv), so every
  // store of the wave covers 512 consecutive bytes
  T (*const Ts)[kTri + 1] = M0;
#pragma unroll
  for (int hh = 0; hh < 2; ++hh) {
    wave_sync();  // the tile's last readers are done
    if (h == hh) {
#pragma unroll
      for (int i = 0; i < kTri; ++i) Ts[c][i] = hh ? x[kTri - 1 - i] : x[i];
    }
    wave_sync();
#pragma unroll
    for (int q = 0; q < kTri * kTri / 64; ++q) {
      const int idx = lane + 64 * q;
      out[hh * kTri * kTri + idx] = Ts[idx / kTri][idx % kTri];
    }
  }
}

// Inverses of the 32x32 diagonal blocks of L (unit lower) and U (upper),
// one wave per block.
template <typename T>
__global__ void __launch_bounds__(64)
k_trtri_diag(const T* __restrict__ LU0, int64_t lda, int64_t bstride, int n,
             T* __restrict__ dinv0, const int* __restrict__ only = nullptr) {
  __shared__ T Ms[2][kTri][kTri + 1];
  const int nblk = ntri(n);
  const int b = blockIdx.x / nblk, ib = blockIdx.x % nblk;
  if (only && !only[b]) return;  // (optional) blocks b with only[b] != 0
  trtri_block<T>(LU0 + (int64_t)b * bstride, lda, n, ib,
                 dinv0 + (int64_t)b * dinv_elems(n) + (int64_t)ib * 2 * kTri * kTri, Ms[0], Ms[1],
                 (int)threadIdx.x);
}

hipError_t launch_trtri_diag(int n, const double* LU, int64_t lda, int64_t bstride, int batch,
                             double* dinv, hipStream_t st) {
  if (batch <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_trtri_diag<double>, dim3(batch * ntri(n)), dim3(64), 0, st, LU, lda,
                     bstride, n, dinv);
  return hipGetLastError();
}

hipError_t launch_trtri_diag_masked(int n, const double* LU, int64_t lda, int64_t bstride,
                                    int batch, double* dinv, const int* only, hipStream_t st) {
  if (batch <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_trtri_diag<double>, dim3(batch * ntri(n)), dim3(64), 0, st, LU, lda,
                     bstride, n, dinv, only);
  return hipGetLastError();
}

// =====================================================================
// Batched solve with a stored LU (dgetrs), CW right-hand-side columns per
// workgroup held in LDS.  Blocked substitution in 32-row blocks: the
// diagonal block is applied as a product with its stored inverse, the
// off-diagonal part as an MFMA update of the remaining rows.
//   TRANS = false: X <- U^-1 L^-1 P^T X   (gather rows by perm)
//   TRANS = true : X <- P U^-T ... i.e. solves A^T X = B (scatter by perm)
// =====================================================================
template <int CW, bool TRANS>
__global__ void __launch_bounds__(256)
k_getrs(const double* __restrict__ LU0, int64_t lda, int64_t bstride, int n,
        const int* __restrict__ perm0, const double* __restrict__ dinv0, ColSeg X,
        int64_t xbstride, int ncols, int nchunk) {
  extern __shared__ double lds[];
  const int npad = ntri(n) * kTri;
  double* Xl = lds;  // [CW][npad]
  const int nblk = ntri(n);
  // XCD-aware order: consecutive logical ids (chunks of one item) stay on
  // one XCD so the item's LU factors are shared through that XCD's L2.
  // Workgroup w runs on XCD w % 8; XCD x receives q + (x < rem) of the G
  // workgroups, numbered contiguously -- a bijection of [0, G).
  const int G = gridDim.x;
  const int q = G >> 3, rem = G & 7;
  const int x = blockIdx.x & 7;
  const int logical = x * q + min(x, rem) + (blockIdx.x >> 3);
  const int b = logical / nchunk, chunk = logical % nchunk;
  const int c0 = chunk * CW;
  const double* LU = LU0 + (int64_t)b * bstride;
  const int* perm = perm0 + (int64_t)b * n;
  const double* dinv = dinv0 + (int64_t)b * dinv_elems(n);
  const int64_t xoff = (int64_t)b * xbstride;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;

  // gather
  for (int idx = tid; idx < CW * npad; idx += 256) {
    const int c = idx / npad, i = idx - c * npad;
    const int col = c0 + c;
    double v = 0.0;
    if (i < n && col < ncols) {
      const int src = TRANS ? i : perm[i];
      v = seg_in(X, col)[xoff + src];
    }
    Xl[c * npad + i] = v;
  }
  __syncthreads();

  constexpr int NCT = CW / 16;
  // two triangular sweeps
  for (int sweep = 0; sweep < 2; ++sweep) {
    // sweep 0: lower (TRANS: U^T), forward; sweep 1: upper (TRANS: L^T), backward
    const bool fwd = (sweep == 0);
    // non-TRANS: sweep0 uses Linv (slot 0) & L; sweep1 uses Uinv (slot 1) & U
    // TRANS    : sweep0 uses Uinv^T & U^T;         sweep1 uses Linv^T & L^T
    const int islot = TRANS ? (fwd ? 1 : 0) : (fwd ? 0 : 1);
    for (int step = 0; step < nblk; ++step) {
      const int ib = fwd ? step : nblk - 1 - step;
      const int i0 = ib * kTri;
      const double* Di = dinv + (int64_t)ib * 2 * kTri * kTri + islot * kTri * kTri;
      // (a) x_blk <- D * x_blk  (D = inverse of the diagonal block, or its
      // transpose).  The 32 x CW block is 2*NCT <= 4 MFMA tiles: one per wave.
      const bool has_tile = wid < 2 * NCT;
      const int rt_a = wid & 1, ct_a = wid >> 1;
      d4 a4 = d4zero();
      if (has_tile) {
#pragma unroll
        for (int kk = 0; kk < kTri / 4; ++kk) {
          const int k = kk * 4 + lk;
          const int row = rt_a * 16 + lr;
          const double af = TRANS ? Di[k + row * kTri] : Di[row + k * kTri];
          const double bf = Xl[(ct_a * 16 + lr) * npad + i0 + k];
          a4 = mfma(af, bf, a4);
        }
      }
      __syncthreads();
      if (has_tile) {
#pragma unroll
        for (int v = 0; v < 4; ++v)
          Xl[(ct_a * 16 + lr) * npad + i0 + rt_a * 16 + crow(lane, v)] = a4[v];
      }
      __syncthreads();
      // (b) rows outside the block: fwd -> rows [i0+32, npad), bwd -> rows [0, i0)
      const int rlo = fwd ? i0 + kTri : 0;
      const int rhi = fwd ? npad : i0;
      const int nrt = (rhi - rlo) >> 4;
      for (int t = wid; t < nrt * NCT; t += 4) {
        const int rt = t % nrt, ct = t / nrt;
        const int rbase = rlo + rt * 16;
        d4 a4;
#pragma unroll
        for (int v = 0; v < 4; ++v) a4[v] = Xl[(ct * 16 + lr) * npad + rbase + crow(lane, v)];
#pragma unroll
        for (int kk = 0; kk < kTri / 4; ++kk) {
          const int k = kk * 4 + lk;
          const int row = rbase + lr;      // row of the operand block
          const int kcol = i0 + k;         // column of the operand block
          double af = 0.0;
          if (row < n && kcol < n) {
            // non-TRANS: L/U[row][kcol];  TRANS: (U/L)^T[row][kcol] = LU[kcol][row]
            af = TRANS ? LU[(int64_t)kcol + (int64_t)row * lda] : LU[(int64_t)row + (int64_t)kcol * lda];
          }
          const double bf = Xl[(ct * 16 + lr) * npad + i0 + k];
          a4 = mfma(-af, bf, a4);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) Xl[(ct * 16 + lr) * npad + rbase + crow(lane, v)] = a4[v];
      }
      __syncthreads();
    }
  }

  // scatter
  for (int idx = tid; idx < CW * npad; idx += 256) {
    const int c = idx / npad, i = idx - c * npad;
    const int col = c0 + c;
    if (i < n && col < ncols) {
      const int dst = TRANS ? perm[i] : i;
      seg_out(X, col)[xoff + dst] = Xl[c * npad + i];
    }
  }
}

// =====================================================================
// Batched solve with a stored LU for n <= 512: one workgroup of 512 threads
// per (block, column chunk of NCT*16 columns).  The right-hand sides stay
// in MFMA accumulators for the whole solve; wave w owns the 16-row tiles
// w, w+8, w+16, w+24 (interleaved, so the triangular sweeps stay balanced).
// Per 32-row diagonal block: the owners stage the block in LDS, the block is
// multiplied by its stored inverse (one tile per wave), and every wave
// updates its own tiles on the far side of the block with L (or U) read from
// L2 -- each factor element is read once per chunk, two barriers per block.
//   TRANS = false: X <- U^-1 L^-1 P^T X;  TRANS = true: X <- P L^-T U^-T X.
// =====================================================================
template <int NCT, bool TRANS, typename T>
__global__ void __launch_bounds__(512)
k_getrs_reg(const T* __restrict__ LU0, int64_t lda, int64_t bstride, int n,
            const int* __restrict__ perm0, const T* __restrict__ dinv0, ColSeg X,
            int64_t xbstride, int ncols, int nchunk, const int* __restrict__ only) {
  constexpr int W = NCT * 16;
  typedef typename Elem<T>::V4 V4;
  __shared__ T xa[kTri * (W + 1)];  // raw block (owners -> all)
  __shared__ T xs[kTri * (W + 1)];  // solved block
  // the 32x32 diagonal inverse of each step, double-buffered: copied from
  // global memory straight into LDS (global_load_lds, 16 bytes per lane)
  // during the previous step's updates
  __shared__ __align__(16) T Ds[2][kTri * kTri];
  const int G = gridDim.x;
  const int gq = G >> 3, grem = G & 7;
  const int xcd = blockIdx.x & 7;
  const int logical = xcd * gq + min(xcd, grem) + (blockIdx.x >> 3);
  const int b = logical / nchunk, chunk = logical % nchunk;
  if (only && !only[b]) return;  // (optional) blocks b with only[b] != 0
  const int c0 = chunk * W;
  const T* LU = LU0 + (int64_t)b * bstride;
  const int* perm = perm0 + (int64_t)b * n;
  const T* dinv = dinv0 + (int64_t)b * dinv_elems(n);
  const int64_t xoff = (int64_t)b * xbstride;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int nblk = ntri(n);
  const int nrt = nblk * 2;  // 16-row tiles
  // factor fragment addressing (bytes): lane part and the step of kk;
  // 31 * lda * sizeof(T) < 2^32 is checked by the launcher
  const uint32_t ldb = (uint32_t)lda * sizeof(T);
  const uint32_t lane_off = TRANS ? lk * (uint32_t)sizeof(T) + lr * ldb : lr * (uint32_t)sizeof(T) + lk * ldb;
  const uint32_t kstep = TRANS ? 4 * (uint32_t)sizeof(T) : 4 * ldb;

  // Di of global step gs (sweep-major) into Ds[gs & 1]: waves w <
  // kTri*kTri*sizeof(T)/1024 copy 1 KB each.  A barrier does not wait for
  // these copies (they count on the vector-memory counter, not the LDS
  // one): the copying waves wait for them explicitly before the barrier
  // that precedes the first read (step (a1))
  constexpr int DSW = kTri * kTri * (int)sizeof(T) / 1024;
  static_assert(DSW <= 8 && DSW * 1024 == kTri * kTri * (int)sizeof(T), "Di copy split");
  auto prefetch_di = [&](int gs) {
    const bool f = gs < nblk;
    const int ibn = f ? gs : nblk - 1 - (gs - nblk);
    const int isl = TRANS ? (f ? 1 : 0) : (f ? 0 : 1);
    if (wid < DSW) {
      const char* src = (const char*)(dinv + (int64_t)ibn * 2 * kTri * kTri + isl * kTri * kTri) + wid * 1024 + lane * 16;
      char* dst = (char*)&Ds[gs & 1][0] + wid * 1024;
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                       (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
    }
  };
  prefetch_di(0);

  V4 acc[4][NCT];
  // gather (rows by perm for the forward solve)
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) {
      const int col = c0 + ct * 16 + lr;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = (wid + 8 * q) * 16 + Elem<T>::crow(lane, v);
        T val = 0.0;
        if (row < n && col < ncols) val = seg_in(X, col)[xoff + (TRANS ? row : perm[row])];
        acc[q][ct][v] = val;
      }
    }

  for (int sweep = 0; sweep < 2; ++sweep) {
    const bool fwd = (sweep == 0);
    for (int step = 0; step < nblk; ++step) {
      const int ib = fwd ? step : nblk - 1 - step;
      const int i0 = ib * kTri;
      const int gs = sweep * nblk + step;
      const T* Di = Ds[gs & 1];
      // (a1) owners of row tiles 2ib, 2ib+1 stage the raw block
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int rt = wid + 8 * q;
        if ((rt >> 1) == ib) {
#pragma unroll
          for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v)
              xa[((rt & 1) * 16 + Elem<T>::crow(lane, v)) * (W + 1) + ct * 16 + lr] = acc[q][ct][v];
        }
      }
      // this step's Di (copied into LDS during the previous step) has landed
      if (wid < DSW) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      __syncthreads();
      // (a2) xs = D * xa, 2*NCT tiles spread over the waves
      for (int tt = wid; tt < 2 * NCT; tt += 8) {
        const int rtl = tt & 1, ct = tt >> 1;
        V4 a4 = zero4<T>();
#pragma unroll
        for (int kk = 0; kk < kTri / 4; ++kk) {
          const int k = kk * 4 + lk;
          const int row = rtl * 16 + lr;
          const T af = TRANS ? Di[k + row * kTri] : Di[row + k * kTri];
          a4 = mfma(af, xa[k * (W + 1) + ct * 16 + lr], a4);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) xs[(rtl * 16 + Elem<T>::crow(lane, v)) * (W + 1) + ct * 16 + lr] = a4[v];
      }
      __syncthreads();
      if (gs + 1 < 2 * nblk) prefetch_di(gs + 1);
      // (b) owners take the solved block back; other tiles on the far side
      // of the block are updated with the off-diagonal factor block.  The
      // tiles to update form a run of q; the factor fragments of the next
      // tile are fetched under the MFMAs of the current one.  Fragment
      // (kk, lane) of tile rt is F[rt*16 + lr][i0 + kk*4 + lk] with F = L or
      // U (F^T from the transposed storage when TRANS): a per-tile scalar
      // base, a lane offset and a per-kk scalar offset, all 32-bit.
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int rt = wid + 8 * q;
        if (rt < nrt && (rt >> 1) == ib) {
#pragma unroll
          for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v)
              acc[q][ct][v] = xs[((rt & 1) * 16 + Elem<T>::crow(lane, v)) * (W + 1) + ct * 16 + lr];
        }
      }
      auto upd = [&](int q) {
        const int rt = wid + 8 * q;
        return rt < nrt && (fwd ? (rt >> 1) > ib : (rt >> 1) < ib);
      };
      const uint32_t vo = opaque(lane_off);
      auto fetch = [&](int q, T (&f)[kTri / 4]) {
        const int rt = wid + 8 * q;
        const T* base = TRANS ? LU + i0 + (int64_t)(rt * 16) * lda : LU + (int64_t)i0 * lda + rt * 16;
        const __amdgpu_buffer_rsrc_t rs = buf_rsrc(base);
        if (rt * 16 + 16 <= n && i0 + kTri <= n) {
#pragma unroll
          for (int kk = 0; kk < kTri / 4; ++kk) f[kk] = buf_load<T>(rs, vo, kk * kstep);
        } else {
#pragma unroll
          for (int kk = 0; kk < kTri / 4; ++kk)
            f[kk] = (rt * 16 + lr < n && i0 + kk * 4 + lk < n) ? buf_load<T>(rs, vo, kk * kstep) : T(0);
        }
      };
      T af[kTri / 4], afn[kTri / 4];
      bool have = false;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (upd(q)) {
          if (!have) {
            fetch(q, af);
          } else {
#pragma unroll
            for (int kk = 0; kk < kTri / 4; ++kk) af[kk] = afn[kk];
          }
          have = true;
          if (q + 1 < 4 && upd(q + 1)) fetch(q + 1, afn);
#pragma unroll
          for (int kk = 0; kk < kTri / 4; ++kk) {
            const int k = kk * 4 + lk;
#pragma unroll
            for (int ct = 0; ct < NCT; ++ct)
              acc[q][ct] = mfma(-af[kk], xs[k * (W + 1) + ct * 16 + lr], acc[q][ct]);
          }
        }
      }
    }
  }
  // scatter
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) {
      const int col = c0 + ct * 16 + lr;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = (wid + 8 * q) * 16 + Elem<T>::crow(lane, v);
        if (row < n && col < ncols) seg_out(X, col)[xoff + (TRANS ? perm[row] : row)] = acc[q][ct][v];
      }
    }
}

// the register solve addresses factor fragments with 32-bit byte offsets
template <typename T> static bool getrs_reg_ld_ok(int64_t lda) {
  return lda > 0 && lda * (int64_t)sizeof(T) * 32 < (int64_t)1 << 32;
}

template <int NCT, bool TRANS, typename T>
static void launch_getrs_reg(int n, const T* LU, int64_t lda, int64_t bstride, int batch,
                             const int* perm, const T* dinv, const ColSeg& X,
                             int64_t xbstride, int ncols, hipStream_t st,
                             const int* only = nullptr) {
  const int nchunk = (ncols + NCT * 16 - 1) / (NCT * 16);
  hipLaunchKernelGGL((k_getrs_reg<NCT, TRANS, T>), dim3(batch * nchunk), dim3(512), 0, st, LU, lda,
                     bstride, n, perm, dinv, X, xbstride, ncols, nchunk, only);
}

int g_getrs_tile = 0;
void set_getrs_tile(int nct) { g_getrs_tile = nct; }
int getrs_tile() { return g_getrs_tile; }

template <bool TRANS>
static hipError_t launch_getrs_impl(int n, const double* LU, int64_t lda, int64_t bstride,
                                    int batch, const int* perm, const double* dinv,
                                    const ColSeg& X, int64_t xbstride, int ncols,
                                    hipStream_t st, const int* only = nullptr) {
  if (batch <= 0 || n <= 0 || ncols <= 0) return hipSuccess;
  if (n > kMaxBlock) return hipErrorInvalidValue;
  // the node mask exists on the register path only (blocks of up to 512)
  if (only && !(n <= 512 && getrs_reg_ld_ok<double>(lda))) return hipErrorInvalidValue;
  if (n <= 512 && getrs_reg_ld_ok<double>(lda)) {
    // column tile: 16, 32 or 80 right-hand sides per workgroup, by width
    // (16 for blocks of up to 256 rows -- the 2r x 2r cores and the smaller
    // leaves: their steps are latency, and more, narrower workgroups
    // shorten them, while a block's factor re-read per workgroup is cheap
    // at that size; measured 11 % faster per system at rank 97-98) unless
    // a tuner fixed it
    int nct = (ncols <= 16 || n <= 256) ? 1 : (ncols <= 32 ? 2 : 5);
    if (g_getrs_tile == 1 || g_getrs_tile == 2 || g_getrs_tile == 5) nct = g_getrs_tile;
    if (nct == 1)
      launch_getrs_reg<1, TRANS, double>(n, LU, lda, bstride, batch, perm, dinv, X, xbstride, ncols, st, only);
    else if (nct == 2)
      launch_getrs_reg<2, TRANS, double>(n, LU, lda, bstride, batch, perm, dinv, X, xbstride, ncols, st, only);
    else
      launch_getrs_reg<5, TRANS, double>(n, LU, lda, bstride, batch, perm, dinv, X, xbstride, ncols, st, only);
    return hipGetLastError();
  }
  const int npad = ntri(n) * kTri;
  const size_t lim = 160 * 1024;
  // 32 columns per workgroup when they fit, else 16
  if ((size_t)32 * npad * sizeof(double) <= lim && ncols > 16) {
    const int nchunk = (ncols + 31) / 32;
    static bool attr = false;
    if (!attr) {
      (void)hipFuncSetAttribute((const void*)k_getrs<32, TRANS>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lim);
      attr = true;
    }
    hipLaunchKernelGGL((k_getrs<32, TRANS>), dim3(batch * nchunk), dim3(256),
                       (size_t)32 * npad * sizeof(double), st, LU, lda, bstride, n, perm, dinv,
                       X, xbstride, ncols, nchunk);
  } else {
    const int nchunk = (ncols + 15) / 16;
    if ((size_t)16 * npad * sizeof(double) > lim) return hipErrorInvalidValue;
    static bool attr = false;
    if (!attr) {
      (void)hipFuncSetAttribute((const void*)k_getrs<16, TRANS>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lim);
      attr = true;
    }
    hipLaunchKernelGGL((k_getrs<16, TRANS>), dim3(batch * nchunk), dim3(256),
                       (size_t)16 * npad * sizeof(double), st, LU, lda, bstride, n, perm, dinv,
                       X, xbstride, ncols, nchunk);
  }
  return hipGetLastError();
}

hipError_t launch_getrs(int n, const double* LU, int64_t lda, int64_t bstride, int batch,
                        const int* perm, const double* dinv, const ColSeg& X, int64_t xbstride,
                        int ncols, hipStream_t st) {
  return launch_getrs_impl<false>(n, LU, lda, bstride, batch, perm, dinv, X, xbstride, ncols, st);
}
hipError_t launch_getrs_t(int n, const double* LU, int64_t lda, int64_t bstride, int batch,
                          const int* perm, const double* dinv, const ColSeg& X, int64_t xbstride,
                          int ncols, hipStream_t st) {
  return launch_getrs_impl<true>(n, LU, lda, bstride, batch, perm, dinv, X, xbstride, ncols, st);
}
hipError_t launch_getrs_masked(int n, const double* LU, int64_t lda, int64_t bstride, int batch,
                               const int* perm, const double* dinv, const ColSeg& X,
                               int64_t xbstride, int ncols, const int* only, hipStream_t st) {
  return launch_getrs_impl<false>(n, LU, lda, bstride, batch, perm, dinv, X, xbstride, ncols, st,
                                  only);
}

// =====================================================================
// Core solve with the structured LU of S = [I, A; B, I] (r = 32 or 64):
// with pivots [1..r, r + ipiv2] the factors are
//   L = [I, 0; P2 B, L2],  U = [I, A; 0, U2],
// so S x = R is y2 = L2^-1 (P2 R2 - (P2 B) R1), x2 = U2^-1 y2,
// x1 = R1 - A x2 (the first r rows of y are R1).  L2 and U2 are applied by
// 32-row blocks with the stored inverses of S's diagonal blocks r/32, ...,
// the couplings and the two GEMMs on MFMA.  One workgroup per node and 16
// right-hand-side columns, four waves; every factor fragment a wave uses
// (56 values per lane at r = 64) is loaded at the start, so the stages
// wait only on LDS and the MFMA chains (the 16-row tiles: the GEMMs one
// per wave; block 0 of L2 / U2 waves 0-1, block 1 waves 2-3, each coupling
// the other pair).  R (2r x nc per node, ld 2r, node stride rs) is
// overwritten by the solution.  Cores flagged in `general` (factored by the
// unstructured LU) are skipped: the masked generic solve takes them.
// =====================================================================
template <int R>
__global__ void __launch_bounds__(256)
k_core_solve_s(const double* __restrict__ S0, const int* __restrict__ perm0,
               const double* __restrict__ dinv0, const int* __restrict__ general,
               double* __restrict__ R0, int64_t rs, int nc, double* __restrict__ PS,
               double* __restrict__ PR) {
  constexpr int N2 = 2 * R, MT = R / 16, NB = R / 32, Q = R / 4;
  static_assert(R == 32 || R == 64, "structured core solve: r = 32 or 64");
  __shared__ double Xt[R][17];  // R1, then x1
  __shared__ double Zb[R][17];  // P2 R2, then Z (consumed block by block)
  __shared__ double Yb[R][17];  // y2
  __shared__ double X2[R][17];  // x2
  const int b = blockIdx.y;
  const bool gen = general[b] != 0;  // solved by the generic kernel before this one
  if (gen && !PR) return;
  const int c0 = blockIdx.x * 16;
  const int sc = min(16, nc - c0);
  const double* S = S0 + (int64_t)b * N2 * N2;
  const int* perm = perm0 + (int64_t)b * N2;
  const double* dv = dinv0 + (int64_t)b * dinv_elems(N2);
  double* Rn = R0 + (int64_t)b * rs;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 15, lk = lane >> 4;
  const bool gw = wid < MT;  // a GEMM tile (rows 16 wid ..)
  // block k of L2 / U2 is applied by waves 2k, 2k+1 (tile 2k + (wid & 1));
  // the coupling of block k to the other block by the other pair
  const int kb = wid >> 1, half = wid & 1;
  const bool dw = kb < NB;
  if (!gen) {
    double fz[Q], fx[Q], fli[8], fui[8], fc[8];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int k = 4 * q + lk;
      fz[q] = gw ? S[(R + 16 * wid + lr) + (int64_t)k * N2] : 0.0;      // P2 B (row, k)
      fx[q] = gw ? S[(16 * wid + lr) + (int64_t)(R + k) * N2] : 0.0;    // A (row, k)
    }
    const double* Di = dv + (int64_t)(NB + kb) * 2 * kTri * kTri;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int k = 4 * q + lk;
      fli[q] = dw ? Di[(16 * half + lr) + k * 32] : 0.0;
      fui[q] = dw ? Di[kTri * kTri + (16 * half + lr) + k * 32] : 0.0;
      // NB == 2: waves 2-3 take L2(block 1, block 0) (rows of tile 2 + half,
      // columns of block 0), waves 0-1 U2(block 0, block 1)
      fc[q] = (NB == 2) ? (kb == 1 ? S[(R + 32 + 16 * half + lr) + (int64_t)(R + k) * N2]
                                   : S[(R + 16 * half + lr) + (int64_t)(R + 32 + k) * N2])
                        : 0.0;
    }
    for (int e = tid; e < N2 * 16; e += 256) {
      const int i = e % N2, j = e / N2;
      const double v = j < sc ? Rn[(i < R ? i : perm[i]) + (int64_t)(c0 + j) * N2] : 0.0;
      if (i < R)
        Xt[i][j] = v;
      else
        Zb[i - R][j] = v;
    }
    __syncthreads();
    auto tile_mm = [&](const double (&af)[8], double (*Bm)[17], int brow0) {
      d4 acc = d4zero();
#pragma unroll
      for (int q = 0; q < 8; ++q) acc = mfma(af[q], Bm[brow0 + 4 * q + lk][lr], acc);
      return acc;
    };
    // Z = P2 R2 - (P2 B) R1
    if (gw) {
      d4 acc = d4zero();
#pragma unroll
      for (int q = 0; q < Q; ++q) acc = mfma(fz[q], Xt[4 * q + lk][lr], acc);
#pragma unroll
      for (int v = 0; v < 4; ++v) Zb[16 * wid + crow(lane, v)][lr] -= acc[v];
    }
    lds_barrier();
    // forward: y_0 = L2_00^-1 Z_0; Z_1 -= L2_10 y_0; y_1 = L2_11^-1 Z_1
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      if (kb == k) {
        const d4 acc = tile_mm(fli, Zb, 32 * k);
#pragma unroll
        for (int v = 0; v < 4; ++v) Yb[32 * k + 16 * half + crow(lane, v)][lr] = acc[v];
      }
      lds_barrier();
      if (k + 1 < NB) {
        if (kb == k + 1) {
          const d4 acc = tile_mm(fc, Yb, 32 * k);
#pragma unroll
          for (int v = 0; v < 4; ++v) Zb[32 * (k + 1) + 16 * half + crow(lane, v)][lr] -= acc[v];
        }
        lds_barrier();
      }
    }
    // backward: x2_1 = U2_11^-1 y_1; y_0 -= U2_01 x2_1; x2_0 = U2_00^-1 y_0
#pragma unroll
    for (int k = NB - 1; k >= 0; --k) {
      if (kb == k) {
        const d4 acc = tile_mm(fui, Yb, 32 * k);
#pragma unroll
        for (int v = 0; v < 4; ++v) X2[32 * k + 16 * half + crow(lane, v)][lr] = acc[v];
      }
      lds_barrier();
      if (k > 0) {
        if (kb == k - 1) {
          const d4 acc = tile_mm(fc, X2, 32 * k);
#pragma unroll
          for (int v = 0; v < 4; ++v) Yb[32 * (k - 1) + 16 * half + crow(lane, v)][lr] -= acc[v];
        }
        lds_barrier();
      }
    }
    // x1 = R1 - A x2
    if (gw) {
      d4 acc = d4zero();
#pragma unroll
      for (int q = 0; q < Q; ++q) acc = mfma(fx[q], X2[4 * q + lk][lr], acc);
#pragma unroll
      for (int v = 0; v < 4; ++v) Xt[16 * wid + crow(lane, v)][lr] -= acc[v];
    }
  lds_barrier();
  for (int e = tid; e < N2 * 16; e += 256) {
    const int i = e % N2, j = e / N2;
    if (j < sc) Rn[i + (int64_t)(c0 + j) * N2] = i < R ? Xt[i][j] : X2[i - R][j];
  }
  } else {
    // a general core: its solution E_b, written by the generic solve
    for (int e = tid; e < N2 * 16; e += 256) {
      const int i = e % N2, j = e / N2;
      const double v = j < sc ? Rn[i + (int64_t)(c0 + j) * N2] : 0.0;
      if (i < R)
        Xt[i][j] = v;
      else
        X2[i - R][j] = v;
    }
    __syncthreads();
  }
  if (PR) {
    // The parent's core from this node's solution (node-sum form): the
    // sum this child hands up is T = E_b(0:r) + E_b(r:2r).  Parent p = b/2
    // takes R_p = [T_c1; T_c0] and, in its first r columns (those of Y),
    // S_p = [I, T_c1; T_c0, I] -- what k_assemble_pair forms, written here
    // so the level above needs no assembly launch.
    const int pn = b >> 1, odd = b & 1;
    double* Rp = PR + (int64_t)pn * rs;
    double* Sp = PS ? PS + (int64_t)pn * N2 * N2 : nullptr;
    for (int e = tid; e < R * 16; e += 256) {
      const int a = e % R, j = e / R;
      if (j < sc) {
        const int col = c0 + j;
        const double t = Xt[a][j] + X2[a][j];
        Rp[(odd ? a : R + a) + (int64_t)col * N2] = t;
        if (Sp && col < R) {
          if (odd) {
            Sp[a + (int64_t)(R + col) * N2] = t;
            Sp[(R + a) + (int64_t)(R + col) * N2] = (a == col) ? 1.0 : 0.0;
          } else {
            Sp[(R + a) + (int64_t)col * N2] = t;
            Sp[a + (int64_t)col * N2] = (a == col) ? 1.0 : 0.0;
          }
        }
      }
    }
  }
}

bool core_solve_structured(int r) { return r == 32 || r == 64; }

hipError_t launch_core_solve_s(int r, const double* S, int batch, const int* perm,
                               const double* dinv, const int* general, double* R, int64_t rs,
                               int nc, double* PS, double* PR, hipStream_t st) {
  if (batch <= 0 || nc <= 0) return hipSuccess;
  if (!core_solve_structured(r) || !general) return hipErrorInvalidValue;
  // the parent's assembly pairs the nodes: a whole number of parents
  if (PR && (batch & 1)) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((nc + 15) / 16), (unsigned)batch);
  if (r == 32)
    hipLaunchKernelGGL(k_core_solve_s<32>, grid, dim3(256), 0, st, S, perm, dinv, general, R, rs, nc,
                       PS, PR);
  else
    hipLaunchKernelGGL(k_core_solve_s<64>, grid, dim3(256), 0, st, S, perm, dinv, general, R, rs, nc,
                       PS, PR);
  return hipGetLastError();
}

// FP32 leaf kernels for the mixed-precision path (blocks up to 512).  The
// right-hand sides stay FP64 in memory; the solve converts on load/store.
hipError_t launch_getrf_f32(int n, float* A, int64_t lda, int64_t bstride, int batch, int* ipiv,
                            int* perm, int* info, hipStream_t st) {
  if (batch <= 0 || n <= 0) return hipSuccess;
  if (n > 512) return hipErrorInvalidValue;
  if (n <= 256)
    hipLaunchKernelGGL((k_getrf_reg<GETRF_NB, 256, 2, float>), dim3(batch), dim3(256), 0, st, A,
                       lda, bstride, n, ipiv, perm, info, (const int*)nullptr, (const double*)nullptr,
                       (int64_t)0, (int64_t)0);
  else
    hipLaunchKernelGGL((k_getrf_reg<GETRF_NB, 512, 2, float>), dim3(batch), dim3(512), 0, st, A,
                       lda, bstride, n, ipiv, perm, info, (const int*)nullptr, (const double*)nullptr,
                       (int64_t)0, (int64_t)0);
  return hipGetLastError();
}

hipError_t launch_trtri_diag_f32(int n, const float* LU, int64_t lda, int64_t bstride, int batch,
                                 float* dinv, hipStream_t st) {
  if (batch <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_trtri_diag<float>, dim3(batch * ntri(n)), dim3(64), 0, st, LU, lda,
                     bstride, n, dinv);
  return hipGetLastError();
}


// Block copy with conversion: D_b (ld ldd, stride dbs) <- S_b (FP64, ld lds,
// stride sbs), n x n per block; grid (batch, n / 64 column groups).
template <typename T>
__global__ void __launch_bounds__(256)
k_copy_blocks(int n, const double* __restrict__ S0, int64_t lds, int64_t sbs, T* __restrict__ D0,
              int64_t ldd, int64_t dbs) {
  const double* S = S0 + (int64_t)blockIdx.x * sbs;
  T* D = D0 + (int64_t)blockIdx.x * dbs;
  const int c1 = min(n, (int)(blockIdx.y + 1) * 64);
  for (int c = blockIdx.y * 64; c < c1; ++c)
    for (int r = threadIdx.x; r < n; r += 256) D[r + (int64_t)c * ldd] = (T)S[r + (int64_t)c * lds];
}

// Leaf factorization out of place: LU_b (ld ldl, stride lbs, element type
// T) <- LU of K_b (FP64, ld ldk, stride kbs) for every block b, with the
// stored 32x32 diagonal inverses in dinv.  Blocks of n = 128, 256, 384 or
// 512 take k_getrf_ll; the blocks it flags (retry[b] = 1, a multiplier
// beyond the pivot margin) and blocks of any other n are copied from K and
// factored by the pivoted kernels, then their diagonal blocks inverted.
template <typename T>
static hipError_t fallback_getrf(int n, const double* K, int64_t ldk, int64_t kbs, T* LU,
                                 int64_t ldl, int64_t lbs, int batch, int* ipiv, int* perm,
                                 int* info, const int* only, hipStream_t st);

template <>
hipError_t fallback_getrf<double>(int n, const double* K, int64_t ldk, int64_t kbs, double* LU,
                                  int64_t ldl, int64_t lbs, int batch, int* ipiv, int* perm,
                                  int* info, const int* only, hipStream_t st) {
  if (n <= 128) {
    hipLaunchKernelGGL(k_getrf_w64<32>, dim3(batch), dim3(64), 0, st, LU, ldl, lbs, n, ipiv, perm,
                       info, only, K, ldk, kbs);
  } else if (n <= 256) {
    hipLaunchKernelGGL((k_getrf_reg<GETRF_NB, 256, 2, double>), dim3(batch), dim3(256), 0, st, LU,
                       ldl, lbs, n, ipiv, perm, info, only, K, ldk, kbs);
  } else if (n <= 512) {
    hipLaunchKernelGGL((k_getrf_reg<GETRF_NB, 512, 2, double>), dim3(batch), dim3(512), 0, st, LU,
                       ldl, lbs, n, ipiv, perm, info, only, K, ldk, kbs);
  } else {
    // larger blocks: every block (never masked), copied then factored in place
    hipLaunchKernelGGL(k_copy_blocks<double>, dim3(batch, (n + 63) / 64), dim3(256), 0, st, n, K,
                       ldk, kbs, LU, ldl, lbs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_getrf(n, LU, ldl, lbs, batch, ipiv, perm, info, st);
  }
  return hipGetLastError();
}

template <>
hipError_t fallback_getrf<float>(int n, const double* K, int64_t ldk, int64_t kbs, float* LU,
                                 int64_t ldl, int64_t lbs, int batch, int* ipiv, int* perm,
                                 int* info, const int* only, hipStream_t st) {
  if (n > 512) return hipErrorInvalidValue;
  if (n <= 256)
    hipLaunchKernelGGL((k_getrf_reg<GETRF_NB, 256, 2, float>), dim3(batch), dim3(256), 0, st, LU,
                       ldl, lbs, n, ipiv, perm, info, only, K, ldk, kbs);
  else
    hipLaunchKernelGGL((k_getrf_reg<GETRF_NB, 512, 2, float>), dim3(batch), dim3(512), 0, st, LU,
                       ldl, lbs, n, ipiv, perm, info, only, K, ldk, kbs);
  return hipGetLastError();
}

template <typename T>
static hipError_t getrf_leaf_impl(int n, const double* K, int64_t ldk, int64_t kbs, T* LU,
                                  int64_t ldl, int64_t lbs, int batch, T* dinv, int* ipiv,
                                  int* perm, int* info, int* retry, hipStream_t st,
                                  bool fallback_inverses) {
  if (batch <= 0 || n <= 0) return hipSuccess;
  // the fast kernel addresses a slab of K and a block of LU with 32-bit
  // byte offsets
  const bool fast = (n % 128 == 0) && n <= 512 && ldk * 8 * 64 < ((int64_t)1 << 32) &&
                    ldl * (int64_t)sizeof(T) * n < ((int64_t)1 << 32);
  if (fast) {
    const dim3 g(batch), blk(512);
    switch (n / 128) {
      case 1: hipLaunchKernelGGL((k_getrf_ll<1, T>), g, blk, 0, st, K, ldk, kbs, LU, ldl, lbs, dinv, ipiv, perm, info, retry); break;
      case 2: hipLaunchKernelGGL((k_getrf_ll<2, T>), g, blk, 0, st, K, ldk, kbs, LU, ldl, lbs, dinv, ipiv, perm, info, retry); break;
      case 3: hipLaunchKernelGGL((k_getrf_ll<3, T>), g, blk, 0, st, K, ldk, kbs, LU, ldl, lbs, dinv, ipiv, perm, info, retry); break;
      default: hipLaunchKernelGGL((k_getrf_ll<4, T>), g, blk, 0, st, K, ldk, kbs, LU, ldl, lbs, dinv, ipiv, perm, info, retry); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const int* only = fast ? retry : nullptr;
  hipError_t e = fallback_getrf<T>(n, K, ldk, kbs, LU, ldl, lbs, batch, ipiv, perm, info, only, st);
  if (e != hipSuccess || !fallback_inverses) return e;
  hipLaunchKernelGGL(k_trtri_diag<T>, dim3(batch * ntri(n)), dim3(64), 0, st, LU, ldl, lbs, n,
                     dinv, only);
  return hipGetLastError();
}

hipError_t launch_getrf_leaf(int n, const double* K, int64_t ldk, int64_t kbs, double* LU,
                             int64_t ldl, int64_t lbs, int batch, double* dinv, int* ipiv,
                             int* perm, int* info, int* retry, hipStream_t st) {
  return getrf_leaf_impl<double>(n, K, ldk, kbs, LU, ldl, lbs, batch, dinv, ipiv, perm, info,
                                 retry, st, true);
}
hipError_t launch_getrf_leaf_f32(int n, const double* K, int64_t ldk, int64_t kbs, float* LU,
                                 int64_t ldl, int64_t lbs, int batch, float* dinv, int* ipiv,
                                 int* perm, int* info, int* retry, hipStream_t st) {
  return getrf_leaf_impl<float>(n, K, ldk, kbs, LU, ldl, lbs, batch, dinv, ipiv, perm, info,
                                retry, st, true);
}

hipError_t launch_getrs_f32(int n, const float* LU, int64_t lda, int64_t bstride, int batch,
                            const int* perm, const float* dinv, const ColSeg& X, int64_t xbstride,
                            int ncols, bool trans, hipStream_t st) {
  if (batch <= 0 || n <= 0 || ncols <= 0) return hipSuccess;
  if (n > 512 || !getrs_reg_ld_ok<float>(lda)) return hipErrorInvalidValue;
  if (trans) {
    if (ncols <= 16) launch_getrs_reg<1, true, float>(n, LU, lda, bstride, batch, perm, dinv, X, xbstride, ncols, st);
    else if (ncols <= 32) launch_getrs_reg<2, true, float>(n, LU, lda, bstride, batch, perm, dinv, X, xbstride, ncols, st);
    else launch_getrs_reg<5, true, float>(n, LU, lda, bstride, batch, perm, dinv, X, xbstride, ncols, st);
  } else {
    if (ncols <= 16) launch_getrs_reg<1, false, float>(n, LU, lda, bstride, batch, perm, dinv, X, xbstride, ncols, st);
    else if (ncols <= 32) launch_getrs_reg<2, false, float>(n, LU, lda, bstride, batch, perm, dinv, X, xbstride, ncols, st);
    else launch_getrs_reg<5, false, float>(n, LU, lda, bstride, batch, perm, dinv, X, xbstride, ncols, st);
  }
  return hipGetLastError();
}

__global__ void k_convert_f32(int64_t rows, int cols, const double* __restrict__ src, int64_t lds,
                              float* __restrict__ dst, int64_t ldd) {
  const int64_t total = rows * (int64_t)cols;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    const int64_t i = e % rows, j = e / rows;
    dst[i + j * ldd] = (float)src[i + j * lds];
  }
}

hipError_t launch_convert_f32(int64_t rows, int cols, const double* src, int64_t lds, float* dst,
                              int64_t ldd, hipStream_t st) {
  const int64_t total = rows * (int64_t)cols;
  if (total <= 0) return hipSuccess;
  int64_t g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_convert_f32, dim3((unsigned)g), dim3(256), 0, st, rows, cols, src, lds, dst,
                     ldd);
  return hipGetLastError();
}

// Column max-norms: out[j] = max_i |A(i, j)| (one workgroup per column).
// max |A(:, j)| per column, a NaN propagating (max is exact, so the order of
// the comparisons does not matter).  Stage one: kColmaxParts blocks per
// column each reduce a strided share of the rows into part[j * G + g];
// stage two: one wave per column reduces the G partials.
__device__ __forceinline__ double max_nan(double a, double m) { return (a > m || a != a) ? a : m; }

__global__ void __launch_bounds__(256) k_colmax_part(int64_t rows, const double* __restrict__ A,
                                                     int64_t lda, double* __restrict__ part) {
  __shared__ double wm[4];
  const int j = blockIdx.y, g = blockIdx.x, G = gridDim.x;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const double* Aj = A + (int64_t)j * lda;
  double m = 0.0;
  for (int64_t i = (int64_t)g * 256 + threadIdx.x; i < rows; i += (int64_t)G * 256)
    m = max_nan(fabs(Aj[i]), m);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = max_nan(__shfl_xor(m, off, 64), m);
  if (lane == 0) wm[wid] = m;
  __syncthreads();
  if (threadIdx.x == 0)
    part[(int64_t)j * G + g] = max_nan(max_nan(wm[0], wm[1]), max_nan(wm[2], wm[3]));
}

__global__ void __launch_bounds__(64) k_colmax_final(const double* __restrict__ part, int G,
                                                     double* __restrict__ out) {
  const int j = blockIdx.x, lane = threadIdx.x;
  double m = 0.0;
  for (int g = lane; g < G; g += 64) m = max_nan(part[(int64_t)j * G + g], m);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = max_nan(__shfl_xor(m, off, 64), m);
  if (lane == 0) out[j] = m;
}

hipError_t launch_colmax(int64_t rows, int cols, const double* A, int64_t lda, double* out,
                         double* part, hipStream_t st) {
  if (cols <= 0) return hipSuccess;
  const int64_t chunks = (rows + 2047) / 2048;  // at least 2048 rows per block
  const int G = (int)std::max<int64_t>(1, std::min<int64_t>(kColmaxParts, chunks));
  hipLaunchKernelGGL(k_colmax_part, dim3(G, cols), dim3(256), 0, st, rows, A, lda, part);
  hipLaunchKernelGGL(k_colmax_final, dim3(cols), dim3(64), 0, st, part, G, out);
  return hipGetLastError();
}

// Y += X (rows x cols).
__global__ void k_axpy_cols(int64_t rows, int cols, const double* __restrict__ X, int64_t ldx,
                            double* __restrict__ Y, int64_t ldy) {
  const int64_t total = rows * (int64_t)cols;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    const int64_t i = e % rows, j = e / rows;
    Y[i + j * ldy] += X[i + j * ldx];
  }
}

hipError_t launch_add_cols(int64_t rows, int cols, const double* X, int64_t ldx, double* Y,
                           int64_t ldy, hipStream_t st) {
  const int64_t total = rows * (int64_t)cols;
  if (total <= 0) return hipSuccess;
  int64_t g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_axpy_cols, dim3((unsigned)g), dim3(256), 0, st, rows, cols, X, ldx, Y, ldy);
  return hipGetLastError();
}

// =====================================================================
// Level reduce ("gemmRed", gemm_reduce.cc:58-76): split-K V^T X over row
// chunks that never straddle a child, one partial per chunk, summed later
// in fixed chunk order (deterministic, unlike the reference's REDOP_ADD
// fold).  Workgroup tile: 64 rows of V^T (one 16-row MFMA tile per wave)
// x NCT*16 columns of X; K staged through LDS 32 rows at a time with
// coalesced column loads.
// =====================================================================
template <int NCT>
__global__ void __launch_bounds__(256)
k_reduce(const double* __restrict__ V, int64_t ldv, int r, ColSeg X, int ncols,
         int64_t n_child, int cpc, int64_t rc, double* __restrict__ part) {
  constexpr int KS = 32;
  __shared__ double Vs[64 * (KS + 1)];
  __shared__ double Xs[NCT * 16 * (KS + 1)];
  const int chunk = blockIdx.x;
  const int child = chunk / cpc, q = chunk % cpc;
  const int64_t rbeg = (int64_t)child * n_child + (int64_t)q * rc;
  const int64_t cend = (int64_t)(child + 1) * n_child;
  const int64_t rend = rbeg + rc < cend ? rbeg + rc : cend;
  const int j0 = blockIdx.y * NCT * 16;
  const int a0 = blockIdx.z * 64;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  d4 acc[NCT];
#pragma unroll
  for (int ct = 0; ct < NCT; ++ct) acc[ct] = d4zero();

  // row chunks of KS; the next chunk's V and X values are fetched into
  // registers under the current chunk's MFMAs
  constexpr int VPT = 64 * KS / 256, XPT = NCT * 16 * KS / 256;
  static_assert(VPT * 256 == 64 * KS && XPT * 256 == NCT * 16 * KS, "staging split");
  double vr[VPT], xr[XPT];
  auto fetch = [&](int64_t k0) {
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
      const int idx = tid + 256 * i, a = idx / KS, k = idx % KS;
      const int64_t row = k0 + k;
      vr[i] = (a0 + a < r && row < rend) ? V[row + (int64_t)(a0 + a) * ldv] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < XPT; ++i) {
      const int idx = tid + 256 * i, j = idx / KS, k = idx % KS;
      const int64_t row = k0 + k;
      xr[i] = (j0 + j < ncols && row < rend) ? seg_in(X, j0 + j)[row] : 0.0;
    }
  };
  if (rbeg < rend) fetch(rbeg);
  for (int64_t k0 = rbeg; k0 < rend; k0 += KS) {
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
      const int idx = tid + 256 * i;
      Vs[(idx / KS) * (KS + 1) + idx % KS] = vr[i];
    }
#pragma unroll
    for (int i = 0; i < XPT; ++i) {
      const int idx = tid + 256 * i;
      Xs[(idx / KS) * (KS + 1) + idx % KS] = xr[i];
    }
    __syncthreads();
    if (k0 + KS < rend) fetch(k0 + KS);
#pragma unroll
    for (int kk = 0; kk < KS / 4; ++kk) {
      const double af = Vs[(wid * 16 + lr) * (KS + 1) + kk * 4 + lk];
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) {
        if (j0 + ct * 16 < ncols) {
          const double bf = Xs[(ct * 16 + lr) * (KS + 1) + kk * 4 + lk];
          acc[ct] = mfma(af, bf, acc[ct]);
        }
      }
    }
    __syncthreads();
  }
  double* P = part + (int64_t)chunk * r * ncols;
#pragma unroll
  for (int ct = 0; ct < NCT; ++ct) {
    const int col = j0 + ct * 16 + lr;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int a = a0 + wid * 16 + crow(lane, v);
      if (a < r && col < ncols) P[a + (int64_t)col * r] = acc[ct][v];
    }
  }
}

int64_t reduce_chunks_per_child(int64_t n_child, int64_t rc) { return (n_child + rc - 1) / rc; }

hipError_t launch_reduce(const double* V, int64_t ldv, int r, const ColSeg& X, int ncols,
                         int64_t n_child, int nchild, int64_t rc, double* part, int* cpc_out,
                         hipStream_t st) {
  const int64_t cpc = reduce_chunks_per_child(n_child, rc);
  if (cpc_out) *cpc_out = (int)cpc;
  if (nchild <= 0 || ncols <= 0 || r <= 0) return hipSuccess;
  dim3 grid((unsigned)(nchild * cpc), (unsigned)((ncols + 127) / 128), (unsigned)((r + 63) / 64));
  hipLaunchKernelGGL(k_reduce<8>, grid, dim3(256), 0, st, V, ldv, r, X, ncols, n_child, (int)cpc,
                     rc, part);
  return hipGetLastError();
}

// =====================================================================
// Core assembly (node_solve.cc:60-81): sum each child's partials in chunk
// order, then S = [I, G_1; G_0, I] and R = [H_1; H_0] (halves swapped).
// =====================================================================
// Sum of the projections of leaves [q0, q1) under `child` at offset `off`.
__device__ __forceinline__ double sum_part(const double* part, int child, int cpc, int64_t pstride,
                                           int64_t off, int q0, int q1) {
  double s = 0.0;
  const double* p = part + (int64_t)child * cpc * pstride + off;
  for (int q = q0; q < q1; ++q) s += p[(int64_t)q * pstride];
  return s;
}

// Core assembly.  A block covers 256 / KS elements of one node; with KS > 1
// the cpc leaves under a child are split into KS contiguous ranges summed
// by separate thread groups and combined in a fixed order (deterministic),
// so the upper levels, where a child spans many leaves and there are few
// nodes, still spread over the chip.
template <int KS>
__global__ void __launch_bounds__(256)
k_assemble(const double* __restrict__ part, int cpc, int r, int ncols_part, int gcol, int hcol,
           int nh, double* __restrict__ S, double* __restrict__ R, int swap) {
  constexpr int EPB = 256 / KS;
  __shared__ double ps[KS > 1 ? KS : 1][EPB];
  const int node = blockIdx.y;
  const int n2 = 2 * r;
  const int g = threadIdx.x / EPB, le = threadIdx.x % EPB;
  const int q0 = (int)((int64_t)cpc * g / KS), q1 = (int)((int64_t)cpc * (g + 1) / KS);
  const int64_t pstride = (int64_t)r * ncols_part;
  const int64_t nS = S ? (int64_t)n2 * n2 : 0;
  const int64_t total = nS + (int64_t)n2 * nh;
  for (int64_t e0 = (int64_t)blockIdx.x * EPB; e0 < total; e0 += (int64_t)gridDim.x * EPB) {
    const int64_t e = e0 + le;
    double v = 0.0;
    bool sum = false;  // v is a partial sum of this group's leaves
    double* dst = nullptr;
    if (e < nS) {
      const int i = (int)(e % n2), j = (int)(e / n2);
      dst = S + (int64_t)node * nS + i + (int64_t)j * n2;
      if ((i < r) == (j < r)) {
        v = (i == j) ? 1.0 : 0.0;
      } else if (i >= r) {  // S(r+a, j) = V0^T u0 (a, j)
        v = sum_part(part, 2 * node, cpc, pstride, (i - r) + (int64_t)(gcol + j) * r, q0, q1);
        sum = true;
      } else {              // S(a, r+j) = V1^T u1 (a, j)
        v = sum_part(part, 2 * node + 1, cpc, pstride, i + (int64_t)(gcol + j - r) * r, q0, q1);
        sum = true;
      }
    } else if (e < total) {
      const int64_t f = e - nS;
      const int i = (int)(f % n2), j = (int)(f / n2);
      const int hi = i >= r;  // lower half of R
      const int child = 2 * node + (swap ? 1 - hi : hi);
      dst = R + (int64_t)node * n2 * nh + i + (int64_t)j * n2;
      v = sum_part(part, child, cpc, pstride, (i - hi * r) + (int64_t)(hcol + j) * r, q0, q1);
      sum = true;
    }
    if (KS > 1) {
      ps[g][le] = v;
      __syncthreads();
      if (g == 0 && dst) {
        if (sum)
#pragma unroll
          for (int h = 1; h < KS; ++h) v += ps[h][le];
        *dst = v;
      }
      __syncthreads();
    } else if (dst) {
      *dst = v;
    }
  }
}

hipError_t launch_assemble(const double* part, int cpc, int r, int ncols_part, int gcol, int hcol,
                           int nh, int nnode, double* S, double* R, hipStream_t st,
                           bool swap) {
  if (nnode <= 0) return hipSuccess;
  const int64_t total = (S ? 4LL * r * r : 0) + 2LL * r * nh;
  if (cpc >= 16) {
    int gx = (int)((total + 63) / 64);
    if (gx > 1024) gx = 1024;
    if (gx < 1) gx = 1;
    hipLaunchKernelGGL(k_assemble<4>, dim3(gx, nnode), dim3(256), 0, st, part, cpc, r, ncols_part,
                       gcol, hcol, nh, S, R, swap ? 1 : 0);
  } else {
    int gx = (int)((total + 255) / 256);
    if (gx > 256) gx = 256;
    if (gx < 1) gx = 1;
    hipLaunchKernelGGL(k_assemble<1>, dim3(gx, nnode), dim3(256), 0, st, part, cpc, r, ncols_part,
                       gcol, hcol, nh, S, R, swap ? 1 : 0);
  }
  return hipGetLastError();
}

// =====================================================================
// The level loop in node-sum form.  Node n of level t solves
//   S_n E_n = R_n,  S_n = [I, T_{c1}^Y; T_{c0}^Y, I],  R_n = [T_{c1}; T_{c0}]
// (c0 = 2n, c1 = 2n + 1 its children, T_c the sums over c's leaves of the
// current projections V_i^T [Y_i | B_i]).  The per-leaf update of a level,
// P_i <- P_i - P_i^Y E_{n,c(i)}, is linear in P_i, so a child's updated sum
// is T_c - T_c^Y E_{n,c}, and the rows of S_n E_n = R_n say exactly
//   T_{c0} - T_{c0}^Y E_n(0:r) = E_n(r:2r),  T_{c1} - T_{c1}^Y E_n(r:2r) = E_n(0:r):
// the sum a node hands to its parent is the sum of the two halves of its
// own core solution.  Every level above the leaves is therefore assembled
// from the level below's solutions alone, and the per-leaf products
// (G_i = sum_t W_t,i E^B_t, W_{t-1} = W_t (I - E^Y_t)) are evaluated once,
// root first, along each leaf's path.  E of node (t, n) is stored at heap
// index 2^t - 1 + n.
// =====================================================================

// Core assembly from the level below: T_c = E_c(0:r, :) + E_c(r:2r, :)
// (E_c at Ec + c*ecs, ld 2r); node n's S = [I, T_{2n+1}(:, gcol..); T_{2n}
// (:, gcol..), I] (skipped when S is null) and R = [T_{2n+1}; T_{2n}]
// (columns hcol .. hcol + nh; at R + n*rs, ld 2r).
__global__ void __launch_bounds__(256)
k_assemble_pair(const double* __restrict__ Ec, int64_t ecs, int r, int gcol, int hcol, int nh,
                double* __restrict__ S, double* __restrict__ R, int64_t rs) {
  const int node = blockIdx.y;
  const int n2 = 2 * r;
  const int64_t nS = S ? (int64_t)n2 * n2 : 0;
  const int64_t total = nS + (int64_t)n2 * nh;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    int c, a, col;
    double* dst;
    if (e < nS) {
      const int i = (int)(e % n2), j = (int)(e / n2);
      dst = S + (int64_t)node * nS + i + (int64_t)j * n2;
      if ((i < r) == (j < r)) {
        *dst = (i == j) ? 1.0 : 0.0;
        continue;
      }
      if (i >= r) { c = 2 * node; a = i - r; col = gcol + j; }           // S(r+a, j) = T_c0^Y
      else        { c = 2 * node + 1; a = i; col = gcol + j - r; }       // S(a, r+j) = T_c1^Y
    } else {
      const int64_t f = e - nS;
      const int i = (int)(f % n2), j = (int)(f / n2);
      const int hi = i >= r;
      dst = R + (int64_t)node * rs + i + (int64_t)j * n2;
      c = 2 * node + 1 - hi; a = i - hi * r; col = hcol + j;
    }
    const double* ep = Ec + (int64_t)c * ecs + a + (int64_t)col * n2;
    *dst = ep[0] + ep[r];
  }
}

hipError_t launch_assemble_pair(const double* Ec, int64_t ecs, int r, int gcol, int hcol, int nh,
                                int nnode, double* S, double* R, int64_t rs, hipStream_t st) {
  if (nnode <= 0) return hipSuccess;
  const int64_t total = (S ? 4LL * r * r : 0) + 2LL * r * nh;
  int gx = (int)((total + 255) / 256);
  if (gx > 256) gx = 256;
  if (gx < 1) gx = 1;
  hipLaunchKernelGGL(k_assemble_pair, dim3(gx, nnode), dim3(256), 0, st, Ec, ecs, r, gcol, hcol,
                     nh, S, R, rs);
  return hipGetLastError();
}

// Top-down pass, one workgroup per leaf slot i (leaves i0, i0 + 1, ...):
//   H = 0;  for t = 0 .. L-1:  H <- E^B + H - E^Y H;   G_i = H
// with E^Y (r x r) and E^B (r x s) the rows of node (t, i >> (L-t))'s core
// solution that belong to its child holding leaf i (rows [0, r) for the
// even child, [r, 2r) for the odd one): G_i = sum_t W_t,i E^B_t evaluated
// by Horner from the root.  Node g's blocks at EY + g*eys and EB + g*ebs
// (ld 2r); G_i at G + i*r*s (ld r).  H lives in LDS, double-buffered, in
// chunks of cw columns (cw * (r + ...) sized by the launcher); the E^Y
// fragments of a row tile (up to 128 k) are loaded together before its
// MFMA chain.
__global__ void __launch_bounds__(256)
k_path_down(const double* __restrict__ EY, int64_t eys, const double* __restrict__ EB,
            int64_t ebs, int L, int r, int s, int cw, int64_t i0, double* __restrict__ G) {
  extern __shared__ double hsm[];
  const int64_t i = i0 + blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int n2 = 2 * r, mt = (r + 15) >> 4;
  const int hld = cw + 1;  // LDS row stride of H (row-major, r x cw)
  double* Hb[2] = {hsm, hsm + (int64_t)r * hld};
  for (int c0 = 0; c0 < s; c0 += cw) {
    const int sc = min(cw, s - c0);
    for (int e = tid; e < r * hld; e += 256) Hb[0][e] = 0.0;
    __syncthreads();
    int cur = 0;
    for (int t = 0; t < L; ++t) {
      const int64_t ch = i >> (L - 1 - t);
      const int64_t g = ((int64_t)1 << t) - 1 + (ch >> 1);
      const int64_t hoff = (ch & 1) * (int64_t)r;
      const double* ey = EY + g * eys + hoff;
      const double* eb = EB + g * ebs + hoff + (int64_t)c0 * n2;
      const double* hc = Hb[cur];
      double* hn = Hb[cur ^ 1];
      for (int tm = wid; tm < mt; tm += 4) {
        const int arow = tm * 16 + lr;
        double bv[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int orow = tm * 16 + crow(lane, v);
          bv[v] = (orow < r && lr < sc) ? eb[orow + (int64_t)lr * n2] : 0.0;
        }
        d4 acc = d4zero();
        for (int kb = 0; kb < r; kb += 128) {
          double af[32];
#pragma unroll
          for (int q = 0; q < 32; ++q) {
            const int k = kb + q * 4 + lk;
            af[q] = (arow < r && k < r) ? ey[arow + (int64_t)k * n2] : 0.0;
          }
#pragma unroll
          for (int q = 0; q < 32; ++q) {
            if (kb + q * 4 < r) {  // wave-uniform
              const int k = kb + q * 4 + lk;
              const double hv = (k < r && lr < cw) ? hc[k * hld + lr] : 0.0;
              acc = mfma(af[q], hv, acc);
            }
          }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int orow = tm * 16 + crow(lane, v);
          if (orow < r && lr < cw) hn[orow * hld + lr] = bv[v] + hc[orow * hld + lr] - acc[v];
        }
      }
      __syncthreads();
      cur ^= 1;
    }
    for (int e = tid; e < r * sc; e += 256) {
      const int a = e % r, j = e / r;
      G[i * r * s + a + (int64_t)(c0 + j) * r] = Hb[cur][a * hld + j];
    }
    __syncthreads();
  }
}

// k_path_down for r <= 64 (one 16-row tile per wave, the inner dimension
// in 16 k-steps): the next level's E^Y fragments and E^B values are loaded
// under the current level's MFMA chain, and the level barrier orders LDS
// only, so those loads stay in flight across it.  Same arithmetic and order
// as k_path_down.
__global__ void __launch_bounds__(256)
k_path_down64(const double* __restrict__ EY, int64_t eys, const double* __restrict__ EB,
              int64_t ebs, int L, int r, int s, int cw, int64_t i0, double* __restrict__ G) {
  extern __shared__ double hsm[];
  const int64_t i = i0 + blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int n2 = 2 * r, mt = (r + 15) >> 4;
  const int hld = cw + 1;
  double* Hb[2] = {hsm, hsm + (int64_t)r * hld};
  const bool act = wid < mt;
  const int arow = wid * 16 + lr;
  for (int c0 = 0; c0 < s; c0 += cw) {
    const int sc = min(cw, s - c0);
    auto fetch = [&](int t, double (&af)[16], double (&bv)[4]) {
      const int64_t ch = i >> (L - 1 - t);
      const int64_t g = ((int64_t)1 << t) - 1 + (ch >> 1);
      const int64_t hoff = (ch & 1) * (int64_t)r;
      const double* ey = EY + g * eys + hoff;
      const double* eb = EB + g * ebs + hoff + (int64_t)c0 * n2;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int k = q * 4 + lk;
        af[q] = (act && arow < r && k < r) ? ey[arow + (int64_t)k * n2] : 0.0;
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int orow = wid * 16 + crow(lane, v);
        bv[v] = (act && orow < r && lr < sc) ? eb[orow + (int64_t)lr * n2] : 0.0;
      }
    };
    for (int e = tid; e < r * hld; e += 256) Hb[0][e] = 0.0;
    double af[16], bv[4], an[16], bn[4];
    fetch(0, af, bv);
    __syncthreads();
    int cur = 0;
    for (int t = 0; t < L; ++t) {
      if (t + 1 < L) fetch(t + 1, an, bn);
      const double* hc = Hb[cur];
      double* hn = Hb[cur ^ 1];
      if (act) {
        d4 acc = d4zero();
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          if (q * 4 < r) {  // wave-uniform
            const int k = q * 4 + lk;
            const double hv = (k < r && lr < cw) ? hc[k * hld + lr] : 0.0;
            acc = mfma(af[q], hv, acc);
          }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int orow = wid * 16 + crow(lane, v);
          if (orow < r && lr < cw) hn[orow * hld + lr] = bv[v] + hc[orow * hld + lr] - acc[v];
        }
      }
      lds_barrier();
      cur ^= 1;
#pragma unroll
      for (int q = 0; q < 16; ++q) af[q] = an[q];
#pragma unroll
      for (int v = 0; v < 4; ++v) bv[v] = bn[v];
    }
    for (int e = tid; e < r * sc; e += 256) {
      const int a = e % r, j = e / r;
      G[i * r * s + a + (int64_t)(c0 + j) * r] = Hb[cur][a * hld + j];
    }
    __syncthreads();
  }
}

// LDS of k_path_down: two r x (cw + 1) blocks, cw = 16, 8 or 4 so that they
// stay within 64 KiB.
static int path_cw(int r) {
  for (int cw = 16; cw >= 4; cw >>= 1)
    if ((size_t)2 * r * (cw + 1) * sizeof(double) <= 65536) return cw;
  return 0;
}

hipError_t launch_path_down(const double* EY, int64_t eys, const double* EB, int64_t ebs, int L,
                            int r, int s, int64_t i0, int64_t count, double* G, hipStream_t st) {
  if (count <= 0 || s <= 0 || L <= 0) return hipSuccess;
  const int cw = path_cw(r);
  if (cw == 0) return hipErrorInvalidValue;
  const size_t lds = (size_t)2 * r * (cw + 1) * sizeof(double);
  if (r <= 64)
    hipLaunchKernelGGL(k_path_down64, dim3((unsigned)count), dim3(256), lds, st, EY, eys, EB, ebs,
                       L, r, s, cw, i0, G);
  else
    hipLaunchKernelGGL(k_path_down, dim3((unsigned)count), dim3(256), lds, st, EY, eys, EB, ebs, L,
                       r, s, cw, i0, G);
  return hipGetLastError();
}

// Adjoint of the top-down pass, one level (one workgroup per node n):
//   Hn_n = Hc_{2n} + Hc_{2n+1} - (E_n^Y)^T [Hc_{2n}; Hc_{2n+1}]
// E_n^Y: the first r columns of node n's core solution (2r x r at
// EY + n*eys, ld 2r); Hc_c at Hc + c*r*s and Hn_n at Hn + n*r*s (r x s,
// ld r).
__global__ void __launch_bounds__(256)
k_adj_up(const double* __restrict__ EY, int64_t eys, const double* __restrict__ Hc, int r, int s,
         double* __restrict__ Hn) {
  const int64_t n = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int n2 = 2 * r;
  const int64_t hs = (int64_t)r * s;
  const double* ey = EY + n * eys;
  const double* h0 = Hc + 2 * n * hs;
  const double* h1 = h0 + hs;
  double* out = Hn + n * hs;
  const int mt = (r + 15) >> 4, nt = (s + 15) >> 4;
  for (int task = wid; task < mt * nt; task += 4) {
    const int tm = task % mt, tn = task / mt;
    const int arow = tm * 16 + lr, bcol = tn * 16 + lr;
    d4 acc = d4zero();
    for (int k0 = 0; k0 < n2; k0 += 4) {
      const int k = k0 + lk;
      const double av = (arow < r && k < n2) ? ey[k + (int64_t)arow * n2] : 0.0;
      double bv = 0.0;
      if (bcol < s && k < n2) bv = k < r ? h0[k + (int64_t)bcol * r] : h1[k - r + (int64_t)bcol * r];
      acc = mfma(av, bv, acc);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int orow = tm * 16 + crow(lane, v), col = tn * 16 + lr;
      if (orow < r && col < s) {
        const int64_t o = orow + (int64_t)col * r;
        out[o] = h0[o] + h1[o] - acc[v];
      }
    }
  }
}

hipError_t launch_adj_up(const double* EY, int64_t eys, const double* Hc, int r, int s, int nnode,
                         double* Hn, hipStream_t st) {
  if (nnode <= 0 || s <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_up, dim3((unsigned)nnode), dim3(256), 0, st, EY, eys, Hc, r, s, Hn);
  return hipGetLastError();
}

// Adjoint of the bottom-up assembly, one level: node n's right-hand side
//   [Hc_{2n} + Tbar_n; Hc_{2n+1} + Tbar_n]   (2r x s at R + n*rs, ld 2r)
// where Tbar_n is the half of the parent's solution that the parent's R
// took from node n: rows [r, 2r) of Rp_{n/2} for an even n, [0, r) for an
// odd one (Rp null at the root: Tbar = 0).  Hc_c at Hc + c*r*s (ld r).
__global__ void __launch_bounds__(256)
k_adj_assemble(const double* __restrict__ Hc, const double* __restrict__ Rp, int64_t rps, int r,
               int s, double* __restrict__ R, int64_t rs) {
  const int64_t node = blockIdx.y;
  const int n2 = 2 * r;
  const int64_t total = (int64_t)n2 * s;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    const int i = (int)(e % n2), j = (int)(e / n2);
    const int hi = i >= r, a = i - hi * r;
    double v = Hc[(2 * node + hi) * (int64_t)r * s + a + (int64_t)j * r];
    if (Rp) v += Rp[(node >> 1) * rps + (1 - (node & 1)) * (int64_t)r + a + (int64_t)j * n2];
    R[node * rs + i + (int64_t)j * n2] = v;
  }
}

hipError_t launch_adj_assemble(const double* Hc, const double* Rp, int64_t rps, int r, int s,
                               int nnode, double* R, int64_t rs, hipStream_t st) {
  if (nnode <= 0 || s <= 0) return hipSuccess;
  int gx = (int)((2LL * r * s + 255) / 256);
  if (gx > 256) gx = 256;
  hipLaunchKernelGGL(k_adj_assemble, dim3(gx, nnode), dim3(256), 0, st, Hc, Rp, rps, r, s, R, rs);
  return hipGetLastError();
}

// =====================================================================
// Level update ("gemmBro", gemm_broadcast.cc:58-82):
//   X_c[:, j] += alpha * u_c * eta_c[:, j]
// 64-row groups that never straddle a child (one 16-row MFMA tile per
// wave) x NCT*16 columns; u streamed as the A operand, eta as B.
// =====================================================================
__global__ void __launch_bounds__(256)
k_update(const double* __restrict__ u, int64_t ldu, int r, const double* __restrict__ eta,
         int64_t eta_ld, int64_t eta_ns, int csz, int64_t n_child, int gpc, double alpha,
         ColSeg X, int ncols, int half_xor) {
  // block: 4 waves over 256 rows of one child and one 16-column tile; a
  // wave owns 64 rows (four 16-row tiles), so each eta fragment feeds four
  // MFMAs, and the u fragments of 4 k-steps (16 loads) are in flight together
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int child = blockIdx.x / gpc, g = blockIdx.x % gpc;
  const int64_t cend = (int64_t)(child + 1) * n_child;
  const int64_t rbase = (int64_t)child * n_child + (int64_t)g * 256 + wid * 64;
  const int col = blockIdx.y * 16 + lr;
  const bool cok = col < ncols;
  const int slot = child / csz;
  const double* E = eta + (int64_t)(slot >> 1) * eta_ns + (int64_t)((slot & 1) ^ half_xor) * r;
  d4 acc[4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t row = rbase + a * 16 + crow(lane, v);
      acc[a][v] = (row < cend && cok) ? seg_in(X, col)[row] : 0.0;
    }
  bool rok[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) rok[a] = rbase + a * 16 + lr < cend;
  for (int k0 = 0; k0 < r; k0 += 16) {
    double af[4][4], bf[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const int k = k0 + kk * 4 + lk;
      const bool kok = k < r;
#pragma unroll
      for (int a = 0; a < 4; ++a)
        af[kk][a] = (kok && rok[a]) ? alpha * u[rbase + a * 16 + lr + (int64_t)k * ldu] : 0.0;
      bf[kk] = (kok && cok) ? E[k + (int64_t)col * eta_ld] : 0.0;
    }
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
#pragma unroll
      for (int a = 0; a < 4; ++a) acc[a] = mfma(af[kk][a], bf[kk], acc[a]);
  }
  if (!cok) return;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t row = rbase + a * 16 + crow(lane, v);
      if (row < cend) seg_out(X, col)[row] = acc[a][v];
    }
}

hipError_t launch_update(const double* u, int64_t ldu, int r, const double* eta, int64_t eta_ld,
                         int64_t eta_node_stride, int csz, int64_t n_child, int nchild,
                         double alpha, const ColSeg& X, int ncols, hipStream_t st,
                         int half_xor) {
  if (nchild <= 0 || ncols <= 0) return hipSuccess;
  const int64_t gpc = (n_child + 255) / 256;
  dim3 grid((unsigned)(nchild * gpc), (unsigned)((ncols + 15) / 16));
  hipLaunchKernelGGL(k_update, grid, dim3(256), 0, st, u, ldu, r, eta, eta_ld,
                     eta_node_stride, csz, n_child, (int)gpc, alpha, X, ncols, half_xor & 1);
  return hipGetLastError();
}

// LMatrix::node_solve operands (node_solve.cc:60-81): per node slot k,
// S = [I, VTu(2kr+r : 2kr+2r, :); VTu(2kr : 2kr+r, :), I] and R = VTd node
// rows with the halves swapped.
__global__ void k_node_assemble(int r, const double* __restrict__ VTu, int64_t ldu,
                                const double* __restrict__ VTd, int64_t ldd, int nrhs,
                                double* __restrict__ S, double* __restrict__ R) {
  const int node = blockIdx.y;
  const int n2 = 2 * r;
  const int64_t nS = (int64_t)n2 * n2;
  const int64_t total = nS + (int64_t)n2 * nrhs;
  const int64_t base = (int64_t)node * n2;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    if (e < nS) {
      const int i = (int)(e % n2), j = (int)(e / n2);
      double v;
      if ((i < r) == (j < r)) v = (i == j) ? 1.0 : 0.0;
      else if (i >= r) v = VTu[base + (i - r) + (int64_t)j * ldu];
      else v = VTu[base + r + i + (int64_t)(j - r) * ldu];
      S[(int64_t)node * nS + i + (int64_t)j * n2] = v;
    } else {
      const int64_t f = e - nS;
      const int i = (int)(f % n2), j = (int)(f / n2);
      const int src = (i < r) ? r + i : i - r;
      R[(int64_t)node * n2 * nrhs + i + (int64_t)j * n2] = VTd[base + src + (int64_t)j * ldd];
    }
  }
}

hipError_t launch_node_assemble(int nnode, int r, const double* VTu, int64_t ldu, const double* VTd,
                                int64_t ldd, int nrhs, double* S, double* R, hipStream_t st) {
  if (nnode <= 0) return hipSuccess;
  const int64_t total = 4LL * r * r + 2LL * r * nrhs;
  int gx = (int)((total + 255) / 256);
  if (gx > 64) gx = 64;
  hipLaunchKernelGGL(k_node_assemble, dim3(gx, nnode), dim3(256), 0, st, r, VTu, ldu, VTd, ldd,
                     nrhs, S, R);
  return hipGetLastError();
}

// Cores of sibling ranks r0 != r1 (NodeSolveRegionTask,
// node_solve_region.cc:77-96), n = r0 + r1, node k's operands at element
// offsets k*s*:  S = I_n with S(i, r0+j) = VTu1(i, j), S(r0+j, i) = VTu0(j, i)
// (i < r0, j < r1), and R = [VTd1; VTd0] (n x nrhs, ld n).
__global__ void k_region_assemble(int r0, int r1, int nrhs, const double* __restrict__ VTu0,
                                  int64_t ldu0, int64_t su0, const double* __restrict__ VTu1,
                                  int64_t ldu1, int64_t su1, const double* __restrict__ VTd0,
                                  int64_t ldd0, int64_t sd0, const double* __restrict__ VTd1,
                                  int64_t ldd1, int64_t sd1, double* __restrict__ S,
                                  double* __restrict__ R) {
  const int64_t node = blockIdx.y;
  const int n = r0 + r1;
  const int64_t nS = (int64_t)n * n;
  const int64_t total = nS + (int64_t)n * nrhs;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    if (e < nS) {
      const int i = (int)(e % n), j = (int)(e / n);
      double v;
      if ((i < r0) == (j < r0)) v = (i == j) ? 1.0 : 0.0;
      else if (i < r0) v = VTu1[node * su1 + i + (int64_t)(j - r0) * ldu1];
      else v = VTu0[node * su0 + (i - r0) + (int64_t)j * ldu0];
      S[node * nS + e] = v;
    } else {
      const int64_t f = e - nS;
      const int i = (int)(f % n), j = (int)(f / n);
      R[node * n * nrhs + f] = (i < r0) ? VTd1[node * sd1 + i + (int64_t)j * ldd1]
                                        : VTd0[node * sd0 + (i - r0) + (int64_t)j * ldd0];
    }
  }
}

// Solution back into the slots (node_solve_region.cc:99-107): VTd1 <- rows
// [0, r0) of R, VTd0 <- rows [r0, r0+r1).
__global__ void k_region_scatter(int r0, int r1, int nrhs, const double* __restrict__ R,
                                 double* __restrict__ VTd0, int64_t ldd0, int64_t sd0,
                                 double* __restrict__ VTd1, int64_t ldd1, int64_t sd1) {
  const int64_t node = blockIdx.y;
  const int n = r0 + r1;
  const int64_t total = (int64_t)n * nrhs;
  for (int64_t f = (int64_t)blockIdx.x * 256 + threadIdx.x; f < total;
       f += (int64_t)gridDim.x * 256) {
    const int i = (int)(f % n), j = (int)(f / n);
    const double v = R[node * total + f];
    if (i < r0) VTd1[node * sd1 + i + (int64_t)j * ldd1] = v;
    else VTd0[node * sd0 + (i - r0) + (int64_t)j * ldd0] = v;
  }
}

hipError_t launch_region_assemble(int nnode, int r0, int r1, int nrhs, const double* VTu0,
                                  int64_t ldu0, int64_t su0, const double* VTu1, int64_t ldu1,
                                  int64_t su1, const double* VTd0, int64_t ldd0, int64_t sd0,
                                  const double* VTd1, int64_t ldd1, int64_t sd1, double* S,
                                  double* R, hipStream_t st) {
  if (nnode <= 0) return hipSuccess;
  const int64_t n = (int64_t)r0 + r1;
  int64_t gx = (n * n + n * nrhs + 255) / 256;
  if (gx > 64) gx = 64;
  hipLaunchKernelGGL(k_region_assemble, dim3((unsigned)gx, nnode), dim3(256), 0, st, r0, r1, nrhs,
                     VTu0, ldu0, su0, VTu1, ldu1, su1, VTd0, ldd0, sd0, VTd1, ldd1, sd1, S, R);
  return hipGetLastError();
}

hipError_t launch_region_scatter(int nnode, int r0, int r1, int nrhs, const double* R,
                                 double* VTd0, int64_t ldd0, int64_t sd0, double* VTd1,
                                 int64_t ldd1, int64_t sd1, hipStream_t st) {
  const int64_t total = ((int64_t)r0 + r1) * nrhs;
  if (nnode <= 0 || total <= 0) return hipSuccess;
  int64_t gx = (total + 255) / 256;
  if (gx > 64) gx = 64;
  hipLaunchKernelGGL(k_region_scatter, dim3((unsigned)gx, nnode), dim3(256), 0, st, r0, r1, nrhs,
                     R, VTd0, ldd0, sd0, VTd1, ldd1, sd1);
  return hipGetLastError();
}

// Slot sums for LMatrix::gemmRed: C[slot*ka + a, j] = alpha * sum of the
// nper consecutive partials (ka x kb, ld ka) that belong to the slot.
__global__ void k_sum_slots(const double* __restrict__ part, int nper, int nslot, int ka, int kb,
                            double alpha, double* __restrict__ C, int64_t ldc) {
  const int64_t total = (int64_t)nslot * ka * kb;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    const int a = (int)(e % ka);
    const int64_t t = e / ka;
    const int j = (int)(t % kb), slot = (int)(t / kb);
    const double* p = part + (int64_t)slot * nper * ka * kb + a + (int64_t)j * ka;
    double s = 0.0;
    for (int q = 0; q < nper; ++q) s += p[(int64_t)q * ka * kb];
    C[(int64_t)slot * ka + a + (int64_t)j * ldc] = alpha * s;
  }
}

hipError_t launch_sum_slots(const double* part, int nper, int nslot, int ka, int kb, double alpha,
                            double* C, int64_t ldc, hipStream_t st) {
  const int64_t total = (int64_t)nslot * ka * kb;
  if (total <= 0) return hipSuccess;
  int g = (int)((total + 255) / 256);
  if (g > 1024) g = 1024;
  hipLaunchKernelGGL(k_sum_slots, dim3(g), dim3(256), 0, st, part, nper, nslot, ka, kb, alpha, C,
                     ldc);
  return hipGetLastError();
}

// =====================================================================
// HODLR matvec: part_i = V_i^T x_i (k_reduce with one chunk per leaf),
// tot = sum_i part_i (fixed order), then per leaf
//   y_i = alpha * (K_i x_i + U_i (tot - part_i)) + beta * y_i.
// =====================================================================
// 16 elements x 16 leaf chunks per block: thread (c, e) sums leaves c,
// c + 16, ... of element e, then the 16 chunk sums are added in chunk order
// (a fixed order: the result does not depend on the launch)
__global__ void __launch_bounds__(256) k_mv_total(const double* __restrict__ part, int nleaf,
                                                  int64_t elems, double* __restrict__ tot) {
  __shared__ double red[16][17];
  const int el = threadIdx.x & 15, c = threadIdx.x >> 4;
  const int64_t e = (int64_t)blockIdx.x * 16 + el;
  double s = 0.0;
  if (e < elems)
    for (int i = c; i < nleaf; i += 16) s += part[(int64_t)i * elems + e];
  red[c][el] = s;
  __syncthreads();
  if (c == 0 && e < elems) {
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) t += red[q][el];
    tot[e] = t;
  }
}

// Block: 4 waves over 256 rows of one leaf and one 16-column tile of the
// right-hand sides; a wave owns 64 rows (four 16-row tiles), so each X
// fragment feeds four MFMAs, and the K fragments of 4 k-steps (16 loads) are
// in flight together.
template <bool TRANS>
__global__ void __launch_bounds__(256)
k_mv_apply(int leaf, int r, int nrhs, const double* __restrict__ K, int64_t ldk,
           const double* __restrict__ U, int64_t ldu, const double* __restrict__ X, int64_t ldx,
           const double* __restrict__ part, const double* __restrict__ tot, double alpha,
           double beta, double* __restrict__ Y, int64_t ldy, int gpl) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int li = blockIdx.x / gpl, g = blockIdx.x % gpl;
  const int64_t l0 = (int64_t)li * leaf;
  const int rloc = g * 256 + wid * 64;  // first row of this wave within the leaf
  const int col = blockIdx.y * 16 + lr;
  const bool cok = col < nrhs;
  const double* P = part + (int64_t)li * r * nrhs;
  d4 acc[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) acc[a] = d4zero();
  bool rok[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) rok[a] = rloc + a * 16 + lr < leaf;
  for (int k0 = 0; k0 < leaf; k0 += 16) {
    double af[4][4], bf[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const int k = k0 + kk * 4 + lk;
      const bool kok = k < leaf;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int64_t arow = rloc + a * 16 + lr;
        af[kk][a] = (kok && rok[a]) ? (TRANS ? K[l0 + k + arow * ldk] : K[l0 + arow + (int64_t)k * ldk])
                                    : 0.0;
      }
      bf[kk] = (kok && cok) ? X[l0 + k + (int64_t)col * ldx] : 0.0;
    }
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
#pragma unroll
      for (int a = 0; a < 4; ++a) acc[a] = mfma(af[kk][a], bf[kk], acc[a]);
  }
  for (int k0 = 0; k0 < r; k0 += 4) {
    const int k = k0 + lk;
    const bool kok = k < r;
    const double bf = (kok && cok) ? tot[k + (int64_t)col * r] - P[k + (int64_t)col * r] : 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int64_t arow = rloc + a * 16 + lr;
      const double af = (kok && rok[a]) ? U[l0 + arow + (int64_t)k * ldu] : 0.0;
      acc[a] = mfma(af, bf, acc[a]);
    }
  }
  if (!cok) return;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = rloc + a * 16 + crow(lane, v);
      if (row < leaf) {
        double* y = Y + l0 + row + (int64_t)col * ldy;
        const double prev = (beta == 0.0) ? 0.0 : beta * *y;
        *y = alpha * acc[a][v] + prev;
      }
    }
}

int64_t matvec_work_elems(int64_t n, int leaf, int rank, int nrhs) {
  const int64_t nleaf = n / leaf;
  return (nleaf + 1) * (int64_t)rank * nrhs;
}

hipError_t launch_leaf_projections(int64_t n, int leaf, int rank, int nrhs, const double* V,
                                   int64_t ldv, const double* X, int64_t ldx, double* part,
                                   double* tot, hipStream_t st) {
  const int64_t nleaf = n / leaf;
  ColSeg xs = seg_single(const_cast<double*>(X), ldx);
  hipError_t e = launch_reduce(V, ldv, rank, xs, nrhs, leaf, (int)nleaf, leaf, part, nullptr, st);
  if (e != hipSuccess) return e;
  const int64_t elems = (int64_t)rank * nrhs;
  hipLaunchKernelGGL(k_mv_total, dim3((unsigned)((elems + 15) / 16)), dim3(256), 0, st, part,
                     (int)nleaf, elems, tot);
  return hipGetLastError();
}

// =====================================================================
// Per-leaf outer products for the gradients: C_i = alpha * A_i * B_i^T
// with a short inner dimension k (the number of right-hand sides).
// Workgroup tile: 64 rows (one 16-row MFMA tile per wave) x NCT*16
// columns; the k loop runs in steps of 4 with zero padding.
// =====================================================================
template <int NCT>
__global__ void __launch_bounds__(256)
k_leaf_outer(int leaf, int nc, int k, double alpha, const double* __restrict__ A, int64_t lda,
             const double* __restrict__ B, int64_t ldb, const double* __restrict__ part,
             const double* __restrict__ tot, double* __restrict__ C, int64_t ldc, int gpl) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int li = blockIdx.x / gpl, g = blockIdx.x % gpl;
  const int64_t l0 = (int64_t)li * leaf;
  const int rloc = g * 64 + wid * 16;
  const int j0 = blockIdx.y * NCT * 16;
  const double* P = tot ? part + (int64_t)li * nc * k : nullptr;
  d4 acc[NCT];
#pragma unroll
  for (int ct = 0; ct < NCT; ++ct) acc[ct] = d4zero();
  for (int k0 = 0; k0 < k; k0 += 4) {
    const int kk = k0 + lk;
    const int arow = rloc + lr;
    const double af = (kk < k && arow < leaf) ? alpha * A[l0 + arow + (int64_t)kk * lda] : 0.0;
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) {
      if (j0 + ct * 16 < nc) {
        const int col = j0 + ct * 16 + lr;
        double bf = 0.0;
        if (kk < k && col < nc)
          bf = tot ? tot[col + (int64_t)kk * nc] - P[col + (int64_t)kk * nc]
                   : B[l0 + col + (int64_t)kk * ldb];
        acc[ct] = mfma(af, bf, acc[ct]);
      }
    }
  }
#pragma unroll
  for (int ct = 0; ct < NCT; ++ct) {
    const int col = j0 + ct * 16 + lr;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = rloc + crow(lane, v);
      if (row < leaf && col < nc) C[l0 + row + (int64_t)col * ldc] = acc[ct][v];
    }
  }
}

hipError_t launch_leaf_outer(int64_t n, int leaf, int nc, int k, double alpha, const double* A,
                             int64_t lda, const double* B, int64_t ldb, const double* part,
                             const double* tot, double* C, int64_t ldc, hipStream_t st) {
  const int64_t nleaf = n / leaf;
  if (nleaf <= 0 || nc <= 0) return hipSuccess;
  const int gpl = (leaf + 63) / 64;
  dim3 grid((unsigned)(nleaf * gpl), (unsigned)((nc + 63) / 64));
  hipLaunchKernelGGL(k_leaf_outer<4>, grid, dim3(256), 0, st, leaf, nc, k, alpha, A, lda, B, ldb,
                     part, tot, C, ldc, gpl);
  return hipGetLastError();
}

// Fixed scratch of the allocation-free matvec (hodlr_matvec_noalloc): the
// per-leaf projections V_i^T x and their total, kNoAllocScratchElems
// doubles.  Every device the library runs on holds its own copy until the
// process ends.  In host code `scratch` is only a symbol: the device address
// of the current device's copy comes from hipGetSymbolAddress, looked up
// once per device and kept in g_scratch_ptr.
static __device__ __align__(256) double scratch[kNoAllocScratchElems];

namespace {
std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;
}  // namespace

hipError_t noalloc_scratch(double** out) {
  *out = nullptr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

hipError_t launch_matvec_runs(const LeafRun* runs, int nrun, int64_t nleaf, int rank, int nrhs,
                              const double* K, int64_t ldk, const double* U, int64_t ldu,
                              const double* V, int64_t ldv, const double* X, int64_t ldx,
                              double alpha, double beta, double* Y, int64_t ldy, double* work,
                              hipStream_t st, bool trans) {
  if (trans) {
    // A^T = blockdiag(K_i^T) + (V U^T off the leaf blocks): the same product
    // with U and V exchanged and the leaf blocks read transposed
    std::swap(U, V);
    std::swap(ldu, ldv);
  }
  const int64_t elems = (int64_t)rank * nrhs;
  double* part = work;
  double* tot = work + nleaf * elems;
  hipError_t e = hipSuccess;
  for (int k = 0; k < nrun; ++k) {
    const LeafRun& R = runs[k];
    if (R.m == 0) {  // empty leaf slots project to zero
      e = launch_fill_scale(R.count * elems, 1, part + R.first * elems, R.count * elems, 0, 0.0, st);
    } else {
      ColSeg xs = seg_single(const_cast<double*>(X) + R.row0, ldx);
      e = launch_reduce(V + R.row0, ldv, rank, xs, nrhs, R.m, (int)R.count, R.m,
                        part + R.first * elems, nullptr, st);
    }
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_mv_total, dim3((unsigned)((elems + 15) / 16)), dim3(256), 0, st, part,
                     (int)nleaf, elems, tot);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  for (int k = 0; k < nrun; ++k) {
    const LeafRun& R = runs[k];
    if (R.m == 0) continue;
    const int gpl = (R.m + 255) / 256;
    dim3 grid((unsigned)(R.count * gpl), (unsigned)((nrhs + 15) / 16));
    const double* Kr = K + R.row0;
    const double* Ur = U + R.row0;
    const double* Xr = X + R.row0;
    double* Yr = Y + R.row0;
    const double* pr = part + R.first * elems;
    if (trans)
      hipLaunchKernelGGL((k_mv_apply<true>), grid, dim3(256), 0, st, R.m, rank, nrhs, Kr, ldk, Ur,
                         ldu, Xr, ldx, pr, tot, alpha, beta, Yr, ldy, gpl);
    else
      hipLaunchKernelGGL((k_mv_apply<false>), grid, dim3(256), 0, st, R.m, rank, nrhs, Kr, ldk, Ur,
                         ldu, Xr, ldx, pr, tot, alpha, beta, Yr, ldy, gpl);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_matvec(int64_t n, int leaf, int rank, int nrhs, const double* K, int64_t ldk,
                         const double* U, int64_t ldu, const double* V, int64_t ldv,
                         const double* X, int64_t ldx, double alpha, double beta, double* Y,
                         int64_t ldy, double* work, hipStream_t st, bool trans) {
  LeafRun run;
  run.first = 0;
  run.count = n / leaf;
  run.m = leaf;
  return launch_matvec_runs(&run, 1, run.count, rank, nrhs, K, ldk, U, ldu, V, ldv, X, ldx, alpha,
                            beta, Y, ldy, work, st, trans);
}

// =====================================================================
// Workload generator: Philox4x64-10 (Salmon et al., SC'11), one counter
// block (4 words) per thread.
// =====================================================================
__device__ __forceinline__ void philox4x64(uint64_t c[4], uint64_t k0, uint64_t k1) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ULL, M1 = 0xCA5A826395121157ULL;
  const uint64_t W0 = 0x9E3779B97F4A7C15ULL, W1 = 0xBB67AE8584CAA73BULL;
#pragma unroll
  for (int rnd = 0; rnd < 10; ++rnd) {
    if (rnd > 0) { k0 += W0; k1 += W1; }
    const uint64_t hi0 = __umul64hi(M0, c[0]), lo0 = M0 * c[0];
    const uint64_t hi1 = __umul64hi(M1, c[2]), lo1 = M1 * c[2];
    const uint64_t n0 = hi1 ^ c[1] ^ k0, n1 = lo1, n2 = hi0 ^ c[3] ^ k1, n3 = lo0;
    c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
  }
}

__global__ void k_generate(uint64_t seed, uint32_t sid, uint32_t system, int kind, int64_t rows,
                           int cols, double scale, double* __restrict__ out, int64_t ld) {
  const int64_t total = rows * (int64_t)cols;
  const int64_t nblk = (total + 3) / 4;
  for (int64_t blk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; blk < nblk;
       blk += (int64_t)gridDim.x * blockDim.x) {
    uint64_t w[4] = {(uint64_t)blk, sid, system, 0};
    philox4x64(w, seed, 0);
    double z[4];
    if (kind == 0) {
#pragma unroll
      for (int q = 0; q < 4; ++q) z[q] = (double)(w[q] >> 11) * 0x1.0p-53;
    } else {
      const double two_pi = 6.283185307179586476925286766559;
#pragma unroll
      for (int pr = 0; pr < 2; ++pr) {
        const double u1 = (double)((w[2 * pr] >> 11) + 1) * 0x1.0p-53;
        const double u2 = (double)(w[2 * pr + 1] >> 11) * 0x1.0p-53;
        const double rad = sqrt(-2.0 * log(u1));
        z[2 * pr] = rad * cos(two_pi * u2);
        z[2 * pr + 1] = rad * sin(two_pi * u2);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t e = blk * 4 + q;
      if (e < total) out[(e % rows) + (e / rows) * ld] = z[q] * scale;
    }
  }
}

hipError_t launch_generate(uint64_t seed, uint32_t stream_id, uint32_t system, int kind,
                           int64_t rows, int cols, double scale, double* out, int64_t ld,
                           hipStream_t st) {
  const int64_t nblk = (rows * (int64_t)cols + 3) / 4;
  if (nblk <= 0) return hipSuccess;
  int64_t g = (nblk + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_generate, dim3((unsigned)g), dim3(256), 0, st, seed, stream_id, system,
                     kind, rows, cols, scale, out, ld);
  return hipGetLastError();
}

__global__ void k_shift_diag(int64_t n, int leaf, double shift, double* __restrict__ K,
                             int64_t ldk) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) K[i + (i % leaf) * ldk] += shift;
}

hipError_t launch_shift_diag(int64_t n, int leaf, double shift, double* K, int64_t ldk,
                             hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_shift_diag, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, leaf,
                     shift, K, ldk);
  return hipGetLastError();
}

// DenseBlockTask (dense_block.cc:54-75): K_i = U_i V_i^T + diag(D_i + offset)
// for every leaf i (K is n x leaf, leaf i = rows [i*leaf, (i+1)*leaf)).
// Setup work, one element per thread, the rank-r sum in index order.
__global__ void k_dense_blocks(int64_t n, int leaf, int rank, const double* __restrict__ U,
                               int64_t ldu, const double* __restrict__ V, int64_t ldv,
                               const double* __restrict__ D, double offset,
                               double* __restrict__ K, int64_t ldk) {
#pragma clang fp contract(off)
  const int64_t total = n * (int64_t)leaf;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    const int64_t row = e % n;
    const int col = (int)(e / n);
    const int64_t vrow = (row / leaf) * leaf + col;
    // unfused multiply and add in k order: the products and sums of the
    // reference's dgemm('N','T') (dgemm.f, alpha = 1, beta = 0), so K is
    // the reference generator's bit for bit
    // (the pragma at the top of the kernel keeps these operations unfused)
    double acc = 0.0;
    for (int k = 0; k < rank; ++k) {
      const double prod = U[row + (int64_t)k * ldu] * V[vrow + (int64_t)k * ldv];
      acc = acc + prod;
    }
    if ((int)(row % leaf) == col) acc = acc + (D[row] + offset);
    K[row + (int64_t)col * ldk] = acc;
  }
}

hipError_t launch_dense_blocks(int64_t n, int leaf, int rank, const double* U, int64_t ldu,
                               const double* V, int64_t ldv, const double* D, double offset,
                               double* K, int64_t ldk, hipStream_t st) {
  const int64_t total = n * (int64_t)leaf;
  if (total <= 0) return hipSuccess;
  int64_t g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_dense_blocks, dim3((unsigned)g), dim3(256), 0, st, n, leaf, rank, U, ldu,
                     V, ldv, D, offset, K, ldk);
  return hipGetLastError();
}

// The reference generators (Matrix::rand / Vector::rand, matrix.cc:250-297
// and :79-102, with PtrMatrix::rand, ptr_matrix.cc:62-71) on the device:
// partition k's seed is the k-th lrand48 of srand48(master) (the master
// standing for the reference's time(NULL)), and its rows_per_part x cols
// block is that seed's drand48 stream visited rows outer, columns inner,
// plus offset.  The POSIX LCG x <- (0x5DEECE66D x + 0xB) mod 2^48 is jumped
// ahead in O(log n) by composing affine maps, so every thread starts at its
// own element; drand48 = x * 2^-48 exactly, lrand48 = x >> 17.
__device__ __forceinline__ uint64_t lcg48_jump(uint64_t x, uint64_t n) {
  const uint64_t M = (1ULL << 48) - 1;
  uint64_t A = 1, C = 0, a = 0x5DEECE66DULL, c = 0xBULL;
  while (n) {
    if (n & 1) {
      A = (a * A) & M;
      C = (a * C + c) & M;
    }
    c = ((a + 1) * c) & M;
    a = (a * a) & M;
    n >>= 1;
  }
  return (A * x + C) & M;
}
__device__ __forceinline__ uint64_t lcg48_seed(uint64_t seed) {
  return ((seed & 0xffffffffULL) << 16) | 0x330EULL;
}

constexpr int kRefChunk = 16;  // consecutive stream elements per thread

__global__ void k_ref_fill(uint64_t master, int64_t rows, int cols, double offset,
                           double* __restrict__ out, int64_t ld) {
  const int64_t part = blockIdx.y;
  const uint64_t seed = lcg48_jump(lcg48_seed(master), (uint64_t)part + 1) >> 17;
  const int64_t total = rows * (int64_t)cols;
  const int64_t e0 = ((int64_t)blockIdx.x * 256 + threadIdx.x) * kRefChunk;
  if (e0 >= total) return;
  uint64_t x = lcg48_jump(lcg48_seed(seed), (uint64_t)e0);
  double* blk = out + part * rows;
  const int64_t e1 = e0 + kRefChunk < total ? e0 + kRefChunk : total;
  for (int64_t e = e0; e < e1; ++e) {
    x = (0x5DEECE66DULL * x + 0xBULL) & ((1ULL << 48) - 1);
    const double v = (double)x * (1.0 / 281474976710656.0);
    blk[e / cols + (e % cols) * ld] = v + offset;
  }
}

hipError_t launch_ref_fill(uint64_t master, int npart, int64_t rows, int cols, double offset,
                           double* out, int64_t ld, hipStream_t st) {
  const int64_t total = rows * (int64_t)cols;
  if (npart <= 0 || total <= 0) return hipSuccess;
  const int64_t threads = (total + kRefChunk - 1) / kRefChunk;
  hipLaunchKernelGGL(k_ref_fill, dim3((unsigned)((threads + 255) / 256), (unsigned)npart),
                     dim3(256), 0, st, master, rows, cols, offset, out, ld);
  return hipGetLastError();
}

// LMatrix::clear / LMatrix::scale (clear_matrix.cc, scale_matrix.cc,
// ptr_matrix.cc:50-60): mode 0 stores v, mode 1 multiplies by v (so a NaN
// survives a scale by 0, as in the reference).
__global__ void k_fill_scale(int64_t rows, int cols, double* __restrict__ A, int64_t ld, int mode,
                             double v) {
  const int64_t total = rows * (int64_t)cols;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    double* a = A + (e % rows) + (e / rows) * ld;
    *a = mode == 0 ? v : *a * v;
  }
}

hipError_t launch_fill_scale(int64_t rows, int cols, double* A, int64_t ld, int mode, double v,
                             hipStream_t st) {
  const int64_t total = rows * (int64_t)cols;
  if (total <= 0) return hipSuccess;
  int64_t g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_fill_scale, dim3((unsigned)g), dim3(256), 0, st, rows, cols, A, ld, mode, v);
  return hipGetLastError();
}

__global__ void k_copy_cols(int64_t rows, int cols, const double* __restrict__ src, int64_t lds,
                            double* __restrict__ dst, int64_t ldd) {
  const int64_t total = rows * (int64_t)cols;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    const int64_t i = e % rows, j = e / rows;
    dst[i + j * ldd] = src[i + j * lds];
  }
}

// =====================================================================
// Batched small GEMM over leaves (one workgroup per item i):
//   C_i(:, j) = alpha * (op(A_i) B_i)(:, j) + beta * D_i(:, j)
// written to (C1, D1, alpha1, beta1) for j < csplit and to (C2, D2,
// alpha2, beta2) at column j - csplit otherwise; a null C skips its
// columns.  A_i = A + i*sa (M x K, or K x M read transposed; null A is the
// identity; a null D with a nonzero beta is the identity too); B_i = B + (c>>1)*bs2 + ((c&1)^bflip)*bs1 with c = (i + ioff) >> bshift,
// which selects the half of a node's core solution that belongs to the
// leaf's child at a given level.  16 x 16 output tiles are dealt to the
// four waves.  D may alias C; A and B must not.
// =====================================================================
template <int GM, int GN>
__global__ void __launch_bounds__(256) k_bgemm(BGemmArgs g0, BGemmArgs g1) {
  // blockIdx.y selects one of two independent products of the same shape
  // (one launch for both of a level's per-leaf updates)
  const BGemmArgs& g = blockIdx.y ? g1 : g0;
  const int i = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const double* A = g.A ? g.A + (int64_t)i * g.sa : nullptr;
  const int c = (i + g.ioff) >> g.bshift;
  const double* Bp = g.B + (int64_t)(c >> 1) * g.bs2 + (int64_t)((c & 1) ^ g.bflip) * g.bs1;
  const int mt = (g.M + 15) >> 4, nt = (g.Nc + 15) >> 4;
  // each wave task covers GM x GN tiles of 16 x 16: GM*GN independent MFMA
  // accumulation chains, one A value per row tile and one B value per
  // column tile loaded per k-step
  const int tgm = (mt + GM - 1) / GM, tgn = (nt + GN - 1) / GN;
  for (int task = wid; task < tgm * tgn; task += 4) {
    const int tm0 = (task % tgm) * GM, tn0 = (task / tgm) * GN;
    bool on[GN];
#pragma unroll
    for (int j = 0; j < GN; ++j) {
      const int col0 = (tn0 + j) * 16;
      on[j] = tn0 + j < nt && !(col0 + 16 <= g.csplit && !g.C1) && !(col0 >= g.csplit && !g.C2);
    }
    d4 acc[GM][GN];
#pragma unroll
    for (int a = 0; a < GM; ++a)
#pragma unroll
      for (int j = 0; j < GN; ++j) acc[a][j] = d4zero();
    // the inner dimension in chunks of 16 (four k-steps); the next chunk's
    // A and B values are fetched under the current chunk's MFMAs
    double afc[4][GM], bfc[4][GN], afn[4][GM], bfn[4][GN];
    auto fetch = [&](int k0, double (&af)[4][GM], double (&bf)[4][GN]) {
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const int k = k0 + kk * 4 + lk;
#pragma unroll
        for (int a = 0; a < GM; ++a) {
          const int arow = (tm0 + a) * 16 + lr;
          double v = 0.0;
          if (k < g.K && arow < g.M)
            v = !A ? (arow == k ? 1.0 : 0.0)
                   : (g.transA ? A[k + (int64_t)arow * g.lda] : A[arow + (int64_t)k * g.lda]);
          af[kk][a] = v;
        }
#pragma unroll
        for (int j = 0; j < GN; ++j) {
          const int bcol = (tn0 + j) * 16 + lr;
          bf[kk][j] = (on[j] && k < g.K && bcol < g.Nc) ? Bp[k + (int64_t)bcol * g.ldb] : 0.0;
        }
      }
    };
    fetch(0, afc, bfc);
    for (int k0 = 0; k0 < g.K; k0 += 16) {
      if (k0 + 16 < g.K) fetch(k0 + 16, afn, bfn);
#pragma unroll
      for (int kk = 0; kk < 4; ++kk)
#pragma unroll
        for (int a = 0; a < GM; ++a)
#pragma unroll
          for (int j = 0; j < GN; ++j) acc[a][j] = mfma(afc[kk][a], bfc[kk][j], acc[a][j]);
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
#pragma unroll
        for (int a = 0; a < GM; ++a) afc[kk][a] = afn[kk][a];
#pragma unroll
        for (int j = 0; j < GN; ++j) bfc[kk][j] = bfn[kk][j];
      }
    }
    // epilogue: each lane's output column picks its side (C1 or C2); all D
    // values of the task are read before any C is written (C may alias D
    // element for element), so the reads are in flight together
    double dv[GM][GN][4];
    double* cp[GN];
    int64_t co[GN], dof[GN];
    double al[GN], be[GN];
    bool colok[GN];
#pragma unroll
    for (int j = 0; j < GN; ++j) {
      const int ocol = (tn0 + j) * 16 + lr;
      const bool s1 = ocol < g.csplit;
      const int oc = s1 ? ocol : ocol - g.csplit;
      cp[j] = s1 ? g.C1 : g.C2;
      al[j] = s1 ? g.alpha1 : g.alpha2;
      be[j] = s1 ? g.beta1 : g.beta2;
      co[j] = (int64_t)i * (s1 ? g.sc1 : g.sc2) + (int64_t)oc * (s1 ? g.ldc1 : g.ldc2);
      dof[j] = (int64_t)i * (s1 ? g.sd1 : g.sd2) + (int64_t)oc * (s1 ? g.ldd1 : g.ldd2);
      colok[j] = on[j] && ocol < g.Nc && cp[j] != nullptr;
    }
#pragma unroll
    for (int a = 0; a < GM; ++a)
#pragma unroll
      for (int j = 0; j < GN; ++j) {
        const int ocol = (tn0 + j) * 16 + lr;
        const double* dp = ocol < g.csplit ? g.D1 : g.D2;
        const int dcol = ocol < g.csplit ? ocol : ocol - g.csplit;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int orow = (tm0 + a) * 16 + crow(lane, v);
          // a null D with beta != 0 is the identity
          dv[a][j][v] = (colok[j] && tm0 + a < mt && orow < g.M && be[j] != 0.0)
                            ? (dp ? dp[dof[j] + orow] : (orow == dcol ? 1.0 : 0.0))
                            : 0.0;
        }
      }
#pragma unroll
    for (int a = 0; a < GM; ++a)
#pragma unroll
      for (int j = 0; j < GN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int orow = (tm0 + a) * 16 + crow(lane, v);
          if (colok[j] && tm0 + a < mt && orow < g.M)
            cp[j][co[j] + orow] = al[j] * acc[a][j][v] + be[j] * dv[a][j][v];
        }
  }
}

// One launch for g0 and (when non-null) g1, which must have g0's batch and
// tile shape.
static hipError_t launch_bgemm_pair(const BGemmArgs& g0, const BGemmArgs* g1, hipStream_t st) {
  const BGemmArgs& g = g0;
  if (g.batch <= 0 || g.M <= 0 || g.Nc <= 0) return hipSuccess;
  const int mt = (g.M + 15) >> 4, nt = (g.Nc + 15) >> 4;
  const dim3 grid((unsigned)g.batch, g1 ? 2u : 1u);
  const BGemmArgs& h = g1 ? *g1 : g0;
  // group tiles only while every wave keeps a task
  if (nt == 5 && mt >= 4)  // r + s in (64, 80]: one task per row tile, no idle column tiles
    hipLaunchKernelGGL((k_bgemm<1, 5>), grid, dim3(256), 0, st, g, h);
  else if (mt * nt >= 16 && nt >= 4)
    hipLaunchKernelGGL((k_bgemm<1, 4>), grid, dim3(256), 0, st, g, h);
  else if (mt * nt >= 16 && mt >= 4 && nt == 1)
    hipLaunchKernelGGL((k_bgemm<4, 1>), grid, dim3(256), 0, st, g, h);
  else if (mt * nt >= 8 && nt >= 2)
    hipLaunchKernelGGL((k_bgemm<1, 2>), grid, dim3(256), 0, st, g, h);
  else if (mt * nt >= 8)
    hipLaunchKernelGGL((k_bgemm<2, 1>), grid, dim3(256), 0, st, g, h);
  else
    hipLaunchKernelGGL((k_bgemm<1, 1>), grid, dim3(256), 0, st, g, h);
  return hipGetLastError();
}

hipError_t launch_bgemm(const BGemmArgs& g, hipStream_t st) { return launch_bgemm_pair(g, nullptr, st); }

hipError_t launch_bgemm2(const BGemmArgs& g0, const BGemmArgs& g1, hipStream_t st) {
  const bool same = g0.batch == g1.batch && ((g0.M + 15) >> 4) == ((g1.M + 15) >> 4) &&
                    ((g0.Nc + 15) >> 4) == ((g1.Nc + 15) >> 4) && g1.M > 0 && g1.Nc > 0;
  if (!same) {
    hipError_t e = launch_bgemm_pair(g0, nullptr, st);
    return e != hipSuccess ? e : launch_bgemm_pair(g1, nullptr, st);
  }
  return launch_bgemm_pair(g0, &g1, st);
}

// W_i = I (r x r, ld r, stride `sw` >= r*r between blocks) for i < batch.
__global__ void k_eye(int batch, int r, int64_t sw, double* __restrict__ W) {
  const int64_t total = (int64_t)batch * r * r;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    const int64_t i = e / ((int64_t)r * r), q = e % ((int64_t)r * r);
    W[i * sw + q] = (q % r == q / r) ? 1.0 : 0.0;
  }
}

hipError_t launch_eye(int batch, int r, int64_t sw, double* W, hipStream_t st) {
  const int64_t total = (int64_t)batch * r * r;
  if (total <= 0) return hipSuccess;
  if (sw < (int64_t)r * r) return hipErrorInvalidValue;
  int64_t g = (total + 255) / 256;
  if (g > 16384) g = 16384;
  hipLaunchKernelGGL(k_eye, dim3((unsigned)g), dim3(256), 0, st, batch, r, sw, W);
  return hipGetLastError();
}

// Lane-layout self-test of v_mfma_f64_16x16x4_f64 (one wave, in bounds):
// A[i][k] = i + 100 k, B[k][j] = (k == j % 4), so D[i][j] = i + 100 (j % 4);
// out[4*lane + v] receives accumulator element v of each lane.
__global__ void k_selftest_mfma(double* __restrict__ out) {
  const int l = threadIdx.x;
  const double a = (double)(l & 15) + 100.0 * (double)(l >> 4);
  const double b = ((l >> 4) == ((l & 15) & 3)) ? 1.0 : 0.0;
  d4 c = mfma(a, b, d4zero());
#pragma unroll
  for (int v = 0; v < 4; ++v) out[4 * l + v] = c[v];
}

hipError_t launch_selftest_mfma(double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_selftest_mfma, dim3(1), dim3(64), 0, st, out);
  return hipGetLastError();
}

// Same probe for v_mfma_f32_16x16x4_f32; out[4*lane + v] (as double).
__global__ void k_selftest_mfma_f32(double* __restrict__ out) {
  typedef float f4_ __attribute__((ext_vector_type(4)));
  const int l = threadIdx.x;
  const float a = (float)(l & 15) + 100.0f * (float)(l >> 4);
  const float b = ((l >> 4) == ((l & 15) & 3)) ? 1.0f : 0.0f;
  f4_ z = {0.0f, 0.0f, 0.0f, 0.0f};
  f4_ c = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, z, 0, 0, 0);
#pragma unroll
  for (int v = 0; v < 4; ++v) out[4 * l + v] = (double)c[v];
}

hipError_t launch_selftest_mfma_f32(double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_selftest_mfma_f32, dim3(1), dim3(64), 0, st, out);
  return hipGetLastError();
}

hipError_t launch_copy_cols(int64_t rows, int cols, const double* src, int64_t lds, double* dst,
                            int64_t ldd, hipStream_t st) {
  const int64_t total = rows * (int64_t)cols;
  if (total <= 0) return hipSuccess;
  int64_t g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_copy_cols, dim3((unsigned)g), dim3(256), 0, st, rows, cols, src, lds, dst,
                     ldd);
  return hipGetLastError();
}

}  // namespace hodlr
