/*
 * hodlr_oracle.c -- CPU oracle (checker) for the HODLR factor-and-solve path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Nothing in fastsolver2_amd/ links or calls it.
 *
 * Each function restates the reference's algorithm (Charles-Chao-Chen/
 * fastSolver2, files cited per function) or the published netlib LAPACK 3.x
 * algorithm the reference calls through `-llapack -lblas`
 * (spmd_driver/Makefile:52; vendored Fortran text under netlib_blas).
 *
 * Pinning (which independent check holds each function):
 *   drand48 family      -- glibc drand48_r/lrand48_r via ctypes (tests).
 *   philox4x64          -- numpy.random.Philox raw output + Random123 KAT.
 *   dgemm/dtrsm/dgetrf  -- numpy (OpenBLAS) products/solves; brute-force
 *     /dgetrs/dgesv        pivot sequence on tiny matrices; P*L*U == A.
 *   hsolve / driver     -- dense LU of the assembled matrix (numpy), the
 *                          L=0, U=0 and L=1 (Sherman-Morrison-Woodbury)
 *                          special cases, HODLR-matvec residual.
 *   gemm_red/gemm_bro/  -- the reference's own known-answer relations
 *     node_solve           (tests/test.cc:316-541), restated as tests.
 *   matvec / assemble   -- product with the dense assembly; A = UV^T+diag(D)
 *                          for the reference generator (test.cc:255-262).
 */
#define _GNU_SOURCE
#include "hodlr_oracle.h"

#include <dlfcn.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define IDX(r, c, ld) ((int64_t)(r) + (int64_t)(c) * (int64_t)(ld))

/* ======================================================================
 * drand48 family: X <- (0x5DEECE66D * X + 0xB) mod 2^48.
 * srand48(seed): X = (seed & 0xffffffff) << 16 | 0x330E;
 * drand48 = X / 2^48 (after stepping); lrand48 = X >> 17.
 * ==================================================================== */
void orc_srand48(orc_rand48 *s, long seed) {
  uint64_t v = (uint64_t)seed & 0xffffffffULL;
  s->x = (v << 16) | 0x330EULL;
}
static void rand48_step(orc_rand48 *s) {
  s->x = (0x5DEECE66DULL * s->x + 0xBULL) & ((1ULL << 48) - 1);
}
double orc_drand48(orc_rand48 *s) {
  rand48_step(s);
  return (double)s->x * (1.0 / 281474976710656.0); /* 2^-48, exact */
}
long orc_lrand48(orc_rand48 *s) {
  rand48_step(s);
  return (long)(s->x >> 17);
}

/* matrix.cc:250-264 / lmatrix.cc:327-360: one seed per partition, drawn
 * from an lrand48 stream seeded with the master seed. */
void orc_ref_seeds(long master, int npart, long *seeds) {
  orc_rand48 st;
  orc_srand48(&st, master);
  for (int i = 0; i < npart; ++i) seeds[i] = orc_lrand48(&st);
}

/* ptr_matrix.cc:62-71: rows outer, columns inner, += offset. */
void orc_ptr_rand(long seed, int offset, int rows, int cols, int ld, double *a) {
  orc_rand48 st;
  orc_srand48(&st, seed);
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j) {
      double v = orc_drand48(&st);
      a[IDX(i, j, ld)] = v + offset;
    }
}

/* matrix.cc:266-297: partition k is a (rows_per_part x cols) view at row
 * k*rows_per_part of the column-major (npart*rows_per_part) x cols matrix. */
void orc_ref_matrix(long master, int npart, int rows_per_part, int cols, double *a) {
  long *seeds = (long *)malloc(sizeof(long) * (size_t)npart);
  orc_ref_seeds(master, npart, seeds);
  int ld = npart * rows_per_part;
  for (int k = 0; k < npart; ++k)
    orc_ptr_rand(seeds[k], 0, rows_per_part, cols, ld, a + (int64_t)k * rows_per_part);
  free(seeds);
}

/* matrix.cc:79-102: sequential fill of each partition, += offset. */
void orc_ref_vector(long master, int npart, int rows_per_part, int offset, double *a) {
  long *seeds = (long *)malloc(sizeof(long) * (size_t)npart);
  orc_ref_seeds(master, npart, seeds);
  int64_t cnt = 0;
  for (int k = 0; k < npart; ++k) {
    orc_rand48 st;
    orc_srand48(&st, seeds[k]);
    for (int j = 0; j < rows_per_part; ++j) a[cnt++] = orc_drand48(&st) + offset;
  }
  free(seeds);
}

/* dense_block.cc:54-75 with ptr_matrix.cc:141-164: regenerate U_i, V_i
 * (m x rank, ld m) and D_i (m x 1, +offset) from the leaf seeds, then
 * K_i = U_i V_i^T (dgemm N,T), K_i(j,j) += D_i(j). */
void orc_ref_dense_blocks(int npart, int m, int rank, int offset,
                          const long *useeds, const long *vseeds, const long *dseeds,
                          double *K, int64_t ldk) {
  double *u = (double *)malloc(sizeof(double) * (size_t)m * rank);
  double *v = (double *)malloc(sizeof(double) * (size_t)m * rank);
  double *d = (double *)malloc(sizeof(double) * (size_t)m);
  for (int p = 0; p < npart; ++p) {
    orc_ptr_rand(useeds[p], 0, m, rank, m, u);
    orc_ptr_rand(vseeds[p], 0, m, rank, m, v);
    orc_ptr_rand(dseeds[p], offset, m, 1, m, d);
    double *Kp = K + (int64_t)p * m;
    orc_dgemm('n', 't', m, m, rank, 1.0, u, m, v, m, 0.0, Kp, (int)ldk);
    for (int j = 0; j < m; ++j) Kp[IDX(j, j, ldk)] += d[j];
  }
  free(u); free(v); free(d);
}

/* ======================================================================
 * Philox4x64-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11).  Multipliers and Weyl constants as published.
 * ==================================================================== */
static void mulhilo64(uint64_t a, uint64_t b, uint64_t *hi, uint64_t *lo) {
  __uint128_t p = (__uint128_t)a * (__uint128_t)b;
  *hi = (uint64_t)(p >> 64);
  *lo = (uint64_t)p;
}
void orc_philox4x64(const uint64_t ctr_in[4], const uint64_t key_in[2], uint64_t out[4]) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ULL, M1 = 0xCA5A826395121157ULL;
  const uint64_t W0 = 0x9E3779B97F4A7C15ULL, W1 = 0xBB67AE8584CAA73BULL;
  uint64_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint64_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += W0; k1 += W1; }
    uint64_t hi0, lo0, hi1, lo1;
    mulhilo64(M0, c0, &hi0, &lo0);
    mulhilo64(M1, c2, &hi1, &lo1);
    uint64_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void philox_block(uint64_t seed, uint32_t stream, uint32_t system, uint64_t blk,
                         uint64_t w[4]) {
  uint64_t ctr[4] = {blk, stream, system, 0};
  uint64_t key[2] = {seed, 0};
  orc_philox4x64(ctr, key, w);
}

void orc_gen_uniform(uint64_t seed, uint32_t stream, uint32_t system,
                     int64_t n, double scale, double *out) {
  for (int64_t b = 0; b * 4 < n; ++b) {
    uint64_t w[4];
    philox_block(seed, stream, system, (uint64_t)b, w);
    for (int q = 0; q < 4 && b * 4 + q < n; ++q)
      out[b * 4 + q] = (double)(w[q] >> 11) * 0x1.0p-53 * scale;
  }
}

/* Box-Muller: u1 = ((w0>>11)+1)*2^-53 in (0,1], u2 = (w1>>11)*2^-53;
 * z0 = sqrt(-2 ln u1) cos(2 pi u2), z1 = sqrt(-2 ln u1) sin(2 pi u2). */
void orc_gen_normal(uint64_t seed, uint32_t stream, uint32_t system,
                    int64_t n, double scale, double *out) {
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t b = 0; b * 4 < n; ++b) {
    uint64_t w[4];
    philox_block(seed, stream, system, (uint64_t)b, w);
    double z[4];
    for (int pr = 0; pr < 2; ++pr) {
      double u1 = (double)((w[2 * pr] >> 11) + 1) * 0x1.0p-53;
      double u2 = (double)(w[2 * pr + 1] >> 11) * 0x1.0p-53;
      double rad = sqrt(-2.0 * log(u1));
      z[2 * pr] = rad * cos(two_pi * u2);
      z[2 * pr + 1] = rad * sin(two_pi * u2);
    }
    for (int q = 0; q < 4 && b * 4 + q < n; ++q) out[b * 4 + q] = z[q] * scale;
  }
}

/* ======================================================================
 * netlib BLAS/LAPACK restatements (loop order as in the netlib_blas Fortran files).
 * ==================================================================== */
/* idamax.f: first index (1-based) of the largest |x_i|. */
int orc_idamax(int n, const double *x, int incx) {
  if (n < 1 || incx <= 0) return 0;
  int imax = 1;
  double dmax = fabs(x[0]);
  for (int i = 2; i <= n; ++i) {
    double v = fabs(x[(int64_t)(i - 1) * incx]);
    if (v > dmax) { imax = i; dmax = v; }
  }
  return imax;
}

/* dgemm.f:278-380 (no zero tests in the NN update loop in this version). */
void orc_dgemm(char ta, char tb, int m, int n, int k, double alpha,
               const double *A, int lda, const double *B, int ldb,
               double beta, double *C, int ldc) {
  int nota = (ta == 'n' || ta == 'N'), notb = (tb == 'n' || tb == 'N');
  if (m == 0 || n == 0) return;
  if (alpha == 0.0 || k == 0) {
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < m; ++i)
        C[IDX(i, j, ldc)] = (beta == 0.0) ? 0.0 : beta * C[IDX(i, j, ldc)];
    return;
  }
  if (notb) {
    if (nota) {
      for (int j = 0; j < n; ++j) {
        if (beta == 0.0) for (int i = 0; i < m; ++i) C[IDX(i, j, ldc)] = 0.0;
        else if (beta != 1.0) for (int i = 0; i < m; ++i) C[IDX(i, j, ldc)] *= beta;
        for (int l = 0; l < k; ++l) {
          double temp = alpha * B[IDX(l, j, ldb)];
          for (int i = 0; i < m; ++i) C[IDX(i, j, ldc)] += temp * A[IDX(i, l, lda)];
        }
      }
    } else {
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
          double temp = 0.0;
          for (int l = 0; l < k; ++l) temp += A[IDX(l, i, lda)] * B[IDX(l, j, ldb)];
          C[IDX(i, j, ldc)] = (beta == 0.0) ? alpha * temp : alpha * temp + beta * C[IDX(i, j, ldc)];
        }
    }
  } else {
    if (nota) {
      for (int j = 0; j < n; ++j) {
        if (beta == 0.0) for (int i = 0; i < m; ++i) C[IDX(i, j, ldc)] = 0.0;
        else if (beta != 1.0) for (int i = 0; i < m; ++i) C[IDX(i, j, ldc)] *= beta;
        for (int l = 0; l < k; ++l) {
          double temp = alpha * B[IDX(j, l, ldb)];
          for (int i = 0; i < m; ++i) C[IDX(i, j, ldc)] += temp * A[IDX(i, l, lda)];
        }
      }
    } else {
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
          double temp = 0.0;
          for (int l = 0; l < k; ++l) temp += A[IDX(l, i, lda)] * B[IDX(j, l, ldb)];
          C[IDX(i, j, ldc)] = (beta == 0.0) ? alpha * temp : alpha * temp + beta * C[IDX(i, j, ldc)];
        }
    }
  }
}

/* dtrsm.f, SIDE='L', alpha = 1: the four (uplo, trans) cases. */
void orc_dtrsm_left(char uplo, char trans, char diag, int m, int n,
                    const double *A, int lda, double *B, int ldb) {
  int upper = (uplo == 'u' || uplo == 'U');
  int notr = (trans == 'n' || trans == 'N');
  int nounit = (diag == 'n' || diag == 'N');
  if (notr) {
    if (upper) {                               /* dtrsm.f:279-293 */
      for (int j = 0; j < n; ++j)
        for (int k = m - 1; k >= 0; --k)
          if (B[IDX(k, j, ldb)] != 0.0) {
            if (nounit) B[IDX(k, j, ldb)] /= A[IDX(k, k, lda)];
            double bk = B[IDX(k, j, ldb)];
            for (int i = 0; i < k; ++i) B[IDX(i, j, ldb)] -= bk * A[IDX(i, k, lda)];
          }
    } else {                                   /* dtrsm.f:295-309 */
      for (int j = 0; j < n; ++j)
        for (int k = 0; k < m; ++k)
          if (B[IDX(k, j, ldb)] != 0.0) {
            if (nounit) B[IDX(k, j, ldb)] /= A[IDX(k, k, lda)];
            double bk = B[IDX(k, j, ldb)];
            for (int i = k + 1; i < m; ++i) B[IDX(i, j, ldb)] -= bk * A[IDX(i, k, lda)];
          }
    }
  } else {
    if (upper) {                               /* dtrsm.f:316-325: B := inv(A^T) B */
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
          double temp = B[IDX(i, j, ldb)];
          for (int k = 0; k < i; ++k) temp -= A[IDX(k, i, lda)] * B[IDX(k, j, ldb)];
          if (nounit) temp /= A[IDX(i, i, lda)];
          B[IDX(i, j, ldb)] = temp;
        }
    } else {                                   /* dtrsm.f:327-336 */
      for (int j = 0; j < n; ++j)
        for (int i = m - 1; i >= 0; --i) {
          double temp = B[IDX(i, j, ldb)];
          for (int k = i + 1; k < m; ++k) temp -= A[IDX(k, i, lda)] * B[IDX(k, j, ldb)];
          if (nounit) temp /= A[IDX(i, i, lda)];
          B[IDX(i, j, ldb)] = temp;
        }
    }
  }
}

/* dlaswp.f: row interchanges k1..k2 (1-based), incx = +1 forward, -1 backward. */
void orc_dlaswp(int n, double *A, int lda, int k1, int k2, const int *ipiv, int incx) {
  if (incx > 0) {
    for (int i = k1; i <= k2; ++i) {
      int ip = ipiv[i - 1];
      if (ip != i)
        for (int c = 0; c < n; ++c) {
          double t = A[IDX(i - 1, c, lda)];
          A[IDX(i - 1, c, lda)] = A[IDX(ip - 1, c, lda)];
          A[IDX(ip - 1, c, lda)] = t;
        }
    }
  } else if (incx < 0) {
    for (int i = k2; i >= k1; --i) {
      int ip = ipiv[i - 1];
      if (ip != i)
        for (int c = 0; c < n; ++c) {
          double t = A[IDX(i - 1, c, lda)];
          A[IDX(i - 1, c, lda)] = A[IDX(ip - 1, c, lda)];
          A[IDX(ip - 1, c, lda)] = t;
        }
    }
  }
}

/* dgetrf2.f:172-266: recursive LU with partial pivoting (column split n1 =
 * min(m,n)/2).  ipiv is 1-based relative to this block.  Returns INFO. */
int orc_dgetrf2(int m, int n, double *A, int lda, int *ipiv) {
  if (m == 0 || n == 0) return 0;
  int info = 0;
  if (m == 1) {
    ipiv[0] = 1;
    if (A[0] == 0.0) info = 1;
    return info;
  }
  if (n == 1) {
    const double sfmin = 2.2250738585072014e-308; /* dlamch('S') */
    int i = orc_idamax(m, A, 1);
    ipiv[0] = i;
    if (A[i - 1] != 0.0) {
      if (i != 1) { double t = A[0]; A[0] = A[i - 1]; A[i - 1] = t; }
      if (fabs(A[0]) >= sfmin) {
        double rcp = 1.0 / A[0];
        for (int k = 1; k < m; ++k) A[k] *= rcp;
      } else {
        for (int k = 1; k < m; ++k) A[k] /= A[0];
      }
    } else {
      info = 1;
    }
    return info;
  }
  int mn = m < n ? m : n;
  int n1 = mn / 2, n2 = n - n1;
  int iinfo = orc_dgetrf2(m, n1, A, lda, ipiv);
  if (info == 0 && iinfo > 0) info = iinfo;
  double *A12 = A + IDX(0, n1, lda), *A21 = A + IDX(n1, 0, lda), *A22 = A + IDX(n1, n1, lda);
  orc_dlaswp(n2, A12, lda, 1, n1, ipiv, 1);
  orc_dtrsm_left('L', 'N', 'U', n1, n2, A, lda, A12, lda);
  orc_dgemm('N', 'N', m - n1, n2, n1, -1.0, A21, lda, A12, lda, 1.0, A22, lda);
  iinfo = orc_dgetrf2(m - n1, n2, A22, lda, ipiv + n1);
  if (info == 0 && iinfo > 0) info = iinfo + n1;
  for (int i = n1; i < mn; ++i) ipiv[i] += n1;
  orc_dlaswp(n1, A, lda, n1 + 1, mn, ipiv, 1);
  return info;
}

/* dgetrf.f:167-220 with NB = 64 (ilaenv.f:270-277 for xGETRF). */
int orc_dgetrf(int m, int n, double *A, int lda, int *ipiv) {
  if (m == 0 || n == 0) return 0;
  const int nb = 64;
  int mn = m < n ? m : n;
  if (nb <= 1 || nb >= mn) return orc_dgetrf2(m, n, A, lda, ipiv);
  int info = 0;
  for (int j = 1; j <= mn; j += nb) {
    int jb = mn - j + 1 < nb ? mn - j + 1 : nb;
    int iinfo = orc_dgetrf2(m - j + 1, jb, A + IDX(j - 1, j - 1, lda), lda, ipiv + j - 1);
    if (info == 0 && iinfo > 0) info = iinfo + j - 1;
    int hi = m < j + jb - 1 ? m : j + jb - 1;
    for (int i = j; i <= hi; ++i) ipiv[i - 1] += j - 1;
    orc_dlaswp(j - 1, A, lda, j, j + jb - 1, ipiv, 1);
    if (j + jb <= n) {
      orc_dlaswp(n - j - jb + 1, A + IDX(0, j + jb - 1, lda), lda, j, j + jb - 1, ipiv, 1);
      orc_dtrsm_left('L', 'N', 'U', jb, n - j - jb + 1, A + IDX(j - 1, j - 1, lda), lda,
                     A + IDX(j - 1, j + jb - 1, lda), lda);
      if (j + jb <= m)
        orc_dgemm('N', 'N', m - j - jb + 1, n - j - jb + 1, jb, -1.0,
                  A + IDX(j + jb - 1, j - 1, lda), lda, A + IDX(j - 1, j + jb - 1, lda), lda,
                  1.0, A + IDX(j + jb - 1, j + jb - 1, lda), lda);
    }
  }
  return info;
}

/* dgetrs.f:191-210. */
int orc_dgetrs(char trans, int n, int nrhs, const double *A, int lda,
               const int *ipiv, double *B, int ldb) {
  if (n == 0 || nrhs == 0) return 0;
  if (trans == 'n' || trans == 'N') {
    orc_dlaswp(nrhs, B, ldb, 1, n, ipiv, 1);
    orc_dtrsm_left('L', 'N', 'U', n, nrhs, A, lda, B, ldb);
    orc_dtrsm_left('U', 'N', 'N', n, nrhs, A, lda, B, ldb);
  } else {
    orc_dtrsm_left('U', 'T', 'N', n, nrhs, A, lda, B, ldb);
    orc_dtrsm_left('L', 'T', 'U', n, nrhs, A, lda, B, ldb);
    orc_dlaswp(nrhs, B, ldb, 1, n, ipiv, -1);
  }
  return 0;
}

/* dgesv.f:167-173. */
int orc_dgesv(int n, int nrhs, double *A, int lda, int *ipiv, double *B, int ldb) {
  int info = orc_dgetrf(n, n, A, lda, ipiv);
  if (info == 0) orc_dgetrs('N', n, nrhs, A, lda, ipiv, B, ldb);
  return info;
}

/* ======================================================================
 * BLAS backend switch (Fortran ABI, everything by pointer).
 * ==================================================================== */
typedef void (*fdgemm_t)(const char *, const char *, const int *, const int *, const int *,
                         const double *, const double *, const int *, const double *,
                         const int *, const double *, double *, const int *);
typedef void (*fdgesv_t)(const int *, const int *, double *, const int *, int *, double *,
                         const int *, int *);
typedef void (*fsetthreads_t)(int);

static fdgemm_t g_fdgemm = NULL;
static fdgesv_t g_fdgesv = NULL;
static char g_blas_name[512] = "netlib-restated";

int orc_blas_load(const char *path, const char *prefix) {
  void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!h) return -1;
  char nm[128];
  snprintf(nm, sizeof nm, "%sdgemm_", prefix ? prefix : "");
  fdgemm_t gm = (fdgemm_t)dlsym(h, nm);
  snprintf(nm, sizeof nm, "%sdgesv_", prefix ? prefix : "");
  fdgesv_t gs = (fdgesv_t)dlsym(h, nm);
  if (!gm || !gs) return -2;
  /* each oracle thread calls BLAS sequentially, as the reference requires
   * (spmd_driver/Makefile:44-49). */
  snprintf(nm, sizeof nm, "%sopenblas_set_num_threads", prefix ? prefix : "");
  fsetthreads_t st = (fsetthreads_t)dlsym(h, nm);
  if (st) st(1);
  g_fdgemm = gm;
  g_fdgesv = gs;
  snprintf(g_blas_name, sizeof g_blas_name, "%s", path);
  return 0;
}
void orc_blas_use_netlib(void) {
  g_fdgemm = NULL;
  g_fdgesv = NULL;
  snprintf(g_blas_name, sizeof g_blas_name, "netlib-restated");
}
const char *orc_blas_name(void) { return g_blas_name; }

static void b_dgemm(char ta, char tb, int m, int n, int k, double alpha, const double *A,
                    int lda, const double *B, int ldb, double beta, double *C, int ldc) {
  if (g_fdgemm) {
    if (m == 0 || n == 0) return;
    g_fdgemm(&ta, &tb, &m, &n, &k, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);
  } else {
    orc_dgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  }
}
static int b_dgesv(int n, int nrhs, double *A, int lda, int *ipiv, double *B, int ldb) {
  if (g_fdgesv) {
    int info = 0;
    g_fdgesv(&n, &nrhs, A, &lda, ipiv, B, &ldb, &info);
    return info;
  }
  return orc_dgesv(n, nrhs, A, lda, ipiv, B, ldb);
}

/* ======================================================================
 * hsolve restatement (src/tasks/leaf_solve.cc:73-154).
 * X columns [0,nrhs) = d of this node, [nrhs, nrhs+rank) = u of this node.
 * ==================================================================== */
int orc_hsolve(int nrow, int nrhs, int rank, int nPart, int LD,
               double *K, double *X, const double *V) {
  if (nrow <= 0 || nrhs < 0 || rank <= 0 || nPart <= 0 || LD < nrow) return -1;
  if (nPart == 1) {                                        /* :80-92 */
    int *ipiv = (int *)malloc(sizeof(int) * (size_t)nrow);
    int info = b_dgesv(nrow, nrhs, K, LD, ipiv, X, LD);
    free(ipiv);
    return info;
  }
  if (nrow % 2 || nPart % 2) return -2;                    /* :95-96 */
  int h = nrow / 2;
  double *d0 = X, *d1 = X + h;
  const double *V0 = V, *V1 = V + h;
  double *u0 = d0 + (int64_t)nrhs * LD, *u1 = d1 + (int64_t)nrhs * LD;
  int info = orc_hsolve(h, nrhs + rank, rank, nPart / 2, LD, K, d0, V0);       /* :103 */
  if (info) return info;
  info = orc_hsolve(h, nrhs + rank, rank, nPart / 2, LD, K + h, d1, V1);       /* :104 */
  if (info) return info;
  int S_size = 2 * rank;                                  /* :119-126 */
  double *S = (double *)calloc((size_t)S_size * S_size, sizeof(double));
  double *RHS = (double *)malloc(sizeof(double) * (size_t)S_size * (nrhs > 0 ? nrhs : 1));
  for (int i = 0; i < S_size; ++i) S[S_size * i + i] = 1.0;
  double *V0Tu0 = S + rank;                               /* :129-132 */
  double *V1Tu1 = S + (int64_t)rank * S_size;
  double *V0Td0 = RHS + rank;
  double *V1Td1 = RHS;
  b_dgemm('t', 'n', rank, rank, h, 1.0, V0, LD, u0, LD, 0.0, V0Tu0, S_size);   /* :134-137 */
  b_dgemm('t', 'n', rank, rank, h, 1.0, V1, LD, u1, LD, 0.0, V1Tu1, S_size);
  b_dgemm('t', 'n', rank, nrhs, h, 1.0, V0, LD, d0, LD, 0.0, V0Td0, S_size);
  b_dgemm('t', 'n', rank, nrhs, h, 1.0, V1, LD, d1, LD, 0.0, V1Td1, S_size);
  int *ipiv = (int *)malloc(sizeof(int) * (size_t)S_size);
  info = b_dgesv(S_size, nrhs, S, S_size, ipiv, RHS, S_size);                  /* :139-141 */
  free(ipiv);
  if (info == 0) {                                        /* :143-153 */
    double *eta0 = V1Td1, *eta1 = V0Td0;
    b_dgemm('n', 'n', h, nrhs, rank, -1.0, u0, LD, eta0, S_size, 1.0, d0, LD);
    b_dgemm('n', 'n', h, nrhs, rank, -1.0, u1, LD, eta1, S_size, 1.0, d1, LD);
  }
  free(S);
  free(RHS);
  return info ? (1 << 30) : 0;  /* singular core: report as INFO > 0 */
}

/* ======================================================================
 * LMatrix level ops (slot mapping of the Contraction projector,
 * projector.cc:17-39: slot = p / colorSize, colorSize = nchunk / nslot).
 * ==================================================================== */
void orc_gemm_red(int rows, int nchunk, int nslot, double alpha,
                  const double *A, int lda, int ka, const double *B, int ldb, int kb,
                  double *C, int ldc) {
  int blk = rows / nchunk, csz = nchunk / nslot;
  for (int s = 0; s < nslot; ++s)                          /* scale(0), lmatrix.cc:719 */
    for (int j = 0; j < kb; ++j)
      for (int i = 0; i < ka; ++i) C[IDX(s * ka + i, j, ldc)] = 0.0;
  for (int p = 0; p < nchunk; ++p) {                       /* gemm_reduce.cc:58-76 */
    int slot = p / csz;
    b_dgemm('t', 'n', ka, kb, blk, alpha, A + (int64_t)p * blk, lda, B + (int64_t)p * blk, ldb,
            1.0, C + (int64_t)slot * ka, ldc);
  }
}

void orc_gemm_bro(int rows, int nchunk, int nslot, double alpha,
                  const double *u, int ldu, int k, const double *eta, int lde,
                  double *d, int ldd, int nd) {
  int blk = rows / nchunk, csz = nchunk / nslot;
  for (int p = 0; p < nchunk; ++p) {                       /* gemm_broadcast.cc:58-82 */
    int slot = p / csz;
    b_dgemm('n', 'n', blk, nd, k, alpha, u + (int64_t)p * blk, ldu, eta + (int64_t)slot * k, lde,
            1.0, d + (int64_t)p * blk, ldd);
  }
}

int orc_node_solve(int nnode, int r, const double *VTu, int ldu, double *VTd, int ldd, int nrhs) {
  int n2 = 2 * r;
  double *S = (double *)malloc(sizeof(double) * (size_t)n2 * n2);
  double *B = (double *)malloc(sizeof(double) * (size_t)n2 * (nrhs > 0 ? nrhs : 1));
  int *ipiv = (int *)malloc(sizeof(int) * (size_t)n2);
  int info = 0;
  for (int p = 0; p < nnode && info == 0; ++p) {
    const double *A = VTu + (int64_t)p * n2;
    double *Bm = VTd + (int64_t)p * n2;
    memset(S, 0, sizeof(double) * (size_t)n2 * n2);        /* S.identity(), :61 */
    for (int i = 0; i < n2; ++i) S[IDX(i, i, n2)] = 1.0;
    for (int i = 0; i < r; ++i)                            /* :66-71 */
      for (int j = 0; j < r; ++j) {
        S[IDX(r + i, j, n2)] = A[IDX(i, j, ldu)];
        S[IDX(i, r + j, n2)] = A[IDX(r + i, j, ldu)];
      }
    for (int j = 0; j < nrhs; ++j)                         /* :73-80 swap halves */
      for (int i = 0; i < r; ++i) {
        double t = Bm[IDX(i, j, ldd)];
        Bm[IDX(i, j, ldd)] = Bm[IDX(r + i, j, ldd)];
        Bm[IDX(r + i, j, ldd)] = t;
      }
    info = b_dgesv(n2, nrhs, S, n2, ipiv, Bm, ldd);        /* S.solve(BMat), :81 */
    if (info) info += p * n2;
  }
  free(S); free(B); free(ipiv);
  return info;
}

int orc_node_solve_region(int r0, int r1, int nrhs,
                          const double *VTu0, int ld0, const double *VTu1, int ld1,
                          double *VTd0, int ldd0, double *VTd1, int ldd1) {
  int n = r0 + r1;
  double *S = (double *)calloc((size_t)n * n, sizeof(double));
  double *B = (double *)malloc(sizeof(double) * (size_t)n * (nrhs > 0 ? nrhs : 1));
  int *ipiv = (int *)malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; ++i) S[IDX(i, i, n)] = 1.0;
  for (int i = 0; i < r0; ++i)                             /* node_solve_region.cc:77-82 */
    for (int j = 0; j < r1; ++j) {
      S[IDX(i, r0 + j, n)] = VTu1[IDX(i, j, ld1)];
      S[IDX(r0 + j, i, n)] = VTu0[IDX(j, i, ld0)];
    }
  for (int j = 0; j < nrhs; ++j) {                         /* :84-91 */
    for (int i = 0; i < r0; ++i) B[IDX(i, j, n)] = VTd1[IDX(i, j, ldd1)];
    for (int i = 0; i < r1; ++i) B[IDX(r0 + i, j, n)] = VTd0[IDX(i, j, ldd0)];
  }
  int info = b_dgesv(n, nrhs, S, n, ipiv, B, n);
  if (info == 0)
    for (int j = 0; j < nrhs; ++j) {                       /* :95-102 */
      for (int i = 0; i < r0; ++i) VTd1[IDX(i, j, ldd1)] = B[IDX(i, j, n)];
      for (int i = 0; i < r1; ++i) VTd0[IDX(i, j, ldd0)] = B[IDX(r0 + i, j, n)];
    }
  free(S); free(B); free(ipiv);
  return info;
}

/* ======================================================================
 * Threaded driver (solver_irregular_part.cc:104-140).  Index launches
 * become pthread-parallel loops; the REDOP_ADD fold of gemmRed is made
 * deterministic: per-chunk products summed in chunk order.
 * ==================================================================== */
typedef struct {
  int kind;        /* 0 leaf tasks, 1 reduce chunks, 2 node solves, 3 broadcast chunks */
  int64_t N;
  int m, L, r, s, T, level;
  double *K, *X;
  const double *V;
  double *VTu, *VTd;   /* slots */
  double *part;        /* per-chunk partials (reduce) */
  int next;            /* shared work counter */
  int count;
  int info;
  pthread_mutex_t mu;
} drv_t;

static int drv_take(drv_t *d) {
  pthread_mutex_lock(&d->mu);
  int i = d->next < d->count ? d->next++ : -1;
  pthread_mutex_unlock(&d->mu);
  return i;
}
static void drv_err(drv_t *d, int info) {
  pthread_mutex_lock(&d->mu);
  if (d->info == 0) d->info = info;
  pthread_mutex_unlock(&d->mu);
}

static void *drv_worker(void *arg) {
  drv_t *d = (drv_t *)arg;
  int64_t N = d->N;
  int ntask = 1 << d->T;
  int64_t rblk = N / ntask;
  int r = d->r, s = d->s, i = d->level;
  int wd = s + r * i;                 /* d = cols [0, s + r i)         */
  int64_t ucol = (int64_t)(s + r * i);  /* u = cols [s + r i, s + r(i+1)) */
  int nslot = 1 << (i + 1), csz = ntask / nslot;
  for (;;) {
    int p = drv_take(d);
    if (p < 0) break;
    if (d->kind == 0) {
      /* LeafSolveTask (leaf_solve.cc:38-71): nPart = 2^(L-T) leaves, nrhs
       * passed = total cols - level*rank = s + r*T. */
      int nPart = 1 << (d->L - d->T);
      int info = orc_hsolve((int)rblk, s + r * d->T, r, nPart, (int)N,
                            d->K + p * rblk, d->X + p * rblk, d->V + p * rblk);
      if (info) drv_err(d, info);
    } else if (d->kind == 1) {
      /* GemmRedTask: chunk p's V_p^T [u_p | d_p] into its own partial. */
      double *pp = d->part + (int64_t)p * r * (r + wd);
      b_dgemm('t', 'n', r, r, (int)rblk, 1.0, d->V + p * rblk, (int)N,
              d->X + p * rblk + ucol * N, (int)N, 0.0, pp, r);
      if (wd > 0)
        b_dgemm('t', 'n', r, wd, (int)rblk, 1.0, d->V + p * rblk, (int)N,
                d->X + p * rblk, (int)N, 0.0, pp + (int64_t)r * r, r);
    } else if (d->kind == 2) {
      /* NodeSolveTask on node slot p (2r rows). */
      int info = orc_node_solve(1, r, d->VTu + (int64_t)p * 2 * r, nslot * r,
                                d->VTd + (int64_t)p * 2 * r, nslot * r, wd);
      if (info) drv_err(d, info + (1 << 20));
    } else {
      /* GemmBroTask: d_p -= u_p * eta_slot(p). */
      int slot = p / csz;
      if (wd > 0)
        b_dgemm('n', 'n', (int)rblk, wd, r, -1.0, d->X + p * rblk + ucol * N, (int)N,
                d->VTd + (int64_t)slot * r, nslot * r, 1.0, d->X + p * rblk, (int)N);
    }
  }
  return NULL;
}

static void drv_run(drv_t *d, int kind, int count, int nthreads) {
  d->kind = kind;
  d->next = 0;
  d->count = count;
  if (nthreads <= 1 || count <= 1) {
    drv_worker(d);
    return;
  }
  int nt = nthreads < count ? nthreads : count;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nt);
  for (int t = 0; t < nt; ++t) pthread_create(&th[t], NULL, drv_worker, d);
  for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
  free(th);
}

int orc_hodlr_solve(int64_t N, int m, int L, int r, int s, int task_level,
                    int nthreads, double *K, double *X, const double *V) {
  if (N <= 0 || m <= 0 || L < 0 || r <= 0 || s < 0) return -1;
  if (N != ((int64_t)m << L)) return -2;
  if (task_level < 0 || task_level > L) return -3;
  drv_t d;
  memset(&d, 0, sizeof d);
  pthread_mutex_init(&d.mu, NULL);
  d.N = N; d.m = m; d.L = L; d.r = r; d.s = s; d.T = task_level;
  d.K = K; d.X = X; d.V = V;
  int ntask = 1 << task_level;
  /* leaf phase: kTree.solve (solver_irregular_part.cc:114) */
  drv_run(&d, 0, ntask, nthreads);
  if (d.info) { pthread_mutex_destroy(&d.mu); return d.info; }
  /* level loop (solver_irregular_part.cc:117-140) */
  for (int i = task_level - 1; i >= 0; --i) {
    int nslot = 1 << (i + 1), csz = ntask / nslot;
    int wd = s + r * i;
    d.level = i;
    d.VTu = (double *)calloc((size_t)nslot * r * r, sizeof(double));
    d.VTd = (double *)calloc((size_t)nslot * r * (wd > 0 ? wd : 1), sizeof(double));
    d.part = (double *)malloc(sizeof(double) * (size_t)ntask * r * (r + wd));
    drv_run(&d, 1, ntask, nthreads);
    /* fold partials into child slots in chunk order (REDOP_ADD) */
    for (int p = 0; p < ntask; ++p) {
      int slot = p / csz;
      const double *pp = d.part + (int64_t)p * r * (r + wd);
      for (int j = 0; j < r; ++j)
        for (int a = 0; a < r; ++a) d.VTu[IDX(slot * r + a, j, nslot * r)] += pp[IDX(a, j, r)];
      for (int j = 0; j < wd; ++j)
        for (int a = 0; a < r; ++a)
          d.VTd[IDX(slot * r + a, j, nslot * r)] += pp[(int64_t)r * r + IDX(a, j, r)];
    }
    drv_run(&d, 2, 1 << i, nthreads);
    if (d.info == 0) drv_run(&d, 3, ntask, nthreads);
    free(d.VTu); free(d.VTd); free(d.part);
    if (d.info) break;
  }
  pthread_mutex_destroy(&d.mu);
  return d.info;
}

void orc_build_x(int64_t N, int r, int s, int L, const double *B, int64_t ldb,
                 const double *U, int64_t ldu, double *X) {
  for (int j = 0; j < s; ++j) memcpy(X + j * N, B + j * ldb, sizeof(double) * (size_t)N);
  for (int t = 0; t < L; ++t)
    for (int j = 0; j < r; ++j)
      memcpy(X + (int64_t)(s + r * t + j) * N, U + j * ldu, sizeof(double) * (size_t)N);
}

/* ======================================================================
 * HODLR matvec and dense assembly.
 * ==================================================================== */
void orc_hodlr_matvec(int64_t N, int m, int r, int nrhs,
                      const double *K, int64_t ldk, const double *U, int64_t ldu,
                      const double *V, int64_t ldv, const double *x, int64_t ldx,
                      double *y, int64_t ldy) {
  int64_t P = N / m;
  double *vtx = (double *)malloc(sizeof(double) * (size_t)r);
  double *vtxi = (double *)malloc(sizeof(double) * (size_t)r);
  for (int c = 0; c < nrhs; ++c) {
    const double *xc = x + c * ldx;
    double *yc = y + c * ldy;
    for (int a = 0; a < r; ++a) {
      double acc = 0.0;
      for (int64_t i = 0; i < N; ++i) acc += V[i + a * ldv] * xc[i];
      vtx[a] = acc;
    }
    for (int64_t p = 0; p < P; ++p) {
      int64_t r0 = p * m;
      for (int a = 0; a < r; ++a) {
        double acc = 0.0;
        for (int i = 0; i < m; ++i) acc += V[r0 + i + a * ldv] * xc[r0 + i];
        vtxi[a] = vtx[a] - acc;
      }
      for (int i = 0; i < m; ++i) {
        double acc = 0.0;
        for (int j = 0; j < m; ++j) acc += K[r0 + i + j * ldk] * xc[r0 + j];
        for (int a = 0; a < r; ++a) acc += U[r0 + i + a * ldu] * vtxi[a];
        yc[r0 + i] = acc;
      }
    }
  }
  free(vtx); free(vtxi);
}

void orc_dense_assemble(int64_t N, int m, int r, const double *K, int64_t ldk,
                        const double *U, int64_t ldu, const double *V, int64_t ldv,
                        double *A) {
  for (int64_t j = 0; j < N; ++j)
    for (int64_t i = 0; i < N; ++i) {
      double v;
      if (i / m == j / m) {
        v = K[i + (j % m) * ldk];
      } else {
        v = 0.0;
        for (int a = 0; a < r; ++a) v += U[i + a * ldu] * V[j + a * ldv];
      }
      A[i + j * N] = v;
    }
}
