"""ctypes wrapper of the CPU oracle (oracle/libhodlr_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product package
(``fastsolver2_amd``) never imports it.

All arrays are numpy float64, column-major ("F" order), as the reference's
storage (src/ptr_matrix.cc:29-37).
"""
from __future__ import annotations

import ctypes as C
import glob
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libhodlr_oracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)
_i64 = C.c_int64


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc only)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_drand48.restype = C.c_double
        L.orc_lrand48.restype = C.c_long
        L.orc_blas_name.restype = C.c_char_p
        L.orc_philox4x64.argtypes = [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                     C.POINTER(C.c_uint64)]
        L.orc_gen_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, _i64,
                                      C.c_double, _dp]
        L.orc_gen_normal.argtypes = L.orc_gen_uniform.argtypes
        L.orc_ref_dense_blocks.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _lp, _lp,
                                           _lp, _dp, _i64]
        L.orc_hodlr_solve.argtypes = [_i64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_int, _dp, _dp, _dp]
        L.orc_build_x.argtypes = [_i64, C.c_int, C.c_int, C.c_int, _dp, _i64, _dp, _i64,
                                  _dp]
        L.orc_hodlr_matvec.argtypes = [_i64, C.c_int, C.c_int, C.c_int, _dp, _i64, _dp,
                                       _i64, _dp, _i64, _dp, _i64, _dp, _i64]
        L.orc_dense_assemble.argtypes = [_i64, C.c_int, C.c_int, _dp, _i64, _dp, _i64,
                                         _dp, _i64, _dp]
        L.orc_dgemm.argtypes = [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, C.c_double,
                                _dp, C.c_int, _dp, C.c_int, C.c_double, _dp, C.c_int]
        L.orc_dtrsm_left.argtypes = [C.c_char, C.c_char, C.c_char, C.c_int, C.c_int, _dp,
                                     C.c_int, _dp, C.c_int]
        L.orc_gemm_red.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_int,
                                   C.c_int, _dp, C.c_int, C.c_int, _dp, C.c_int]
        L.orc_gemm_bro.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_int,
                                   C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_int]
        L.orc_node_solve.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_int]
        L.orc_node_solve_region.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_int, _dp,
                                            C.c_int, _dp, C.c_int, _dp, C.c_int]
        L.orc_blas_load.argtypes = [C.c_char_p, C.c_char_p]
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1)
    return a.ctypes.data_as(_dp)


def _ld(a):
    return a.shape[0]


# ---------------------------------------------------------------- generators
class Rand48:
    """drand48_r / lrand48_r stream (restated LCG)."""

    def __init__(self, seed: int):
        self._st = (C.c_uint64 * 1)()
        lib().orc_srand48(self._st, C.c_long(seed))

    def drand48(self) -> float:
        return lib().orc_drand48(self._st)

    def lrand48(self) -> int:
        return lib().orc_lrand48(self._st)


def philox4x64(ctr, key):
    c = (C.c_uint64 * 4)(*ctr)
    k = (C.c_uint64 * 2)(*key)
    o = (C.c_uint64 * 4)()
    lib().orc_philox4x64(c, k, o)
    return list(o)


def gen_uniform(seed, stream, system, n, scale=1.0):
    out = np.empty(n, dtype=np.float64)
    lib().orc_gen_uniform(seed, stream, system, n, scale, _p(out))
    return out


def gen_normal(seed, stream, system, n, scale=1.0):
    out = np.empty(n, dtype=np.float64)
    lib().orc_gen_normal(seed, stream, system, n, scale, _p(out))
    return out


def ref_seeds(master: int, npart: int):
    s = (C.c_long * npart)()
    lib().orc_ref_seeds(C.c_long(master), npart, s)
    return list(s)


def ref_matrix(master, npart, rows_per_part, cols):
    a = np.zeros((npart * rows_per_part, cols), dtype=np.float64, order="F")
    lib().orc_ref_matrix(C.c_long(master), npart, rows_per_part, cols, _p(a))
    return a


def ref_vector(master, npart, rows_per_part, offset):
    a = np.zeros(npart * rows_per_part, dtype=np.float64)
    lib().orc_ref_vector(C.c_long(master), npart, rows_per_part, offset, _p(a))
    return a


def ref_dense_blocks(npart, m, rank, offset, useeds, vseeds, dseeds):
    K = np.zeros((npart * m, m), dtype=np.float64, order="F")
    ar = lambda s: (C.c_long * npart)(*s)
    lib().orc_ref_dense_blocks(npart, m, rank, offset, ar(useeds), ar(vseeds), ar(dseeds),
                               _p(K), K.shape[0])
    return K


# ---------------------------------------------------------------- LAPACK
def dgemm(ta, tb, alpha, A, B, beta, Cm):
    m = Cm.shape[0]
    n = Cm.shape[1]
    k = A.shape[1] if ta in "nN" else A.shape[0]
    lib().orc_dgemm(ta.encode(), tb.encode(), m, n, k, alpha, _p(A), _ld(A), _p(B), _ld(B),
                    beta, _p(Cm), _ld(Cm))
    return Cm


def dgetrf(A):
    m, n = A.shape
    ipiv = (C.c_int * min(m, n))()
    info = lib().orc_dgetrf(m, n, _p(A), _ld(A), ipiv)
    return info, np.array(list(ipiv), dtype=np.int32)


def dgetrf2(A):
    m, n = A.shape
    ipiv = (C.c_int * min(m, n))()
    info = lib().orc_dgetrf2(m, n, _p(A), _ld(A), ipiv)
    return info, np.array(list(ipiv), dtype=np.int32)


def dgetrs(trans, A, ipiv, B):
    n = A.shape[0]
    ip = (C.c_int * n)(*[int(x) for x in ipiv])
    lib().orc_dgetrs(C.c_char(trans.encode()), n, B.shape[1], _p(A), _ld(A), ip, _p(B), _ld(B))
    return B


def dgesv(A, B):
    n = A.shape[0]
    ipiv = (C.c_int * n)()
    info = lib().orc_dgesv(n, B.shape[1], _p(A), _ld(A), ipiv, _p(B), _ld(B))
    return info, np.array(list(ipiv), dtype=np.int32)


def dtrsm_left(uplo, trans, diag, A, B):
    lib().orc_dtrsm_left(uplo.encode(), trans.encode(), diag.encode(), B.shape[0], B.shape[1],
                         _p(A), _ld(A), _p(B), _ld(B))
    return B


# ---------------------------------------------------------------- BLAS backend
def find_openblas():
    """Locate a Fortran-ABI OpenBLAS shipped with scipy (symbols scipy_*)."""
    try:
        import scipy
        base = os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs")
        cands = sorted(glob.glob(os.path.join(base, "libscipy_openblas*.so")))
        if cands:
            return cands[0], "scipy_"
    except Exception:
        pass
    return None, None


def use_openblas() -> str:
    path, prefix = find_openblas()
    if path is None:
        lib().orc_blas_use_netlib()
        return "netlib-restated"
    rc = lib().orc_blas_load(path.encode(), prefix.encode())
    if rc != 0:
        lib().orc_blas_use_netlib()
        return "netlib-restated"
    return "OpenBLAS (" + os.path.basename(path) + ")"


def use_netlib():
    lib().orc_blas_use_netlib()


# ---------------------------------------------------------------- HODLR
def hsolve(nrow, nrhs, rank, nPart, K, X, V):
    """Reference hsolve on F-ordered K (LD rows), X, V (same LD), in place."""
    assert K.shape[0] == X.shape[0] == V.shape[0]
    return lib().orc_hsolve(nrow, nrhs, rank, nPart, K.shape[0], _p(K), _p(X), _p(V))


def build_x(B, U, L):
    N, s = B.shape
    r = U.shape[1]
    X = np.zeros((N, s + r * L), dtype=np.float64, order="F")
    lib().orc_build_x(N, r, s, L, _p(B), _ld(B), _p(U), _ld(U), _p(X))
    return X


def hodlr_solve(K, U, V, B, leaf, task_level=None, nthreads=1, return_x=False):
    """The reference driver (solver_irregular_part.cc:104-140) on copies.

    Returns the solution (N x s); with return_x also the full column block
    X = [x | u_0 | ... | u_{L-1}] after the solve.
    """
    N = K.shape[0]
    L = int(round(np.log2(N // leaf)))
    assert leaf << L == N
    r = U.shape[1]
    s = B.shape[1]
    if task_level is None:
        task_level = min(L, max(0, int(np.floor(np.log2(max(1, nthreads))))))
    Kc = np.array(K, dtype=np.float64, order="F", copy=True)
    X = build_x(np.asfortranarray(B), np.asfortranarray(U), L)
    Vc = np.asfortranarray(V)
    info = lib().orc_hodlr_solve(N, leaf, L, r, s, task_level, nthreads, _p(Kc), _p(X), _p(Vc))
    if info != 0:
        raise RuntimeError(f"oracle hodlr_solve INFO={info}")
    if return_x:
        return X[:, :s].copy(order="F"), X
    return X[:, :s].copy(order="F")


def hodlr_matvec(K, U, V, x, leaf):
    N = K.shape[0]
    x = np.asfortranarray(x.reshape(N, -1))
    y = np.zeros_like(x, order="F")
    lib().orc_hodlr_matvec(N, leaf, U.shape[1], x.shape[1], _p(K), _ld(K), _p(U), _ld(U),
                           _p(V), _ld(V), _p(x), _ld(x), _p(y), _ld(y))
    return y


def dense_assemble(K, U, V, leaf):
    N = K.shape[0]
    A = np.zeros((N, N), dtype=np.float64, order="F")
    lib().orc_dense_assemble(N, leaf, U.shape[1], _p(K), _ld(K), _p(U), _ld(U), _p(V), _ld(V),
                             _p(A))
    return A


# ---------------------------------------------------------- uneven trees
# A tree of 2^L leaf slots of sizes leaf_sizes[i] (0 = empty slot): leaf i
# covers rows [row0_i, row0_i + m_i), its block K_i is columns [0, m_i) of
# those rows of K (N x max m), and every entry outside the leaf blocks is
# U(i,:) V(j,:)^T (global U, V), as for balanced trees.  The reference
# solves balanced trees only; the uneven tree is solved here by embedding it
# in a balanced tree of 2^L leaves of size M = max m whose leaf i is
# [K_i, 0; 0, I] with zero U, V and B rows below row m_i: that matrix is a
# permutation of blockdiag(A, I), so its solution restricted to the real
# rows is A^-1 B exactly.


def uneven_rows(leaf_sizes):
    return np.concatenate([[0], np.cumsum(leaf_sizes)]).astype(np.int64)


def embed_uneven(K, U, V, B, leaf_sizes):
    sizes = [int(x) for x in leaf_sizes]
    M = max(sizes)
    P = len(sizes)
    row0 = uneven_rows(sizes)
    r, s = U.shape[1], B.shape[1]
    Kp = np.zeros((P * M, M), order="F")
    Up = np.zeros((P * M, r), order="F")
    Vp = np.zeros((P * M, r), order="F")
    Bp = np.zeros((P * M, s), order="F")
    idx = []
    for i, m in enumerate(sizes):
        a, b = i * M, i * M + m
        Kp[a:b, :m] = K[row0[i]:row0[i] + m, :m]
        for j in range(m, M):
            Kp[a + j, j] = 1.0
        Up[a:b] = U[row0[i]:row0[i] + m]
        Vp[a:b] = V[row0[i]:row0[i] + m]
        Bp[a:b] = B[row0[i]:row0[i] + m]
        idx.append(np.arange(a, b))
    return Kp, Up, Vp, Bp, M, np.concatenate(idx)


def hodlr_solve_uneven(K, U, V, B, leaf_sizes, nthreads=1):
    """The restated reference solve on the balanced embedding of an uneven
    tree; returns the solution rows of the uneven system."""
    Kp, Up, Vp, Bp, M, idx = embed_uneven(K, U, V, B, leaf_sizes)
    xp = hodlr_solve(Kp, Up, Vp, Bp, M, nthreads=nthreads)
    return np.asfortranarray(xp[idx])


def dense_assemble_uneven(K, U, V, leaf_sizes):
    A = U @ V.T
    row0 = uneven_rows(leaf_sizes)
    for i, m in enumerate(leaf_sizes):
        a = row0[i]
        A[a:a + m, a:a + m] = K[a:a + m, :m]
    return np.asfortranarray(A)


def pair_solve(K, U, V, W, B, leaf_sizes, nthreads=1):
    """Two shards with ranks r0 != r1 (spmd_driver/solver_irregular_ranks.cc:
    146-251): per shard c, [d_c | Y_c] = A_c^-1 [B_c | U_c] (the leaf task and
    in-machine level loop, restated solve), T_c = W_c^T [d_c | Y_c]
    (LMatrix::gemm, :181-182), the (r0+r1)^2 root core by NodeSolveRegionTask
    (:218-220, node_solve_region), and x_c = d_c - Y_c eta_c
    (LMatrix::gemm_inplace, :250-251).  K, U, V, W are pairs (shard 0,
    shard 1), leaf_sizes the shards' leaf-slot sizes; returns x."""
    s = B.shape[1]
    n0 = U[0].shape[0]
    rows = (slice(0, n0), slice(n0, n0 + U[1].shape[0]))
    dY, T = [], []
    for c in (0, 1):
        Xc = np.asfortranarray(np.hstack([B[rows[c]], U[c]]))
        Xc = hodlr_solve_uneven(K[c], U[c], V[c], Xc, leaf_sizes[c], nthreads=nthreads)
        dY.append(Xc)
        T.append(gemm_red(np.asfortranarray(W[c]), Xc, 1, 1))
    VTd0, VTu0 = np.asfortranarray(T[0][:, :s]), np.asfortranarray(T[0][:, s:])
    VTd1, VTu1 = np.asfortranarray(T[1][:, :s]), np.asfortranarray(T[1][:, s:])
    info = node_solve_region(VTu0, VTu1, VTd0, VTd1)
    if info:
        raise RuntimeError(f"node_solve_region INFO={info}")
    eta = (VTd1, VTd0)
    out = []
    for c in (0, 1):
        d = np.asfortranarray(dY[c][:, :s])
        Y = np.asfortranarray(dY[c][:, s:])
        out.append(gemm_bro(Y, eta[c], d, 1, 1, alpha=-1.0))
    return np.asfortranarray(np.vstack(out))


def dense_assemble_pair(K, U, V, W, leaf_sizes):
    A0 = dense_assemble_uneven(K[0], U[0], V[0], leaf_sizes[0])
    A1 = dense_assemble_uneven(K[1], U[1], V[1], leaf_sizes[1])
    return np.asfortranarray(np.block([[A0, U[0] @ W[1].T], [U[1] @ W[0].T, A1]]))


def gemm_red(A, B, nchunk, nslot, alpha=1.0):
    rows, ka = A.shape
    kb = B.shape[1]
    Cm = np.zeros((nslot * ka, kb), dtype=np.float64, order="F")
    lib().orc_gemm_red(rows, nchunk, nslot, alpha, _p(A), _ld(A), ka, _p(B), _ld(B), kb,
                       _p(Cm), _ld(Cm))
    return Cm


def gemm_bro(u, eta, d, nchunk, nslot, alpha=-1.0):
    rows, k = u.shape
    lib().orc_gemm_bro(rows, nchunk, nslot, alpha, _p(u), _ld(u), k, _p(eta), _ld(eta), _p(d),
                       _ld(d), d.shape[1])
    return d


def node_solve(VTu, VTd, r):
    nnode = VTu.shape[0] // (2 * r)
    info = lib().orc_node_solve(nnode, r, _p(VTu), _ld(VTu), _p(VTd), _ld(VTd), VTd.shape[1])
    return info


def node_solve_region(VTu0, VTu1, VTd0, VTd1):
    r1, r0 = VTu0.shape
    return lib().orc_node_solve_region(r0, r1, VTd0.shape[1], _p(VTu0), _ld(VTu0), _p(VTu1),
                                       _ld(VTu1), _p(VTd0), _ld(VTd0), _p(VTd1), _ld(VTd1))


# ---------------------------------------------------------------- workloads
def workload_gaussian(N, leaf, r, s, seed=0, system=0, shift=None, rhs="uniform"):
    """Host restatement of the Gaussian workload generator (Philox streams
    0..3 = K, U, V, B), identical in definition to the product's device
    generator: K_i = G_i + shift*I, U,V ~ N(0,1)/sqrt(r), B ~ U[0,1) or N(0,1)."""
    shift = float(leaf) if shift is None else float(shift)
    K = gen_normal(seed, 0, system, N * leaf).reshape((N, leaf), order="F")
    K = np.asfortranarray(K)
    for i in range(leaf):
        K[np.arange(i, N, leaf), i] += shift
    sc = 1.0 / np.sqrt(float(r))
    U = np.asfortranarray(gen_normal(seed, 1, system, N * r, sc).reshape((N, r), order="F"))
    V = np.asfortranarray(gen_normal(seed, 2, system, N * r, sc).reshape((N, r), order="F"))
    gb = gen_uniform if rhs == "uniform" else gen_normal
    B = np.asfortranarray(gb(seed, 3, system, N * s).reshape((N, s), order="F"))
    return K, U, V, B
