"""Pins the CPU oracle (oracle/) against independent references.

Every oracle function is checked against something that is not the oracle:
glibc, numpy/OpenBLAS, brute force, closed forms, or the reference's own
known-answer relations (tests/test.cc of the reference).  These run on the
CPU (no GPU marker).
"""
import ctypes
import ctypes.util
import itertools

import numpy as np
import pytest


def F(a):
    return np.asfortranarray(a, dtype=np.float64)


# ------------------------------------------------------------ drand48
def test_drand48_matches_glibc(orc):
    """orc_drand48/lrand48 vs glibc srand48_r/drand48_r/lrand48_r (the
    generator behind PtrMatrix::rand, src/ptr_matrix.cc:62-71)."""
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    buf = ctypes.create_string_buffer(64)  # struct drand48_data
    for seed in (0, 1, 12345, 2**31 - 1, 987654321987):
        libc.srand48_r(ctypes.c_long(seed), buf)
        st = orc.Rand48(seed)
        for _ in range(50):
            d = ctypes.c_double()
            libc.drand48_r(buf, ctypes.byref(d))
            assert st.drand48() == d.value
        libc.srand48_r(ctypes.c_long(seed), buf)
        st = orc.Rand48(seed)
        for _ in range(20):
            v = ctypes.c_long()
            libc.lrand48_r(buf, ctypes.byref(v))
            assert st.lrand48() == v.value


def test_ref_matrix_fill_order(orc):
    """Matrix::rand fills each partition rows-outer / cols-inner from its
    own seed (matrix.cc:266-297, ptr_matrix.cc:62-71)."""
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    buf = ctypes.create_string_buffer(64)
    npart, rpp, cols = 4, 5, 3
    a = orc.ref_matrix(77, npart, rpp, cols)
    seeds = orc.ref_seeds(77, npart)
    for k in range(npart):
        libc.srand48_r(ctypes.c_long(seeds[k]), buf)
        for i in range(rpp):
            for j in range(cols):
                d = ctypes.c_double()
                libc.drand48_r(buf, ctypes.byref(d))
                assert a[k * rpp + i, j] == d.value


def test_ref_dense_blocks_kat(orc):
    """test_lmatrix_init KAT (tests/test.cc:249-262): a dense block equals
    (U V^T + diag(D))(block)."""
    npart, m, r, off = 4, 6, 2, 100
    U = orc.ref_matrix(11, npart, m, r)
    V = orc.ref_matrix(12, npart, m, r)
    D = orc.ref_vector(13, npart, m, off)
    K = orc.ref_dense_blocks(npart, m, r, off, orc.ref_seeds(11, npart), orc.ref_seeds(12, npart),
                             orc.ref_seeds(13, npart))
    A = U @ V.T + np.diag(D)
    for p in range(npart):
        blk = A[p * m:(p + 1) * m, p * m:(p + 1) * m]
        assert np.linalg.norm(K[p * m:(p + 1) * m, :] - blk) < 1e-13


# ------------------------------------------------------------ philox
def test_philox_kat_and_numpy(orc):
    """Philox4x64-10: the published all-zero known-answer vector, and
    numpy.random.Philox raw output (numpy increments the counter before
    each block)."""
    assert orc.philox4x64([0, 0, 0, 0], [0, 0]) == [
        0x16554d9eca36314c, 0xdb20fe9d672d0fdc, 0xd7e772cee186176b, 0x7e68b68aec7ba23b]
    for key, ctr in [((7, 0), (5, 1, 2, 0)), ((2**64 - 1, 3), (1, 0, 9, 4)),
                     ((0, 0), (2**40, 3, 0, 0))]:
        c = sum(v << (64 * i) for i, v in enumerate(ctr)) - 1
        c %= 2**256
        ctr_np = np.array([(c >> (64 * i)) & (2**64 - 1) for i in range(4)], dtype=np.uint64)
        g = np.random.Philox(key=np.array(key, dtype=np.uint64), counter=ctr_np)
        assert list(map(int, g.random_raw(4))) == orc.philox4x64(list(ctr), list(key))


def test_generators_definition(orc):
    """uniform = (w >> 11) 2^-53 and Box-Muller normals, from raw Philox words."""
    seed, stream, system = 5, 3, 2
    u = orc.gen_uniform(seed, stream, system, 10, 2.0)
    n = orc.gen_normal(seed, stream, system, 10, 0.5)
    for e in range(10):
        w = orc.philox4x64([e // 4, stream, system, 0], [seed, 0])
        assert u[e] == (w[e % 4] >> 11) * 2.0**-53 * 2.0
        pr = (e % 4) // 2
        u1 = ((w[2 * pr] >> 11) + 1) * 2.0**-53
        u2 = (w[2 * pr + 1] >> 11) * 2.0**-53
        rad = np.sqrt(-2.0 * np.log(u1))
        z = rad * (np.cos(2 * np.pi * u2) if e % 2 == 0 else np.sin(2 * np.pi * u2))
        assert abs(n[e] - 0.5 * z) <= 1e-15 * max(1.0, abs(z))
    big = orc.gen_normal(1, 0, 0, 200000)
    assert abs(big.mean()) < 0.01 and abs(big.std() - 1.0) < 0.01


# ------------------------------------------------------------ LAPACK restatements
@pytest.mark.parametrize("ta,tb", list(itertools.product("NT", "NT")))
def test_dgemm_vs_numpy(orc, ta, tb):
    rng = np.random.default_rng(0)
    m, n, k = 7, 5, 9
    A = F(rng.standard_normal((m, k) if ta == "N" else (k, m)))
    B = F(rng.standard_normal((k, n) if tb == "N" else (n, k)))
    C0 = F(rng.standard_normal((m, n)))
    Cm = C0.copy(order="F")
    orc.dgemm(ta, tb, 0.7, A, B, -1.3, Cm)
    opA = A if ta == "N" else A.T
    opB = B if tb == "N" else B.T
    assert np.allclose(Cm, 0.7 * opA @ opB - 1.3 * C0, rtol=1e-14, atol=1e-14)


def _brute_force_lu(A):
    """Textbook Gaussian elimination with partial pivoting (first max)."""
    A = A.copy()
    n = A.shape[0]
    piv = []
    for j in range(n):
        p = j + int(np.argmax(np.abs(A[j:, j])))
        piv.append(p + 1)
        A[[j, p]] = A[[p, j]]
        if A[j, j] != 0:
            A[j + 1:, j] /= A[j, j]
            A[j + 1:, j + 1:] -= np.outer(A[j + 1:, j], A[j, j + 1:])
    return A, np.array(piv)


@pytest.mark.parametrize("n", [1, 2, 3, 8, 65, 130])
def test_dgetrf_pivots_and_factors(orc, n):
    rng = np.random.default_rng(n)
    A = F(rng.standard_normal((n, n)))
    LU = A.copy(order="F")
    info, ipiv = orc.dgetrf(LU)
    assert info == 0
    LU2, piv2 = _brute_force_lu(np.array(A))
    assert np.array_equal(ipiv, piv2)
    assert np.allclose(LU, LU2, rtol=1e-10, atol=1e-12)
    # P L U == A
    L = np.tril(LU, -1) + np.eye(n)
    Uu = np.triu(LU)
    PA = A.copy()
    for i, p in enumerate(ipiv):
        PA[[i, p - 1]] = PA[[p - 1, i]]
    assert np.allclose(L @ Uu, PA, atol=1e-12)


def test_dgetrf2_equals_dgetrf_small(orc):
    rng = np.random.default_rng(3)
    A = F(rng.standard_normal((40, 40)))
    a1, a2 = A.copy(order="F"), A.copy(order="F")
    i1, p1 = orc.dgetrf(a1)
    i2, p2 = orc.dgetrf2(a2)
    assert i1 == i2 == 0 and np.array_equal(p1, p2)
    assert np.allclose(a1, a2, atol=1e-13)


def test_dgesv_and_singular(orc):
    rng = np.random.default_rng(5)
    A = F(rng.standard_normal((50, 50)))
    B = F(rng.standard_normal((50, 3)))
    X = B.copy(order="F")
    info, _ = orc.dgesv(A.copy(order="F"), X)
    assert info == 0
    assert np.allclose(X, np.linalg.solve(A, B), rtol=1e-11, atol=1e-12)
    S = A.copy(order="F")
    S[:, 7] = 0.0                       # exactly singular column -> INFO = 8
    info, _ = orc.dgesv(S, B.copy(order="F"))
    assert info == 8


def test_dgetrs_transposed(orc):
    rng = np.random.default_rng(6)
    A = F(rng.standard_normal((30, 30)))
    B = F(rng.standard_normal((30, 2)))
    LU = A.copy(order="F")
    info, ipiv = orc.dgetrf(LU)
    X = orc.dgetrs("T", LU, ipiv, B.copy(order="F"))
    assert np.allclose(A.T @ X, B, atol=1e-12)


# ------------------------------------------------------------ HODLR restatement
def _problem(orc, N, m, r, s, seed=0, shift=None):
    return orc.workload_gaussian(N, m, r, s, seed=seed, shift=shift)


@pytest.mark.parametrize("N,m,r,s", [(64, 16, 4, 1), (512, 64, 8, 3), (2048, 128, 16, 2),
                                     (1024, 128, 30, 1), (768, 96, 5, 4)])
def test_hodlr_solve_vs_dense_lu(orc, N, m, r, s):
    """Oracle driver vs numpy dense LU of the assembled matrix, 1e-10 rel."""
    K, U, V, B = _problem(orc, N, m, r, s, seed=N + r)
    A = orc.dense_assemble(K, U, V, m)
    xd = np.linalg.solve(A, B)
    for tl, nt in [(0, 1), (min(2, int(np.log2(N // m))), 4)]:
        x = orc.hodlr_solve(K, U, V, B, m, task_level=tl, nthreads=nt)
        assert np.abs(x - xd).max() <= 1e-10 * np.abs(xd).max()


def test_hodlr_special_cases(orc):
    rng = np.random.default_rng(9)
    # L = 0: a single dgesv
    m = 32
    K = F(rng.standard_normal((m, m)) + 10 * np.eye(m))
    U = F(rng.standard_normal((m, 3)))
    V = F(rng.standard_normal((m, 3)))
    B = F(rng.standard_normal((m, 2)))
    x = orc.hodlr_solve(K, U, V, B, m)
    assert np.allclose(x, np.linalg.solve(K, B), rtol=1e-12, atol=1e-13)
    # U = 0: block diagonal
    N = 256
    K, U, V, B = _problem(orc, N, 32, 4, 2, seed=1)
    U0 = F(np.zeros_like(U))
    x = orc.hodlr_solve(K, U0, V, B, 32)
    for p in range(N // 32):
        sl = slice(p * 32, (p + 1) * 32)
        assert np.allclose(x[sl], np.linalg.solve(K[sl], B[sl]), rtol=1e-11, atol=1e-13)
    # L = 1: Sherman-Morrison-Woodbury closed form
    K, U, V, B = _problem(orc, 128, 64, 5, 1, seed=2)
    A0 = np.zeros((128, 128))
    A0[:64, :64] = K[:64]
    A0[64:, 64:] = K[64:]
    Z = np.zeros((128, 10))
    W = np.zeros((128, 10))
    Z[:64, :5] = U[:64]
    W[64:, :5] = V[64:]
    Z[64:, 5:] = U[64:]
    W[:64, 5:] = V[:64]
    Ai = np.linalg.inv(A0)
    xs = Ai @ B - Ai @ Z @ np.linalg.solve(np.eye(10) + W.T @ Ai @ Z, W.T @ Ai @ B)
    x = orc.hodlr_solve(K, U, V, B, 64)
    assert np.abs(x - xs).max() <= 1e-11 * np.abs(xs).max()


def test_reference_generator_is_uvt_plus_d(orc):
    """With the reference generator, A = U V^T + diag(D) exactly, and the
    oracle solve matches a dense solve of that matrix."""
    npart, m, r, off = 8, 16, 3, 1000
    N = npart * m
    U = orc.ref_matrix(21, npart, m, r)
    V = orc.ref_matrix(22, npart, m, r)
    D = orc.ref_vector(23, npart, m, off)
    K = orc.ref_dense_blocks(npart, m, r, off, orc.ref_seeds(21, npart), orc.ref_seeds(22, npart),
                             orc.ref_seeds(23, npart))
    A = orc.dense_assemble(K, U, V, m)
    assert np.abs(A - (U @ V.T + np.diag(D))).max() < 1e-12
    b = F(orc.ref_matrix(24, npart, m, 1))
    x = orc.hodlr_solve(K, U, V, b, m)
    assert np.abs(x - np.linalg.solve(A, b)).max() < 1e-12


def test_hsolve_u_blocks_are_level_solves(orc):
    """After the solve, X's u_t block holds Y_t = (A restricted to the level-t
    children)^-1 U_c -- the quantity the device plan stores per level."""
    N, m, r, s = 512, 64, 4, 1
    K, U, V, B = _problem(orc, N, m, r, s, seed=3)
    L = 3
    x, X = orc.hodlr_solve(K, U, V, B, m, return_x=True)
    A = orc.dense_assemble(K, U, V, m)
    for t in range(L):
        nc = N >> (t + 1)
        Yt = X[:, s + r * t: s + r * (t + 1)]
        for c in range(2 ** (t + 1)):
            sl = slice(c * nc, (c + 1) * nc)
            ref = np.linalg.solve(A[sl, sl], U[sl])
            assert np.abs(Yt[sl] - ref).max() <= 1e-11 * np.abs(ref).max()


def test_gemm_reduce_kat(orc):
    """test_gemm_reduce (tests/test.cc:316-344): W = [V_top^T U_top; V_bot^T U_bot]."""
    rng = np.random.default_rng(1)
    m, n = 16, 3
    U = F(rng.random((m, n)))
    V = F(rng.random((m, n)))
    W = orc.gemm_red(V, U, nchunk=4, nslot=2)
    assert np.linalg.norm(W[:n] - V[:8].T @ U[:8]) < 1e-13
    assert np.linalg.norm(W[n:] - V[8:].T @ U[8:]) < 1e-13


def test_gemm_broadcast_kat(orc):
    """test_gemm_broadcast (tests/test.cc:346-372) with alpha = 1, D = 0."""
    rng = np.random.default_rng(2)
    m, n = 16, 3
    U = F(rng.random((m, n)))
    E = F(rng.random((2 * n, n)))
    W = F(np.zeros((m, n)))
    orc.gemm_bro(U, E, W, nchunk=8, nslot=2, alpha=1.0)
    assert np.linalg.norm(W[:8] - U[:8] @ E[:n]) < 1e-13
    assert np.linalg.norm(W[8:] - U[8:] @ E[n:]) < 1e-13


def test_node_solve_kat(orc):
    """test_node_solve (tests/test.cc:374-425): explicit S assembly, swapped
    RHS, dense solve."""
    rng = np.random.default_rng(4)
    r, level = 3, 2
    VTu = F(rng.random((level * 2 * r, r)))
    VTd = F(rng.random((level * 2 * r, 2)))
    expect = []
    for l in range(level):
        o = l * 2 * r
        S = np.eye(2 * r)
        S[r:, :r] = VTu[o:o + r]
        S[:r, r:] = VTu[o + r:o + 2 * r]
        rhs = np.vstack([VTd[o + r:o + 2 * r], VTd[o:o + r]])
        expect.append(np.linalg.solve(S, rhs))
    got = VTd.copy(order="F")
    assert orc.node_solve(VTu, got, r) == 0
    for l in range(level):
        assert np.linalg.norm(got[l * 2 * r:(l + 1) * 2 * r] - expect[l]) < 1e-13


def test_node_solve_region_kat(orc):
    """NodeSolveRegionTask (node_solve_region.cc:77-107), r0 != r1."""
    rng = np.random.default_rng(8)
    r0, r1, s = 3, 5, 2
    VTu0 = F(rng.random((r1, r0)))
    VTu1 = F(rng.random((r0, r1)))
    VTd0 = F(rng.random((r1, s)))
    VTd1 = F(rng.random((r0, s)))
    S = np.eye(r0 + r1)
    S[:r0, r0:] = VTu1
    S[r0:, :r0] = VTu0
    sol = np.linalg.solve(S, np.vstack([VTd1, VTd0]))
    d0, d1 = VTd0.copy(order="F"), VTd1.copy(order="F")
    assert orc.node_solve_region(VTu0, VTu1, d0, d1) == 0
    assert np.allclose(d1, sol[:r0], atol=1e-13) and np.allclose(d0, sol[r0:], atol=1e-13)


def test_matvec_vs_dense(orc):
    K, U, V, B = _problem(orc, 512, 64, 6, 3, seed=7)
    A = orc.dense_assemble(K, U, V, 64)
    y = orc.hodlr_matvec(K, U, V, B, 64)
    assert np.abs(y - A @ B).max() < 1e-12 * np.abs(A @ B).max()


def test_singular_leaf_reports_info(orc):
    K, U, V, B = _problem(orc, 256, 32, 4, 1, seed=8)
    K[32 * 3:32 * 4, 5] = 0.0            # leaf 3, column 5 exactly zero
    with pytest.raises(RuntimeError, match="INFO=6"):
        orc.hodlr_solve(K, U, V, B, 32)


def test_openblas_backend_agrees(orc):
    """The CPU baseline's BLAS backend (dlopen'ed OpenBLAS) gives the same
    solution as the restated netlib routines."""
    K, U, V, B = _problem(orc, 1024, 64, 8, 2, seed=11)
    orc.use_netlib()
    x1 = orc.hodlr_solve(K, U, V, B, 64, task_level=2, nthreads=4)
    name = orc.use_openblas()
    try:
        x2 = orc.hodlr_solve(K, U, V, B, 64, task_level=2, nthreads=4)
    finally:
        orc.use_netlib()
    assert np.abs(x1 - x2).max() <= 1e-12 * np.abs(x1).max(), name


def test_uneven_embedding_matches_dense(orc):
    """The balanced embedding of an uneven tree (hodlr_solve_uneven: every
    leaf padded to the largest with an identity block and zero U, V, B rows)
    solves the uneven system: against a dense solve of the assembled matrix,
    including empty leaf slots; and the dense assembly of a uniform tree
    equals the balanced one."""
    rng = np.random.default_rng(17)
    for sizes, r, s in (([40, 40, 400, 40], 6, 2), ([64, 0, 96, 32, 0, 0, 80, 48], 5, 3),
                        ([50] * 5 + [0] * 3, 4, 1)):
        N, M = sum(sizes), max(sizes)
        K = np.asfortranarray(rng.standard_normal((N, M)))
        row0 = orc.uneven_rows(sizes)
        for i, m in enumerate(sizes):
            K[row0[i] + np.arange(m), np.arange(m)] += m
        U = np.asfortranarray(rng.standard_normal((N, r)) / np.sqrt(r))
        V = np.asfortranarray(rng.standard_normal((N, r)) / np.sqrt(r))
        B = np.asfortranarray(rng.random((N, s)))
        A = orc.dense_assemble_uneven(K, U, V, sizes)
        x = orc.hodlr_solve_uneven(K, U, V, B, sizes)
        xd = np.linalg.solve(A, B)
        assert np.abs(x - xd).max() <= 1e-12 * np.abs(xd).max(), sizes
    K, U, V, B = _problem(orc, 512, 64, 6, 2, seed=19)
    Au, Ab = orc.dense_assemble_uneven(K, U, V, [64] * 8), orc.dense_assemble(K, U, V, 64)
    assert np.abs(Au - Ab).max() <= 1e-15 * np.abs(Ab).max()   # U V^T summation order only


def test_pair_solve_matches_dense(orc):
    """The restated two-shard driver with ranks r0 != r1
    (solver_irregular_ranks.cc:146-251) solves A = [A_0, U_0 W_1^T;
    U_1 W_0^T, A_1]: against a dense solve, shards balanced and uneven."""
    rng = np.random.default_rng(23)
    for sizes, ranks, s in ((([48] * 4, [48] * 4), (6, 11), 2),
                            (([40, 400, 40, 40], [64, 0, 64, 64]), (9, 4), 3)):
        K, U, V, W, Bs = [], [], [], [], []
        for c in (0, 1):
            sz, r = sizes[c], ranks[c]
            N, M = sum(sz), max(sz)
            Kc = np.asfortranarray(rng.standard_normal((N, M)))
            row0 = orc.uneven_rows(sz)
            for i, m in enumerate(sz):
                Kc[row0[i] + np.arange(m), np.arange(m)] += m
            K.append(Kc)
            U.append(np.asfortranarray(rng.standard_normal((N, r)) / np.sqrt(r)))
            V.append(np.asfortranarray(rng.standard_normal((N, r)) / np.sqrt(r)))
            W.append(np.asfortranarray(rng.standard_normal((N, ranks[1 - c])) / np.sqrt(r)))
            Bs.append(rng.random((N, s)))
        B = np.asfortranarray(np.vstack(Bs))
        x = orc.pair_solve(K, U, V, W, B, sizes)
        A = orc.dense_assemble_pair(K, U, V, W, sizes)
        xd = np.linalg.solve(A, B)
        assert np.abs(x - xd).max() <= 1e-12 * np.abs(xd).max(), ranks
