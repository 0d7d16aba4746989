"""Parity of the HIP path with the CPU oracle (gpu marker).

Every test calls the device library through the C ABI (via the ctypes layer
in fastsolver2_amd) and compares with the oracle -- the restated reference
solve (oracle/hodlr_oracle.c) -- on the same inputs.  FP64 tolerance: the
solution matches the oracle to 1e-10 relative (max-norm), the bar the
reference comparison is held to.
"""
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_dev(F, a):
    a = np.asfortranarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a.reshape(-1, 1, order="F")
    t = F.fortran_empty(a.shape[0], a.shape[1])
    t.copy_(torch.from_numpy(a))
    return t


def to_host(t):
    torch.cuda.synchronize()
    return np.asfortranarray(t.cpu().numpy())


def relerr(x, ref):
    return float(np.abs(x - ref).max() / max(np.abs(ref).max(), 1e-300))


@pytest.fixture(scope="module", autouse=True)
def _blas(orc):
    orc.use_openblas()
    yield
    orc.use_netlib()


# ------------------------------------------------------------------ MFMA layout
def test_mfma_f64_lane_layout(gpu):
    """The FP64 matrix-core accumulator layout every kernel relies on."""
    say("case: mfma f64 16x16x4 lane layout")
    F = gpu
    out = torch.zeros(256, dtype=torch.float64, device="cuda")
    assert F.lib().hodlr_selftest_mfma(F._ptr(out), F._stream()) == 0
    got = out.cpu().numpy()
    for l in range(64):
        for v in range(4):
            row, col = l // 16 + 4 * v, l % 16
            assert got[4 * l + v] == row + 100 * (col % 4), (l, v, got[4 * l + v])


# ------------------------------------------------------------------ generators
def test_device_generator_matches_oracle(gpu, orc):
    say("case: generator uniform/normal 1000x7")
    F = gpu
    u = to_host(F.generate(9, 3, 1, 0, 1000, 7, 1.5))
    n = to_host(F.generate(9, 4, 1, 1, 1000, 7, 0.25))
    uo = orc.gen_uniform(9, 3, 1, 7000, 1.5).reshape((1000, 7), order="F")
    no = orc.gen_normal(9, 4, 1, 7000, 0.25).reshape((1000, 7), order="F")
    assert np.array_equal(u, uo)                       # integer-exact stream
    assert np.abs(n - no).max() <= 4e-16 * np.abs(no).max() * 8


def test_workload_generator_definition(gpu, orc):
    say("case: workload_gaussian N=512 m=64 r=8 s=3")
    F = gpu
    K, U, V, B = F.workload_gaussian(512, 64, 8, 3, seed=4)
    Ko, Uo, Vo, Bo = orc.workload_gaussian(512, 64, 8, 3, seed=4)
    for d, o in ((K, Ko), (U, Uo), (V, Vo), (B, Bo)):
        h = to_host(d)
        assert h.shape == o.shape
        assert np.abs(h - o).max() <= 1e-14 * max(1.0, np.abs(o).max())


# ------------------------------------------------------------------ batched LAPACK
@pytest.mark.parametrize("n", [1, 3, 16, 31, 32, 33, 64, 100, 128, 257, 512, 600, 1000])
def test_dgesv_batched_pivoting(gpu, orc, n):
    """Batched LU with partial pivoting on unshifted Gaussian blocks (pivots
    active) vs numpy, and pivot vector vs the restated dgetrf."""
    say(f"case: dgesv_batched n={n}")
    F = gpu
    batch, nrhs = 3, 5
    rng = np.random.default_rng(n)
    As = [rng.standard_normal((n, n)) for _ in range(batch)]
    Bs = [rng.standard_normal((n, nrhs)) for _ in range(batch)]
    A = to_dev(F, np.hstack(As))
    B = to_dev(F, np.hstack(Bs))
    info, ipiv = F.dgesv_batched(A, B, n, batch)
    assert not info.any()
    X = to_host(B)
    LU = to_host(A)
    ipiv = ipiv.cpu().numpy().reshape(batch, n)
    for b in range(batch):
        xb = X[:, b * nrhs:(b + 1) * nrhs]
        ref = np.linalg.solve(As[b], Bs[b])
        cond = np.linalg.cond(As[b])
        assert relerr(xb, ref) <= 1e-15 * cond * n + 1e-13, (b, cond)
        if n <= 600:  # every LU kernel: the w64 cores, the register panel, the LDS panel
            LUo = np.asfortranarray(As[b].copy())
            _, po = orc.dgetrf(LUo)
            assert np.array_equal(ipiv[b], po)
            assert np.abs(LU[:, b * n:(b + 1) * n] - LUo).max() <= 1e-11 * np.abs(LUo).max()


def _oop_vs_netlib(F, orc, blocks, n):
    """Out-of-place batched LU of `blocks` against the restated netlib dgetrf:
    identical IPIV and INFO, factors within 1e-13 (relative to the largest
    factor entry), K untouched."""
    batch = len(blocks)
    Kh = np.vstack(blocks)
    K = to_dev(F, Kh)
    LU, ipiv, info = F.dgetrf_batched_oop(K, n, batch)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(K), Kh)               # K is only read
    LU = LU.cpu().numpy()
    ipiv = ipiv.cpu().numpy()
    info = info.cpu().numpy()
    out = []
    for b in range(batch):
        LUo = np.asfortranarray(blocks[b].copy())
        io, po = orc.dgetrf(LUo)
        assert info[b] == io, (b, info[b], io)
        assert np.array_equal(ipiv[b], po), (b, np.nonzero(ipiv[b] != po)[0][:5])
        got = LU[b, :, :n].T
        if io == 0:
            assert np.abs(got - LUo).max() <= 1e-13 * np.abs(LUo).max(), b
        out.append(po)
    return out


@pytest.mark.parametrize("n", [128, 256, 384, 512])
def test_leaf_lu_fast_path_matches_netlib(gpu, orc, n):
    """Diagonally dominant blocks G + n I: dgetrf keeps every diagonal pivot
    and the no-interchange path runs; IPIV must be the identity exactly as
    netlib's, the factors equal to rounding."""
    say(f"case: leaf LU fast path n={n}")
    F = gpu
    rng = np.random.default_rng(n)
    blocks = [rng.standard_normal((n, n)) + n * np.eye(n) for _ in range(5)]
    for po in _oop_vs_netlib(F, orc, blocks, n):
        assert np.array_equal(po, np.arange(1, n + 1))


@pytest.mark.parametrize("n", [256, 512])
def test_leaf_lu_late_pivot_and_ties(gpu, orc, n):
    """Blocks that need an interchange only at a late column (rows j and
    j + 5 of a dominant block swapped: dgetrf swaps them back at column j),
    an exact tie |a_30| = |a_00| at column 0 (dgetrf keeps the diagonal: the
    first of equal magnitudes), a near tie inside the pivot margin and one
    outside it, a Gaussian block, and a block with a zero column: pivots and
    INFO equal the restated netlib dgetrf's in every case."""
    say(f"case: leaf LU pivots n={n}")
    F = gpu
    rng = np.random.default_rng(100 + n)
    blocks = []
    for j in (1, 70, n - 10, n - 6):
        d = rng.standard_normal((n, n)) + n * np.eye(n)
        d[[j, j + 5]] = d[[j + 5, j]]
        blocks.append(d)
    for f in (1.0, 1.0 - 1e-9, 1.0 - 1e-3):
        d = rng.standard_normal((n, n)) + n * np.eye(n)
        d[3, 0] = -f * d[0, 0]
        blocks.append(d)
    blocks.append(rng.standard_normal((n, n)))
    d = rng.standard_normal((n, n)) + n * np.eye(n)
    d[:, n // 2 + 3] = 0.0
    blocks.append(d)
    pos = _oop_vs_netlib(F, orc, blocks, n)
    for k, j in enumerate((1, 70, n - 10, n - 6)):
        want = np.arange(1, n + 1)
        want[j] = j + 6
        assert np.array_equal(pos[k], want), j


@pytest.mark.parametrize("n,lda", [(300, (1 << 22) + 64), (40, (1 << 24) + 64)])
def test_dgesv_large_leading_dimension(gpu, n, lda):
    """Leading dimensions past the 32-bit byte-offset limits of the fast
    paths (the LDS-staged trailing update needs lda <= 2^22, the register
    solve 32 * lda * 8 < 2^32) take the 64-bit-address paths; their results
    must match numpy (the operand is a strided view into a large buffer)."""
    say(f"case: dgesv_batched n={n} lda={lda}")
    F = gpu
    rng = np.random.default_rng(n)
    Ah = rng.standard_normal((n, n))
    Bh = rng.standard_normal((n, 3))
    buf = torch.empty(lda * (n - 1) + n, dtype=torch.float64, device="cuda")
    A = torch.as_strided(buf, (n, n), (1, lda))
    A.copy_(torch.from_numpy(np.asfortranarray(Ah)))
    B = to_dev(F, Bh)
    info, ipiv = F.dgesv_batched(A, B, n, 1)
    assert not info.any()
    X = to_host(B)
    ref = np.linalg.solve(Ah, Bh)
    assert relerr(X, ref) <= 1e-15 * np.linalg.cond(Ah) * n + 1e-13
    del A, buf
    torch.cuda.empty_cache()


def test_dgesv_batched_singular(gpu):
    say("case: dgesv_batched singular")
    F = gpu
    n = 40
    rng = np.random.default_rng(1)
    A0 = rng.standard_normal((n, n))
    A1 = rng.standard_normal((n, n))
    A1[:, 9] = 0.0
    A = to_dev(F, np.hstack([A0, A1]))
    B = to_dev(F, rng.standard_normal((n, 2)))
    info, _ = F.dgesv_batched(A, B, n, 2)
    assert info.tolist() == [0, 10]        # INFO 10 in batch item 1


def test_dgesv_batched_singular_item_past_2048(gpu):
    """Per-item INFO for a large batch: a singular item at index 2050 and
    one at the last index are reported in their own slots."""
    say("case: dgesv_batched singular items 2050 and 2999 of 3000")
    F = gpu
    n, batch = 8, 3000
    rng = np.random.default_rng(7)
    As = rng.standard_normal((batch, n, n))
    As[2050][:, 3] = 0.0
    As[2999][:, 0] = 0.0
    A = to_dev(F, np.hstack(list(As)))
    B = to_dev(F, rng.standard_normal((n, batch)))
    info, _ = F.dgesv_batched(A, B, n, batch)
    info = info.cpu().numpy()
    assert info[2050] == 4 and info[2999] == 1
    assert np.count_nonzero(info) == 2


# ------------------------------------------------------------------ full solve
SHAPES = [
    # N, leaf, rank, nrhs
    (256, 256, 8, 2),        # L = 0: a single dgesv
    (512, 64, 8, 3),
    (2048, 128, 16, 2),
    (800, 100, 7, 5),        # leaf not a multiple of 16 / 32
    (1600, 400, 20, 1),      # the drivers' default leaf size 400
    (16384, 256, 30, 1),     # config 1 of the benchmark set
    (8192, 512, 64, 8),
    (4096, 64, 50, 3),       # 2r = 100
]


@pytest.mark.parametrize("N,m,r,s", SHAPES)
def test_factor_solve_matches_oracle(gpu, orc, N, m, r, s):
    say(f"case: factor_solve N={N} m={m} r={r} s={s}")
    F = gpu
    L = int(np.log2(N // m))
    K, U, V, B = F.workload_gaussian(N, m, r, s, seed=N + r)
    Kh, Uh, Vh, Bh = to_host(K), to_host(U), to_host(V), to_host(B)
    xo = orc.hodlr_solve(Kh, Uh, Vh, Bh, m, task_level=min(L, 3), nthreads=8)
    plan = F.HodlrPlan(N, m, L, r, nrhs_max=s)
    plan.factor_solve(K, U, V, B)
    assert plan.status()[0] == 0
    x = to_host(B)
    assert relerr(x, xo) <= RTOL
    # factor reuse: a second right-hand side through hodlr_solve
    B2h = np.asfortranarray(np.random.default_rng(1).standard_normal((N, s)))
    B2 = to_dev(F, B2h)
    plan.solve(B2)
    xo2 = orc.hodlr_solve(Kh, Uh, Vh, B2h, m, task_level=min(L, 3), nthreads=8)
    assert relerr(to_host(B2), xo2) <= RTOL
    plan.close()



def test_config1_seed0_matches_oracle(gpu, orc):
    """The CPU-runnable workload exactly as specified: N=16384, 64 leaves of
    256, rank 30, 1 RHS U[0,1), seed 0."""
    say("case: config 1 N=16384 m=256 r=30 s=1 seed 0")
    F = gpu
    N, m, r, s, L = 16384, 256, 30, 1, 6
    K, U, V, B = F.workload_gaussian(N, m, r, s, seed=0)
    Kh, Uh, Vh, Bh = to_host(K), to_host(U), to_host(V), to_host(B)
    xo = orc.hodlr_solve(Kh, Uh, Vh, Bh, m, task_level=3, nthreads=8)
    plan = F.HodlrPlan(N, m, L, r, nrhs_max=s)
    plan.factor_solve(K, U, V, B)
    assert plan.status()[0] == 0
    e = relerr(to_host(B), xo)
    say(f"  rel err vs oracle {e:.2e}")
    assert e <= RTOL
    plan.close()


def test_padded_leading_dimension_is_bitwise_identical(gpu):
    """Operands with leading dimension N + 8 (the benchmark layout) give
    bit-identical results to ld = N: addressing only, same arithmetic."""
    F = gpu
    N, m, r, s = 8192, 256, 32, 4
    L = 5
    say(f"case: padded ld N={N} m={m} r={r} s={s}")
    outs = []
    for pad in (0, 8):
        K, U, V, B = F.workload_gaussian(N, m, r, s, seed=4, pad=pad)
        assert F._ld(K) == N + pad
        plan = F.HodlrPlan(N, m, L, r, nrhs_max=s)
        plan.factor_solve(K, U, V, B)
        assert plan.status()[0] == 0
        X2 = F.fortran_clone(B)
        plan.solve(X2)  # factor reuse on the solution (solves A y = x)
        outs.append((to_host(B), to_host(X2)))
        plan.close()
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("scale,r", [(1.0, 16), (200.0, 16), (1.0, 100), (200.0, 100),
                                     (1.0, 32), (200.0, 32), (200.0, 64), (1.0, 96),
                                     (200.0, 96)])
def test_core_structured_and_general_lu(gpu, orc, scale, r):
    """Cores S = [I, A; B, I] take the structured LU when max|B| <= 1 (plain
    workload) and the general pivoted LU otherwise (U scaled by 200 makes
    |B| > 1 at every level); both match the oracle to 1e-10.  Ranks 96 and
    100 take the wide-core path (64 < r <= 128: the Schur complement factored
    by the 128-row leaf LU).  At ranks 64 and 96 (multiples of 32) the core
    solves take the structured solve, with the generic solve masked to the
    cores the general LU factored -- in the fused factor+solve and in a
    factor-reuse solve, which reads every level's flags.  At ranks 32 and
    64 the structured solve also forms the level above's cores from every
    node's solution, the general cores' included."""
    F = gpu
    N, m, s = 2048, 128, 2
    say(f"case: core LU path, U scale {scale}, rank {r}")
    K, U, V, B = F.workload_gaussian(N, m, r, s, seed=13)
    U.mul_(scale)
    Kh, Uh, Vh, Bh = (to_host(t) for t in (K, U, V, B))
    ref = orc.hodlr_solve(Kh, Uh, Vh, Bh, m)
    plan = F.HodlrPlan(N, m, 4, r, nrhs_max=s)
    plan.factor_solve(K, U, V, B)
    assert plan.status()[0] == 0
    x = to_host(B)
    e = relerr(x, ref)
    # the tolerance: 1e-10, or 4x the oracle's own distance from a dense
    # solve where the scaled system amplifies rounding more (rank 100, U x 200:
    # cond 7e4, oracle vs dense 2.4e-10)
    xd = np.linalg.solve(orc.dense_assemble(Kh, Uh, Vh, m), Bh)
    tol = max(RTOL, 4.0 * relerr(ref, xd))
    say(f"  rel err {e:.2e} (tol {tol:.1e})")
    assert e <= tol and relerr(x, xd) <= tol
    X2 = to_dev(F, Bh)
    plan.solve(X2)
    assert relerr(to_host(X2), ref) <= tol
    plan.close()

@pytest.mark.parametrize("r", [16, 30, 64, 96, 128])
def test_core_schur_complement_pivots(gpu, orc, r):
    """A root core whose structured LU pivots inside the Schur complement:
    identity leaves (Y = U), V_c = [I; 0], U_0 = [B; 0] with |B| <= 1 (the
    structured path) and U_1 = [A; 0] with large A, so S = [I, A; B, I] and
    C = I - B A needs row interchanges in every panel.  The solution must
    match the oracle (restated dgetrf pivoting) and a dense solve to the
    core's condition number times rounding."""
    F = gpu
    m = 160
    N = 2 * m
    say(f"case: Schur-complement pivoting r={r}")
    rng = np.random.default_rng(300 + r)
    Bm = rng.uniform(-1.0, 1.0, (r, r))
    Am = 5.0 * rng.standard_normal((r, r))
    K = np.zeros((N, m))
    K[:m] = np.eye(m)
    K[m:] = np.eye(m)
    U = np.zeros((N, r))
    V = np.zeros((N, r))
    U[:r] = Bm
    U[m:m + r] = Am
    V[:r] = np.eye(r)
    V[m:m + r] = np.eye(r)
    b = rng.standard_normal((N, 3))
    K, U, V, b = (np.asfortranarray(a) for a in (K, U, V, b))
    S = np.block([[np.eye(r), Am], [Bm, np.eye(r)]])
    tol = max(1e-13, 4e-16 * np.linalg.cond(S) * r)
    xo = orc.hodlr_solve(K, U, V, b, m)
    A = orc.dense_assemble(K, U, V, m)
    xd = np.linalg.solve(A, b)
    assert relerr(xo, xd) <= tol
    Kd, Ud, Vd, Bd = to_dev(F, K), to_dev(F, U), to_dev(F, V), to_dev(F, b)
    plan = F.HodlrPlan(N, m, 1, r, nrhs_max=3)
    plan.factor_solve(Kd, Ud, Vd, Bd)
    assert plan.status()[0] == 0
    x = to_host(Bd)
    assert relerr(x, xo) <= tol, (relerr(x, xo), tol)
    assert relerr(x, xd) <= tol
    # factor reuse: the stored core factorization solves a new right-hand side
    b2 = np.asfortranarray(rng.standard_normal((N, 2)))
    B2 = to_dev(F, b2)
    plan.solve(B2)
    assert relerr(to_host(B2), np.linalg.solve(A, b2)) <= tol
    plan.close()


def test_factor_then_solve_many_rhs(gpu, orc):
    """hodlr_factor + hodlr_solve with more columns than nrhs_max (column
    batches), against the oracle."""
    say("case: factor + solve 40 rhs N=4096 m=128 r=16")
    F = gpu
    N, m, r = 4096, 128, 16
    K, U, V, _ = F.workload_gaussian(N, m, r, 1, seed=3)
    Kh, Uh, Vh = to_host(K), to_host(U), to_host(V)
    plan = F.HodlrPlan(N, m, 5, r, nrhs_max=4)
    plan.factor(K, U, V)
    Bh = np.asfortranarray(np.random.default_rng(2).random((N, 40)))
    B = to_dev(F, Bh)
    plan.solve(B)
    xo = orc.hodlr_solve(Kh, Uh, Vh, Bh, m, task_level=3, nthreads=8)
    assert relerr(to_host(B), xo) <= RTOL
    plan.close()


def test_reference_generator_solve(gpu, orc):
    """The reference's own generator (drand48 seeds, K = U V^T + diag(D+1e3),
    dense_block.cc:54-75) through the device solve."""
    say("case: reference generator N=4096 m=256 r=16")
    F = gpu
    npart, m, r = 16, 256, 16
    N = npart * m
    U = orc.ref_matrix(101, npart, m, r)
    V = orc.ref_matrix(102, npart, m, r)
    K = orc.ref_dense_blocks(npart, m, r, 1000, orc.ref_seeds(101, npart),
                             orc.ref_seeds(102, npart), orc.ref_seeds(103, npart))
    b = np.asfortranarray(orc.ref_matrix(104, npart, m, 1))
    xo = orc.hodlr_solve(K, U, V, b, m)
    Kd, Ud, Vd, Bd = to_dev(F, K), to_dev(F, U), to_dev(F, V), to_dev(F, b)
    plan = F.HodlrPlan(N, m, 4, r, nrhs_max=1)
    plan.factor_solve(Kd, Ud, Vd, Bd)
    assert relerr(to_host(Bd), xo) <= RTOL
    plan.close()


def test_singular_leaf_status(gpu):
    say("case: singular leaf N=1024 m=128")
    F = gpu
    K, U, V, B = F.workload_gaussian(1024, 128, 8, 1, seed=5)
    K[3 * 128:4 * 128, 17] = 0.0
    plan = F.HodlrPlan(1024, 128, 3, 8, nrhs_max=1)
    plan.factor_solve(K, U, V, B)
    st, where = plan.status()
    assert st == 1 and where == 3
    plan.close()


def test_bad_shapes_rejected(gpu):
    F = gpu
    with pytest.raises(F.HodlrError):
        F.HodlrPlan(1000, 64, 4, 8)        # N != leaf * 2^levels
    with pytest.raises(F.HodlrError):
        F.HodlrPlan(1024, 8, 7, 8)         # leaf <= rank (matrix.cc:225)


# ------------------------------------------------------------------ reference layout
@pytest.mark.parametrize("N,m,r,s", [(512, 64, 4, 2), (2048, 256, 16, 1), (1200, 150, 10, 3)])
def test_hsolve_ref_full_layout(gpu, orc, N, m, r, s):
    """Device hsolve on the reference's column layout X = [b | u_0..u_{L-1}]:
    every column (solution and the per-level u blocks) vs the oracle hsolve."""
    say(f"case: hsolve_ref N={N} m={m} r={r} s={s}")
    F = gpu
    L = int(np.log2(N // m))
    K, U, V, B = orc.workload_gaussian(N, m, r, s, seed=7)
    X = orc.build_x(B, U, L)
    Ko, Xo = K.copy(order="F"), X.copy(order="F")
    assert orc.hsolve(N, s, r, N // m, Ko, Xo, V) == 0
    Kd, Xd, Vd = to_dev(F, K), to_dev(F, X), to_dev(F, V)
    assert F.hsolve(N, s, r, N // m, Kd, Xd, Vd) == 0
    Xg = to_host(Xd)
    assert relerr(Xg[:, :s], Xo[:, :s]) <= RTOL
    for t in range(L):
        blk = slice(s + r * t, s + r * (t + 1))
        assert relerr(Xg[:, blk], Xo[:, blk]) <= RTOL, t


# ------------------------------------------------------------------ LMatrix operators
def test_gemm_red_bro_node_solve(gpu, orc):
    say("case: gemm_red / gemm_bro / node_solve")
    F = gpu
    rng = np.random.default_rng(3)
    rows, ka, kb = 4096, 12, 7
    A = np.asfortranarray(rng.standard_normal((rows, ka)))
    Bm = np.asfortranarray(rng.standard_normal((rows, kb)))
    for nchunk, nslot in [(8, 2), (16, 16), (4, 1)]:
        Cg = to_host(F.gemm_red(to_dev(F, A), to_dev(F, Bm), nchunk, nslot, alpha=0.5))
        Co = orc.gemm_red(A, Bm, nchunk, nslot, alpha=0.5)
        assert relerr(Cg, Co) <= 1e-13
    k, nd = 6, 9
    u = np.asfortranarray(rng.standard_normal((rows, k)))
    for nchunk, nslot in [(8, 2), (4, 4)]:
        E = np.asfortranarray(rng.standard_normal((nslot * k, nd)))
        D = np.asfortranarray(rng.standard_normal((rows, nd)))
        Dg = to_dev(F, D)
        F.gemm_bro(to_dev(F, u), to_dev(F, E), Dg, nchunk, nslot, alpha=-1.0)
        Do = orc.gemm_bro(u, E, D.copy(order="F"), nchunk, nslot, alpha=-1.0)
        assert relerr(to_host(Dg), Do) <= 1e-13
    r, nnode, nrhs = 5, 4, 3
    VTu = np.asfortranarray(rng.random((nnode * 2 * r, r)))
    VTd = np.asfortranarray(rng.random((nnode * 2 * r, nrhs)))
    VTd_o = VTd.copy(order="F")
    assert orc.node_solve(VTu, VTd_o, r) == 0
    VTd_g = to_dev(F, VTd)
    assert not F.node_solve(to_dev(F, VTu), VTd_g, r).any()
    assert relerr(to_host(VTd_g), VTd_o) <= 1e-12


def test_operators_bitwise_on_integer_inputs(gpu, orc):
    """gemm_red / gemm_bro on small integers (every product and partial sum
    exact in FP64, so any summation order gives the same bits) and
    node_solve on cores whose elimination is exact (S = [I, A; 0, I]:
    x2 = R2, x1 = R1 - A x2 in integers) equal the oracle bit for bit --
    the test.cc known-answer relations (:325-343, :353-371, :386-424)."""
    say("case: integer-exact gemm_red / gemm_bro / node_solve")
    F = gpu
    rng = np.random.default_rng(11)
    rows, ka, kb = 2048, 9, 5
    A = np.asfortranarray(rng.integers(-4, 5, (rows, ka)).astype(np.float64))
    Bm = np.asfortranarray(rng.integers(-4, 5, (rows, kb)).astype(np.float64))
    for nchunk, nslot in [(8, 2), (16, 16), (4, 1)]:
        Cg = to_host(F.gemm_red(to_dev(F, A), to_dev(F, Bm), nchunk, nslot, alpha=2.0))
        Co = orc.gemm_red(A, Bm, nchunk, nslot, alpha=2.0)
        assert np.array_equal(Cg, Co), (nchunk, nslot)
    k, nd = 6, 7
    u = np.asfortranarray(rng.integers(-3, 4, (rows, k)).astype(np.float64))
    for nchunk, nslot in [(8, 2), (4, 4)]:
        E = np.asfortranarray(rng.integers(-3, 4, (nslot * k, nd)).astype(np.float64))
        D = np.asfortranarray(rng.integers(-9, 10, (rows, nd)).astype(np.float64))
        Dg = to_dev(F, D)
        F.gemm_bro(to_dev(F, u), to_dev(F, E), Dg, nchunk, nslot, alpha=-1.0)
        Do = orc.gemm_bro(u, E, D.copy(order="F"), nchunk, nslot, alpha=-1.0)
        assert np.array_equal(to_host(Dg), Do), (nchunk, nslot)
    r, nnode, nrhs = 6, 8, 4
    VTu = np.asfortranarray(rng.integers(-3, 4, (nnode * 2 * r, r)).astype(np.float64))
    for k2 in range(nnode):
        # S = [I, VTu(r:2r); VTu(0:r), I]: a zero lower-left block makes
        # S = [I, A; 0, I], factored without pivoting or rounding
        VTu[k2 * 2 * r:k2 * 2 * r + r, :] = 0.0
    VTd = np.asfortranarray(rng.integers(-5, 6, (nnode * 2 * r, nrhs)).astype(np.float64))
    VTd_o = VTd.copy(order="F")
    assert orc.node_solve(VTu, VTd_o, r) == 0
    VTd_g = to_dev(F, VTd)
    assert not F.node_solve(to_dev(F, VTu), VTd_g, r).any()
    assert np.array_equal(to_host(VTd_g), VTd_o)


@pytest.mark.parametrize("r0,r1,nrhs,nnode,scale", [
    (5, 9, 3, 1, 1.0),       # the reference's uniform [0,1) operands, one node
    (64, 100, 17, 6, 0.5),   # Gaussian cores that pivot, batched
    (30, 1, 4, 3, 2.0),
    (1, 30, 2, 3, 2.0),
    (300, 200, 9, 2, 0.1),   # largest case: a 500 x 500 core
])
def test_node_solve_region_matches_oracle(gpu, orc, r0, r1, nrhs, nnode, scale):
    """hodlr_node_solve_region (NodeSolveRegionTask, node_solve_region.cc:
    77-107: S(i, r0+j) = VTu1(i,j), S(r0+j, i) = VTu0(j,i), B = [VTd1; VTd0])
    per node against orc_node_solve_region, and the core equations S x = B
    themselves."""
    say(f"case: node_solve_region r0={r0} r1={r1} nrhs={nrhs} nnode={nnode}")
    F = gpu
    rng = np.random.default_rng(r0 * 1000 + r1)
    if scale == 1.0:
        mk = lambda *s: rng.random(s)  # noqa: E731
    else:
        mk = lambda *s: scale * rng.standard_normal(s)  # noqa: E731
    VTu0, VTu1 = mk(nnode, r1, r0), mk(nnode, r0, r1)
    VTd0, VTd1 = mk(nnode, r1, nrhs), mk(nnode, r0, nrhs)

    def dev3(a):
        t = torch.empty((a.shape[0], a.shape[2], a.shape[1]), dtype=torch.float64,
                        device="cuda").transpose(1, 2)
        t.copy_(torch.from_numpy(a))
        return t

    g = [dev3(a) for a in (VTu0, VTu1, VTd0, VTd1)]
    if nnode == 1:
        g = [t[0] for t in g]
    info = F.node_solve_region(*g).cpu().numpy()
    assert not info.any(), info
    torch.cuda.synchronize()
    d0g = g[2].cpu().numpy().reshape(nnode, r1, nrhs)
    d1g = g[3].cpu().numpy().reshape(nnode, r0, nrhs)
    for k in range(nnode):
        d0 = np.array(VTd0[k], order="F", copy=True)
        d1 = np.array(VTd1[k], order="F", copy=True)
        assert orc.node_solve_region(np.asfortranarray(VTu0[k]), np.asfortranarray(VTu1[k]),
                                     d0, d1) == 0
        S = np.eye(r0 + r1)
        S[:r0, r0:] = VTu1[k]
        S[r0:, :r0] = VTu0[k]
        # same pivots as the oracle; the two differ by summation order only,
        # amplified by at most the core's condition number
        tol = max(1e-13, 2e-15 * np.linalg.cond(S))
        assert relerr(d1g[k], d1) <= tol and relerr(d0g[k], d0) <= tol, k
        x = np.vstack([d1g[k], d0g[k]])
        rhs = np.vstack([VTd1[k], VTd0[k]])
        assert np.abs(S @ x - rhs).max() <= 1e-12 * np.abs(S).max() * np.abs(x).max() * (r0 + r1)


def test_node_solve_region_exact_and_singular(gpu, orc):
    """Integer cores with VTu0 = 0 (S = [I, A; 0, I]: x = [R1 - A R2; R2]
    in exact integer arithmetic, bit-equal to the oracle), and a singular
    core (I - VTu0 VTu1 singular) reporting the oracle's LAPACK INFO."""
    say("case: node_solve_region exact / singular")
    F = gpu
    rng = np.random.default_rng(21)
    r0, r1, nrhs = 7, 4, 3
    VTu0 = np.zeros((r1, r0))
    VTu1 = rng.integers(-3, 4, (r0, r1)).astype(np.float64)
    VTd0 = rng.integers(-5, 6, (r1, nrhs)).astype(np.float64)
    VTd1 = rng.integers(-5, 6, (r0, nrhs)).astype(np.float64)
    g = [to_dev(F, a) for a in (VTu0, VTu1, VTd0, VTd1)]
    assert not F.node_solve_region(*g).cpu().numpy().any()
    d0, d1 = VTd0.copy(order="F"), VTd1.copy(order="F")
    assert orc.node_solve_region(F_(VTu0), F_(VTu1), d0, d1) == 0
    assert np.array_equal(to_host(g[3]), d1) and np.array_equal(to_host(g[2]), d0)
    assert np.array_equal(d0, VTd0) and np.array_equal(d1, VTd1 - VTu1 @ VTd0)
    # singular: r0 = 1, r1 = 2, I - VTu0 VTu1 = diag(0, 1)
    VTu0 = np.array([[1.0], [0.0]])
    VTu1 = np.array([[1.0, 0.0]])
    VTd0, VTd1 = np.ones((2, 1)), np.ones((1, 1))
    g = [to_dev(F, a) for a in (VTu0, VTu1, VTd0, VTd1)]
    info = int(F.node_solve_region(*g).cpu().numpy()[0])
    info_o = orc.node_solve_region(F_(VTu0), F_(VTu1), F_(VTd0), F_(VTd1))
    assert info > 0 and info == info_o, (info, info_o)


def F_(a):
    return np.asfortranarray(a, dtype=np.float64)


def test_matvec_matches_oracle(gpu, orc):
    say("case: matvec N=4096 m=128 r=12 s=5")
    F = gpu
    N, m, r, s = 4096, 128, 12, 5
    K, U, V, B = F.workload_gaussian(N, m, r, s, seed=8)
    y = to_host(F.matvec(K, U, V, B, m))
    yo = orc.hodlr_matvec(to_host(K), to_host(U), to_host(V), to_host(B), m)
    assert relerr(y, yo) <= 1e-13


# ------------------------------------------------------------------ full size
def test_config2_size_residual_and_oracle_sample(gpu, orc):
    """N = 2^18, leaf 512, rank 64, 8 RHS (the benchmark workload): residual
    through the device HODLR matvec, and the solution on the first leaves
    against the oracle run on the same arrays."""
    say("case: full size N=262144 m=512 r=64 s=8")
    F = gpu
    N, m, r, s, L = 262144, 512, 64, 8, 9
    K, U, V, B = F.workload_gaussian(N, m, r, s, seed=0)
    K0, B0 = K.clone(), B.clone()
    plan = F.HodlrPlan(N, m, L, r, nrhs_max=s)
    plan.factor_solve(K, U, V, B)
    assert plan.status()[0] == 0
    R = F.matvec(K0, U, V, B, m, alpha=-1.0)
    R.add_(B0)
    torch.cuda.synchronize()
    rel = (R.abs().max() / (B0.abs().max())).item()
    say(f"  relative residual {rel:.3e}")
    assert rel <= 1e-12
    # oracle on the same arrays (reference algorithm, OpenBLAS, 16 threads)
    xo = orc.hodlr_solve(to_host(K0), to_host(U), to_host(V), to_host(B0), m, task_level=4,
                         nthreads=16)
    assert relerr(to_host(B), xo) <= RTOL
    plan.close()


@pytest.mark.parametrize("r", [200, 300])
def test_large_rank_plan(gpu, orc, r):
    """Ranks above 128: cores of 2r rows take the general pivoted LU and
    the generic solve (2r up to 1216), and the top-down pass keeps H in
    narrower LDS column chunks (16 columns at r = 200, 8 at r = 300).  The
    fused factor+solve and a factor-reuse solve against the oracle at
    1e-10, the adjoint solve against the dense A^-T."""
    F = gpu
    m, L, s = r + 64, 2, 20
    N = m << L
    say(f"case: large rank r={r} N={N} leaf={m} nrhs={s}")
    K, U, V, B = F.workload_gaussian(N, m, r, s, seed=21)
    Kh, Uh, Vh, Bh = (to_host(t) for t in (K, U, V, B))
    ref = orc.hodlr_solve(Kh, Uh, Vh, Bh, m)
    plan = F.HodlrPlan(N, m, L, r, nrhs_max=s)
    X = F.fortran_clone(B)
    plan.factor_solve(K, U, V, X)
    assert plan.status()[0] == 0
    assert relerr(to_host(X), ref) <= RTOL
    X2 = F.fortran_clone(B)
    plan.solve(X2)
    assert relerr(to_host(X2), ref) <= RTOL
    A = orc.dense_assemble(Kh, Uh, Vh, m)
    G = F.fortran_clone(B)
    plan.solve_adjoint(G)
    assert relerr(to_host(G), np.linalg.solve(A.T, Bh)) <= RTOL
    plan.close()
