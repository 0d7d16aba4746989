"""Trees of uneven leaves (gpu marker): the shapes of the reference's
irregular drivers -- shards with 10x leaves (spmd_driver/
solver_irregular_part.cc:73-75) -- and systems of P leaves with P not a
power of two (empty trailing leaf slots instead of padding rows).

Parity: the plan's solution against the restated reference solve of the
balanced embedding (oracle.hodlr_solve_uneven: every leaf padded to the
largest with an identity block and zero U, V, B rows, a permutation of
blockdiag(A, I)) at 1e-10 relative, and against a dense solve of the
assembled matrix at small N; the residual through the plan's own HODLR
matvec (which must equal a dense product)."""
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_host(t):
    torch.cuda.synchronize()
    return np.asfortranarray(t.detach().cpu().numpy())


def relerr(x, ref):
    return float(np.abs(x - ref).max() / max(np.abs(ref).max(), 1e-300))


@pytest.fixture(scope="module", autouse=True)
def _blas(orc):
    orc.use_openblas()
    yield
    orc.use_netlib()


CASES = [
    # (name, leaf sizes as (spmd_level, matrix_level, leaf, heavy) or a list, rank, nrhs)
    ("part-defaults", (2, 2, 40, 1), 10, 1),        # the driver's defaults: leaf 40, rank 10
    ("part-two-heavy", (2, 2, 100, 2), 30, 3),      # heavy leaves of 1000 rows
    ("part-fast-lu", (1, 3, 48, 1), 16, 2),         # 480-row heavy leaves, 48-row light
    ("trailing-5", [128] * 5 + [0] * 3, 20, 2),     # 5 leaves in 8 slots
    ("trailing-11", [64] * 11 + [0] * 5, 12, 1),    # 11 leaves in 16 slots
    ("mixed-sizes", [256, 0, 384, 128, 512, 512, 0, 200], 24, 4),
]


def sizes_of(spec):
    from fastsolver2_amd.workloads import shard_leaf_sizes
    if isinstance(spec, tuple):
        g, ml, leaf, heavy = spec
        return shard_leaf_sizes(g, ml, leaf, heavy), g + ml
    levels = int(round(np.log2(len(spec))))
    return list(spec), levels


@pytest.mark.parametrize("name,spec,r,s", CASES, ids=[c[0] for c in CASES])
def test_uneven_factor_solve_matches_oracle(gpu, orc, name, spec, r, s):
    F = gpu
    from fastsolver2_amd.workloads import workload_uneven
    sizes, L = sizes_of(spec)
    say(f"case: uneven {name} levels={L} N={sum(sizes)} rank={r} nrhs={s} sizes={sizes}")
    K, U, V, B = workload_uneven(sizes, r, s, seed=5)
    Kh, Uh, Vh, Bh = (to_host(t) for t in (K, U, V, B))
    plan = F.HodlrPlan(0, 0, L, r, nrhs_max=s, leaf_sizes=sizes)
    X = F.fortran_clone(B)
    plan.factor_solve(K, U, V, X)
    assert plan.status() == (0, -1)
    x = to_host(X)
    xo = orc.hodlr_solve_uneven(Kh, Uh, Vh, Bh, sizes, nthreads=4)
    assert relerr(x, xo) <= RTOL, relerr(x, xo)
    # residual through the plan's matvec, and that matvec against dense
    R = F.fortran_clone(B)
    plan.matvec(K, U, V, X, alpha=-1.0, beta=1.0, Y=R)
    assert np.abs(to_host(R)).max() <= 1e-12 * np.abs(Bh).max()
    if sum(sizes) <= 6000:
        A = orc.dense_assemble_uneven(Kh, Uh, Vh, sizes)
        assert relerr(x, np.linalg.solve(A, Bh)) <= RTOL
        y = to_host(plan.matvec(K, U, V, B))
        assert relerr(y, A @ Bh) <= 1e-13
        yt = to_host(plan.matvec(K, U, V, B, trans=True))
        assert relerr(yt, A.T @ Bh) <= 1e-13
    # factor reuse: the stored factorization solves another right-hand side
    B2 = F.generate(9, 3, 0, 1, sum(sizes), s, 1.0)
    B2h = to_host(B2)
    plan.solve(B2)
    assert relerr(to_host(B2), orc.hodlr_solve_uneven(Kh, Uh, Vh, B2h, sizes, nthreads=4)) <= RTOL
    plan.close()


def test_uneven_adjoint_and_refinement(gpu, orc):
    """The adjoint solve on an uneven tree against the dense A^-T, and a
    mixed-precision plan (FP32 leaves) refined to 1e-12 through the plan's
    matvec."""
    F = gpu
    from fastsolver2_amd.workloads import shard_leaf_sizes, workload_uneven
    sizes = shard_leaf_sizes(2, 2, 48, 1)  # heavy leaves of 480
    L, r, s = 4, 16, 3
    say(f"case: uneven adjoint + mixed IR sizes={sizes}")
    K, U, V, B = workload_uneven(sizes, r, s, seed=7)
    Kh, Uh, Vh, Bh = (to_host(t) for t in (K, U, V, B))
    A = orc.dense_assemble_uneven(Kh, Uh, Vh, sizes)
    plan = F.HodlrPlan(0, 0, L, r, nrhs_max=s, leaf_sizes=sizes)
    plan.factor(K, U, V)
    G = F.fortran_clone(B)
    plan.solve_adjoint(G)
    assert relerr(to_host(G), np.linalg.solve(A.T, Bh)) <= RTOL
    plan.close()
    mp = F.HodlrPlan(0, 0, L, r, nrhs_max=s, leaf_sizes=sizes, precision=F.MIXED_IR)
    mp.factor(K, U, V)
    X, st = mp.solve_ir(K, U, V, B, tol=1e-12, maxit=10)
    assert st["converged"] and st["residual"] <= 1e-12, st
    assert st["history"][0] > 1e-10  # the first pass is FP32-accurate only
    assert relerr(to_host(X), np.linalg.solve(A, Bh)) <= RTOL
    mp.close()


def test_uneven_shapes_rejected(gpu):
    F = gpu
    with pytest.raises(F.HodlrError):   # a non-empty leaf not above the rank
        F.HodlrPlan(0, 0, 2, 16, leaf_sizes=[64, 16, 64, 64])
    with pytest.raises(F.HodlrError):   # larger than the leaf kernels take
        F.HodlrPlan(0, 0, 1, 8, leaf_sizes=[2000, 64])
    with pytest.raises(F.HodlrError):   # all slots empty
        F.HodlrPlan(0, 0, 1, 8, leaf_sizes=[0, 0])
    with pytest.raises(F.HodlrError):   # more leaves than slots
        F.HodlrPlan(9 * 64, 64, 3, 8)


PAIRS = [
    # (shard-0 sizes, rank0, shard-1 sizes, rank1, nrhs): solver_irregular_ranks.cc
    ([96] * 8, 12, [96] * 8, 30, 2),                  # equal shards, rank0 != rank1
    ([400, 40, 40, 40], 9, [64] * 4, 20, 3),           # a heavy leaf in shard 0
    ([256] * 4, 64, [128, 128, 0, 0], 100, 1),         # empty slots in shard 1
]


@pytest.mark.parametrize("sz0,r0,sz1,r1,s", PAIRS)
def test_pair_matches_oracle(gpu, orc, sz0, r0, sz1, r1, s):
    """Two shards with ranks r0 != r1 (the reference's irregular-ranks
    driver): the GPU pair solve against the oracle's restatement of that
    driver (oracle.pair_solve, pinned by a dense solve) at 1e-10, and
    against a dense solve of the assembled matrix."""
    F = gpu
    from fastsolver2_amd.workloads import workload_uneven
    say(f"case: pair r0={r0} r1={r1} sizes={sz0}/{sz1} nrhs={s}")
    K, U, V, W, Bs = [], [], [], [], []
    for c, (sz, r, rp) in enumerate(((sz0, r0, r1), (sz1, r1, r0))):
        Kc, Uc, Vc, Bc = workload_uneven(sz, r, s, seed=11, system=c)
        K.append(Kc); U.append(Uc); V.append(Vc); Bs.append(Bc)
        W.append(F.generate(11, 4, c, 1, sum(sz), rp, 1.0 / np.sqrt(rp)))
    n0, n1 = sum(sz0), sum(sz1)
    B = F.fortran_empty(n0 + n1, s)
    B[:n0].copy_(Bs[0])
    B[n0:].copy_(Bs[1])
    host = [[to_host(t) for t in pr] for pr in (K, U, V, W)]
    Bh = to_host(B)
    pair = F.HodlrPair(sz0, r0, sz1, r1, nrhs_max=s)
    X = F.fortran_clone(B)
    pair.factor_solve(K, U, V, W, X)
    assert pair.status()[0] == 0
    x = to_host(X)
    xo = orc.pair_solve(*host, Bh, (sz0, sz1), nthreads=4)
    assert relerr(x, xo) <= RTOL, relerr(x, xo)
    A = orc.dense_assemble_pair(*host, (sz0, sz1))
    assert relerr(x, np.linalg.solve(A, Bh)) <= RTOL
    pair.close()
