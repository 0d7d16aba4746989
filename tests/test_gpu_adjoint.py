"""Adjoint solve and gradients of the HODLR solve (gpu marker).

The reference has no gradient path; these results are pinned by
mathematics that fixes them independently of the implementation:
  - A^-T g against a dense LU solve of the assembled A^T (numpy);
  - dK, dU, dV, dB of l = 0.5 ||x - x*||^2 against the dense-LU gradient
    dA = -Lam x^T projected on the HODLR structure (leaf blocks for K;
    dA with the leaf blocks masked out, times V or U^T, for U and V);
  - central finite differences (h = 1e-6) of the same loss through the
    device factor + solve on 64 random entries of K, U, V and B, at the
    differentiable-solve workload N = 65536, leaf 256, rank 32, 16 RHS.
Tolerances are stated per test.
"""
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


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
    return np.asfortranarray(t.detach().cpu().numpy())


def relerr(x, ref):
    return float(np.abs(x - ref).max() / max(np.abs(ref).max(), 1e-300))


def dense_grads(A, X, Xstar, U, V, leaf):
    """Dense-LU gradient of 0.5 ||X - X*||^2, X = A^-1 B, on the HODLR
    structure A = blockdiag(K_i) + (U V^T off the leaf blocks)."""
    N = A.shape[0]
    G = X - Xstar
    Lam = np.linalg.solve(A.T, G)
    dA = -Lam @ X.T
    dK = np.empty((N, leaf))
    M = np.ones((N, N))
    for i in range(N // leaf):
        sl = slice(i * leaf, (i + 1) * leaf)
        dK[sl, :] = dA[sl, sl]
        M[sl, sl] = 0.0
    dAm = dA * M
    return dK, dAm @ V, dAm.T @ U, Lam


@pytest.mark.parametrize("N,leaf,r,s", [(1024, 64, 8, 3), (2048, 128, 16, 20), (4096, 256, 32, 1)])
def test_adjoint_solve_vs_dense(gpu, orc, N, leaf, r, s):
    """A^-T G with the stored factorization equals the dense solve of A^T
    to 1e-10 relative (max-norm)."""
    say(f"case: adjoint solve N={N} leaf={leaf} r={r} s={s}")
    F = gpu
    K, U, V, _ = orc.workload_gaussian(N, leaf, r, 1, seed=11)
    G = np.asfortranarray(np.random.default_rng(5).standard_normal((N, s)))
    A = orc.dense_assemble(K, U, V, leaf)
    ref = np.linalg.solve(A.T, G)
    L = int(round(np.log2(N // leaf)))
    plan = F.HodlrPlan(N, leaf, L, r, nrhs_max=max(s, 1))
    Kd, Ud, Vd, Gd = to_dev(F, K), to_dev(F, U), to_dev(F, V), to_dev(F, G)
    plan.factor(Kd, Ud, Vd)
    plan.solve_adjoint(Gd)
    assert plan.status()[0] == 0
    e = relerr(to_host(Gd), ref)
    say(f"  rel err {e:.2e}")
    assert e <= 1e-10
    plan.close()


def test_grad_vs_dense_lu(gpu, orc):
    """Gradients of 0.5||x - x*||^2 through torch.autograd (hand-written
    adjoint) against the dense-LU gradient at N = 4096 (leaf 256, rank 32,
    16 RHS, workload generators of the differentiable-solve case), 1e-9
    relative (max-norm per gradient)."""
    N, leaf, r, s = 4096, 256, 32, 16
    say(f"case: autograd vs dense LU N={N} leaf={leaf} r={r} s={s}")
    F = gpu
    K, U, V, B = orc.workload_gaussian(N, leaf, r, s, seed=3)
    Xs = np.asfortranarray(np.random.default_rng(9).standard_normal((N, s)))
    A = orc.dense_assemble(K, U, V, leaf)
    X = np.linalg.solve(A, B)
    dK, dU, dV, dB = dense_grads(A, X, Xs, U, V, leaf)
    Kt, Ut, Vt, Bt = (to_dev(F, a).requires_grad_(True) for a in (K, U, V, B))
    Xt = F.hodlr_solve(Kt, Ut, Vt, Bt, leaf)
    loss = 0.5 * ((Xt - to_dev(F, Xs)) ** 2).sum()
    loss.backward()
    assert relerr(to_host(Xt), X) <= 1e-10
    for name, got, ref in (("dK", Kt.grad, dK), ("dU", Ut.grad, dU), ("dV", Vt.grad, dV),
                           ("dB", Bt.grad, dB)):
        e = relerr(to_host(got), ref)
        say(f"  {name}: rel err {e:.2e}")
        assert e <= 1e-9, name


def test_grad_finite_differences(gpu):
    """Central finite differences (h = 1e-6) of l = 0.5||x - x*||^2 on 64
    random entries of K, U, V, B at N = 65536, leaf 256, rank 32, 16 RHS.
    The loss difference is formed as 0.5 * sum((x+ - x-) * (x+ + x- - 2x*))
    so that it carries no cancellation of the two O(|l|) loss values.
    Tolerance: |fd - g| <= 1e-5 * max|g| + 1e-5 * |g| per entry.  The first
    term is the rounding floor of the difference quotient: both solves carry
    errors of about eps * |x| per entry (~1e-19 here), which summed over
    N * s = 2^20 entries against |x+ + x- - 2x*| ~ 1 and divided by 2h give
    ~1e-9, about 1e-5 of the largest gradient entry (~3e-4 for dK)."""
    N, leaf, r, s = 65536, 256, 32, 16
    h = 1e-6
    say(f"case: finite differences N={N} leaf={leaf} r={r} s={s}")
    F = gpu
    K, U, V, B = F.workload_gaussian(N, leaf, r, s, seed=0)
    g = torch.Generator(device="cpu").manual_seed(1)
    Xs = F.fortran(torch.randn(s, N, generator=g, dtype=torch.float64).t().to("cuda"))
    Kt, Ut, Vt, Bt = (t.clone().requires_grad_(True) for t in (K, U, V, B))
    Xt = F.hodlr_solve(Kt, Ut, Vt, Bt, leaf)
    (0.5 * ((Xt - Xs) ** 2).sum()).backward()
    grads = {"K": Kt.grad, "U": Ut.grad, "V": Vt.grad, "B": Bt.grad}
    base = {"K": K, "U": U, "V": V, "B": B}
    L = int(round(np.log2(N // leaf)))
    plan = F.HodlrPlan(N, leaf, L, r, nrhs_max=s)
    Kw = F.fortran_empty(N, leaf)
    Xw = F.fortran_empty(N, s)

    def solve(ops):
        Kw.copy_(ops["K"])
        Xw.copy_(ops["B"])
        plan.factor_solve(Kw, ops["U"], ops["V"], Xw)
        return Xw.clone()

    rng = np.random.default_rng(2)
    names = ["K", "U", "V", "B"]
    worst = 0.0
    for q in range(64):
        nm = names[q % 4]
        t = base[nm]
        i, j = int(rng.integers(0, t.shape[0])), int(rng.integers(0, t.shape[1]))
        ops = dict(base)
        pert = t.clone()
        pert[i, j] += h
        ops[nm] = pert
        xp = solve(ops)
        pert[i, j] -= 2 * h
        xm = solve(ops)
        fd = (0.5 * ((xp - xm) * (xp + xm - 2 * Xs)).sum() / (2 * h)).item()
        an = grads[nm][i, j].item()
        scale = grads[nm].abs().max().item()
        err = abs(fd - an)
        worst = max(worst, err / (1e-5 * scale + 1e-5 * abs(an)))
        assert err <= 1e-5 * scale + 1e-5 * abs(an), (nm, i, j, fd, an)
    say(f"  worst error / tolerance over 64 entries: {worst:.3f}")
    assert plan.status()[0] == 0
    plan.close()
