"""Synthetic HODLR workloads beyond the plain Gaussian one.

The mixed-precision workload asks for leaf blocks K_i = G_i + s*I (G iid
N(0,1), m x m) with s set so that cond(A) ~ 1e4.
  - ``condition_estimate`` estimates cond_2 of the whole HODLR matrix:
    sigma_max by power iteration on A^T A (FP64 HODLR matvecs), sigma_min
    by inverse iteration on (A^T A)^-1 = A^-1 A^-T (an FP64 factorization,
    solve and adjoint solve);
  - ``conditioned_shift_global`` bisects s on that estimate;
  - ``conditioned_shift`` is the cheaper per-leaf version (max cond_2 over
    the first few leaf blocks), used to bracket the global search.
The blocks are the device generator's (same seed, stream 0 = K), so the
calibration sees exactly the matrices the solver will factor.
"""
from __future__ import annotations

import numpy as np
import torch

from . import (HodlrPlan, fortran_clone, fortran_empty, generate, lib, matvec, workload_gaussian,
               _call, _ld, _ptr, _stream)

__all__ = ["conditioned_shift", "condition_estimate", "conditioned_shift_global", "batch_systems",
           "batch_operands", "batch_leaf_sizes", "shard_leaf_sizes", "workload_uneven"]


def shard_leaf_sizes(spmd_level, matrix_level, leaf, heavy, factor=10):
    """Leaf sizes of the reference's irregular driver
    (spmd_driver/solver_irregular_part.cc:73-75): 2^spmd_level shards, each a
    balanced subtree of 2^matrix_level leaves; shards with index < heavy
    have `factor` times the leaf size.  The tree has spmd_level +
    matrix_level levels."""
    out = []
    for shard in range(1 << spmd_level):
        m = leaf * factor if shard < heavy else leaf
        out += [m] * (1 << matrix_level)
    return out


def workload_uneven(leaf_sizes, r, s, seed=0, system=0, shift=None, rhs="uniform"):
    """Gaussian workload on a tree of uneven leaves: K (N x max m) holds leaf
    i's block G_i + shift_i I in columns [0, m_i) of its rows (shift_i = m_i
    unless given), U, V iid N(0,1)/sqrt(r), B iid U[0,1) (or N(0,1));
    streams 0..3 of the device Philox generator as the balanced workload."""
    sizes = [int(x) for x in leaf_sizes]
    N, M = sum(sizes), max(sizes)
    K = generate(seed, 0, system, 1, N, M, 1.0)
    rows, cols, vals = [], [], []
    row0 = 0
    for m in sizes:
        if m:
            j = torch.arange(m, dtype=torch.int64)
            rows.append(row0 + j)
            cols.append(j)
            vals.append(torch.full((m,), float(m if shift is None else shift), dtype=torch.float64))
        row0 += m
    ri = torch.cat(rows).to(K.device)
    ci = torch.cat(cols).to(K.device)
    K[ri, ci] += torch.cat(vals).to(K.device)
    sc = 1.0 / float(np.sqrt(r))
    U = generate(seed, 1, system, 1, N, r, sc)
    V = generate(seed, 2, system, 1, N, r, sc)
    B = generate(seed, 3, system, 0 if rhs == "uniform" else 1, N, s, 1.0)
    return K, U, V, B


def conditioned_shift(N, leaf, target_cond=1e4, seed=0, system=0, samples=2, iters=30):
    """Shift s with max_i cond_2(G_i + s I) ~ target_cond over the first
    `samples` leaves of the stream-0 generator (bisection on s)."""
    G = generate(seed, 0, system, 1, N, leaf, 1.0)  # N x leaf, leaf i = rows i*leaf..
    blocks = [G[i * leaf:(i + 1) * leaf, :].cpu().numpy().astype(np.float64)
              for i in range(min(samples, N // leaf))]
    del G
    torch.cuda.synchronize()
    eye = np.eye(leaf)

    def cond(s):
        return max(np.linalg.cond(b + s * eye) for b in blocks)

    # beyond the spectrum's edge (radius ~ sqrt(leaf)) cond falls as s
    # grows: scan down from 4 sqrt(leaf) to the first s above the target,
    # then bisect that bracket
    step = 0.1 * np.sqrt(leaf)
    hi = 4.0 * np.sqrt(leaf)
    while cond(hi) > target_cond:
        hi *= 2.0
    lo = hi - step
    while cond(lo) <= target_cond:
        hi, lo = lo, lo - step
        if lo <= 0.0:
            raise ValueError(f"no shift reaches condition {target_cond:g}")
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        if cond(mid) > target_cond:
            lo = mid
        else:
            hi = mid
        if hi - lo < 1e-6 * hi:
            break
    s = 0.5 * (lo + hi)
    return float(s), float(cond(s))


def _normalize(x):
    n = x.norm().item()
    x.div_(n)
    return n


def condition_estimate(K, U, V, leaf, iters=30, seed=0):
    """(sigma_max, sigma_min, cond_2) estimates of the HODLR matrix A(K, U, V)."""
    N, r = U.shape
    L = int(round(np.log2(N // leaf)))
    g = torch.Generator(device="cpu").manual_seed(seed)
    x = fortran_empty(N, 1)
    x.copy_(torch.randn(N, 1, generator=g, dtype=torch.float64))
    _normalize(x)
    lam = 0.0
    for _ in range(iters):
        y = matvec(K, U, V, x, leaf)
        x = matvec(K, U, V, y, leaf, trans=True)
        lam = _normalize(x)
    smax = float(np.sqrt(lam))
    plan = HodlrPlan(N, leaf, L, r, nrhs_max=1)
    Kf = fortran_clone(K)
    plan.factor(Kf, U, V)
    st, where = plan.status()
    if st != 0:
        plan.close()
        return smax, 0.0, float("inf")
    x = fortran_empty(N, 1)
    x.copy_(torch.randn(N, 1, generator=g, dtype=torch.float64))
    _normalize(x)
    mu = 0.0
    for _ in range(iters):
        plan.solve_adjoint(x)
        plan.solve(x)
        mu = _normalize(x)
    plan.close()
    del Kf
    smin = float(1.0 / np.sqrt(mu))
    return smax, smin, smax / smin


def conditioned_shift_global(N, leaf, r, target_cond=1e4, seed=0, iters=12, power_iters=30):
    """Shift s with cond_2(A) ~ target_cond for the Gaussian workload
    (K_i = G_i + s I, U, V iid N(0,1)/sqrt(r), streams 0..2 of `seed`)."""
    K, U, V, _ = workload_gaussian(N, leaf, r, 1, seed=seed, shift=0.0)
    G = fortran_clone(K)

    def cond_at(s):
        K.copy_(G)
        _call(lib().hodlr_shift_leaf_diagonal, N, leaf, float(s), _ptr(K), _ld(K), _stream(),
              where="hodlr_shift_leaf_diagonal")
        return condition_estimate(K, U, V, leaf, iters=power_iters, seed=seed)[2]

    lo = conditioned_shift(N, leaf, target_cond, seed=seed)[0]
    c_lo = cond_at(lo)
    hi = lo
    c_hi = c_lo
    while c_hi > target_cond:
        lo, c_lo = hi, c_hi
        hi = hi * 1.5 + 1.0
        c_hi = cond_at(hi)
    if c_lo <= target_cond:  # the per-leaf shift already meets the target
        return float(lo), float(c_lo)
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        c = cond_at(mid)
        if c > target_cond:
            lo = mid
        else:
            hi, c_hi = mid, c
    del K, U, V, G
    return float(hi), float(c_hi)


def batch_systems(nsys=256, seed=0, n_range=(4096, 65536), leaves=(128, 256, 512),
                  ranks=(16, 100)):
    """The uneven batch: per system N log-uniform in n_range rounded to a
    multiple of the leaf size, leaf size drawn from `leaves`, rank uniform in
    `ranks` (inclusive); numpy's Philox stream of `seed` draws them.  A
    system of P leaves sits in a tree of 2^ceil(log2 P) leaf slots whose
    last slots are empty (HODLR_PLAN_TRAILING_EMPTY): `n` = P * leaf rows,
    no padding."""
    rng = np.random.Generator(np.random.Philox(seed))
    out = []
    for _ in range(nsys):
        nraw = float(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))
        m = int(leaves[int(rng.integers(len(leaves)))])
        r = int(rng.integers(ranks[0], ranks[1] + 1))
        P = max(2, int(round(nraw / m)))
        L = int(np.ceil(np.log2(P)))
        out.append(dict(n=P * m, leaf=m, levels=L, rank=r, nrhs_max=1))
    return out


def batch_leaf_sizes(sy):
    """Leaf-slot sizes of a batch system (empty trailing slots as 0)."""
    P = sy["n"] // sy["leaf"]
    return [sy["leaf"]] * P + [0] * ((1 << sy["levels"]) - P)


def batch_operands(systems, seed=0, pad=0):
    """Device operands of every system (generators of the Gaussian workload
    with the system index in the counter: K_i = G_i + m I, U, V iid
    N(0,1)/sqrt(r), one right-hand side iid U[0,1))."""
    Ks, Us, Vs, Bs = [], [], [], []
    for k, sy in enumerate(systems):
        K, U, V, B = workload_gaussian(sy["n"], sy["leaf"], sy["rank"], 1, seed=seed, system=k,
                                       pad=pad)
        Ks.append(K); Us.append(U); Vs.append(V); Bs.append(B)
    return Ks, Us, Vs, Bs
