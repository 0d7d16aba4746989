"""Synthetic HODLR workloads beyond the plain Gaussian one.

``conditioned_shift`` calibrates the diagonal shift s of the leaf blocks
K_i = G_i + s*I (G iid N(0,1), m x m) so that their 2-norm condition number
is about a target (the mixed-precision workload asks for cond ~ 1e4).  The
blocks are the device generator's (same seed, stream 0 = K), so the
calibration sees exactly the matrices the solver will factor; the condition
number is taken as the maximum over the first few leaves.
"""
from __future__ import annotations

import numpy as np
import torch

from . import generate

__all__ = ["conditioned_shift"]


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
