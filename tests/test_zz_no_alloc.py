"""hodlr_matvec_noalloc: the matvec with a fixed device scratch array and no
device allocation (gpu marker).

Through the C ABI, two calls in a row -- the second on a smaller problem,
so it runs on what the first left in the shared scratch -- each equal to
hodlr_matvec on the same operands bit for bit (the same kernels and the
same summation order); a problem whose scratch exceeds the fixed array is
rejected before anything is enqueued.
"""
import ctypes as C
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _call(F, name, K, U, V, X, Y, leaf, alpha, beta):
    fn = getattr(F.lib(), name)
    return fn(K.shape[0], leaf, U.shape[1], X.shape[1], F._ptr(K), F._ld(K), F._ptr(U), F._ld(U),
              F._ptr(V), F._ld(V), F._ptr(X), F._ld(X), alpha, beta, F._ptr(Y), F._ld(Y),
              F._stream())


def test_noalloc_matvec_matches_matvec(gpu):
    F = gpu
    # (N, leaf, rank, nrhs, seed): the second is smaller than the first
    cases = [(262144, 512, 64, 8, 41), (16384, 256, 30, 3, 42)]
    for N, m, r, s, seed in cases:
        say(f"case: noalloc matvec N={N} leaf={m} rank={r} nrhs={s}")
        K, U, V, X = F.workload_gaussian(N, m, r, s, seed=seed, pad=8)
        Y0 = F.generate(seed, 7, 0, 1, N, s, 1.0, pad=8)
        Ya, Yb = F.fortran_clone(Y0), F.fortran_clone(Y0)
        assert _call(F, "hodlr_matvec", K, U, V, X, Ya, m, -0.5, 2.0) == 0
        assert _call(F, "hodlr_matvec_noalloc", K, U, V, X, Yb, m, -0.5, 2.0) == 0
        torch.cuda.synchronize()
        assert torch.equal(Ya, Yb)
        assert Yb.abs().max().item() > 0.0


def test_noalloc_matvec_rejects_oversize(gpu):
    F = gpu
    say("case: noalloc matvec beyond its scratch")
    N, m, r, s = 1 << 20, 128, 64, 2   # (8192 + 1) * 64 * 2 > 2^20 doubles
    p = C.c_void_p(256)
    st = F.lib().hodlr_matvec_noalloc(N, m, r, s, p, N, p, N, p, N, p, N, 1.0, 0.0, p, N,
                                      F._stream())
    assert st == -3
    assert b"noalloc" in F.lib().hodlr_last_error_detail()
