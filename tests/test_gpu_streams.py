"""Library scratch is kept per (device, stream) (gpu marker).

Entry points without a plan (matvec, gradients) draw scratch from the
library.  Calls issued on two streams at once, with no synchronization
between them and with scratch needs of different sizes, must give exactly
the results of the same calls run one after another on one stream; and a
call that would have to grow its stream's scratch during stream capture
fails cleanly with HODLR_ERR_STATE instead of allocating inside the graph.
"""
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_host(t):
    torch.cuda.synchronize()
    return np.asfortranarray(t.detach().cpu().numpy())


def _work(F, sysA, sysB):
    """matvec and grad of system A and of system B (returns host arrays)."""
    out = []
    for (K, U, V, X, m) in (sysA, sysB):
        Y = F.matvec(K, U, V, X, m)
        gK, gU, gV = F.grad(U, V, X, Y, m)
        out.append((Y, gU, gV))
    return out


def test_two_streams_match_one_stream(gpu):
    F = gpu
    say("case: matvec + grad on two streams vs one stream")
    # different sizes so the two streams' scratch needs differ
    A = F.workload_gaussian(16384, 256, 32, 6, seed=31)
    B = F.workload_gaussian(65536, 512, 64, 3, seed=32)
    sysA = (A[0], A[1], A[2], A[3], 256)
    sysB = (B[0], B[1], B[2], B[3], 512)
    torch.cuda.synchronize()
    ref = [tuple(to_host(t) for t in r) for r in _work(F, sysA, sysB)]
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    for rep in range(4):
        outs = []
        for k in range(3):
            with torch.cuda.stream(s1):
                Y = F.matvec(*sysA[:4], sysA[4])
                gA = F.grad(sysA[1], sysA[2], sysA[3], Y, sysA[4])
            with torch.cuda.stream(s2):
                Z = F.matvec(*sysB[:4], sysB[4])
                gB = F.grad(sysB[1], sysB[2], sysB[3], Z, sysB[4])
            outs.append(((Y, gA[1], gA[2]), (Z, gB[1], gB[2])))
        s1.synchronize()
        s2.synchronize()
        for (oa, ob) in outs:
            for got, want in zip(oa, ref[0]):
                assert np.array_equal(to_host(got), want), rep
            for got, want in zip(ob, ref[1]):
                assert np.array_equal(to_host(got), want), rep
    for s in (s1, s2):
        with torch.cuda.stream(s):
            assert F.lib().hodlr_workspace_release(F._stream()) == 0
    torch.cuda.synchronize()


def test_scratch_growth_refused_during_capture(gpu):
    """A fresh stream has no scratch: a matvec captured on it cannot grow it
    (HODLR_ERR_STATE, no allocation node); after one eager call on the same
    stream the captured call works and its replays match the eager result."""
    F = gpu
    say("case: scratch growth during capture")
    K, U, V, X = F.workload_gaussian(8192, 256, 16, 2, seed=33)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    Y = F.fortran_empty(8192, 2)
    with torch.cuda.stream(s):
        g = torch.cuda.CUDAGraph()
        with pytest.raises(F.HodlrError) as ei:
            with torch.cuda.graph(g, stream=s):
                F.matvec(K, U, V, X, 256, Y=Y)
        assert ei.value.status == -6
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        ref = to_host(F.matvec(K, U, V, X, 256))
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            F.matvec(K, U, V, X, 256, Y=Y)
    for rep in range(3):
        Y.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(to_host(Y), ref), rep
        junk = F.matvec(K, U, V, X, 256)  # eager work between replays
        torch.cuda.synchronize()
        del junk
