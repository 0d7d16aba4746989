"""Golden fixtures (tests/golden/): generator parameters, input checksums and
expected solutions written by tests/golden/make_fixtures.py from the oracle
after a dense-LU cross-check.  The CPU tests pin the generators (checksums)
and the oracle (solutions); the GPU test holds the device solve to the same
solutions at the FP64 parity bar (1e-10 relative)."""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
sys.path.insert(0, GOLD)

import make_fixtures as MF  # noqa: E402

FIXTURES = json.load(open(os.path.join(GOLD, "fixtures.json")))
IDS = [f["name"] for f in FIXTURES]


@pytest.fixture(scope="module")
def solutions():
    with np.load(os.path.join(GOLD, "solutions.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _inputs(f):
    return MF.inputs(f["mode"], f["n"], f["leaf"], f["rank"], f["nrhs"], f["seed"])


@pytest.mark.parametrize("f", FIXTURES, ids=IDS)
def test_generators_reproduce_fixture_inputs(orc, f):
    assert MF.digest(*_inputs(f)) == f["input_sha256"]


@pytest.mark.parametrize("f", FIXTURES, ids=IDS)
def test_oracle_reproduces_fixture_solution(orc, solutions, f):
    K, U, V, B = _inputs(f)
    x = orc.hodlr_solve(K, U, V, B, f["leaf"])
    ref = solutions[f["name"]]
    assert np.abs(x - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.gpu
@pytest.mark.parametrize("f", FIXTURES, ids=IDS)
def test_device_solve_matches_fixture(gpu, solutions, f):
    import torch
    F = gpu
    print(f"case: golden {f['name']}", flush=True)
    K, U, V, B = _inputs(f)

    def dev(a):
        t = F.fortran_empty(a.shape[0], a.shape[1])
        t.copy_(torch.from_numpy(np.ascontiguousarray(a)))
        return t

    Kd, Ud, Vd, Bd = dev(K), dev(U), dev(V), dev(B)
    N, m, r, s = f["n"], f["leaf"], f["rank"], f["nrhs"]
    plan = F.HodlrPlan(N, m, int(round(np.log2(N // m))), r, nrhs_max=s)
    plan.factor_solve(Kd, Ud, Vd, Bd)
    assert plan.status()[0] == 0
    torch.cuda.synchronize()
    x = Bd.cpu().numpy()
    ref = solutions[f["name"]]
    assert np.abs(x - ref).max() <= 1e-10 * np.abs(ref).max()
    plan.close()


REF_FIXTURES = [f for f in FIXTURES if f["mode"] == "reference"]


@pytest.mark.gpu
@pytest.mark.parametrize("f", REF_FIXTURES, ids=[f["name"] for f in REF_FIXTURES])
def test_device_reference_generator_reproduces_fixture(gpu, solutions, f):
    """The reference-mode fixtures generated on the device from their seeds
    alone (hodlr_generate_ref: drand48 streams, hodlr_dense_blocks:
    K_i = U_i V_i^T + diag(D_i)): the same bytes as the fixture -- its
    SHA-256 over (K, U, V, B) -- then solved to the fixture's solution."""
    import torch
    F = gpu
    print(f"case: device reference generator {f['name']}", flush=True)
    N, m, r, s, seed = f["n"], f["leaf"], f["rank"], f["nrhs"], f["seed"]
    npart = N // m
    U = F.generate_ref(seed, npart, m, r)
    V = F.generate_ref(seed + 1, npart, m, r)
    D = F.generate_ref(seed + 2, npart, m, 1, offset=10 * m)
    B = F.generate_ref(seed + 3, npart, m, s)
    K = F.dense_blocks(U, V, D, m)
    torch.cuda.synchronize()
    host = [np.asfortranarray(t.cpu().numpy()) for t in (K, U, V, B)]
    want = MF.inputs("reference", N, m, r, s, seed)
    for name, got, exp in zip("KUVB", host, want):
        assert np.array_equal(got, exp), (name, float(np.abs(got - exp).max()))
    assert MF.digest(*host) == f["input_sha256"]
    plan = F.HodlrPlan(N, m, int(round(np.log2(npart))), r, nrhs_max=s)
    plan.factor_solve(K, U, V, B)
    assert plan.status()[0] == 0
    torch.cuda.synchronize()
    ref = solutions[f["name"]]
    assert np.abs(B.cpu().numpy() - ref).max() <= 1e-10 * np.abs(ref).max()
    plan.close()
