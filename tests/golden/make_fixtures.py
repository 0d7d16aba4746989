"""Write the golden fixtures of the HODLR solve (run from the repo root:
``python tests/golden/make_fixtures.py``).

Each fixture is a small system defined by generator parameters -- the
Gaussian workload (Philox4x64-10 streams) or the reference generator
(drand48 seeds, K_i = U_i V_i^T + diag(D_i + offset), src/dense_block.cc:54-75)
-- together with the SHA-256 of the generated inputs and the expected
solution.  The expected solution is the oracle's restated reference solve
(oracle/hodlr_oracle.c) and is written only if it agrees with numpy's dense
LU solve of the assembled matrix to 1e-12 relative.  The fixtures are data:
parameters, checksums and solutions (fixtures.json, solutions.npz).
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

CASES = [
    # (mode, N, leaf, rank, nrhs, seed)
    ("gaussian", 64, 16, 4, 1, 0),
    ("gaussian", 512, 64, 8, 2, 1),
    ("gaussian", 2048, 128, 16, 3, 2),
    ("gaussian", 2048, 64, 16, 1, 3),
    ("reference", 64, 16, 4, 1, 10),
    ("reference", 512, 64, 8, 2, 11),
    ("reference", 2048, 128, 16, 1, 12),
]


def inputs(mode, N, m, r, s, seed):
    if mode == "gaussian":
        return O.workload_gaussian(N, m, r, s, seed=seed)
    npart = N // m
    off = 10 * m  # diagonal offset D + off keeps the leaf blocks well conditioned
    U = O.ref_matrix(seed, npart, m, r)
    V = O.ref_matrix(seed + 1, npart, m, r)
    K = O.ref_dense_blocks(npart, m, r, off, O.ref_seeds(seed, npart),
                           O.ref_seeds(seed + 1, npart), O.ref_seeds(seed + 2, npart))
    B = np.asfortranarray(O.ref_matrix(seed + 3, npart, m, s))
    return K, U, V, B


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.asfortranarray(a, dtype=np.float64).tobytes(order="F"))
    return h.hexdigest()


def main():
    O.lib()
    meta, sols = [], {}
    for k, (mode, N, m, r, s, seed) in enumerate(CASES):
        K, U, V, B = inputs(mode, N, m, r, s, seed)
        x = O.hodlr_solve(K, U, V, B, m)
        A = O.dense_assemble(K, U, V, m)
        xd = np.linalg.solve(A, B)
        err = float(np.abs(x - xd).max() / np.abs(xd).max())
        assert err <= 1e-12, (k, err)
        name = f"{mode}_n{N}_leaf{m}_rank{r}_nrhs{s}_seed{seed}"
        meta.append(dict(name=name, mode=mode, n=N, leaf=m, rank=r, nrhs=s, seed=seed,
                         input_sha256=digest(K, U, V, B), dense_lu_rel_err=err))
        sols[name] = x
        print(f"{name}: dense-LU agreement {err:.1e}")
    with open(os.path.join(HERE, "fixtures.json"), "w") as f:
        json.dump(meta, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "solutions.npz"), **sols)


if __name__ == "__main__":
    main()
