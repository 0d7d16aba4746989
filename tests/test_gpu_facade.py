"""The C++ facade tests (tests/cpp/test_facade.cpp, built by build()): the
reference's tests/test.cc checks -- matrix initialization and dense blocks,
gemm reduce, gemm broadcast, node solve, their two-level-partition forms,
the leaf-task + level-loop solver of test_solver -- plus an end-to-end
HMatrix solve, through hodlr::LMatrix / hodlr::HMatrix on the device (gpu
marker).  The program also regenerates one reference-mode golden fixture
through hodlr::Matrix::rand / Vector::rand with the generator clock set; its
arrays are compared here with the oracle's drand48 generators (bit for bit)
and its solution with the fixture's."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "test_facade")
GOLD = os.path.join(HERE, "golden")


def test_cpp_facade(gpu, orc, tmp_path):
    assert os.path.exists(BIN), "build() compiles tests/cpp/test_facade"
    print("case: C++ facade tests", flush=True)
    res = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0
    for what in ("legion matrix initialization", "gemm reduce", "gemm broadcast", "node solve",
                 "two-level gemm reduce", "two-level gemm broadcast", "two-level node solve",
                 "solver (rank 8, tree level 4, launch level 2)",
                 "solver (rank 5, tree level 3, launch level 3)",
                 "solver (rank 6, tree level 3, launch level 0)", "HMatrix solve"):
        assert f"Test for {what} passed!" in res.stdout, what

    # reference_n512_leaf64_rank8_nrhs2_seed11 (tests/golden/make_fixtures.py)
    seed, npart, m, r, s = 11, 8, 64, 8, 2
    N = npart * m

    def load(name, cols):
        a = np.fromfile(os.path.join(tmp_path, name + ".f64"), dtype=np.float64)
        return a.reshape((N, cols), order="F") if cols else a

    assert np.array_equal(load("U", r), orc.ref_matrix(seed, npart, m, r))
    assert np.array_equal(load("V", r), orc.ref_matrix(seed + 1, npart, m, r))
    assert np.array_equal(load("D", 0), orc.ref_vector(seed + 2, npart, m, 10 * m).ravel())
    assert np.array_equal(load("b", s), orc.ref_matrix(seed + 3, npart, m, s))
    name = f"reference_n{N}_leaf{m}_rank{r}_nrhs{s}_seed{seed}"
    assert any(f["name"] == name for f in json.load(open(os.path.join(GOLD, "fixtures.json"))))
    xo = np.load(os.path.join(GOLD, "solutions.npz"))[name]
    x = load("x", s)
    assert np.abs(x - xo).max() <= 1e-12 * np.abs(xo).max()
