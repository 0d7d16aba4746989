"""The C++ facade tests (tests/cpp/test_facade.cpp, built by build()): the
reference's tests/test.cc checks -- matrix initialization and dense blocks,
gemm reduce, gemm broadcast, node solve -- plus an end-to-end HMatrix solve,
through hodlr::LMatrix / hodlr::HMatrix on the device (gpu marker)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_facade")


def test_cpp_facade(gpu):
    assert os.path.exists(BIN), "build() compiles tests/cpp/test_facade"
    print("case: C++ facade tests", flush=True)
    res = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0
    for what in ("legion matrix initialization", "gemm reduce", "gemm broadcast", "node solve",
                 "HMatrix solve"):
        assert f"Test for {what} passed!" in res.stdout
