import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


def pytest_collection_modifyitems(config, items):
    # GPU tests fail loudly (not skip) when run with -m gpu on a box without
    # the library; without -m gpu they are deselected by the caller.
    pass


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def gpu():
    import torch
    import fastsolver2_amd as F
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    F.lib()  # raises if the HIP library is not built
    torch.cuda.init()
    return F
