"""C-ABI checks that need no GPU: the library loads, exports every entry
point include/hodlr.h declares, and rejects bad arguments before touching a
device (the reference aborts via assert; the library returns a status)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(hodlr_[a-z0-9_]+)\s*\(", txt))
    return names


@pytest.fixture(scope="module")
def L():
    import fastsolver2_amd as F
    return F.lib()


def test_exports_every_declared_symbol(L):
    names = _declared()
    assert len(names) >= 18
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_strerror_and_version(L):
    import fastsolver2_amd as F
    assert b"gfx950" in L.hodlr_version()
    for code in (0, -1, -2, -3, -4, -5, -6, 1, 2):
        assert L.hodlr_strerror(code) not in (None, b"")
    assert L.hodlr_strerror(12345) == b"unknown status"


def test_plan_create_validates_shape(L):
    import fastsolver2_amd as F
    h = C.c_void_p()
    bad = [
        F._Desc(1000, 64, 4, 8, 1, 0, 0, 0),   # N != leaf * 2^levels  -> SHAPE
        F._Desc(1024, 8, 7, 8, 1, 0, 0, 0),    # leaf <= rank          -> SHAPE
        F._Desc(1024, 0, 4, 8, 1, 0, 0, 0),    # leaf <= 0             -> ARG
        F._Desc(1024, 64, 4, 8, 1, 0, 0, 1),   # flags != 0            -> ARG
        F._Desc(4096 << 1, 4096, 1, 8, 1, 0, 0, 0),  # leaf too large  -> UNSUPPORTED
    ]
    want = [-2, -2, -1, -1, -3]
    for d, w in zip(bad, want):
        assert L.hodlr_plan_create(C.byref(h), C.byref(d)) == w
        assert not h.value


def test_operator_argument_checks(L):
    nul = C.c_void_p(0)
    # hsolve: nPart not a power of two / nrow not divisible (leaf_solve.cc:64,95-96)
    assert L.hodlr_hsolve_ref(100, 1, 4, 3, 100, C.c_void_p(8), C.c_void_p(8), C.c_void_p(8),
                              nul) == -2
    assert L.hodlr_hsolve_ref(100, 1, 4, 2, 50, C.c_void_p(8), C.c_void_p(8), C.c_void_p(8),
                              nul) == -1
    # gemmRed: rows % nchunk, nchunk % nslot (lmatrix.cc:725-726)
    p = C.c_void_p(8)
    assert L.hodlr_gemm_red(100, 3, 1, 1.0, p, 100, 2, p, 100, 2, p, 2, nul) == -2
    assert L.hodlr_gemm_red(96, 6, 4, 1.0, p, 96, 2, p, 96, 2, p, 8, nul) == -2
    assert L.hodlr_gemm_bro(96, 6, 4, -1.0, p, 96, 2, p, 8, p, 96, 2, nul) == -2
    assert L.hodlr_node_solve(0, 3, p, 6, p, 6, 1, p, nul) == -1
    assert L.hodlr_matvec(100, 30, 2, 1, p, 100, p, 100, p, 100, p, 100, 1.0, 0.0, p, 100,
                          nul) == -1
    assert L.hodlr_generate(0, 0, 0, 7, 10, 1, 1.0, p, 10, nul) == -1
