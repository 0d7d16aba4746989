"""fastsolver2_amd -- MI355X-native HODLR factor-and-solve.

Python host side over the C ABI of ``libhodlr_amd.so`` (include/hodlr.h).
Device memory and streams come from PyTorch (ROCm); every numerical step
runs in the library's HIP kernels.  There is no CPU fallback: importing
works without a GPU, but every compute call raises if the HIP library or a
GPU is missing.

Matrices are FP64 and column-major, as in the reference
(src/ptr_matrix.cc:29-37): an N x k operand is a torch tensor of shape
(N, k) with strides (1, ld), e.g. ``fortran_empty(N, k)``.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import torch

__all__ = [
    "lib", "lib_path", "available", "fortran_empty", "fortran", "HodlrPlan", "hsolve",
    "gemm_red", "gemm_bro", "node_solve", "node_solve_region", "matvec", "generate_ref",
    "dense_blocks", "generate", "workload_gaussian",
    "dgesv_batched", "HodlrError", "grad", "HodlrSolveFunction", "hodlr_solve", "fortran_clone",
    "FP64", "MIXED_IR", "PLAN_TRAILING_EMPTY", "IRStats", "HodlrBatch", "HodlrPair",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
lib_path = os.path.join(_HERE, "libhodlr_amd.so")
_lib = None

HODLR_OK = 0
HODLR_SINGULAR_LEAF = 1
HODLR_SINGULAR_CORE = 2


class HodlrError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = status
        msg = lib().hodlr_strerror(status).decode()
        detail = lib().hodlr_last_error_detail().decode()
        super().__init__(f"{where}: status {status} ({msg}){' - ' + detail if detail else ''}")


class IRStats(C.Structure):
    """hodlr_ir_stats: refinement outcome (include/hodlr.h)."""
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("residual", C.c_double),
                ("history", C.c_double * 33)]

    def as_dict(self):
        return {"iterations": self.iterations, "converged": bool(self.converged),
                "residual": self.residual,
                "history": [self.history[i] for i in range(self.iterations + 1)]}


FP64 = 0
MIXED_IR = 1
PLAN_TRAILING_EMPTY = 1


class _Desc(C.Structure):
    _fields_ = [("n", C.c_int64), ("leaf", C.c_int), ("levels", C.c_int), ("rank", C.c_int),
                ("nrhs_max", C.c_int), ("precision", C.c_int), ("device", C.c_int),
                ("flags", C.c_int)]


def lib():
    """Load libhodlr_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(lib_path):
        raise RuntimeError(f"{lib_path} is missing: build it with `python -m fastsolver2_amd.build`")
    L = C.CDLL(lib_path)
    vp, i64, dbl, ci = C.c_void_p, C.c_int64, C.c_double, C.c_int
    sig = {
        "hodlr_version": ([], C.c_char_p),
        "hodlr_strerror": ([ci], C.c_char_p),
        "hodlr_last_error_detail": ([], C.c_char_p),
        "hodlr_plan_create": ([C.POINTER(vp), C.POINTER(_Desc)], ci),
        "hodlr_plan_create_uneven": ([C.POINTER(vp), C.POINTER(_Desc), C.POINTER(ci)], ci),
        "hodlr_plan_matvec": ([vp, ci, ci, vp, i64, vp, i64, vp, i64, vp, i64, dbl, dbl, vp, i64,
                               vp], ci),
        "hodlr_pair_create": ([C.POINTER(vp), C.POINTER(_Desc), vp, C.POINTER(_Desc), vp, ci], ci),
        "hodlr_pair_destroy": ([vp], ci),
        "hodlr_pair_plan": ([vp, ci], vp),
        "hodlr_pair_factor_solve": ([vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i64, ci, vp], ci),
        "hodlr_pair_status": ([vp, vp, C.POINTER(ci), C.POINTER(ci), C.POINTER(ci)], ci),
        "hodlr_plan_destroy": ([vp], ci),
        "hodlr_plan_device_bytes": ([vp], i64),
        "hodlr_factor": ([vp, vp, i64, vp, i64, vp, i64, vp], ci),
        "hodlr_solve": ([vp, vp, i64, ci, vp], ci),
        "hodlr_factor_solve": ([vp, vp, i64, vp, i64, vp, i64, vp, i64, ci, vp], ci),
        "hodlr_plan_status": ([vp, vp, C.POINTER(ci), C.POINTER(ci)], ci),
        "hodlr_workspace_release": ([vp], ci),
        "hodlr_solve_adjoint": ([vp, vp, i64, ci, vp], ci),
        "hodlr_batch_create": ([C.POINTER(vp), ci, vp], ci),
        "hodlr_batch_destroy": ([vp], ci),
        "hodlr_batch_size": ([vp], ci),
        "hodlr_batch_plan": ([vp, ci], vp),
        "hodlr_batch_factor_solve": ([vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp], ci),
        "hodlr_batch_status": ([vp, vp, C.POINTER(ci), C.POINTER(ci), C.POINTER(ci)], ci),
        "hodlr_batch_set_tuning": ([vp, ci, ci], ci),
        "hodlr_batch_get_tuning": ([vp, C.POINTER(ci), C.POINTER(ci)], ci),
        "hodlr_batch_tune": ([vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(dbl)], ci),
        "hodlr_solve_ir": ([vp, vp, i64, vp, i64, vp, i64, vp, i64, vp, i64, ci, dbl, ci, ci,
                            C.POINTER(IRStats), vp], ci),
        "hodlr_matvec_t": ([i64, ci, ci, ci, vp, i64, vp, i64, vp, i64, vp, i64, dbl, dbl, vp, i64,
                            vp], ci),
        "hodlr_grad": ([i64, ci, ci, ci, vp, i64, vp, i64, vp, i64, vp, i64, vp, i64, vp, i64,
                        vp, i64, vp], ci),
        "hodlr_matvec": ([i64, ci, ci, ci, vp, i64, vp, i64, vp, i64, vp, i64, dbl, dbl, vp, i64,
                          vp], ci),
        "hodlr_matvec_noalloc": ([i64, ci, ci, ci, vp, i64, vp, i64, vp, i64, vp, i64, dbl, dbl, vp,
                                  i64, vp], ci),
        "hodlr_hsolve_ref": ([ci, ci, ci, ci, ci, vp, vp, vp, vp], ci),
        "hodlr_gemm_red": ([i64, ci, ci, dbl, vp, i64, ci, vp, i64, ci, vp, i64, vp], ci),
        "hodlr_gemm_bro": ([i64, ci, ci, dbl, vp, i64, ci, vp, i64, vp, i64, ci, vp], ci),
        "hodlr_node_solve": ([ci, ci, vp, i64, vp, i64, ci, vp, vp], ci),
        "hodlr_node_solve_region": ([ci, ci, ci, ci, vp, i64, i64, vp, i64, i64, vp, i64, i64,
                                     vp, i64, i64, vp, vp], ci),
        "hodlr_dgetrf_batched": ([ci, vp, i64, i64, ci, vp, vp, vp], ci),
        "hodlr_dgetrf_batched_oop": ([ci, vp, i64, i64, vp, i64, i64, ci, vp, vp, vp], ci),
        "hodlr_dgesv_batched": ([ci, ci, vp, i64, i64, ci, vp, vp, vp, i64, i64, vp], ci),
        "hodlr_generate": ([C.c_uint64, C.c_uint32, C.c_uint32, ci, i64, ci, dbl, vp, i64, vp],
                           ci),
        "hodlr_shift_leaf_diagonal": ([i64, ci, dbl, vp, i64, vp], ci),
        "hodlr_generate_ref": ([C.c_long, ci, i64, ci, dbl, vp, i64, vp], ci),
        "hodlr_selftest_mfma": ([vp, vp], ci),
        "hodlr_selftest_mfma_f32": ([vp, vp], ci),
        "hodlr_dense_blocks": ([i64, ci, ci, vp, i64, vp, i64, vp, dbl, vp, i64, vp], ci),
        "hodlr_profile_enable": ([ci], ci),
        "hodlr_profile_mask": ([C.c_uint], ci),
        "hodlr_profile_reset": ([], ci),
        "hodlr_profile_read": ([ci, C.POINTER(i64), C.POINTER(dbl), C.POINTER(dbl),
                                C.POINTER(dbl)], ci),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def available() -> bool:
    """True when the HIP library is built and a GPU is visible."""
    return os.path.exists(lib_path) and torch.cuda.is_available()


# ---------------------------------------------------------------- tensors
def fortran_empty(rows, cols, device="cuda", dtype=torch.float64, pad=0):
    """Column-major rows x cols tensor with leading dimension rows + pad
    (strides (1, rows + pad)).  A small pad keeps the column stride off a
    power of two, so that a leaf's columns do not alias in the HBM channel
    interleave."""
    if pad <= 0:
        return torch.empty((cols, rows), dtype=dtype, device=device).t()
    return torch.empty((cols, rows + pad), dtype=dtype, device=device).t()[:rows]


def fortran_clone(t: torch.Tensor) -> torch.Tensor:
    """Copy of a column-major operand with the same leading dimension."""
    ld = _ld(t)
    out = fortran_empty(t.shape[0], t.shape[1], device=t.device, dtype=t.dtype,
                        pad=ld - t.shape[0])
    out.copy_(t)
    return out


def fortran(t: torch.Tensor) -> torch.Tensor:
    """Column-major copy of a 2-D tensor (no copy when it already is)."""
    if t.dim() == 1:
        t = t.reshape(-1, 1)
    if t.stride(0) == 1 and (t.shape[1] == 1 or t.stride(1) >= t.shape[0]):
        return t
    return t.t().contiguous().t()


def _ld(t: torch.Tensor) -> int:
    if t.dim() == 1:
        return t.shape[0]
    if t.stride(0) != 1:
        raise ValueError("operand must be column-major (stride(0) == 1)")
    return max(t.stride(1), t.shape[0]) if t.shape[1] > 1 else t.shape[0]


def _check(t: torch.Tensor, name: str):
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if t.dtype != torch.float64:
        raise ValueError(f"{name} must be float64")


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _call(fn, *args, where=""):
    st = fn(*args)
    if st != 0:
        raise HodlrError(st, where or fn.__name__)
    return st


# ---------------------------------------------------------------- plans
class HodlrPlan:
    """Device factorization of one HODLR matrix (the working form of the
    reference's HMatrix, include/hmatrix.hpp:8-34).

    A balanced tree of 2^levels leaves of size `leaf` (n = leaf << levels);
    with n < leaf << levels (a multiple of leaf) the last leaf slots are
    empty (HODLR_PLAN_TRAILING_EMPTY); with `leaf_sizes` (2^levels entries,
    0 for an empty slot) a tree of uneven leaves (hodlr_plan_create_uneven),
    n their sum and `leaf` their maximum."""

    def __init__(self, n, leaf, levels, rank, nrhs_max=1, device=None, precision=FP64,
                 leaf_sizes=None):
        if not torch.cuda.is_available():
            raise RuntimeError("HodlrPlan needs a GPU")
        dev = torch.cuda.current_device() if device is None else int(device)
        h = C.c_void_p()
        if leaf_sizes is not None:
            sizes = [int(x) for x in leaf_sizes]
            if len(sizes) != 1 << levels:
                raise ValueError("leaf_sizes needs 2^levels entries")
            n, leaf = sum(sizes), max(sizes)
            arr = (C.c_int * len(sizes))(*sizes)
            d = _Desc(n, leaf, levels, rank, nrhs_max, precision, dev, 0)
            _call(lib().hodlr_plan_create_uneven, C.byref(h), C.byref(d), arr,
                  where="hodlr_plan_create_uneven")
            self.leaf_sizes = sizes
        else:
            flags = PLAN_TRAILING_EMPTY if n < (leaf << levels) else 0
            d = _Desc(n, leaf, levels, rank, nrhs_max, precision, dev, flags)
            _call(lib().hodlr_plan_create, C.byref(h), C.byref(d), where="hodlr_plan_create")
            nl = n // leaf
            self.leaf_sizes = [leaf] * nl + [0] * ((1 << levels) - nl)
        self.n, self.leaf, self.levels, self.rank, self.nrhs_max = n, leaf, levels, rank, nrhs_max
        self.device = dev
        self.precision = precision
        self._h = h
        self._keep = ()

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().hodlr_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def device_bytes(self):
        return lib().hodlr_plan_device_bytes(self._h)

    def _ops(self, K, U, V):
        for t, nm in ((K, "K"), (U, "U"), (V, "V")):
            _check(t, nm)
        if K.shape != (self.n, self.leaf) or U.shape != (self.n, self.rank) or V.shape != (self.n, self.rank):
            raise ValueError("K must be N x leaf, U and V N x rank")

    def matvec(self, K, U, V, X, alpha=1.0, beta=0.0, Y=None, trans=False):
        """Y = alpha * A X + beta * Y (A^T with trans) for the HODLR matrix
        (K, U, V) on this plan's tree (hodlr_plan_matvec)."""
        self._ops(K, U, V)
        _check(X, "X")
        X2 = X if X.dim() == 2 else X.view(-1, 1)
        if Y is None:
            Y = fortran_empty(X2.shape[0], X2.shape[1], device=X.device)
            beta = 0.0
        Y2 = Y if Y.dim() == 2 else Y.view(-1, 1)
        _call(lib().hodlr_plan_matvec, self._h, 1 if trans else 0, X2.shape[1], _ptr(K), _ld(K),
              _ptr(U), _ld(U), _ptr(V), _ld(V), _ptr(X2), _ld(X2), float(alpha), float(beta),
              _ptr(Y2), _ld(Y2), _stream(), where="hodlr_plan_matvec")
        return Y

    def factor(self, K, U, V):
        """Factor A (K is only read; the leaf factors are kept in the plan,
        and K, U, V stay referenced)."""
        self._ops(K, U, V)
        _call(lib().hodlr_factor, self._h, _ptr(K), _ld(K), _ptr(U), _ld(U), _ptr(V), _ld(V),
              _stream(), where="hodlr_factor")
        self._keep = (K, U, V)

    def factor_solve(self, K, U, V, B):
        """Fused factor + solve; B (N x s) is overwritten by the solution."""
        self._ops(K, U, V)
        _check(B, "B")
        B2 = B if B.dim() == 2 else B.view(-1, 1)
        _call(lib().hodlr_factor_solve, self._h, _ptr(K), _ld(K), _ptr(U), _ld(U), _ptr(V),
              _ld(V), _ptr(B2), _ld(B2), B2.shape[1], _stream(), where="hodlr_factor_solve")
        self._keep = (K, U, V)
        return B

    def solve(self, B):
        """B <- A^-1 B with the stored factorization."""
        _check(B, "B")
        B2 = B if B.dim() == 2 else B.view(-1, 1)
        _call(lib().hodlr_solve, self._h, _ptr(B2), _ld(B2), B2.shape[1], _stream(),
              where="hodlr_solve")
        return B

    def solve_ir(self, K, U, V, B, X=None, tol=1e-12, maxit=10, x_given=False):
        """Iteratively refined solve X = A^-1 B against the ORIGINAL K, U, V
        (FP64 HODLR-matvec residuals); returns (X, stats dict).  With
        x_given, X already holds a first solution (e.g. from factor_solve)."""
        for t, nm in ((K, "K"), (U, "U"), (V, "V"), (B, "B")):
            _check(t, nm)
        B2 = B if B.dim() == 2 else B.view(-1, 1)
        if X is None:
            if x_given:
                raise ValueError("x_given needs X")
            X = fortran_empty(B2.shape[0], B2.shape[1], device=B.device)
        st = IRStats()
        _call(lib().hodlr_solve_ir, self._h, _ptr(K), _ld(K), _ptr(U), _ld(U), _ptr(V), _ld(V),
              _ptr(B2), _ld(B2), _ptr(X), _ld(X), B2.shape[1], float(tol), int(maxit),
              1 if x_given else 0, C.byref(st), _stream(), where="hodlr_solve_ir")
        return X, st.as_dict()

    def solve_adjoint(self, G):
        """G <- A^-T G with the stored factorization."""
        _check(G, "G")
        G2 = G if G.dim() == 2 else G.view(-1, 1)
        _call(lib().hodlr_solve_adjoint, self._h, _ptr(G2), _ld(G2), G2.shape[1], _stream(),
              where="hodlr_solve_adjoint")
        return G

    def status(self, level=False):
        """(status, where) of the last factorization, or (status, where,
        level) with level=True: the leaf index for a singular leaf, the node
        index within tree level `level` (root 0) for a singular core."""
        w, lv = C.c_int(-1), C.c_int(-1)
        st = lib().hodlr_plan_status(self._h, _stream(), C.byref(w), C.byref(lv))
        return (st, w.value, lv.value) if level else (st, w.value)


class HodlrPair:
    """Two sibling shards with different off-diagonal ranks (include/hodlr.h,
    "pairs"; spmd_driver/solver_irregular_ranks.cc): A = [A_0, U_0 W_1^T;
    U_1 W_0^T, A_1].  Each shard is given by its leaf-slot sizes
    (2^levels entries) and rank."""

    def __init__(self, leaf_sizes0, rank0, leaf_sizes1, rank1, nrhs_max=1, device=None):
        if not torch.cuda.is_available():
            raise RuntimeError("HodlrPair needs a GPU")
        dev = torch.cuda.current_device() if device is None else int(device)
        descs, arrs = [], []
        for sizes, r in ((leaf_sizes0, rank0), (leaf_sizes1, rank1)):
            sizes = [int(x) for x in sizes]
            L = int(round(math.log2(len(sizes))))
            if 1 << L != len(sizes):
                raise ValueError("a shard needs 2^levels leaf-slot sizes")
            descs.append(_Desc(sum(sizes), max(sizes), L, r, 0, FP64, dev, 0))
            arrs.append((C.c_int * len(sizes))(*sizes))
        h = C.c_void_p()
        _call(lib().hodlr_pair_create, C.byref(h), C.byref(descs[0]), C.cast(arrs[0], C.c_void_p),
              C.byref(descs[1]), C.cast(arrs[1], C.c_void_p), int(nrhs_max),
              where="hodlr_pair_create")
        self._h = h
        self.n = (descs[0].n, descs[1].n)
        self.rank = (rank0, rank1)
        self.leaf_sizes = ([int(x) for x in leaf_sizes0], [int(x) for x in leaf_sizes1])

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().hodlr_pair_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def factor_solve(self, K, U, V, W, B):
        """K, U, V, W: pairs of device operands (shard 0, shard 1); B
        ((n_0 + n_1) x s) is overwritten by the solution."""
        ops = (K, U, V, W)
        for pr, nm in zip(ops, "KUVW"):
            for t in pr:
                _check(t, nm)
        _check(B, "B")
        B2 = B if B.dim() == 2 else B.view(-1, 1)
        ptrs = [(C.c_void_p * 2)(pr[0].data_ptr(), pr[1].data_ptr()) for pr in ops]
        lds = [(C.c_int64 * 2)(_ld(pr[0]), _ld(pr[1])) for pr in ops]
        args = []
        for p_, l_ in zip(ptrs, lds):
            args += [C.cast(p_, C.c_void_p), C.cast(l_, C.c_void_p)]
        _call(lib().hodlr_pair_factor_solve, self._h, *args, _ptr(B2), _ld(B2), B2.shape[1],
              _stream(), where="hodlr_pair_factor_solve")
        self._keep = (K, U, V, W)
        return B

    def status(self):
        """(status, shard, where, level) of the last factorization."""
        sh, w, lv = C.c_int(-1), C.c_int(-1), C.c_int(-1)
        st = lib().hodlr_pair_status(self._h, _stream(), C.byref(sh), C.byref(w), C.byref(lv))
        return st, sh.value, w.value, lv.value


class HodlrBatch:
    """Independent HODLR systems solved together (include/hodlr.h, batches):
    the batch is captured once as one HIP graph over concurrent lanes and
    replayed while the operand pointers stay the same.  `systems` is a list
    of dicts with keys n, leaf, levels, rank and optionally nrhs_max; a
    system with n < leaf << levels has empty trailing leaf slots."""

    def __init__(self, systems, device=None, precision=FP64):
        if not torch.cuda.is_available():
            raise RuntimeError("HodlrBatch needs a GPU")
        dev = torch.cuda.current_device() if device is None else int(device)
        n = len(systems)
        descs = (_Desc * n)()
        for i, sy in enumerate(systems):
            flags = PLAN_TRAILING_EMPTY if sy["n"] < (sy["leaf"] << sy["levels"]) else 0
            descs[i] = _Desc(sy["n"], sy["leaf"], sy["levels"], sy["rank"], sy.get("nrhs_max", 1),
                             precision, dev, flags)
        h = C.c_void_p()
        _call(lib().hodlr_batch_create, C.byref(h), n, C.cast(descs, C.c_void_p),
              where="hodlr_batch_create")
        self._h = h
        self.n = n
        self.systems = [dict(sy) for sy in systems]

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().hodlr_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _arrays(self, Ks, Us, Vs, Bs):
        n = self.n
        if not (len(Ks) == len(Us) == len(Vs) == len(Bs) == n):
            raise ValueError("one K, U, V, B per system")
        P = C.c_void_p * n
        L = C.c_int64 * n
        I = C.c_int * n
        for t in list(Ks) + list(Us) + list(Vs) + list(Bs):
            _check(t, "operand")
        B2 = [b if b.dim() == 2 else b.view(-1, 1) for b in Bs]
        arr = (P(*[t.data_ptr() for t in Ks]), L(*[_ld(t) for t in Ks]),
               P(*[t.data_ptr() for t in Us]), L(*[_ld(t) for t in Us]),
               P(*[t.data_ptr() for t in Vs]), L(*[_ld(t) for t in Vs]),
               P(*[t.data_ptr() for t in B2]), L(*[_ld(t) for t in B2]),
               I(*[t.shape[1] for t in B2]))
        return [C.cast(a, C.c_void_p) for a in arr], arr

    def factor_solve(self, Ks, Us, Vs, Bs):
        """Every system: K_i is only read (its leaf LU is kept in system i's plan), B_i is overwritten by the solution."""
        ptrs, keep = self._arrays(Ks, Us, Vs, Bs)
        _call(lib().hodlr_batch_factor_solve, self._h, *ptrs, _stream(),
              where="hodlr_batch_factor_solve")
        return Bs

    def tune(self, Ks, Us, Vs, Bs):
        """Autotune (lanes, tile) on scratch operands; returns (lanes, tile, ms)."""
        ptrs, keep = self._arrays(Ks, Us, Vs, Bs)
        best = C.c_double()
        _call(lib().hodlr_batch_tune, self._h, *ptrs, _stream(), C.byref(best),
              where="hodlr_batch_tune")
        lanes, tile = self.get_tuning()
        return lanes, tile, best.value

    def set_tuning(self, lanes, tile=0):
        _call(lib().hodlr_batch_set_tuning, self._h, lanes, tile, where="hodlr_batch_set_tuning")

    def get_tuning(self):
        l, t = C.c_int(), C.c_int()
        _call(lib().hodlr_batch_get_tuning, self._h, C.byref(l), C.byref(t),
              where="hodlr_batch_get_tuning")
        return l.value, t.value

    def status(self):
        """(status, system, where, level) of the first failing system."""
        sy, w, lv = C.c_int(-1), C.c_int(-1), C.c_int(-1)
        st = lib().hodlr_batch_status(self._h, _stream(), C.byref(sy), C.byref(w), C.byref(lv))
        return st, sy.value, w.value, lv.value


# ---------------------------------------------------------------- operators
def hsolve(nrow, nrhs, rank, nPart, K, X, V):
    """Device form of the reference's hsolve (src/tasks/leaf_solve.cc:73-154)
    on column-major device tensors with a common leading dimension."""
    for t, nm in ((K, "K"), (X, "X"), (V, "V")):
        _check(t, nm)
    LD = _ld(K)
    if _ld(X) != LD or _ld(V) != LD:
        raise ValueError("K, X and V must share the leading dimension (leaf_solve.cc:60-61)")
    return lib().hodlr_hsolve_ref(nrow, nrhs, rank, nPart, LD, _ptr(K), _ptr(X), _ptr(V),
                                  _stream())


def gemm_red(A, B, nchunk, nslot, alpha=1.0, C_out=None):
    """LMatrix::gemmRed: C[slot] = alpha * sum_p A_p^T B_p (lmatrix.cc:713-761)."""
    _check(A, "A"); _check(B, "B")
    rows, ka = A.shape
    kb = B.shape[1]
    if C_out is None:
        C_out = fortran_empty(nslot * ka, kb, device=A.device)
    _call(lib().hodlr_gemm_red, rows, nchunk, nslot, alpha, _ptr(A), _ld(A), ka, _ptr(B), _ld(B),
          kb, _ptr(C_out), _ld(C_out), _stream(), where="hodlr_gemm_red")
    return C_out


def gemm_bro(A, E, D, nchunk, nslot, alpha=-1.0):
    """LMatrix::gemmBro: D_p += alpha * A_p E[slot(p)] (lmatrix.cc:814-865)."""
    _check(A, "A"); _check(E, "E"); _check(D, "D")
    rows, k = A.shape
    _call(lib().hodlr_gemm_bro, rows, nchunk, nslot, alpha, _ptr(A), _ld(A), k, _ptr(E), _ld(E),
          _ptr(D), _ld(D), D.shape[1], _stream(), where="hodlr_gemm_bro")
    return D


def node_solve(VTu, VTd, r):
    """LMatrix::node_solve (lmatrix.cc:534-581); returns the per-node LAPACK
    INFO as a device int32 tensor."""
    _check(VTu, "VTu"); _check(VTd, "VTd")
    nnode = VTu.shape[0] // (2 * r)
    info = torch.zeros(nnode, dtype=torch.int32, device=VTu.device)
    _call(lib().hodlr_node_solve, nnode, r, _ptr(VTu), _ld(VTu), _ptr(VTd), _ld(VTd),
          VTd.shape[1], _ptr(info), _stream(), where="hodlr_node_solve")
    return info


def node_solve_region(VTu0, VTu1, VTd0, VTd1):
    """Static LMatrix::node_solve(VTu0, VTu1, VTd0, VTd1) / NodeSolveRegionTask
    (lmatrix.cc:583-636, node_solve_region.cc:35-108) for sibling ranks
    r0 != r1: VTu0 is r1 x r0, VTu1 r0 x r1, VTd0 r1 x nrhs, VTd1 r0 x nrhs;
    VTd1, VTd0 are overwritten by the two halves of the core solution.
    3-D operands (nnode, rows, cols) solve nnode nodes in one call.
    Returns the per-node LAPACK INFO as a device int32 tensor."""
    ops = (VTu0, VTu1, VTd0, VTd1)
    batched = VTu0.dim() == 3
    if batched:
        for t, nm in zip(ops, ("VTu0", "VTu1", "VTd0", "VTd1")):
            if t.dim() != 3 or t.dtype != torch.float64 or not t.is_cuda or t.stride(1) != 1:
                raise ValueError(f"{nm}: expected a CUDA float64 (nnode, rows, cols) tensor "
                                 "with column-major matrices")
        nnode = VTu0.shape[0]
        mats = [(max(t.stride(2), t.shape[1]) if t.shape[2] > 1 else t.shape[1], t.stride(0))
                for t in ops]
    else:
        for t, nm in zip(ops, ("VTu0", "VTu1", "VTd0", "VTd1")):
            _check(t, nm)
        nnode = 1
        mats = [(_ld(t), 0) for t in ops]
    r1, r0 = VTu0.shape[-2], VTu0.shape[-1]
    if (VTu1.shape[-2], VTu1.shape[-1]) != (r0, r1) or VTd0.shape[-2] != r1 or \
            VTd1.shape[-2] != r0 or VTd0.shape[-1] != VTd1.shape[-1]:
        raise ValueError("node_solve_region: VTu0 r1 x r0, VTu1 r0 x r1, VTd0 r1 x nrhs, "
                         "VTd1 r0 x nrhs expected")
    info = torch.zeros(nnode, dtype=torch.int32, device=VTu0.device)
    args = []
    for t, (ld, st) in zip(ops, mats):
        args += [_ptr(t), max(ld, 1), st]
    _call(lib().hodlr_node_solve_region, nnode, r0, r1, VTd0.shape[-1], *args, _ptr(info),
          _stream(), where="hodlr_node_solve_region")
    return info


def matvec(K, U, V, X, leaf, alpha=1.0, beta=0.0, Y=None, trans=False):
    """Y = alpha * A X + beta * Y (A^T with trans) for the HODLR matrix (K, U, V)."""
    for t, nm in ((K, "K"), (U, "U"), (V, "V"), (X, "X")):
        _check(t, nm)
    X2 = X if X.dim() == 2 else X.view(-1, 1)
    if Y is None:
        Y = fortran_empty(X2.shape[0], X2.shape[1], device=X.device)
        beta = 0.0
    Y2 = Y if Y.dim() == 2 else Y.view(-1, 1)
    fn = lib().hodlr_matvec_t if trans else lib().hodlr_matvec
    _call(fn, K.shape[0], leaf, U.shape[1], X2.shape[1], _ptr(K), _ld(K),
          _ptr(U), _ld(U), _ptr(V), _ld(V), _ptr(X2), _ld(X2), alpha, beta, _ptr(Y2), _ld(Y2),
          _stream(), where="hodlr_matvec")
    return Y


def grad(U, V, X, Lam, leaf, dK=True, dU=True, dV=True):
    """Gradients of a loss l(X), X = A^-1 B, given Lam = A^-T dl/dX:
    dK_i = -Lam_i X_i^T, dU_i = -Lam_i (V^T X - V_i^T X_i)^T,
    dV_i = -X_i (U^T Lam - U_i^T Lam_i)^T.  Returns (dK, dU, dV); a gradient
    not asked for is None."""
    for t, nm in ((U, "U"), (V, "V"), (X, "X"), (Lam, "Lam")):
        _check(t, nm)
    N, r = U.shape
    X2 = X if X.dim() == 2 else X.view(-1, 1)
    L2 = Lam if Lam.dim() == 2 else Lam.view(-1, 1)
    gK = fortran_empty(N, leaf, device=U.device) if dK else None
    gU = fortran_empty(N, r, device=U.device) if dU else None
    gV = fortran_empty(N, r, device=U.device) if dV else None
    nul = C.c_void_p(None)
    _call(lib().hodlr_grad, N, leaf, r, X2.shape[1], _ptr(U), _ld(U), _ptr(V), _ld(V), _ptr(X2),
          _ld(X2), _ptr(L2), _ld(L2), _ptr(gK) if dK else nul, _ld(gK) if dK else N,
          _ptr(gU) if dU else nul, _ld(gU) if dU else N, _ptr(gV) if dV else nul,
          _ld(gV) if dV else N, _stream(), where="hodlr_grad")
    return gK, gU, gV


class HodlrSolveFunction(torch.autograd.Function):
    """X = A^-1 B for the HODLR matrix A(K, U, V), differentiable in K, U, V
    and B.  Forward: fused factor + solve on a copy of K; backward: the
    adjoint solve Lam = A^-T dX with the stored factorization, then
    dB = Lam and the rank-k / leaf outer-product gradients (hodlr_grad)."""

    @staticmethod
    def forward(ctx, K, U, V, B, leaf):
        N = K.shape[0]
        levels = int(round(math.log2(N // leaf)))
        s = B.shape[1] if B.dim() == 2 else 1
        Kf = fortran_empty(N, leaf, device=K.device)  # column-major copy; the plan only reads it
        Kf.copy_(K.detach())
        Uf, Vf = fortran(U.detach()), fortran(V.detach())
        X = fortran_empty(N, s, device=B.device)
        X.copy_(B.detach().reshape(N, s))
        plan = HodlrPlan(N, leaf, levels, U.shape[1], nrhs_max=s)
        plan.factor_solve(Kf, Uf, Vf, X)
        st, where = plan.status()
        if st != 0:
            plan.close()
            raise HodlrError(st, f"factor (block {where})")
        ctx.plan = plan
        ctx.leaf = leaf
        ctx.bshape = B.shape
        ctx.save_for_backward(Uf, Vf, X)
        return X if B.dim() == 2 else X.view(-1)

    @staticmethod
    def backward(ctx, gX):
        Uf, Vf, X = ctx.saved_tensors
        N = X.shape[0]
        Lam = fortran_empty(N, X.shape[1], device=X.device)
        Lam.copy_(gX.reshape(N, X.shape[1]))
        ctx.plan.solve_adjoint(Lam)
        need = ctx.needs_input_grad
        gK, gU, gV = grad(Uf, Vf, X, Lam, ctx.leaf, dK=need[0], dU=need[1], dV=need[2])
        gB = Lam.view(ctx.bshape) if need[3] else None
        return gK, gU, gV, gB, None


def hodlr_solve(K, U, V, B, leaf):
    """Differentiable HODLR solve X = A(K, U, V)^-1 B (torch.autograd)."""
    return HodlrSolveFunction.apply(K, U, V, B, leaf)


def dgesv_batched(A, B, n, batch):
    """Batched dgesv on column-major blocks stacked by columns:
    A is n x (n*batch), B is n x (nrhs*batch).  Returns (info, ipiv): the
    per-item LAPACK INFO (batch int32) and the 1-based pivots (batch*n)."""
    _check(A, "A"); _check(B, "B")
    ipiv = torch.empty(batch * n, dtype=torch.int32, device=A.device)
    info = torch.zeros(batch, dtype=torch.int32, device=A.device)
    nrhs = B.shape[1] // batch
    _call(lib().hodlr_dgesv_batched, n, nrhs, _ptr(A), _ld(A), _ld(A) * n, batch, _ptr(ipiv),
          _ptr(info), _ptr(B), _ld(B), _ld(B) * nrhs, _stream(), where="hodlr_dgesv_batched")
    return info, ipiv


def dgetrf_batched_oop(K, n, batch, ldl=None):
    """Out-of-place batched dgetrf of the n x n blocks stacked by rows in K
    (K is (batch*n) x n, block b = rows b*n..); K is not modified.  Returns
    (LU, ipiv, info): LU as a (batch, n, ldl) tensor whose [b, :, :n] is
    block b's factor in column-major order (transpose to index [row, col]),
    the 1-based pivots (batch x n) and the per-block INFO."""
    _check(K, "K")
    ldl = n + 8 if ldl is None else ldl
    LU = torch.zeros(batch, n, ldl, dtype=torch.float64, device=K.device)
    ipiv = torch.zeros(batch, n, dtype=torch.int32, device=K.device)
    info = torch.zeros(batch, dtype=torch.int32, device=K.device)
    _call(lib().hodlr_dgetrf_batched_oop, n, _ptr(K), _ld(K), n, _ptr(LU), ldl, n * ldl, batch,
          _ptr(ipiv), _ptr(info), _stream(), where="hodlr_dgetrf_batched_oop")
    return LU, ipiv, info


def generate(seed, stream_id, system, kind, rows, cols, scale=1.0, out=None, device="cuda",
             pad=0):
    """Philox4x64-10 generator on the device (kind 0 uniform, 1 normal)."""
    if out is None:
        out = fortran_empty(rows, cols, device=device, pad=pad)
    _call(lib().hodlr_generate, seed, stream_id, system, kind, rows, cols, scale, _ptr(out),
          _ld(out), _stream(), where="hodlr_generate")
    return out


def generate_ref(master, npart, rows_per_part, cols, offset=0.0, out=None, device="cuda"):
    """The reference's drand48 generators on the device (hodlr_generate_ref):
    Matrix::rand (cols > 1) / Vector::rand (cols = 1) with `master` in place
    of time(NULL)."""
    if out is None:
        out = fortran_empty(npart * rows_per_part, cols, device=device)
    _call(lib().hodlr_generate_ref, int(master), npart, rows_per_part, cols, float(offset),
          _ptr(out), _ld(out), _stream(), where="hodlr_generate_ref")
    return out


def dense_blocks(U, V, D, leaf, offset=0.0, K=None):
    """DenseBlockTask / LMatrix::init_dense_blocks: K_i = U_i V_i^T +
    diag(D_i + offset) (hodlr_dense_blocks)."""
    for t, nm in ((U, "U"), (V, "V"), (D, "D")):
        _check(t, nm)
    n, r = U.shape
    if K is None:
        K = fortran_empty(n, leaf, device=U.device)
    _call(lib().hodlr_dense_blocks, n, leaf, r, _ptr(U), _ld(U), _ptr(V), _ld(V), _ptr(D),
          float(offset), _ptr(K), _ld(K), _stream(), where="hodlr_dense_blocks")
    return K


def workload_gaussian(N, leaf, r, s, seed=0, system=0, shift=None, rhs="uniform",
                      device="cuda", pad=0):
    """Seeded Gaussian HODLR workload generated on the device:
    K_i = G_i + shift*I (G iid N(0,1)), U, V iid N(0,1)/sqrt(r), B iid U[0,1)
    (rhs='uniform') or N(0,1) (rhs='normal').  Streams 0..3 = K, U, V, B.
    The values do not depend on `pad` (the leading dimension is N + pad)."""
    shift = float(leaf) if shift is None else float(shift)
    K = generate(seed, 0, system, 1, N, leaf, 1.0, device=device, pad=pad)
    _call(lib().hodlr_shift_leaf_diagonal, N, leaf, shift, _ptr(K), _ld(K), _stream(),
          where="hodlr_shift_leaf_diagonal")
    sc = 1.0 / math.sqrt(float(r))
    U = generate(seed, 1, system, 1, N, r, sc, device=device, pad=pad)
    V = generate(seed, 2, system, 1, N, r, sc, device=device, pad=pad)
    B = generate(seed, 3, system, 0 if rhs == "uniform" else 1, N, s, 1.0, device=device,
                 pad=pad)
    return K, U, V, B


# ---------------------------------------------------------------- profiling
PROF_KINDS = ["getrf_leaf", "getrs_leaf", "trtri", "reduce", "assemble", "getrf_core",
              "getrs_core", "update", "leaf_level"]


def profile_enable(on=True, kinds=None):
    """Record launches (of `kinds`, names from PROF_KINDS; default all)."""
    mask = 0xFFFFFFFF if kinds is None else sum(1 << PROF_KINDS.index(k) for k in kinds)
    lib().hodlr_profile_mask(mask)
    lib().hodlr_profile_enable(1 if on else 0)


def profile_reset():
    lib().hodlr_profile_reset()


def profile_read():
    """{kind: dict(count, ms, flops, bytes)} over the recorded launches."""
    out = {}
    for k, name in enumerate(PROF_KINDS):
        c, ms, fl, by = C.c_int64(), C.c_double(), C.c_double(), C.c_double()
        _call(lib().hodlr_profile_read, k, C.byref(c), C.byref(ms), C.byref(fl), C.byref(by),
              where="hodlr_profile_read")
        out[name] = dict(count=c.value, ms=ms.value, flops=fl.value, bytes=by.value)
    return out
