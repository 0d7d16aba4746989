"""The node-sum form of the level loop, restated in numpy with the index
rules of the kernels that run it, against dense solves (CPU).

What the kernels rely on, and what this pins:
  * node n of level t is heap entry 2^t - 1 + n; its children are 2n, 2n+1
    of level t + 1 (the leaves at level L);
  * S_n = [I, T_c1^Y; T_c0^Y, I], R_n = [T_c1; T_c0] (node_solve.cc:60-81:
    the halves swapped), E_n = S_n^-1 R_n, rows [0, r) of E_n belonging to
    the even child and [r, 2r) to the odd one;
  * the sum a node hands to its parent is T = E_n(0:r) + E_n(r:2r)
    (k_assemble_pair), the leaves' sums being their projections
    V_i^T [Y_i | B_i];
  * the top-down pass H <- E^B + H - E^Y H over leaf i's path, child
    ch = i >> (L-1-t) of node (t, ch >> 1), half ch & 1 (k_path_down),
    then x_i = d_i - Y_i H;
  * factor reuse: the right-hand sides' E^B through the stored S_n, with
    the factorization's E^Y in the top-down pass;
  * the adjoint: Hb_i = Y_i^T g_i, Hb_n = Hb_c0 + Hb_c1 - (E_n^Y)^T
    [Hb_c0; Hb_c1] (k_adj_up), node right-hand sides [Hb_c0 + Tb_n;
    Hb_c1 + Tb_n] with Tb_n rows [r, 2r) of the parent's transposed solve
    for an even n, [0, r) for an odd one (k_adj_assemble), g_i -=
    V_i Tb_i, then K^-T.
A mistake in any of these (a swapped half, an off-by-one heap index, a
transposed E^Y) moves the result far from the dense solve.  Empty leaf
slots (zero rows, zero projections) are covered by a tree with trailing
empty slots."""
import numpy as np
import pytest


def heap0(t):
    return (1 << t) - 1


def make_system(rng, sizes, r, s):
    N = sum(sizes)
    mmax = max(sizes)
    K = np.zeros((N, mmax))
    row = 0
    for m in sizes:
        if m:
            K[row:row + m, :m] = 0.1 * rng.standard_normal((m, m)) + 4.0 * np.eye(m)
        row += m
    U = rng.standard_normal((N, r)) / np.sqrt(r)
    V = rng.standard_normal((N, r)) / np.sqrt(r)
    B = rng.standard_normal((N, s))
    return K, U, V, B


def dense(K, U, V, sizes):
    """A = blockdiag(K_i) + the off-diagonal blocks U_c0 V_c1^T, U_c1 V_c0^T
    of every node (global U, V; leaf slots of the given sizes)."""
    N = sum(sizes)
    P = len(sizes)
    L = int(round(np.log2(P)))
    off = np.concatenate([[0], np.cumsum(sizes)])
    A = np.zeros((N, N))
    for i, m in enumerate(sizes):
        A[off[i]:off[i] + m, off[i]:off[i] + m] = K[off[i]:off[i] + m, :m]
    for t in range(L):
        span = P >> t
        for n in range(1 << t):
            a, mid, b = off[n * span], off[n * span + span // 2], off[(n + 1) * span]
            A[a:mid, mid:b] = U[a:mid] @ V[mid:b].T
            A[mid:b, a:mid] = U[mid:b] @ V[a:mid].T
    return A


class NodeSum:
    def __init__(self, K, U, V, sizes):
        self.K, self.U, self.V, self.sizes = K, U, V, sizes
        self.P = len(sizes)
        self.L = int(round(np.log2(self.P)))
        self.off = np.concatenate([[0], np.cumsum(sizes)])
        self.r = U.shape[1]

    def rows(self, i):
        return slice(self.off[i], self.off[i + 1])

    def leaf_solve(self, X, trans=False):
        Y = np.zeros_like(X)
        for i, m in enumerate(self.sizes):
            if m:
                Ki = self.K[self.rows(i), :m]
                Y[self.rows(i)] = np.linalg.solve(Ki.T if trans else Ki, X[self.rows(i)])
        return Y

    def bottom_up(self, part, S=None):
        """Every node's E (heap order); S given: the stored cores (solve)."""
        r, L = self.r, self.L
        E = [None] * (self.P - 1)
        Sn = [None] * (self.P - 1) if S is None else S
        for t in range(L - 1, -1, -1):
            for n in range(1 << t):
                def T(c):
                    if t == L - 1:
                        return part[c]
                    e = E[heap0(t + 1) + c]
                    return e[:r] + e[r:]
                T0, T1 = T(2 * n), T(2 * n + 1)
                g = heap0(t) + n
                if S is None:
                    Sn[g] = np.block([[np.eye(r), T1[:, :r]], [T0[:, :r], np.eye(r)]])
                E[g] = np.linalg.solve(Sn[g], np.vstack([T1, T0]))
        return E, Sn

    def top_down(self, EY, EB, i, ncols):
        r, L = self.r, self.L
        H = np.zeros((r, ncols))
        for t in range(L):
            ch = i >> (L - 1 - t)
            g = heap0(t) + (ch >> 1)
            h = (ch & 1) * r
            H = EB(g)[h:h + r] + H - EY(g)[h:h + r] @ H
        return H

    def factor_solve(self, B):
        r, s = self.r, B.shape[1]
        Yd = self.leaf_solve(np.hstack([self.U, B]))
        self.Y, d = Yd[:, :r], Yd[:, r:]
        part = [self.V[self.rows(i)].T @ Yd[self.rows(i)] for i in range(self.P)]
        self.E, self.S = self.bottom_up(part)
        x = d.copy()
        for i in range(self.P):
            G = self.top_down(lambda g: self.E[g][:, :r], lambda g: self.E[g][:, r:], i, s)
            x[self.rows(i)] -= self.Y[self.rows(i)] @ G
        return x

    def solve(self, B):
        r = self.r
        d = self.leaf_solve(B)
        part = [self.V[self.rows(i)].T @ d[self.rows(i)] for i in range(self.P)]
        EB, _ = self.bottom_up(part, self.S)
        x = d.copy()
        for i in range(self.P):
            G = self.top_down(lambda g: self.E[g][:, :r], lambda g: EB[g], i, B.shape[1])
            x[self.rows(i)] -= self.Y[self.rows(i)] @ G
        return x

    def solve_adjoint(self, Gm):
        r, L, P = self.r, self.L, self.P
        Hb = [None] * (2 * P - 1)
        for i in range(P):
            Hb[heap0(L) + i] = self.Y[self.rows(i)].T @ Gm[self.rows(i)]
        for t in range(L - 1, 0, -1):
            for n in range(1 << t):
                EY = self.E[heap0(t) + n][:, :r]
                h0, h1 = Hb[heap0(t + 1) + 2 * n], Hb[heap0(t + 1) + 2 * n + 1]
                Hb[heap0(t) + n] = h0 + h1 - EY.T @ np.vstack([h0, h1])
        Rb = [None] * (P - 1)
        for t in range(L):
            for n in range(1 << t):
                if t > 0:
                    par = Rb[heap0(t - 1) + (n >> 1)]
                    o = (1 - (n & 1)) * r
                    Tb = par[o:o + r]
                else:
                    Tb = 0.0
                rhs = np.vstack([Hb[heap0(t + 1) + 2 * n] + Tb, Hb[heap0(t + 1) + 2 * n + 1] + Tb])
                Rb[heap0(t) + n] = np.linalg.solve(self.S[heap0(t) + n].T, rhs)
        g = Gm.copy()
        for c in range(P):
            par = Rb[heap0(L - 1) + (c >> 1)]
            o = ((c & 1) ^ 1) * r
            g[self.rows(c)] -= self.V[self.rows(c)] @ par[o:o + r]
        return self.leaf_solve(g, trans=True)


def relerr(x, ref):
    return float(np.abs(x - ref).max() / np.abs(ref).max())


@pytest.mark.parametrize("sizes,r,s", [
    ([32] * 16, 6, 3),                      # balanced, 4 levels
    ([24] * 8, 5, 2),                       # balanced, 3 levels
    ([40] * 5 + [0] * 3, 4, 2),             # 5 leaves in 8 slots (trailing empty)
    ([30, 0, 50, 20, 40, 40, 0, 10], 3, 1),  # uneven, empty slots inside
])
def test_node_sum_model_matches_dense(orc, sizes, r, s):
    rng = np.random.default_rng(17 + len(sizes) + r)
    K, U, V, B = make_system(rng, sizes, r, s)
    A = dense(K, U, V, sizes)
    # the same matrix as the oracle's restated assembly of the reference's
    # HODLR matrix (balanced, or a tree of uneven leaves)
    Kf, Uf, Vf = (np.asfortranarray(a) for a in (K, U, V))
    if len(set(sizes)) == 1:
        ref = orc.dense_assemble(Kf, Uf, Vf, sizes[0])
    else:
        ref = orc.dense_assemble_uneven(Kf, Uf, Vf, sizes)
    assert np.abs(A - ref).max() <= 1e-14 * np.abs(ref).max()
    model = NodeSum(K, U, V, sizes)
    x = model.factor_solve(B)
    assert relerr(x, np.linalg.solve(A, B)) <= 1e-12
    B2 = rng.standard_normal(B.shape)
    assert relerr(model.solve(B2), np.linalg.solve(A, B2)) <= 1e-12
    G = rng.standard_normal(B.shape)
    assert relerr(model.solve_adjoint(G), np.linalg.solve(A.T, G)) <= 1e-12


def test_node_sum_model_detects_a_swapped_half():
    """The check has teeth: the sum handed to the parent taken from one half
    only (a plausible indexing slip) is far from the dense solve."""
    rng = np.random.default_rng(5)
    sizes, r = [32] * 8, 4
    K, U, V, B = make_system(rng, sizes, r, 2)
    A = dense(K, U, V, sizes)
    model = NodeSum(K, U, V, sizes)
    orig = model.bottom_up

    def broken(part, S=None):
        L = model.L
        E, Sn = orig(part, S)
        # re-run with the parent's sum taken from the lower half twice
        E2 = [None] * (model.P - 1)
        Sn2 = [None] * (model.P - 1)
        for t in range(L - 1, -1, -1):
            for n in range(1 << t):
                def T(c):
                    if t == L - 1:
                        return part[c]
                    e = E2[heap0(t + 1) + c]
                    return e[r:] + e[r:]
                T0, T1 = T(2 * n), T(2 * n + 1)
                g = heap0(t) + n
                Sn2[g] = np.block([[np.eye(r), T1[:, :r]], [T0[:, :r], np.eye(r)]])
                E2[g] = np.linalg.solve(Sn2[g], np.vstack([T1, T0]))
        return E2, Sn2

    model.bottom_up = broken
    x = model.factor_solve(B)
    assert relerr(x, np.linalg.solve(A, B)) > 1e-6
