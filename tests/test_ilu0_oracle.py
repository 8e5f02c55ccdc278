"""CPU pinning of the ILU(0) oracle (sparse_blas/spilu0.cpp:11-51).  The
reference's spilu0.cpp includes the un-vendored aggregation/def.h, so it is
not compiled here; the restatement is pinned instead by what the mathematics
fixes: where ILU(0) drops no fill it is the exact LU factorization (checked
against dense Gaussian elimination), the product L U reproduces A on A's
pattern, on an SPD matrix it agrees with the IC(0) factor (itself checked
against the reference's compiled spic0 routines), and a zero pivot is
reported at its row."""
import numpy as np
import pytest

import oracle
from sympiler_amd import workloads as W
from tests.helpers import general_with_diags


def dominant(Ap, Ai, diags, seed):
    """Random values on A's pattern, each diagonal 1 + its row's and column's
    absolute sums (strictly diagonally dominant: every pivot is non-zero)."""
    rng = np.random.default_rng(seed)
    n = len(Ap) - 1
    Ax = rng.uniform(-1.0, 1.0, len(Ai))
    rows = np.repeat(np.arange(n), np.diff(Ap))
    w = np.abs(Ax)
    w[diags] = 0.0
    Ax[diags] = 1.0 + np.bincount(rows, w, n) + np.bincount(np.asarray(Ai), w, n)
    return Ax


def general(n, density, upper, seed):
    Lp, Li, _ = W.random_lower(n, density, 1, seed=seed)
    Ap, Ai, _, diags = general_with_diags(Lp, Li, np.zeros(len(Li)), upper, seed)
    return Ap, Ai, dominant(Ap, Ai, diags, seed), diags


def symmetric_pattern(Lp, Li):
    """CSR of the full symmetric pattern of a lower triangle, with diags."""
    n = len(Lp) - 1
    rows = [[] for _ in range(n)]
    for i in range(n):
        for p in range(Lp[i], Lp[i + 1]):
            j = int(Li[p])
            rows[i].append(j)
            if j != i:
                rows[j].append(i)
    Ap = np.zeros(n + 1, np.int32)
    Ai = []
    diags = np.zeros(n, np.int32)
    for i in range(n):
        cols = sorted(rows[i])
        diags[i] = len(Ai) + cols.index(i)
        Ai.extend(cols)
        Ap[i + 1] = len(Ai)
    return Ap, np.asarray(Ai, np.int32), diags


def dense(Ap, Ai, Ax):
    n = len(Ap) - 1
    D = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(Ap))
    D[rows, Ai] = Ax
    return D


def split(Ap, Ai, l, n):
    D = dense(Ap, Ai, l)
    return np.tril(D, -1) + np.eye(n), np.triu(D)


def test_ilu0_is_exact_lu_without_fill():
    # tridiagonal plus a full last row and column: Gaussian elimination
    # without pivoting creates no entry outside this pattern
    n = 60
    rows = []
    for i in range(n):
        cols = {i}
        if i > 0:
            cols.add(i - 1)
        if i < n - 1:
            cols.add(i + 1)
        cols.add(n - 1)
        if i == n - 1:
            cols.update(range(n))
        rows.append(sorted(cols))
    Ap = np.zeros(n + 1, np.int32)
    Ai = []
    for i, c in enumerate(rows):
        Ai.extend(c)
        Ap[i + 1] = len(Ai)
    Ai = np.asarray(Ai, np.int32)
    diags = np.array([Ap[i] + rows[i].index(i) for i in range(n)], np.int32)
    Ax = dominant(Ap, Ai, diags, 3)
    l, bad = oracle.spilu0_csr(Ap, Ai, Ax, diags)
    assert bad == -1
    # dense Gaussian elimination, no pivoting
    M = dense(Ap, Ai, Ax)
    for k in range(n):
        M[k + 1:, k] /= M[k, k]
        M[k + 1:, k + 1:] -= np.outer(M[k + 1:, k], M[k, k + 1:])
    rows_ = np.repeat(np.arange(n), np.diff(Ap))
    want = M[rows_, Ai]
    assert np.max(np.abs(l - want)) <= 1e-14 * np.max(np.abs(want))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_ilu0_product_reproduces_a_on_its_pattern(seed):
    Ap, Ai, Ax, diags = general(300, 0.02, 4, seed)
    l, bad = oracle.spilu0_csr(Ap, Ai, Ax, diags)
    assert bad == -1
    n = len(Ap) - 1
    L, U = split(Ap, Ai, l, n)
    P = L @ U
    rows = np.repeat(np.arange(n), np.diff(Ap))
    assert np.max(np.abs(P[rows, Ai] - Ax)) <= 1e-13 * np.max(np.abs(Ax))


def test_ilu0_agrees_with_ic0_on_spd():
    # A = L_ic L_ic^T on the pattern; ILU(0) of the same A has
    # U[i,j] = L_ic[j,i] L_ic[i,i] and L[i,j] = L_ic[i,j] / L_ic[j,j]
    Lp, Li, A = W.lap3d_tril(5, 4, 6, 6.01, -1.0)
    n = len(Lp) - 1
    Cp, Ci, Cx, mp = oracle.csr_to_csc(Lp, Li, A)
    pp, ps = oracle.prune_set(Lp, Li)
    lic, ok = oracle.spic0_csc_left(Cp, Ci, Cx, pp, ps)
    assert ok
    Lic = dense(Lp, Li, lic[mp])
    Ap, Ai, diags = symmetric_pattern(Lp, Li)
    Sym = dense(Lp, Li, A)
    Sym = Sym + np.tril(Sym, -1).T
    rows = np.repeat(np.arange(n), np.diff(Ap))
    Ax = Sym[rows, Ai]
    l, bad = oracle.spilu0_csr(Ap, Ai, Ax, diags)
    assert bad == -1
    d = np.diag(Lic)
    want_u = (Lic * d[None, :]).T  # U[i,j] = L_ic[j,i] * L_ic[i,i]
    want_l = Lic / d[None, :]
    want = np.where(Ai < rows, want_l[rows, Ai], want_u[rows, Ai])
    assert np.max(np.abs(l - want)) <= 1e-13 * np.max(np.abs(want))


def test_ilu0_zero_pivot_row_reported():
    # [[1, 1], [1, 1]]: U[1,1] = 1 - 1 * 1 = 0
    Ap = np.array([0, 2, 4], np.int32)
    Ai = np.array([0, 1, 0, 1], np.int32)
    Ax = np.ones(4)
    l, bad = oracle.spilu0_csr(Ap, Ai, Ax, np.array([0, 3], np.int32))
    assert bad == 1 and list(l) == [1.0, 1.0, 1.0, 0.0]
