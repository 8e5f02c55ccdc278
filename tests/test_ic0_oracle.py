"""CPU pinning of the IC(0) oracle (sparse_blas/spic0.cpp): the restatements
against the reference's own compiled routines, the divergence of spic0_csr
from textbook IC(0) on the 4 x 4 grid probe, the per-entry closed forms the
GPU inspector builds its update lists from, and textbook IC(0) against a
dense Cholesky where IC(0) has no dropped fill."""
import math

import numpy as np
import pytest

import oracle
from sympiler_amd import workloads as W
from tests.helpers import bitwise_equal, etree_lower


def spd_random(n, density, seed):
    """Random SPD lower triangle (CSR, diagonal last): off-diagonals in
    [-1, 1], diagonal = 1 + the absolute sum of its row and column."""
    Lp, Li, _ = W.random_lower(n, density, 1, seed=seed)
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1.0, 1.0, len(Li))
    diag = np.asarray(Lp[1:]) - 1
    rows = np.repeat(np.arange(n), np.diff(Lp))
    w = np.abs(A)
    w[diag] = 0.0
    s = np.bincount(rows, w, n) + np.bincount(np.asarray(Li), w, n)
    A[diag] = 1.0 + s
    return Lp, Li, A


def cases():
    yield "lap2d_4x4", W.lap2d_tril(4, 4, 4.0, -1.0)
    yield "lap2d_9x7", W.lap2d_tril(9, 7, 4.0, -1.0)
    yield "lap3d_4x3x5", W.lap3d_tril(4, 3, 5, 6.01, -1.0)
    for s in (1, 2, 3):
        yield f"random{s}", spd_random(80, 0.12, 50 + s)


def closed_form(Lp, Li, A, last_row_only):
    """Per entry (r, c) of the CSR of L: A[r,c] minus L[r,s] * L[c,s] over
    the s < c both rows store, ascending (spic0_csr: only the s whose last
    stored row is r); then sqrt on the diagonal, / L[c,c] elsewhere."""
    n = len(Lp) - 1
    Lx = np.zeros(len(A))
    last = np.full(n, -1)
    for r in range(n):
        for p in range(Lp[r], Lp[r + 1]):
            last[Li[p]] = max(last[Li[p]], r)
    for r in range(n):
        d = Lp[r + 1] - 1
        for p in range(Lp[r], d + 1):
            c = Li[p]
            v = A[p]
            mine = {Li[a]: a for a in range(Lp[r], p)}
            for b in range(Lp[c], Lp[c + 1] - 1):
                s = Li[b]
                if s in mine and (not last_row_only or last[s] == r):
                    v -= Lx[mine[s]] * Lx[b]
            Lx[p] = math.sqrt(v) if p == d else v / Lx[Lp[c + 1] - 1]
    return Lx


def both(Lp, Li, A):
    Cp, Ci, Cx, mp = oracle.csr_to_csc(Lp, Li, A)
    pp, ps = oracle.prune_set(Lp, Li)
    left, ok1 = oracle.spic0_csc_left(Cp, Ci, Cx, pp, ps)
    csr, ok2 = oracle.spic0_csr(Cp, Ci, Cx)
    assert ok1 and ok2
    return (Cp, Ci, Cx, mp), left, csr


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build absent")
@pytest.mark.parametrize("name,m", list(cases()))
def test_spic0_restatements_equal_reference(name, m):
    (Cp, Ci, Cx, _), left, csr = both(*m)
    pp, ps = oracle.prune_set(m[0], m[1])
    ref_left, ok = oracle.ref.spic0_csc_left(Cp, Ci, Cx, pp, ps)
    assert ok and bitwise_equal(left, ref_left)
    assert bitwise_equal(csr, oracle.ref.spic0_csr(Cp, Ci, Cx))


def test_spic0_csr_diverges_from_textbook_on_grid_probe():
    # 4 x 4 grid 2D Laplacian: textbook L[1,1] = sqrt(4 - 1/4); spic0_csr
    # skips that update (row 1 is not the last row of column 0) and keeps 2
    (Cp, _, _, _), left, csr = both(*W.lap2d_tril(4, 4, 4.0, -1.0))
    assert left[Cp[0]] == 2.0 and csr[Cp[0]] == 2.0
    assert left[Cp[1]] == math.sqrt(3.75) == 1.9364916731037085
    assert csr[Cp[1]] == 2.0
    assert 0.083 < np.max(np.abs(left - csr)) < 0.085


@pytest.mark.parametrize("name,m", list(cases()))
def test_closed_forms_equal_the_loops(name, m):
    (_, _, _, mp), left, csr = both(*m)
    assert bitwise_equal(closed_form(*m, last_row_only=False), left[mp])
    assert bitwise_equal(closed_form(*m, last_row_only=True), csr[mp])


def test_textbook_ic0_is_cholesky_without_fill():
    # an elimination-forest pattern (every column stores at most one entry
    # below its diagonal, in the row of its parent) has no Cholesky fill, so
    # IC(0) is the exact Cholesky factor
    rng = np.random.default_rng(7)
    n = 200
    parent = [int(rng.integers(j + 1, n)) if j < n - 1 and rng.random() < 0.9 else -1
              for j in range(n)]
    Lp, Li, _ = etree_lower(parent)
    A = -rng.uniform(0.1, 1.0, len(Li))
    rows = np.repeat(np.arange(n), np.diff(Lp))
    D = np.zeros((n, n))
    D[rows, Li] = A
    D[Li, rows] = A
    diag = np.asarray(Lp[1:]) - 1
    np.fill_diagonal(D, 0.0)
    A[diag] = 1.0 + np.abs(D).sum(axis=1)
    np.fill_diagonal(D, A[diag])
    (_, _, _, mp), left, _ = both(Lp, Li, A)
    want = np.linalg.cholesky(D)[rows, Li]
    assert np.max(np.abs(left[mp] - want)) <= 1e-14 * np.max(np.abs(want))


def test_textbook_ic0_reproduces_a_on_its_pattern():
    Lp, Li, A = W.lap3d_tril(5, 4, 6, 6.01, -1.0)
    (_, _, _, mp), left, _ = both(Lp, Li, A)
    L = left[mp]
    n = len(Lp) - 1
    M = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(Lp))
    M[rows, Li] = L
    P = M @ M.T
    assert np.max(np.abs(P[rows, Li] - A)) <= 1e-13 * np.max(np.abs(A))
