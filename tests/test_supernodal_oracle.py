"""CPU pinning of the supernodal (BCSC) solve oracle: the restatement of
sparse_blas/sptrsv.cpp:176-199 with the dense kernels of dense_blas/BLAS.cpp
(lsolve_BLAS, matvec_BLAS) against the reference's own compiled
sptrsv_csc_block, and the layout against a dense solve of the assembled L."""
import numpy as np
import pytest

import oracle
from sympiler_amd import workloads as W
from tests.helpers import bitwise_equal

CASES = [  # n, widest supernode, most rows below, row window, seed
    (300, 8, 8, 60, 1),
    (2000, 40, 24, 400, 2),
    (3000, 100, 70, 1500, 3),
    (800, 1, 5, 30, 4),
    (1500, 13, 130, 600, 5),
]


def dense(sup, nrows, Lp, Li, Lx, n):
    L = np.zeros((n, n))
    for i in range(len(sup) - 1):
        s, w, nr = sup[i], sup[i + 1] - sup[i], nrows[i]
        rows = Li[Lp[i]:Lp[i] + nr]
        for j in range(w):
            col = Lx[Lp[i] + j * nr: Lp[i] + (j + 1) * nr]
            for r in range(j, nr):
                L[s + r if r < w else rows[r], s + j] = col[r]
    return L


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build absent")
@pytest.mark.parametrize("case", CASES)
def test_block_solve_restatement_equals_reference(case):
    sup, nrows, Lp, Li, Lx = W.bcsc(*case)
    b = W.uniform(case[-1], 0, case[0])
    x = oracle.sptrsv_csc_block(sup, nrows, Lp, Li, Lx, b)
    assert np.all(np.isfinite(x))  # the NaN upper parts are never read
    assert bitwise_equal(x, oracle.ref.sptrsv_csc_block(sup, nrows, Lp, Li, Lx, b))


@pytest.mark.parametrize("case", CASES[:3])
def test_block_solve_layout_against_dense(case):
    n = case[0]
    sup, nrows, Lp, Li, Lx = W.bcsc(*case)
    b = W.uniform(case[-1], 0, n)
    x = oracle.sptrsv_csc_block(sup, nrows, Lp, Li, Lx, b)
    L = dense(sup, nrows, Lp, Li, Lx, n)
    assert np.max(np.abs(L @ x - b)) <= 1e-13 * np.max(np.abs(b))


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
def test_parsy_dense_kernels_equal_sym_lib_kernels():
    """parsy::blockedLsolve (Triangular_BCSC.cpp:24-56) runs
    dlsolve_blas_nonUnit / dmatvec_blas (parsy/triangularSolve/BLAS.cpp:12,
    :123); sptrsv_csc_block runs lsolve_BLAS / matvec_BLAS
    (dense_blas/BLAS.cpp:3, :286).  Both pairs, compiled from the reference,
    give bit-identical results on every block shape (8/4/2/1 column tails,
    empty and long row blocks, padded leading dimensions, a product added to
    a nonzero vector): the two solves differ only in their storage format."""
    rng = np.random.default_rng(0)
    for ncol in list(range(1, 26)) + [63, 64, 100]:
        for nrow in (0, 1, 3, 7, 8, 9, 17, 40):
            ldm = ncol + nrow + int(rng.integers(0, 3))
            M = rng.standard_normal(ldm * ncol)
            M[np.arange(ncol) * (ldm + 1)] += 4.0
            args = (M, ldm, ncol, nrow, rng.standard_normal(ncol),
                    rng.standard_normal(ncol), rng.standard_normal(max(nrow, 1)))
            a = oracle.ref.dense_kernels(*args, parsy=False)
            b = oracle.ref.dense_kernels(*args, parsy=True)
            assert bitwise_equal(a[0], b[0]) and bitwise_equal(a[1], b[1]), (ncol, nrow)


def to_parsy(sup, nrows, Lp, Li, n):
    """The same supernodal factor in parsy's storage (Triangular_BCSC.cpp:
    24-56): per-column value offsets Lp (size_t) and row-list offsets Li_ptr
    (size_t), each supernode's row list (its own columns, then the rows
    below) at Li[Li_ptr[first column]]."""
    m = len(sup) - 1
    vp = np.zeros(n + 1, np.uint64)
    rp = np.zeros(n + 1, np.uint64)
    rows = []
    total_rows = 0
    for i in range(m):
        c0, c1, nr = int(sup[i]), int(sup[i + 1]), int(nrows[i])
        for c in range(c0, c1):
            vp[c] = int(Lp[i]) + (c - c0) * nr
            rp[c] = total_rows
        rows.extend(range(c0, c1))
        rows.extend(int(v) for v in Li[int(Lp[i]) + (c1 - c0):int(Lp[i]) + nr])
        total_rows += nr
    vp[n] = int(Lp[m - 1]) + (int(sup[m]) - int(sup[m - 1])) * int(nrows[m - 1]) if m else 0
    rp[n] = total_rows
    return vp, np.asarray(rows, np.int32), rp


# ------------------------------------------------- transposed (blockedLTsolve)
@pytest.mark.parametrize("case", CASES)
def test_transposed_block_solve_against_dense(case):
    """L^T x = b on the supernodal storage (parsy's blockedLTsolve,
    Triangular_BCSC.cpp:125-186; its BLAS is absent here, so the restatement
    is pinned by the mathematics): the dense transposed solve of the
    assembled L."""
    n = case[0]
    sup, nrows, Lp, Li, Lx = W.bcsc(*case)
    b = W.uniform(case[-1], 1, n)
    x = oracle.sptrsv_csc_block_t(sup, nrows, Lp, Li, Lx, b)
    L = dense(sup, nrows, Lp, Li, Lx, n)
    assert np.max(np.abs(L.T @ x - b)) <= 1e-12 * np.max(np.abs(b))


def test_transposed_block_solve_exact_integers():
    """Small integer entries, unit diagonal and an integer solution: every
    operation is exact, so x must equal x_true exactly -- a wrong entry,
    transposition or order of rows below shows as an integer error."""
    n = 600
    sup, nrows, Lp, Li, Lx = W.bcsc(n, 12, 16, 200, 9)
    rng = np.random.default_rng(1)
    Lx = rng.integers(-2, 3, len(Lx)).astype(np.float64)
    for i in range(len(sup) - 1):  # unit diagonal of every top
        w, nr = sup[i + 1] - sup[i], nrows[i]
        for j in range(w):
            Lx[Lp[i] + j * nr + j] = 1.0
    L = dense(sup, nrows, Lp, Li, Lx, n)
    x_true = rng.integers(-3, 4, n).astype(np.float64)
    b = L.T @ x_true  # exact: small integers
    x = oracle.sptrsv_csc_block_t(sup, nrows, Lp, Li, Lx, b)
    assert np.array_equal(x, x_true)


def test_transposed_is_the_adjoint_of_the_forward_solve():
    """<L^-T b, c> = <b, L^-1 c> (to rounding): the transposed restatement is
    the adjoint of the pinned forward one."""
    sup, nrows, Lp, Li, Lx = W.bcsc(2000, 40, 24, 400, 2)
    b, c = W.uniform(3, 0, 2000), W.uniform(4, 0, 2000)
    u = oracle.sptrsv_csc_block_t(sup, nrows, Lp, Li, Lx, b)
    v = oracle.sptrsv_csc_block(sup, nrows, Lp, Li, Lx, c)
    assert abs(u @ c - b @ v) <= 1e-12 * max(abs(u @ c), 1.0)
