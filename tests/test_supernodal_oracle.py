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
