"""The column-oriented SpMV forms of sparse_blas/spmv.cpp (spmv_csc and its
axpby overloads, spmv_csc_lbc, spmv_csc_block, spmv_csr_block): the
oracle's restatements against the reference's own known answers
(Catch_tests/spmv_tests.cpp) and, bit for bit, against the reference build on
random inputs that include -0.0 in y, empty columns and parts, repeated ids
and rectangular matrices."""
import numpy as np
import pytest

import oracle
from oracle import ref
from tests.helpers import bitwise_equal


def _csc(rng, n, m, density, empty_cols=()):
    cols = []
    for c in range(n):
        k = 0 if c in empty_cols else rng.binomial(m, density)
        cols.append(np.sort(rng.choice(m, size=k, replace=False)))
    Cp = np.concatenate([[0], np.cumsum([len(c) for c in cols])]).astype(np.int32)
    Ci = (np.concatenate(cols) if Cp[-1] else np.zeros(0)).astype(np.int32)
    Cx = rng.uniform(-1, 1, int(Cp[-1]))
    return Cp, Ci, Cx


def _y(rng, m):
    y = rng.uniform(-1, 1, m)
    y[::7] = -0.0
    y[3::11] = 0.0
    return y


def test_known_answers():
    # Catch_tests/spmv_tests.cpp:23-47 (the CSC of the 4 x 4 example) and
    # :51-73 (a symmetric 10 x 10 pattern read as CSC)
    Ap = np.array([0, 2, 3, 4, 6], np.int32)
    Ai = np.array([0, 1, 3, 3, 1, 3], np.int32)
    Ax = np.array([1.0, 2.0, 1.0, 1.0, 1.0, 1.0])
    Cp, Ci, Cx, _ = oracle.csr_to_csc(Ap, Ai, Ax)
    assert np.array_equal(oracle.spmv_csc(Cp, Ci, Cx, np.ones(4), np.zeros(4)), [3, 1, 1, 2])
    Ap = np.array([0, 5, 8, 12, 16, 19, 24, 29, 34, 39, 44], np.int32)
    Ai = np.array([0, 1, 5, 7, 8, 0, 1, 7, 2, 6, 7, 9, 3, 4, 5, 8, 3, 4, 6, 0, 3, 5, 6, 8,
                   2, 4, 5, 6, 9, 0, 1, 2, 7, 9, 0, 3, 5, 8, 9, 2, 6, 7, 8, 9], np.int32)
    got = oracle.spmv_csc(Ap, Ai, np.ones(44), np.ones(10), np.zeros(10))
    assert np.array_equal(got, [5, 3, 4, 4, 3, 5, 5, 5, 5, 5])


def test_block_forms_are_range_restrictions():
    # spmv_csc_block over every supernode is spmv_csc; spmv_csr_block over
    # every supernode is spmv_csr (the loops differ only in their ranges)
    rng = np.random.default_rng(3)
    Cp, Ci, Cx = _csc(rng, 60, 60, 0.1)
    x, y = rng.uniform(-1, 1, 60), _y(rng, 60)
    sup = np.array([0, 7, 8, 30, 60], np.int32)
    assert bitwise_equal(oracle.spmv_csc_block(Cp, Ci, Cx, x, y, sup),
                         oracle.spmv_csc(Cp, Ci, Cx, x, y))
    assert bitwise_equal(oracle.spmv_csr_block(Cp, Ci, Cx, sup, x, y),
                         oracle.spmv_csr(Cp, Ci, Cx, x))


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("seed", range(6))
def test_against_the_reference_build(seed):
    rng = np.random.default_rng(40 + seed)
    n = int(rng.integers(1, 300))
    m = n if seed % 2 else int(rng.integers(1, 400))  # rectangular for the plain form
    Cp, Ci, Cx = _csc(rng, n, m, float(rng.uniform(0.0, 0.2)), empty_cols={0, n // 2})
    x = rng.uniform(-1, 1, n)
    y = _y(rng, m)
    assert bitwise_equal(oracle.spmv_csc(Cp, Ci, Cx, x, y), ref.spmv_csc(Cp, Ci, Cx, x, y))
    k = int(rng.integers(0, 5))
    sup = np.unique(np.concatenate([[0, n], rng.integers(0, n + 1, k)])).astype(np.int32)
    part = sup[: max(1, len(sup) - 1)] if seed % 3 == 0 else sup  # a prefix of the columns
    assert bitwise_equal(oracle.spmv_csc_block(Cp, Ci, Cx, x, y, part),
                         ref.spmv_csc_block(Cp, Ci, Cx, x, y, part))
    if m == n:
        z, yy = _y(rng, n), rng.uniform(-1, 1, n)
        for a, b in ((0.7, -1.3), (rng.uniform(-2, 2, n), rng.uniform(-2, 2, n))):
            assert bitwise_equal(oracle.spmv_csc_axpby(Cp, Ci, Cx, x, yy, z, a, b),
                                 ref.spmv_csc_axpby(Cp, Ci, Cx, x, yy, z, a, b))
        # LBC: parts over a shuffled id list with empty parts and a repeat
        ids = rng.permutation(n).astype(np.int32)
        ids = np.concatenate([ids, ids[:3]]).astype(np.int32)
        cuts = np.sort(rng.integers(0, len(ids) + 1, int(rng.integers(1, 6))))
        sets = np.concatenate([[0], cuts, [len(ids)], [len(ids)]]).astype(np.int32)
        assert bitwise_equal(oracle.spmv_csc_lbc(sets, ids, Cp, Ci, Cx, x, y),
                             ref.spmv_csc_lbc(sets, ids, Cp, Ci, Cx, x, y))
        # spmv_csr_block reads the arrays as CSR rows
        assert bitwise_equal(oracle.spmv_csr_block(Cp, Ci, Cx, part, x, y),
                             ref.spmv_csr_block(Cp, Ci, Cx, part, x, y))


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
def test_lbc_turns_negative_zero_positive():
    # a row no column of any part reaches keeps y's +0.0 added once per part:
    # -0.0 becomes +0.0 (the detail the device form reproduces)
    Cp = np.array([0, 1, 1], np.int32)
    Ci = np.array([0], np.int32)
    Cx = np.array([2.0])
    y = np.array([-0.0, -0.0])
    sets = np.array([0, 1, 2], np.int32)
    ids = np.array([0, 1], np.int32)
    got = oracle.spmv_csc_lbc(sets, ids, Cp, Ci, Cx, np.array([-0.0, 1.0]), y)
    assert bitwise_equal(got, ref.spmv_csc_lbc(sets, ids, Cp, Ci, Cx, np.array([-0.0, 1.0]), y))
    assert np.signbit(got[1]) == False and got[1] == 0.0
