"""The column-oriented SpMV drop-ins (sym_lib::spmv_csc and its axpby
overloads, spmv_csc_lbc, spmv_csc_block, spmv_csr_block and the two
_parallel forms) on the device, called through their C++ mangled names,
bit for bit against the oracle's restatements (themselves pinned against the
reference build), on inputs with -0.0 in y, empty columns and parts, repeated
ids and rectangular matrices."""
import ctypes
import os
from ctypes import c_double, c_int, c_void_p

import numpy as np
import pytest

import oracle
from tests.helpers import bitwise_equal
from tests.test_spmv_oracle import _csc, _y

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "sympiler_amd", "libsympiler_amd.so")
M = {
    "csc": "_ZN7sym_lib8spmv_cscEiPKiS1_PKdS3_Pd",
    "csc_ab": "_ZN7sym_lib8spmv_cscEiPKiS1_PKdS3_S3_Pddd",
    "csc_abv": "_ZN7sym_lib8spmv_cscEiPKiS1_PKdS3_S3_PdS4_S4_",
    "csc_lbc": "_ZN7sym_lib12spmv_csc_lbcEiPKiS1_iS1_S1_PKdS3_Pd",
    "csc_lbc_ptr": "_ZN7sym_lib12spmv_csc_lbcEPKiS1_S1_iS1_S1_PKdS3_Pd",
    "csc_block": "_ZN7sym_lib14spmv_csc_blockEiPKiS1_PKdS3_PdiS1_",
    "csr_block": "_ZN7sym_lib14spmv_csr_blockEiPKiS1_PKdiS1_S3_Pd",
    "csc_par": "_ZN7sym_lib17spmv_csc_parallelEiPKiS1_PKdS3_Pd",
    "csr_par": "_ZN7sym_lib17spmv_csr_parallelEiPKiS1_PKdS3_Pd",
}


@pytest.fixture(scope="module")
def L(dev):
    return ctypes.CDLL(LIB)


def P(a):
    return a.ctypes.data_as(c_void_p)


@pytest.mark.parametrize("seed", range(4))
def test_spmv_csc_forms_bitwise(L, seed):
    rng = np.random.default_rng(70 + seed)
    n = int(rng.integers(1, 3000))
    m = n if seed % 2 else int(rng.integers(1, 4000))
    Cp, Ci, Cx = _csc(rng, n, m, float(rng.uniform(0.0, 0.01)), empty_cols={0, n // 2})
    x = rng.uniform(-1, 1, n)
    y = _y(rng, max(m, int(Ci.max()) + 1 if len(Ci) else 0))
    print(f"[spmv] seed={seed} n={n} m={m} nnz={len(Ci)}", flush=True)
    for name in ("csc", "csc_par"):
        got = y.copy()
        L[M[name]](c_int(n), P(Cp), P(Ci), P(Cx), P(x), P(got))
        assert bitwise_equal(got, oracle.spmv_csc(Cp, Ci, Cx, x, y)), name
    sup = np.unique(np.concatenate([[0, n], rng.integers(0, n + 1, 4)])).astype(np.int32)
    got = y.copy()
    L[M["csc_block"]](c_int(n), P(Cp), P(Ci), P(Cx), P(x), P(got), c_int(len(sup) - 2),
                      P(sup))
    assert bitwise_equal(got, oracle.spmv_csc_block(Cp, Ci, Cx, x, y, sup[:-1]))
    if m != n:
        return
    z, yy = _y(rng, n), rng.uniform(-1, 1, n)
    got = z.copy()
    L[M["csc_ab"]](c_int(n), P(Cp), P(Ci), P(Cx), P(x), P(yy), P(got), c_double(0.7),
                   c_double(-1.3))
    assert bitwise_equal(got, oracle.spmv_csc_axpby(Cp, Ci, Cx, x, yy, z, 0.7, -1.3))
    av, bv = rng.uniform(-2, 2, n), rng.uniform(-2, 2, n)
    got = z.copy()
    L[M["csc_abv"]](c_int(n), P(Cp), P(Ci), P(Cx), P(x), P(yy), P(got), P(av), P(bv))
    assert bitwise_equal(got, oracle.spmv_csc_axpby(Cp, Ci, Cx, x, yy, z, av, bv))
    ids = rng.permutation(n).astype(np.int32)
    ids = np.concatenate([ids, ids[:3]]).astype(np.int32)
    cuts = np.sort(rng.integers(0, len(ids) + 1, 5))
    sets = np.concatenate([[0], cuts, [len(ids)], [len(ids)]]).astype(np.int32)
    want = oracle.spmv_csc_lbc(sets, ids, Cp, Ci, Cx, x, y[:n])
    got = y[:n].copy()
    L[M["csc_lbc"]](c_int(len(sets) - 1), P(sets), P(ids), c_int(n), P(Cp), P(Ci), P(Cx),
                    P(x), P(got))
    assert bitwise_equal(got, want)
    set_ptr = np.array([0, len(sets) - 1], np.int32)
    got = y[:n].copy()
    L[M["csc_lbc_ptr"]](P(set_ptr), P(sets), P(ids), c_int(n), P(Cp), P(Ci), P(Cx), P(x),
                        P(got))
    assert bitwise_equal(got, want)
    # spmv_csr_block / spmv_csr_parallel read the same arrays as CSR rows
    got = y[:n].copy()
    L[M["csr_block"]](c_int(n), P(Cp), P(Ci), P(Cx), c_int(len(sup) - 2), P(sup), P(x),
                      P(got))
    assert bitwise_equal(got, oracle.spmv_csr_block(Cp, Ci, Cx, sup[:-1], x, y[:n]))
    got = np.empty(n)
    L[M["csr_par"]](c_int(n), P(Cp), P(Ci), P(Cx), P(x), P(got))
    assert bitwise_equal(got, oracle.spmv_csr(Cp, Ci, Cx, x))


def test_lbc_negative_zero_and_empty_parts(L):
    # rows reached by no part keep y + (+0.0) per part: -0.0 becomes +0.0
    n = 4
    Cp = np.array([0, 1, 1, 3, 3], np.int32)
    Ci = np.array([0, 0, 2], np.int32)
    Cx = np.array([2.0, -1.0, 0.5])
    x = np.array([1.5, 1.0, 3.0, -2.0])
    y = np.array([-0.0, -0.0, 1.0, -0.0])
    for sets, ids in ((np.array([0, 0, 2, 2, 4], np.int32), np.array([2, 0, 3, 1], np.int32)),
                      (np.array([0, 0], np.int32), np.zeros(1, np.int32))):
        got = y.copy()
        L[M["csc_lbc"]](c_int(len(sets) - 1), P(sets), P(ids), c_int(n), P(Cp), P(Ci), P(Cx),
                        P(x), P(got))
        assert bitwise_equal(got, oracle.spmv_csc_lbc(sets, ids, Cp, Ci, Cx, x, y))
