"""The sym_lib:: drop-in entry points, called through their C++ mangled names
exactly as a program built against the reference's sparse_blas_lib.h would
bind them, on host arrays; each checked against the oracle (bit-identical)
or the reference tests' known answers."""
import ctypes
import os
from ctypes import POINTER, byref, c_double, c_int, c_void_p

import numpy as np
import pytest

import oracle
from sympiler_amd import workloads as W
from tests.helpers import KA_B, KA_CSC_I, KA_CSC_P, bitwise_equal

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "sympiler_amd", "libsympiler_amd.so")


@pytest.fixture(scope="module")
def L(dev):
    return ctypes.CDLL(LIB)


def P(a):
    return a.ctypes.data_as(c_void_p)


def test_sptrsv_csc_known_answer(L):
    # Catch_tests/sptrsv_tests.cpp:17-31
    fn = L["_ZN7sym_lib10sptrsv_cscEiPiS0_PdS1_"]
    Cp, Ci = KA_CSC_P.copy(), KA_CSC_I.copy()
    Cx = np.ones(26)
    x = KA_B.copy()
    fn(c_int(10), P(Cp), P(Ci), P(Cx), P(x))
    assert np.all(x == 1.0)


def test_sptrsv_csr_and_levelset(L):
    Lp, Li, Lx = W.random_lower(3000, 0.002, 1, seed=21)
    b = W.uniform(21, 0, 3000)
    x_ref = oracle.sptrsv_csr(Lp, Li, Lx, b)
    x = b.copy()
    L["_ZN7sym_lib10sptrsv_csrEiPiS0_PdS1_"](c_int(3000), P(Lp), P(Li), P(Lx), P(x))
    assert bitwise_equal(x, x_ref)
    lp, ls, _ = oracle.levelset_csr(Lp, Li)
    x = b.copy()
    L["_ZN7sym_lib19sptrsv_csr_levelsetEiPKiS1_PKdPdiS1_S1_"](
        c_int(3000), P(Lp), P(Li), P(Lx), P(x), c_int(len(lp) - 1), P(lp), P(ls))
    assert bitwise_equal(x, x_ref)


def test_sptrsv_csr_lbc(L):
    Lp, Li, Lx = W.random_lower(2000, 0.003, 1, seed=22)
    b = W.uniform(22, 0, 2000)
    lp, ls, _ = oracle.levelset_csr(Lp, Li)
    nl = len(lp) - 1
    cl = list(range(0, nl, 2)) + [nl]
    level_ptr = np.arange(len(cl), dtype=np.int32)
    par_ptr = np.array([lp[c] for c in cl], np.int32)
    part = ls.copy()
    x = b.copy()
    L["_ZN7sym_lib14sptrsv_csr_lbcEiPiS0_PdS1_iS0_S0_S0_"](
        c_int(2000), P(Lp), P(Li), P(Lx), P(x), c_int(len(level_ptr) - 1), P(level_ptr),
        P(par_ptr), P(part))
    assert bitwise_equal(x, oracle.sptrsv_csr(Lp, Li, Lx, b))


def test_ltsolve_and_csc_levelset(L):
    Lp, Li, Lx = W.ic0_lap3d(12, 0.01)
    n = len(Lp) - 1
    Cp, Ci, Cx, _ = oracle.csr_to_csc(Lp, Li, Lx)
    b = W.uniform(9, 0, n)
    x = b.copy()
    rc = L["_ZN7sym_lib7ltsolveEiPiS0_PdS1_"](c_int(n), P(Cp), P(Ci), P(Cx), P(x))
    assert rc == 1
    assert bitwise_equal(x, oracle.ltsolve(Cp, Ci, Cx, b))
    assert L["_ZN7sym_lib7ltsolveEiPiS0_PdS1_"](c_int(n), None, P(Ci), P(Cx), P(x)) == 0
    lp, ls = oracle.levelset_csc(Cp, Ci)
    x = b.copy()
    L["_ZN7sym_lib19sptrsv_csc_levelsetEiPiS0_PdS1_iPKiS3_"](
        c_int(n), P(Cp), P(Ci), P(Cx), P(x), c_int(len(lp) - 1), P(lp), P(ls))
    assert bitwise_equal(x, oracle.sptrsv_csc(Cp, Ci, Cx, b))


def test_dropins_reuse_an_analysis_only_for_the_same_pattern(L):
    """Drop-in calls keep the analysis of the most recently used patterns:
    a repeated pattern with new values or a new right-hand side, a pattern
    changed in place in the caller's arrays (same n and nnz), more patterns
    than are kept (least recently used evicted), and the transposed solve
    on a kept pattern all match the serial solve bit for bit."""
    fwd = L["_ZN7sym_lib10sptrsv_csrEiPiS0_PdS1_"]
    n = 2500
    Lp, Li, Lx = W.random_lower(n, 0.003, 1, seed=31)
    for rep in range(3):
        Lx_r = Lx * (1.0 + 0.25 * rep)          # same pattern, new values
        b = W.uniform(40 + rep, 0, n)
        x = b.copy()
        fwd(c_int(n), P(Lp), P(Li), P(Lx_r), P(x))
        assert bitwise_equal(x, oracle.sptrsv_csr(Lp, Li, Lx_r, b)), rep
    # move one dependency of some row to another column, in place
    Li2 = Li
    i = next(r for r in range(n) if Lp[r + 1] - Lp[r] == 2 and Li[Lp[r]] > 0)
    old = Li2[Lp[i]]
    Li2[Lp[i]] = old - 1
    b = W.uniform(50, 0, n)
    x = b.copy()
    fwd(c_int(n), P(Lp), P(Li2), P(Lx), P(x))
    assert bitwise_equal(x, oracle.sptrsv_csr(Lp, Li2, Lx, b))
    Li2[Lp[i]] = old
    # more patterns than are kept, then the first again
    mats = [W.random_lower(n + 7 * t, 0.003, 1, seed=60 + t) for t in range(6)]
    for t, (Lq, Lj, Lv) in enumerate(mats + mats[:1]):
        m = len(Lq) - 1
        b = W.uniform(70 + t, 0, m)
        x = b.copy()
        fwd(c_int(m), P(Lq), P(Lj), P(Lv), P(x))
        assert bitwise_equal(x, oracle.sptrsv_csr(Lq, Lj, Lv, b)), t
    # the transposed solve on a kept pattern with new values
    Cp, Ci, Cx, _ = oracle.csr_to_csc(Lp, Li, Lx)
    for rep in range(2):
        Cx_r = Cx * (2.0 - 0.5 * rep)
        b = W.uniform(80 + rep, 0, n)
        x = b.copy()
        assert L["_ZN7sym_lib7ltsolveEiPiS0_PdS1_"](c_int(n), P(Cp), P(Ci), P(Cx_r), P(x)) == 1
        assert bitwise_equal(x, oracle.ltsolve(Cp, Ci, Cx_r, b)), rep


def test_spmv_known_answers(L):
    fn = L["_ZN7sym_lib8spmv_csrEiPKiS1_PKdS3_Pd"]
    Ap = np.array([0, 2, 3, 4, 6], np.int32)
    Ai = np.array([0, 1, 3, 3, 1, 3], np.int32)
    Ax = np.array([1.0, 2.0, 1.0, 1.0, 1.0, 1.0])
    y = np.zeros(4)
    fn(c_int(4), P(Ap), P(Ai), P(Ax), P(np.ones(4)), P(y))
    assert np.array_equal(y, [3, 1, 1, 2])
    fn9 = L["_ZN7sym_lib8spmv_csrEiPKiS1_PKdS3_S3_Pddd"]
    fn9.argtypes = [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                    c_double, c_double]
    Lp, Li, Lx = W.random_lower(4000, 0.003, 1, seed=8)
    x = W.uniform(1, 0, 4000)
    yv = W.uniform(2, 0, 4000)
    z = W.uniform(3, 0, 4000)
    z_ref = oracle.spmv_csr_axpby(Lp, Li, Lx, x, yv, z, 0.25, 3.0)
    fn9(4000, P(Lp), P(Li), P(Lx), P(x), P(yv), P(z), 0.25, 3.0)
    assert bitwise_equal(z, z_ref)


def test_build_levelset_csc(L):
    Lp, Li, Lx = W.lap3d_tril(9, 7, 5)
    n = len(Lp) - 1
    Cp, Ci, _, _ = oracle.csr_to_csc(Lp, Li, Lx)
    lp_o, ls_o = oracle.levelset_csc(Cp, Ci)
    fn = L["_ZN7sym_lib18build_levelSet_CSCEiPKiS1_RPiS3_"]
    fn.restype = c_int
    plp, pls = POINTER(c_int)(), POINTER(c_int)()
    nl = fn(c_int(n), P(Cp), P(Ci), byref(plp), byref(pls))
    assert nl == len(lp_o) - 1 == 9 + 7 + 5 - 2
    assert np.array_equal(np.ctypeslib.as_array(plp, (nl + 1,)), lp_o)
    assert np.array_equal(np.ctypeslib.as_array(pls, (n,)), ls_o)
    # the arrays are new[]'d for the caller (released with the process here)


def test_structure_errors_reported(L):
    # the void drop-ins report bad input by throwing (which must not cross
    # a C frame such as ctypes); the C ABI beneath them returns the status
    import sympiler_amd as sa
    from tests.helpers import to_dev
    Lp, Li, Lx = W.lap2d_tril(3, 3)
    bad = Li.copy()
    r4 = slice(Lp[4], Lp[5])           # row 4 = (1, 1): columns 1, 3, 4
    bad[r4] = [3, 1, 4]                 # not ascending
    with pytest.raises(sa.SymAmdError) as e:
        sa.Plan(*to_dev(Lp, bad, Lx, device="cuda:0"))
    assert e.value.code == 2


def _csc_spd(Lp, Li, A):
    Cp, Ci, Cx, _ = oracle.csr_to_csc(Lp, Li, A)
    pp, ps = oracle.prune_set(Lp, Li)
    return Cp, Ci, Cx, pp, ps


def test_spic0_csc_left_and_lbc(L):
    # sparse_blas/spic0.cpp:11-46 and :48-99 on host CSC arrays
    Lp, Li, A = W.lap3d_tril(9, 8, 7, 6.01, -1.0)
    Cp, Ci, Cx, pp, ps = _csc_spd(Lp, Li, A)
    want, ok = oracle.spic0_csc_left(Cp, Ci, Cx, pp, ps)
    assert ok
    n = len(Cp) - 1
    Lx = np.full(len(Cx), 7.0)
    fn = L["_ZN7sym_lib14spic0_csc_leftEiPKiS1_PKdS1_S1_Pd"]
    fn.restype = ctypes.c_bool
    assert fn(c_int(n), P(Cp), P(Ci), P(Cx), P(pp), P(ps), P(Lx))
    assert bitwise_equal(Lx, want)
    lbc = L["_ZN7sym_lib18spic0_csc_left_LBCEiPKiS1_PKdS1_S1_iS1_S1_S1_Pd"]
    lbc.restype = ctypes.c_bool
    lp, ls, _ = oracle.levelset_csr(Lp, Li)
    par = np.arange(len(lp), dtype=np.int32)  # one partition per level
    Lx = np.zeros(len(Cx))
    assert lbc(c_int(n), P(Cp), P(Ci), P(Cx), P(pp), P(ps), c_int(len(lp) - 1), P(par),
               P(lp), P(ls), P(Lx))
    assert bitwise_equal(Lx, want)


def test_spic0_csc_left_breakdown_writes_the_prefix(L):
    Lp, Li, A = W.lap2d_tril(6, 6, 2.2, -1.0)  # not positive definite
    Cp, Ci, Cx, pp, ps = _csc_spd(Lp, Li, A)
    want, ok = oracle.spic0_csc_left(Cp, Ci, Cx, pp, ps)
    assert not ok
    n = len(Cp) - 1
    jf = int(np.argmax(~(want[Cp[:-1]] > 0)))
    Lx = np.full(len(Cx), 7.0)
    fn = L["_ZN7sym_lib14spic0_csc_leftEiPKiS1_PKdS1_S1_Pd"]
    fn.restype = ctypes.c_bool
    assert not fn(c_int(n), P(Cp), P(Ci), P(Cx), P(pp), P(ps), P(Lx))
    assert bitwise_equal(Lx[:Cp[jf]], want[:Cp[jf]])
    assert np.all(Lx[Cp[jf]:] == 7.0)


def test_spic0_csr_and_lbc(L):
    # sparse_blas/spic0.cpp:106-127 and :130-160, in place on host CSC arrays
    Lp, Li, A = W.lap3d_tril(9, 8, 7, 6.01, -1.0)
    Cp, Ci, Cx, _, _ = _csc_spd(Lp, Li, A)
    want, ok = oracle.spic0_csr(Cp, Ci, Cx)
    assert ok
    n = len(Cp) - 1
    val = Cx.copy()
    L["_ZN7sym_lib9spic0_csrEiPdPiS1_"](c_int(n), P(val), P(Cp.copy()), P(Ci.copy()))
    assert bitwise_equal(val, want)
    lp, ls, _ = oracle.levelset_csr(Lp, Li)
    par = np.arange(len(lp), dtype=np.int32)
    val = Cx.copy()
    L["_ZN7sym_lib13spic0_csr_lbcEiPdPiS1_iS1_S1_S1_"](
        c_int(n), P(val), P(Cp.copy()), P(Ci.copy()), c_int(len(lp) - 1), P(par), P(lp),
        P(ls))
    assert bitwise_equal(val, want)
