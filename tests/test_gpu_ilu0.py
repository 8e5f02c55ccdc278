"""ILU(0) on the device (symamd_ilu0 / sa.ilu0) against the oracle's
restatement of sparse_blas/spilu0.cpp:11-51: bit-identical factors on grid
and random general (structurally non-symmetric) matrices and rows longer
than the register path, zero-pivot and structure errors, in-place use, the
sym_lib drop-ins, and the factor's unit-lower part solved on the device."""
import ctypes
import os
from ctypes import c_int, c_void_p

import numpy as np
import pytest

import oracle
import sympiler_amd as sa
from sympiler_amd import workloads as W
from tests.helpers import bitwise_equal, general_with_diags, to_dev
from tests.test_ilu0_oracle import dominant, general, symmetric_pattern

pytestmark = pytest.mark.gpu


def grid(nx, ny, nz, seed):
    Lp, Li, _ = W.lap3d_tril(nx, ny, nz, 6.0, -1.0)
    Ap, Ai, diags = symmetric_pattern(Lp, Li)
    return Ap, Ai, dominant(Ap, Ai, diags, seed), diags


def wide(n, seed):
    Lp, Li, _ = W.random_lower(n, 0.004, 1, seed=seed)
    Ap, Ai, _, diags = general_with_diags(Lp, Li, np.zeros(len(Li)), 14, seed)
    return Ap, Ai, dominant(Ap, Ai, diags, seed), diags


MATS = {
    "grid_12x10x8": lambda: grid(12, 10, 8, 5),
    "general_4000": lambda: general(4000, 0.001, 3, 6),
    "wide_rows_3000": lambda: wide(3000, 7),
}


@pytest.mark.parametrize("name", list(MATS))
def test_ilu0_bitwise(dev, name):
    Ap, Ai, Ax, diags = MATS[name]()
    print(f"[ilu0] {name} n={len(Ap) - 1} nnz={len(Ai)}", flush=True)
    want, bad = oracle.spilu0_csr(Ap, Ai, Ax, diags)
    assert bad == -1
    dAp, dAi, dAx, dd = to_dev(Ap, Ai, Ax, diags)
    got = sa.ilu0(dAp, dAi, dAx, dd)
    assert bitwise_equal(got.cpu().numpy(), want)
    again = dAx.clone()
    sa.ilu0(dAp, dAi, again, dd, out=again)
    assert bitwise_equal(again.cpu().numpy(), want)
    assert bitwise_equal(dAx.cpu().numpy(), Ax)


def test_ilu0_plan_refactors_new_values(dev):
    Ap, Ai, Ax, diags = grid(9, 8, 7, 13)
    dAp, dAi, dd = to_dev(Ap, Ai, diags)
    plan = sa.Ilu0Plan(dAp, dAi, dd)
    for seed in (14, 15):
        Bx = dominant(Ap, Ai, diags, seed)
        want, bad = oracle.spilu0_csr(Ap, Ai, Bx, diags)
        assert bad == -1
        got = plan.factor(to_dev(Bx))
        assert bitwise_equal(got.cpu().numpy(), want)


def test_ilu0_zero_pivot_and_structure(dev):
    # 2 x 2 ones: U[1,1] = 0
    Ap = np.array([0, 2, 4], np.int32)
    Ai = np.array([0, 1, 0, 1], np.int32)
    d = np.array([0, 3], np.int32)
    dAp, dAi, dAx, dd = to_dev(Ap, Ai, np.ones(4), d)
    with pytest.raises(ValueError, match="breakdown"):
        sa.ilu0(dAp, dAi, dAx, dd)
    l, bad = sa.ilu0(dAp, dAi, dAx, dd, strict=False)
    assert bad == 1 and list(l.cpu().numpy()) == [1.0, 1.0, 1.0, 0.0]
    # a diagonal position that is not on the diagonal
    with pytest.raises(sa.SymAmdError, match="STRUCTURE|diagonal"):
        sa.ilu0(dAp, dAi, dAx, to_dev(np.array([1, 3], np.int32)))


def test_ilu0_then_unit_lower_solve(dev):
    # the factor's L (unit diagonal) is what sptrsv_csr_ilu solves with
    Ap, Ai, Ax, diags = grid(10, 9, 8, 8)
    want, _ = oracle.spilu0_csr(Ap, Ai, Ax, diags)
    dAp, dAi, dAx, dd = to_dev(Ap, Ai, Ax, diags)
    l = sa.ilu0(dAp, dAi, dAx, dd)
    Lp, Li, Lx = sa.ilu_lower(dAp, dAi, l, dd, unit_diag=True)
    plan = sa.Plan(Lp, Li, Lx)
    plan.set_lanes(1)
    n = len(Ap) - 1
    b = W.uniform(12, 0, n)
    x = to_dev(b.copy())
    plan.solve(x, executor="syncfree")
    plan.check()
    assert bitwise_equal(x.cpu().numpy(), oracle.sptrsv_csr_ilu(Ap, Ai, want, diags, b))


LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "sympiler_amd", "libsympiler_amd.so")


def P(a):
    return a.ctypes.data_as(c_void_p)


def test_spilu0_dropins(dev):
    L = ctypes.CDLL(LIB)
    Ap, Ai, Ax, diags = general(2500, 0.002, 3, 9)
    want, bad = oracle.spilu0_csr(Ap, Ai, Ax, diags)
    assert bad == -1
    n, nnz = len(Ap) - 1, len(Ai)
    l = np.zeros(nnz)
    L["_ZN7sym_lib10spilu0_csrEiiPKiS1_PKdS1_Pd"](c_int(n), c_int(nnz), P(Ap), P(Ai), P(Ax),
                                                  P(diags), P(l))
    assert bitwise_equal(l, want)
    lp = np.array([0, 1], np.int32)
    pp = np.array([0, n], np.int32)
    part = np.arange(n, dtype=np.int32)
    tmp = np.full(n * 16, -1, np.int32)
    l = np.zeros(nnz)
    L["_ZN7sym_lib14spilu0_csr_lbcEiiPKiS1_PKdS1_PdiS1_S1_S1_Pi"](
        c_int(n), c_int(nnz), P(Ap), P(Ai), P(Ax), P(diags), P(l), c_int(1), P(lp), P(pp),
        P(part), P(tmp))
    assert bitwise_equal(l, want)
