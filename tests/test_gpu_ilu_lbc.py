"""GPU parity of the unit-lower (ILU) solves and the redundant LBC solve
against the oracle, through the C ABI and through the sym_lib:: drop-ins
(mangled names, host arrays).  All results are bit-identical: one lane per
row or per partition repeats the reference's operation order.

  sptrsv_csr_ilu            sparse_blas/sptrsv.cpp:28-38 (no divide)
  sptrsv_csr_ilu_lbc        sparse_blas/sptrsv.cpp:109-126 (divides)
  sptrsv_csr_lbc_redundant  sparse_blas/sptrsv.cpp:78-106
"""
import ctypes
import os
from ctypes import c_int, c_void_p

import numpy as np
import pytest

import oracle
import sympiler_amd as sa
from sympiler_amd import workloads as W
from tests.helpers import (KA_B, bitwise_equal, general_with_diags, ilu_lower_host,
                           known_answer_csr, lbc_schedules, to_dev)

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "sympiler_amd", "libsympiler_amd.so")


def P(a):
    return a.ctypes.data_as(c_void_p)


def _general(n, density, seed):
    L = W.random_lower(n, density, 1, seed=seed)
    return general_with_diags(*L, upper_per_row=5, seed=seed)


def _schedules(Lp, Li, dev):
    """The two helper schedules plus the library's own coarsening."""
    lp, ls, _ = oracle.levelset_csr(Lp, Li)
    out = dict(lbc_schedules(lp, ls))
    lvl, pp, part = sa.coarsen(*to_dev(Lp, Li, device=dev), inner_parts=8, min_level_dist=4)
    s = (lvl.cpu().numpy(), pp.cpu().numpy(), part.cpu().numpy())
    assert oracle.check_schedule(Lp, Li, *s) == 0
    out["coarsen"] = s
    return out


@pytest.mark.parametrize("unit", [True, False])
def test_ilu_lower_matches_host(dev, unit):
    Ap, Ai, Ax, diags = _general(3000, 0.003, 3)
    Lp, Li, Lx = sa.ilu_lower(*to_dev(Ap, Ai, Ax, diags, device=dev), unit_diag=unit)
    hp, hi, hx = ilu_lower_host(Ap, Ai, Ax, diags, unit)
    assert np.array_equal(Lp.cpu().numpy(), hp)
    assert np.array_equal(Li.cpu().numpy(), hi)
    assert bitwise_equal(Lx.cpu().numpy(), hx)


def test_ilu_lower_rejects_bad_diagonal_positions(dev):
    Ap, Ai, Ax, diags = _general(200, 0.02, 4)
    bad = diags.copy()
    bad[17] = Ap[18]  # first entry of the next row
    with pytest.raises(sa.SymAmdError) as e:
        sa.ilu_lower(*to_dev(Ap, Ai, Ax, bad, device=dev), unit_diag=True)
    assert e.value.code == 1


@pytest.mark.parametrize("n,density,seed", [(3000, 0.003, 5), (20000, 0.0004, 6)])
def test_ilu_solve_bitwise(dev, n, density, seed):
    Ap, Ai, Ax, diags = _general(n, density, seed)
    b = W.uniform(seed, 0, n)
    x_ref = oracle.sptrsv_csr_ilu(Ap, Ai, Ax, diags, b)
    Lp, Li, Lx = sa.ilu_lower(*to_dev(Ap, Ai, Ax, diags, device=dev), unit_diag=True)
    plan = sa.Plan(Lp, Li, Lx)
    plan.set_lanes(1)
    for ex in ("syncfree", "levelset"):
        x = to_dev(b, device=dev)
        plan.solve(x, executor=ex)
        plan.check()
        assert bitwise_equal(x.cpu().numpy(), x_ref), ex


def test_ilu_lbc_bitwise(dev):
    Ap, Ai, Ax, diags = _general(4000, 0.002, 7)
    b = W.uniform(7, 0, 4000)
    dA = to_dev(Ap, Ai, Ax, diags, device=dev)
    Lp, Li, Lx = sa.ilu_lower(*dA, unit_diag=False)
    hp, hi, _ = ilu_lower_host(Ap, Ai, Ax, diags, False)
    for name, (lvl, pp, part) in _schedules(hp, hi, dev).items():
        x_ref = oracle.sptrsv_csr_ilu_lbc(Ap, Ai, Ax, diags, b, lvl, pp, part)
        x = to_dev(b, device=dev)
        sa.solve_lbc(Lp, Li, Lx, x, lvl, *to_dev(pp, part, device=dev))
        assert bitwise_equal(x.cpu().numpy(), x_ref), name


def test_lbc_redundant_bitwise(dev):
    Lp, Li, Lx = W.random_lower(5000, 0.002, 1, seed=8)
    b = W.uniform(8, 0, 5000)
    x0 = W.uniform(9, 0, 5000)
    dL = to_dev(Lp, Li, Lx, device=dev)
    db = to_dev(b, device=dev)
    for name, (lvl, pp, part) in _schedules(Lp, Li, dev).items():
        x_ref = oracle.sptrsv_csr_lbc_redundant(Lp, Li, Lx, x0, b, lvl, pp, part)
        x = to_dev(x0, device=dev)
        sa.solve_lbc_redundant(*dL, x, db, lvl, *to_dev(pp, part, device=dev))
        assert bitwise_equal(x.cpu().numpy(), x_ref), name


# -------------------------------------------------------------- drop-ins
@pytest.fixture(scope="module")
def L(dev):
    return ctypes.CDLL(LIB)


def test_dropin_ilu_and_ilu_lbc(L):
    Ap, Ai, Ax, diags = _general(3000, 0.003, 11)
    n = len(Ap) - 1
    b = W.uniform(11, 0, n)
    x = b.copy()
    L["_ZN7sym_lib14sptrsv_csr_iluEiPKiS1_PKdS1_Pd"](
        c_int(n), P(Ap), P(Ai), P(Ax), P(diags), P(x))
    assert bitwise_equal(x, oracle.sptrsv_csr_ilu(Ap, Ai, Ax, diags, b))
    hp, hi, _ = ilu_lower_host(Ap, Ai, Ax, diags, False)
    lp, ls, _ = oracle.levelset_csr(hp, hi)
    lvl, pp, part = lbc_schedules(lp, ls)["coarse3"]
    x = b.copy()
    L["_ZN7sym_lib18sptrsv_csr_ilu_lbcEiPKiS1_PKdS1_PdiS1_S1_S1_"](
        c_int(n), P(Ap), P(Ai), P(Ax), P(diags), P(x), c_int(len(lvl) - 1), P(lvl),
        P(pp), P(part))
    assert bitwise_equal(x, oracle.sptrsv_csr_ilu_lbc(Ap, Ai, Ax, diags, b, lvl, pp, part))


def test_dropin_lbc_redundant_known_answer(L):
    # Catch_tests/sptrsv_tests.cpp:17-44 data (x == 1 exactly), on both
    # helper schedules; tmp is accepted as in the reference (threads*n)
    Rp, Ri, Rx = known_answer_csr()
    lp, ls, _ = oracle.levelset_csr(Rp, Ri)
    fn = L["_ZN7sym_lib24sptrsv_csr_lbc_redundantEiPiS0_PdS1_S1_iS0_S0_S0_S1_"]
    for lvl, pp, part in lbc_schedules(lp, ls).values():
        x = np.zeros(10)
        b = KA_B.copy()
        tmp = np.zeros(8 * 10)
        fn(c_int(10), P(Rp), P(Ri), P(Rx), P(x), P(b), c_int(len(lvl) - 1), P(lvl),
           P(pp), P(part), P(tmp))
        assert np.all(x == 1.0)
