"""Supernodal (BCSC) solve on the device (symamd_sn_* / SupernodalPlan and
the sym_lib::sptrsv_csc_block drop-ins) against the oracle's restatement of
sparse_blas/sptrsv.cpp:176-199: bit-identical on widths from 1 to past one
wavefront and row blocks longer than a wavefront, repeated solves, and the
structure errors."""
import ctypes
import os
from ctypes import c_int, c_void_p

import numpy as np
import pytest

import oracle
import sympiler_amd as sa
from sympiler_amd import workloads as W
from tests.helpers import bitwise_equal, to_dev
from tests.test_supernodal_oracle import CASES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", CASES + [(60000, 64, 64, 8192, 7)])
def test_supernodal_bitwise(dev, case):
    n = case[0]
    sup, nrows, Lp, Li, Lx = W.bcsc(*case)
    print(f"[sn] case={case} supernodes={len(sup) - 1} values={len(Lx)}", flush=True)
    b = W.uniform(case[-1], 0, n)
    want = oracle.sptrsv_csc_block(sup, nrows, Lp, Li, Lx, b)
    plan = sa.SupernodalPlan(sup, nrows, Lp, Li)
    dLx = to_dev(Lx)
    for _ in range(2):  # the second solve reuses the tags (next epoch)
        x = to_dev(b.copy())
        plan.solve(dLx, x)
        plan.check()
        assert bitwise_equal(x.cpu().numpy(), want)


@pytest.mark.parametrize("k", [1, 3])
def test_supernodal_wide_tops_bitwise(dev, k):
    """Tops from a few columns to 823 wide (five past 512, three in (256, 512]),
    up to 682 rows below: the gather in two 512-row halves, the top's update
    in several passes, four column blocks per product batch (k = 1 kernel) or
    two (block kernel), and the transposed solve with 16 column tiles per
    lane; forward and transposed, every column bit-identical to the oracle."""
    n = 5000
    sup, nrows, Lp, Li, Lx = W.bcsc(n, 1000, 700, 4000, 9)
    widths = np.diff(sup)
    assert widths.max() > 512 and ((widths > 256) & (widths <= 512)).any()
    print(f"[sn] wide tops k={k} supernodes={len(sup) - 1} widest={widths.max()}", flush=True)
    B = W.uniform(21, 0, k * n).reshape(k, n)
    plan = sa.SupernodalPlan(sup, nrows, Lp, Li)
    dLx = to_dev(Lx)
    for solve, ref in ((plan.solve, oracle.sptrsv_csc_block),
                       (plan.solve_transposed, oracle.sptrsv_csc_block_t)):
        X = to_dev(B.copy())
        solve(dLx, X if k > 1 else X[0])
        plan.check()
        got = X.cpu().numpy()
        for c in range(k):
            assert bitwise_equal(got[c], ref(sup, nrows, Lp, Li, Lx, B[c])), (solve, c)


@pytest.mark.parametrize("k", [3, 32])
def test_supernodal_block_columns_bitwise(dev, k):
    n = 3000
    sup, nrows, Lp, Li, Lx = W.bcsc(n, 100, 70, 1500, 3)
    B = W.uniform(11, 0, k * (n + 7)).reshape(k, n + 7)  # column-major, ld = n + 7
    plan = sa.SupernodalPlan(sup, nrows, Lp, Li)
    dLx = to_dev(Lx)
    for _ in range(2):
        X = to_dev(B.copy())
        plan.solve(dLx, X[:, :n])
        plan.check()
        got = X.cpu().numpy()
        for c in range(k):
            want = oracle.sptrsv_csc_block(sup, nrows, Lp, Li, Lx, B[c, :n])
            assert bitwise_equal(got[c, :n], want), c
        assert bitwise_equal(got[:, n:], B[:, n:])  # the padding is untouched
    # the single-column solve after a block one (tags of every column reused)
    x = to_dev(B[0, :n].copy())
    plan.solve(dLx, x)
    plan.check()
    assert bitwise_equal(x.cpu().numpy(), oracle.sptrsv_csc_block(sup, nrows, Lp, Li, Lx,
                                                                  B[0, :n]))


def test_supernodal_structure_errors(dev):
    sup, nrows, Lp, Li, Lx = W.bcsc(300, 8, 8, 60, 1)
    i = int(np.argmax(nrows - np.diff(sup)))  # a supernode with rows below
    bad = Li.copy()
    bad[Lp[i] + (sup[i + 1] - sup[i])] = sup[i]  # a row inside the supernode
    with pytest.raises(sa.SymAmdError, match="block row"):
        sa.SupernodalPlan(sup, nrows, Lp, bad)
    wide = np.array([0, 1100], np.int32)
    with pytest.raises(sa.SymAmdError, match="wide"):
        sa.SupernodalPlan(wide, np.array([1100], np.int32), np.array([0, 1100 * 1100], np.int32),
                          np.arange(1100, dtype=np.int32))


LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "sympiler_amd", "libsympiler_amd.so")


def P(a):
    return a.ctypes.data_as(c_void_p)


def test_sptrsv_csc_block_dropins(dev):
    L = ctypes.CDLL(LIB)
    n = 3000
    sup, nrows, Lp, Li, Lx = W.bcsc(n, 40, 24, 400, 2)
    b = W.uniform(2, 0, n)
    want = oracle.sptrsv_csc_block(sup, nrows, Lp, Li, Lx, b)
    m = len(sup) - 1
    x = b.copy()
    L["_ZN7sym_lib16sptrsv_csc_blockEiPKiS1_S1_PdS2_iS1_"](
        c_int(n), P(Lp), P(Li), P(nrows), P(Lx), P(x), c_int(m), P(sup))
    assert bitwise_equal(x, want)
    # the level-scheduled variant: one level per supernode, sup2node = identity
    levelPtr = np.arange(m + 1, dtype=np.int32)
    levels = sup[:-1].copy()
    sup2node = np.arange(m, dtype=np.int32)
    x = b.copy()
    L["_ZN7sym_lib25sptrsv_csc_levelset_blockEiPKiS1_S1_PdS2_S1_S1_iS1_S1_PS2_"](
        c_int(n), P(Lp), P(Li), P(nrows), P(Lx), P(x), P(levelPtr), P(levels), c_int(m),
        P(sup), P(sup2node), None)
    assert bitwise_equal(x, want)


def test_sptrsv_csc_levelset_block_tree_schedule(dev):
    """sptrsv_csc_levelset_block with several supernodes per level, listed
    in descending order inside each level (sup2node not the identity):
    bit-identical to the serial sptrsv_csc_block."""
    from tests.helpers import postorder_binary_tree
    L = ctypes.CDLL(LIB)
    levels = 8
    sup, nrows, Lp, Li, Lx = W.bcsc_tree(levels, 4, 64, 2.0, 5)
    n, m = int(sup[-1]), len(sup) - 1
    parent = postorder_binary_tree(levels)
    depth = np.zeros(m, np.int64)  # height above the leaves: a valid level
    for i in range(m):  # postorder: children before parents
        if parent[i] >= 0:
            depth[parent[i]] = max(depth[parent[i]], depth[i] + 1)
    order = np.lexsort((-np.arange(m), depth)).astype(np.int32)
    levelPtr = np.searchsorted(depth[order], np.arange(depth.max() + 2)).astype(np.int32)
    b = W.uniform(5, 0, n)
    want = oracle.sptrsv_csc_block(sup, nrows, Lp, Li, Lx, b)
    x = b.copy()
    L["_ZN7sym_lib25sptrsv_csc_levelset_blockEiPKiS1_S1_PdS2_S1_S1_iS1_S1_PS2_"](
        c_int(n), P(Lp), P(Li), P(nrows), P(Lx), P(x), P(levelPtr),
        P(np.ascontiguousarray(sup[order])), c_int(len(levelPtr) - 1), P(sup), P(order), None)
    assert bitwise_equal(x, want)


def test_supernodal_ticket_range_rejected(dev):
    """k columns whose (task, column) tickets would pass 2^31 - 1 are
    refused before any device work."""
    sup, nrows, Lp, Li, Lx = W.bcsc(300, 8, 8, 60, 1)
    plan = sa.SupernodalPlan(sup, nrows, Lp, Li)
    dLx = to_dev(Lx)
    x = to_dev(W.uniform(1, 0, 300))
    k = (2 ** 31 - 1) // 2  # far more tickets than 2^31 for any task count >= 2
    rc = sa.lib().symamd_sn_solve_block(plan.h._h, plan._p, c_void_p(dLx.data_ptr()),
                                        c_void_p(x.data_ptr()), c_int(k), 300)
    assert rc == 7 and b"ticket" in sa.lib().symamd_last_error()


@pytest.mark.parametrize("case", [(3000, 40, 24, 400, 2), (20000, 100, 130, 4096, 5)])
def test_parsy_blocked_lsolve_dropin(dev, case):
    """sym_lib::parsy::blockedLsolve on parsy's storage of the same factor:
    bit-identical to the sptrsv_csc_block restatement (the two reference
    solves share their dense kernels' operation order)."""
    from tests.test_supernodal_oracle import to_parsy
    L = ctypes.CDLL(LIB)
    n = case[0]
    sup, nrows, Lp, Li, Lx = W.bcsc(*case)
    b = W.uniform(case[-1], 0, n)
    want = oracle.sptrsv_csc_block(sup, nrows, Lp, Li, Lx, b)
    vp, rows, rp = to_parsy(sup, nrows, Lp, Li, n)
    col2sup = np.repeat(np.arange(len(sup) - 1, dtype=np.int32), np.diff(sup))
    x = b.copy()
    fn = L["_ZN7sym_lib5parsy13blockedLsolveEiPmPiPdiS1_S2_S2_iS3_"]
    fn.restype = c_int
    r = fn(c_int(n), P(vp), P(rows), P(Lx), c_int(len(Lx)), P(rp), P(col2sup), P(sup),
           c_int(len(sup) - 1), P(x))
    assert r == 1
    assert bitwise_equal(x, want)


def test_parsy_leveled_blocked_lsolve_with_tree_schedule(dev):
    """The tree-level coarsening's schedule of the supernodal tree feeding
    sym_lib::parsy::H2LeveledBlockedLsolve: every supernode's consumers are
    its tree ancestors, and x equals the serial order's bit for bit (the
    reference's atomics reorder the subtractions; it agrees to rounding)."""
    from tests.helpers import postorder_binary_tree, rel_err
    from tests.test_supernodal_oracle import to_parsy
    levels = 9
    sup, nrows, Lp, Li, Lx = W.bcsc_tree(levels, 4, 64, 2.0, 3)
    n, m = int(sup[-1]), len(sup) - 1
    parent = postorder_binary_tree(levels)
    col2sup = np.repeat(np.arange(m, dtype=np.int32), np.diff(sup))
    for i in range(m):
        anc, a = set(), parent[i]
        while a >= 0:
            anc.add(int(a))
            a = parent[a]
        w = sup[i + 1] - sup[i]
        assert {int(col2sup[r]) for r in Li[Lp[i] + w:Lp[i] + nrows[i]]} <= anc
    cost = (np.diff(sup) * nrows).astype(np.float64)
    lp, pp, part = sa.tree_coarsen(parent, cost, 8, 1, 2)
    b = W.uniform(3, 0, n)
    want = oracle.sptrsv_csc_block(sup, nrows, Lp, Li, Lx, b)
    vp, rows, rp = to_parsy(sup, nrows, Lp, Li, n)
    x = b.copy()
    fn = ctypes.CDLL(LIB)["_ZN7sym_lib5parsy22H2LeveledBlockedLsolveEiPmPiPdiS1_S2_S2_iS3_iS2_S2_iS2_S2_i"]
    fn.restype = c_int
    r = fn(c_int(n), P(vp), P(rows), P(Lx), c_int(len(Lx)), P(rp), P(col2sup), P(sup),
           c_int(m), P(x), c_int(len(lp) - 1), P(lp), None, c_int(len(pp) - 1), P(pp),
           P(part), c_int(1))
    assert r == 1
    assert bitwise_equal(x, want)
    assert rel_err(x, want) <= 1e-12


@pytest.mark.parametrize("case", CASES + [(60000, 64, 64, 8192, 7), (20000, 300, 600, 4096, 8)])
def test_supernodal_transposed_bitwise(dev, case):
    """L^T x = b on the same plan (blockedLTsolve's semantics): bit-identical
    to the oracle's restatement (pinned by dense solves and exact integer
    cases), on tops wider than one column chunk and blocks with more rows
    below than one pass of the rows."""
    n = case[0]
    sup, nrows, Lp, Li, Lx = W.bcsc(*case)
    print(f"[snT] case={case} supernodes={len(sup) - 1} values={len(Lx)}", flush=True)
    b = W.uniform(case[-1], 1, n)
    want = oracle.sptrsv_csc_block_t(sup, nrows, Lp, Li, Lx, b)
    plan = sa.SupernodalPlan(sup, nrows, Lp, Li)
    dLx = to_dev(Lx)
    for _ in range(2):  # the second solve reuses the tags (next epoch)
        x = to_dev(b.copy())
        plan.solve_transposed(dLx, x)
        plan.check()
        assert bitwise_equal(x.cpu().numpy(), want)
    # the forward solve after the transposed one, on the same plan
    x = to_dev(b.copy())
    plan.solve(dLx, x)
    plan.check()
    assert bitwise_equal(x.cpu().numpy(), oracle.sptrsv_csc_block(sup, nrows, Lp, Li, Lx, b))


@pytest.mark.parametrize("k", [1, 3, 32])
def test_supernodal_transposed_block_columns_bitwise(dev, k):
    n = 3000
    sup, nrows, Lp, Li, Lx = W.bcsc(n, 100, 70, 1500, 3)
    B = W.uniform(12, 0, k * (n + 7)).reshape(k, n + 7)  # column-major, ld = n + 7
    plan = sa.SupernodalPlan(sup, nrows, Lp, Li)
    dLx = to_dev(Lx)
    for _ in range(2):
        X = to_dev(B.copy())
        plan.solve_transposed(dLx, X[:, :n])
        plan.check()
        got = X.cpu().numpy()
        for c in range(k):
            want = oracle.sptrsv_csc_block_t(sup, nrows, Lp, Li, Lx, B[c, :n])
            assert bitwise_equal(got[c, :n], want), c
        assert bitwise_equal(got[:, n:], B[:, n:])  # the padding is untouched


def _parsy_args(sup, nrows, Lp, Li, Lx, n):
    from tests.test_supernodal_oracle import to_parsy
    vp, rows, rp = to_parsy(sup, nrows, Lp, Li, n)
    col2sup = np.repeat(np.arange(len(sup) - 1, dtype=np.int32), np.diff(sup))
    return [c_int(n), P(vp), P(rows), P(Lx), c_int(len(Lx)), P(rp), P(col2sup), P(sup),
            c_int(len(sup) - 1)], (vp, rows, rp, col2sup)


MANGLED = {
    "blockedLsolve_mrhs": "_ZN7sym_lib5parsy18blockedLsolve_mrhsEiPmPiPdiS1_S2_S2_iS3_ii",
    "blockedLTsolve": "_ZN7sym_lib5parsy14blockedLTsolveEiPmPiPdiS1_S2_S2_iS3_",
    "blockedLTsolve_mrhs": "_ZN7sym_lib5parsy19blockedLTsolve_mrhsEiPmPiPdiS1_S2_S2_iS3_ii",
    "H2LeveledBlockedLsolve_mrhs":
        "_ZN7sym_lib5parsy27H2LeveledBlockedLsolve_mrhsEiPmPiPdiS1_S2_S2_iS3_iiS2_S2_iS2_S2_ii",
    "H2LeveledBlockedLTsolve":
        "_ZN7sym_lib5parsy23H2LeveledBlockedLTsolveEiPmPiPdiS1_S2_S2_iS3_iS2_S2_iS2_S2_i",
    "H2LeveledBlockedLTsolve_mrhs":
        "_ZN7sym_lib5parsy28H2LeveledBlockedLTsolve_mrhsEiPmPiPdiS1_S2_S2_iS3_iiS2_S2_iS2_S2_ii",
}


def _fn(name):
    f = ctypes.CDLL(LIB)[MANGLED[name]]
    f.restype = c_int
    return f


@pytest.mark.parametrize("case", [(3000, 40, 24, 400, 2), (20000, 100, 130, 4096, 5)])
def test_parsy_blocked_transposed_and_multi_rhs_dropins(dev, case):
    """blockedLTsolve on parsy's storage: the transposed restatement bit for
    bit; blockedLsolve_mrhs / blockedLTsolve_mrhs: every column of x[r + j*n]
    as the single-column solves give it (the reference's own multi-column
    builds solve only the first column or the transposed diagonal block)."""
    n = case[0]
    sup, nrows, Lp, Li, Lx = W.bcsc(*case)
    args, keep = _parsy_args(sup, nrows, Lp, Li, Lx, n)
    b = W.uniform(case[-1], 2, n)
    x = b.copy()
    assert _fn("blockedLTsolve")(*args, P(x)) == 1
    assert bitwise_equal(x, oracle.sptrsv_csc_block_t(sup, nrows, Lp, Li, Lx, b))
    k = 4
    B = W.uniform(case[-1], 3, k * n).reshape(k, n)  # column j at x[j*n]
    for name, ref in (("blockedLsolve_mrhs", oracle.sptrsv_csc_block),
                      ("blockedLTsolve_mrhs", oracle.sptrsv_csc_block_t)):
        X = B.copy()
        assert _fn(name)(*args, P(X), c_int(k), c_int(-1)) == 1
        for j in range(k):
            assert bitwise_equal(X[j], ref(sup, nrows, Lp, Li, Lx, B[j])), (name, j)
    # a null right-hand side returns 0, as the reference does
    assert _fn("blockedLTsolve")(*args, None) == 0


def test_parsy_leveled_transposed_and_multi_rhs_with_tree_schedule(dev):
    """The leveled variants on the tree-level coarsening's schedule of the
    supernodal tree: x as the serial solves give it, every column."""
    from tests.helpers import postorder_binary_tree
    levels = 9
    sup, nrows, Lp, Li, Lx = W.bcsc_tree(levels, 4, 64, 2.0, 3)
    n, m = int(sup[-1]), len(sup) - 1
    parent = postorder_binary_tree(levels)
    cost = (np.diff(sup) * nrows).astype(np.float64)
    lp, pp, part = sa.tree_coarsen(parent, cost, 8, 1, 2)
    sched = [c_int(len(lp) - 1), P(lp), None, c_int(len(pp) - 1), P(pp), P(part), c_int(1)]
    args, keep = _parsy_args(sup, nrows, Lp, Li, Lx, n)
    b = W.uniform(5, 0, n)
    x = b.copy()
    assert _fn("H2LeveledBlockedLTsolve")(*args, P(x), *sched) == 1
    assert bitwise_equal(x, oracle.sptrsv_csc_block_t(sup, nrows, Lp, Li, Lx, b))
    k = 3
    B = W.uniform(6, 0, k * n).reshape(k, n)
    for name, ref in (("H2LeveledBlockedLsolve_mrhs", oracle.sptrsv_csc_block),
                      ("H2LeveledBlockedLTsolve_mrhs", oracle.sptrsv_csc_block_t)):
        X = B.copy()
        assert _fn(name)(*args, P(X), c_int(k), *sched[:-1], c_int(1), c_int(-1)) == 1
        for j in range(k):
            assert bitwise_equal(X[j], ref(sup, nrows, Lp, Li, Lx, B[j])), (name, j)
