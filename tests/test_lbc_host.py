"""CPU checks of the coarsened-schedule restatement (tests/lbc_host.py) that
pins the device inspector symamd_coarsen: the reference's schedule invariants
(every row once; every dependency in an earlier coarse level or earlier in
the same part, oracle.check_schedule), bit-exact LBC solves through the
reference build when present, and closed forms on grids and trees."""
import numpy as np
import pytest

import oracle
from sympiler_amd import workloads as W
from tests.helpers import bitwise_equal, etree_lower, postorder_binary_tree
from tests.lbc_host import lbc_schedule

CASES = {
    "random": lambda: W.random_lower(3000, 0.002, 1, seed=12),
    "zipf": lambda: W.zipf_lower(4000, 1.8, 256, 2000, 9),
    "lap2d": lambda: W.lap2d_tril(40, 30),
    "lap3d": lambda: W.lap3d_tril(9, 8, 7),
    "tree": lambda: etree_lower(postorder_binary_tree(9)),
    "diag": lambda: W.random_lower(500, 0.0, 0, seed=1),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("params", [(6, -1, 4000), (1, -1, 4000), (4, 3, 4000),
                                    (32, -1, 8), (3, 0, 1)])
def test_schedule_invariants_and_solve(case, params):
    Lp, Li, Lx = CASES[case]()
    lp, pp, part = lbc_schedule(Lp, Li, None, *params)
    assert oracle.check_schedule(Lp, Li, lp, pp, part) == 0
    # at most `parts` parts per coarse level, none empty
    assert np.all(np.diff(lp) <= max(params[0], 1)) and np.all(np.diff(pp) > 0)
    b = W.uniform(3, 0, len(Lp) - 1)
    x_ref = oracle.sptrsv_csr(Lp, Li, Lx, b)
    assert bitwise_equal(oracle.sptrsv_csr_lbc(Lp, Li, Lx, b, lp, pp, part), x_ref)
    if oracle.ref_available():
        assert bitwise_equal(oracle.ref.sptrsv_csr_lbc(Lp, Li, Lx, b, lp, pp, part), x_ref)


def test_grid_keeps_single_levels():
    # two consecutive levels of a 2D grid in natural order form one component,
    # so with more than one part no merge is usable: one coarse level per level
    Lp, Li, _ = W.lap2d_tril(40, 30)
    lp, pp, part = lbc_schedule(Lp, Li, None, 6, -1, 4000)
    assert len(lp) - 1 == 40 + 30 - 1
    # with one part every block is usable: the dyadic cover of 69 levels
    lp, _, _ = lbc_schedule(Lp, Li, None, 1, -1, 4000)
    assert len(lp) - 1 == 2          # [0, 64), [64, 69)
    lp, _, _ = lbc_schedule(Lp, Li, None, 1, -1, 16)
    assert len(lp) - 1 == 5          # four blocks of 16, then [64, 69)
    lp, _, _ = lbc_schedule(Lp, Li, None, 6, 4, 4000)
    assert len(lp) - 1 == (69 + 3) // 4


def test_binary_tree_closed_form():
    # complete binary tree of 8 levels (255 rows, leaves at level 0): with 4
    # parts, levels [0, 4) hold 16 equal subtrees, [4, 6) four, [6, 8) one
    # (unusable), so levels 6 and 7 stay single
    Lp, Li, _ = etree_lower(postorder_binary_tree(8))
    lp, pp, part = lbc_schedule(Lp, Li, None, 4, -1, 4000)
    assert len(lp) - 1 == 4
    assert list(np.diff(lp)) == [4, 4, 2, 1]
    sizes = np.diff(pp)
    assert list(sizes[:4]) == [60, 60, 60, 60]  # 16 subtrees of 15 rows, 4 per part
    assert list(sizes[4:8]) == [3, 3, 3, 3]     # four subtrees of 3 rows (levels 4, 5)
    assert list(sizes[8:]) == [1, 1, 1]         # level 6 (two rows), level 7 (the root)


def test_costs_steer_the_packing():
    # one costly subtree among equal ones: the worst fit gives it a part alone
    parent = postorder_binary_tree(6)
    Lp, Li, _ = etree_lower(parent)
    n = len(Lp) - 1
    cost = np.ones(n)
    cost[0] = 1000.0  # a leaf of the first subtree
    lp, pp, part = lbc_schedule(Lp, Li, cost, 4, 2, 4000)
    first = part[pp[0]:pp[1]]
    assert 0 in first
    assert oracle.check_schedule(Lp, Li, lp, pp, part) == 0
