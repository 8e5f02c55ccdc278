"""The tree-level coarsening's schedule driving the device LBC solve
(symamd_solve_lbc, the sptrsv_csr_lbc semantics, sparse_blas/sptrsv.cpp:57-75):
bit-identical to the serial solve for lower triangles whose dependencies the
tree covers -- a tree-shaped factor, a random postordered forest, and an
IC(0) factor of a 3D grid with its elimination tree."""
import numpy as np
import pytest

import oracle
import sympiler_amd as sa
from sympiler_amd import workloads as W
from tests.golden.make_tree_lbc import random_forest
from tests.helpers import (bitwise_equal, elimination_tree, etree_lower,
                           postorder_binary_tree, to_dev)

pytestmark = pytest.mark.gpu


def cases():
    rng = np.random.default_rng(21)
    bt = postorder_binary_tree(12)
    yield "binary_tree_4095", bt, etree_lower(bt)
    rf = random_forest(rng, 20000, 8, True)
    yield "random_forest_20000", rf, etree_lower(rf)
    Lp, Li, A = W.lap3d_tril(12, 12, 12)
    yield "ic0_grid12", elimination_tree(Lp, Li), (Lp, Li, W.ic0(Lp, Li, A))


@pytest.mark.parametrize("inner_parts,min_level_dist,div_rate", [(8, 1, 2), (64, 2, 2), (1, 1, 2)])
def test_tree_schedule_solve(inner_parts, min_level_dist, div_rate):
    rng = np.random.default_rng(3)
    for name, parent, (Lp, Li, Lx) in cases():
        lp, pp, part = sa.tree_coarsen(parent, np.diff(Lp).astype(np.float64),
                                       inner_parts, min_level_dist, div_rate)
        assert oracle.check_schedule(Lp, Li, lp, pp, part) == 0, name
        b = rng.standard_normal(len(Lp) - 1)
        want = oracle.sptrsv_csr(Lp, Li, Lx, b)
        dLp, dLi, dLx, x, dpp, dpart = to_dev(Lp, Li, Lx, b, pp, part)
        sa.solve_lbc(dLp, dLi, dLx, x, lp, dpp, dpart)
        assert bitwise_equal(x.cpu().numpy(), want), name
