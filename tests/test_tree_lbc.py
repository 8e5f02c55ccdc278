"""Tree-level coarsening (symamd_tree_coarsen / sym_lib::parsy::
getCoarseLevelSet_6, parsy/symbolic/InspectionLevel_06.cpp:17): host-only,
so every check runs on the CPU.

Pinned by: the reference's own compiled routine (oracle/_ref) on random
forests, ties and degenerate trees; the committed fixtures it wrote
(tests/golden/tree_lbc.npz, for machines without the reference); closed
forms (chain, star, isolated nodes ordered by cost); the schedule invariants
the LBC solves rely on; and bit-exact LBC solves driven by the schedule."""
import ctypes
import os

import numpy as np
import pytest

import oracle
import sympiler_amd as sa
from tests.golden.make_tree_lbc import random_forest
from tests.helpers import (elimination_tree, etree_lower, postorder_binary_tree,
                           bitwise_equal)

HERE = os.path.dirname(os.path.abspath(__file__))
needs_ref = pytest.mark.skipif(not oracle.ref_available(),
                               reason="oracle/_ref not built (needs /root/reference)")


def same(a, b):
    return all(np.array_equal(np.asarray(x), np.asarray(y)) for x, y in zip(a, b))


def check_invariants(parent, lp, pp, part):
    """Each node once; inside a part children first; parts of one coarse
    level share no edge; every other edge runs to a later coarse level."""
    n = len(parent)
    assert lp[0] == 0 and pp[0] == 0 and pp[-1] == n and lp[-1] == len(pp) - 1
    assert np.all(np.diff(lp) >= 0) and np.all(np.diff(pp) >= 0)
    assert sorted(part.tolist()) == list(range(n))
    level_of = np.empty(n, np.int64)
    part_of = np.empty(n, np.int64)
    pos = np.empty(n, np.int64)
    for l in range(len(lp) - 1):
        for p in range(lp[l], lp[l + 1]):
            for k in range(pp[p], pp[p + 1]):
                level_of[part[k]], part_of[part[k]], pos[part[k]] = l, p, k
    for v, p in enumerate(parent):
        if p < 0:
            continue
        if level_of[p] == level_of[v]:
            assert part_of[p] == part_of[v], "edge between parts of one level"
            assert pos[v] < pos[p], "parent before child inside a part"
        else:
            assert level_of[p] > level_of[v], "edge to an earlier level"


@needs_ref
@pytest.mark.parametrize("seed", range(6))
def test_matches_reference_random(seed):
    rng = np.random.default_rng(100 + seed)
    for t in range(60):
        n = int(rng.integers(1, 600))
        roots = int(rng.integers(1, max(2, n // 10)))
        parent = random_forest(rng, n, min(roots, n), bool(t % 2))
        # integer-valued costs: the reference build asserts an integer sum
        cost = (np.ones(n) if t % 3 == 0 else
                rng.integers(1, 5 if t % 3 == 1 else 5000, n).astype(np.float64))
        ip = int(rng.choice([1, 2, 3, 4, 8, 16, 64, 256]))
        mld = int(rng.integers(-1, 6))
        dr = int(rng.choice([2, 3, 4, -2]))
        got = sa.tree_coarsen(parent, cost, ip, mld, dr)
        want = oracle.ref.getCoarseLevelSet_6(parent, cost, ip, mld, dr)
        assert same(got, want), (seed, t, n, ip, mld, dr)
        check_invariants(parent, *got)


@needs_ref
def test_matches_reference_shapes():
    bt = postorder_binary_tree(11)
    Lp, Li, _ = sa.workloads.lap3d_tril(8, 8, 8)
    et = elimination_tree(Lp, Li)
    rng = np.random.default_rng(5)
    for parent in (bt, et):
        for ip, mld, dr in [(1, 1, 2), (2, 0, 2), (8, 2, 2), (32, 1, 4), (5, 3, 3)]:
            cost = rng.integers(1, 50, len(parent)).astype(np.float64)
            got = sa.tree_coarsen(parent, cost, ip, mld, dr)
            assert same(got, oracle.ref.getCoarseLevelSet_6(parent, cost, ip, mld, dr))
            check_invariants(parent, *got)


def test_golden_fixtures():
    f = np.load(os.path.join(HERE, "golden", "tree_lbc.npz"))
    for c in range(int(f["count"][0])):
        ip, mld, dr = (int(v) for v in f[f"params{c}"])
        got = sa.tree_coarsen(f[f"parent{c}"], f[f"cost{c}"], ip, mld, dr)
        want = (f[f"level_ptr{c}"], f[f"part_ptr{c}"], f[f"partition{c}"])
        assert same(got, want), c
        check_invariants(f[f"parent{c}"], *got)


def test_closed_forms():
    n = 50
    chain = np.arange(1, n + 1, dtype=np.int32)
    chain[-1] = -1
    # a chain never has a level wider than the part count: one level, one
    # part, the nodes bottom-up
    lp, pp, part = sa.tree_coarsen(chain, np.ones(n), 4, 1, 2)
    assert lp.tolist() == [0, 1] and pp.tolist() == [0, n]
    assert part.tolist() == list(range(n))
    # a star has two levels: one part, the leaves by id, then the root
    star = np.full(10, 9, np.int32)
    star[-1] = -1
    lp, pp, part = sa.tree_coarsen(star, np.ones(10), 4, 1, 2)
    assert lp.tolist() == [0, 1] and pp.tolist() == [0, 10]
    assert part.tolist() == list(range(10))
    # isolated nodes packed into the one part: largest cost first
    cost = np.array([5.0, 1.0, 9.0, 3.0, 7.0])
    lp, pp, part = sa.tree_coarsen(np.full(5, -1, np.int32), cost, 1, 1, 2)
    assert pp.tolist() == [0, 5] and part.tolist() == [2, 4, 0, 3, 1]
    # one node; no nodes
    assert same(sa.tree_coarsen(np.array([-1], np.int32), [1.0], 4, 1, 2),
                ([0, 1], [0, 1], [0]))
    lp, pp, part = sa.tree_coarsen(np.zeros(0, np.int32), np.zeros(0), 4, 1, 2)
    assert lp.tolist() == [0] and pp.tolist() == [0] and part.size == 0


def test_bins_split_wide_cluster():
    """A complete binary tree of 4 levels, 2 parts: the leaf level (8 > 2
    nodes) is cut off below the two narrow levels next to the root
    (min_level_dist 1), its 8 single-leaf subtrees go worst-fit into two
    bins, then the rest of the tree is one part."""
    bt = postorder_binary_tree(4)   # leaves 0,1,3,4,7,8,10,11
    cost = np.array([8, 7, 0, 6, 5, 0, 0, 4, 3, 0, 2, 1, 0, 0, 0], np.float64)
    lp, pp, part = sa.tree_coarsen(bt, cost, 2, 1, 2)
    assert lp.tolist() == [0, 2, 3]
    # costs 8,7,6,5,4,3,2,1 in that order to the lighter bin (ties: bin 0):
    # bin loads 8|0, 8|7, 8|13, 13|13, 17|13, 17|16, 17|18, 18|18
    bins = [part[pp[0]:pp[1]].tolist(), part[pp[1]:pp[2]].tolist()]
    assert bins == [[0, 4, 7, 11], [1, 3, 8, 10]]
    # the upper cluster's one subtree, walked from node 2: 2, its sibling
    # leaf 5, 6, then 13's leaves 9 and 12, 13, the root
    assert part[pp[2]:pp[3]].tolist() == [2, 5, 6, 9, 12, 13, 14]
    check_invariants(bt, lp, pp, part)


def test_errors():
    with pytest.raises(sa.SymAmdError) as e:
        sa.tree_coarsen(np.array([1, 2, 0], np.int32), np.ones(3), 2, 1, 2)
    assert e.value.code == 2
    with pytest.raises(sa.SymAmdError):
        sa.tree_coarsen(np.array([5, -1], np.int32), np.ones(2), 2, 1, 2)
    with pytest.raises(sa.SymAmdError):
        sa.tree_coarsen(np.array([-1, 0], np.int32), [1.0, np.nan], 2, 1, 2)
    with pytest.raises(sa.SymAmdError):   # the reference would never return
        sa.tree_coarsen(postorder_binary_tree(5), np.ones(31), 4, 1, 1)


def test_lbc_solve_with_tree_schedule():
    """The schedule drives the LBC solve of a lower triangle whose
    dependencies the tree covers, with the serial solve's exact result."""
    rng = np.random.default_rng(9)
    # postordered forests, so that etree_lower gives a lower triangle
    for parent in (postorder_binary_tree(9), random_forest(rng, 700, 3, True)):
        Lp, Li, Lx = etree_lower(parent)
        lp, pp, part = sa.tree_coarsen(parent, np.diff(Lp).astype(np.float64), 8, 1, 2)
        assert oracle.check_schedule(Lp, Li, lp, pp, part) == 0
        b = rng.standard_normal(len(parent))
        assert bitwise_equal(oracle.sptrsv_csr_lbc(Lp, Li, Lx, b, lp, pp, part),
                             oracle.sptrsv_csr(Lp, Li, Lx, b))
    Lp, Li, Lx = sa.workloads.lap3d_tril(6, 6, 6)
    parent = elimination_tree(Lp, Li)
    lp, pp, part = sa.tree_coarsen(parent, np.diff(Lp).astype(np.float64), 4, 1, 2)
    assert oracle.check_schedule(Lp, Li, lp, pp, part) == 0


def test_dropin_symbol():
    """sym_lib::parsy::getCoarseLevelSet_6 called through its mangled name
    (reference arguments by pointer, as the ABI passes them)."""
    L = sa.lib()
    fn = getattr(L, "_ZN7sym_lib5parsy19getCoarseLevelSet_6EmPKiS2_RiRPiS5_S3_S5_S5_iiiPd")
    P = ctypes.POINTER
    ci, pi = ctypes.c_int, P(ctypes.c_int)
    fn.argtypes = [ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, P(ci), P(pi), P(pi),
                   P(ci), P(pi), P(pi), ci, ci, ci, ctypes.c_void_p]
    fn.restype = ci
    parent = postorder_binary_tree(6)
    n = len(parent)
    cost = np.arange(1, n + 1, dtype=np.float64)
    nl, npart = ci(), ci()
    lptr, unused, pptr, node = pi(), pi(), pi(), pi()
    r = fn(n, parent.ctypes.data, None, ctypes.byref(nl), ctypes.byref(lptr),
           ctypes.byref(unused), ctypes.byref(npart), ctypes.byref(pptr),
           ctypes.byref(node), 4, 1, 2, cost.ctypes.data)
    assert r == nl.value
    got = ([lptr[i] for i in range(nl.value + 1)], [pptr[i] for i in range(npart.value + 1)],
           [node[i] for i in range(n)])
    assert same(got, sa.tree_coarsen(parent, cost, 4, 1, 2))
