"""Writes tests/golden/tree_lbc.npz from the REFERENCE'S OWN compiled
parsy::getCoarseLevelSet_6 (parsy/symbolic/InspectionLevel_06.cpp:17, built
into oracle/_ref/libsymref.so by `make -C oracle ref` from the sources under
/root/reference).  Run in the development container:

    python tests/golden/make_tree_lbc.py

Each case stores its inputs (parent array, integer-valued node costs, the
inner-part count, minimum level distance and division rate) and the
reference's three output arrays, so the coarsening is checked against the
reference where the reference is absent.  Cases: random forests in and out
of postorder, the elimination tree of a 3D grid, a postordered complete
binary tree (the supernodal tree shape), a chain, a star and isolated nodes.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from oracle import ref  # noqa: E402
from sympiler_amd import workloads as W  # noqa: E402
from tests.helpers import elimination_tree, postorder_binary_tree  # noqa: E402


def random_forest(rng, n, roots, postorder):
    parent = np.full(n, -1, np.int32)
    if postorder:
        for i in range(n - roots):
            parent[i] = rng.integers(i + 1, n)
    else:
        perm = rng.permutation(n)
        for k in range(roots, n):
            parent[perm[k]] = perm[rng.integers(0, k)]
    return parent


def cases():
    rng = np.random.default_rng(11)
    out = []
    for t in range(8):
        n = int(rng.integers(20, 400))
        parent = random_forest(rng, n, int(rng.integers(1, 6)), t % 2 == 0)
        cost = rng.integers(1, 100, n).astype(np.float64)
        out.append((parent, cost, [2, 4, 8, 16][t % 4], t % 3, [2, 3][t % 2]))
    Lp, Li, _ = W.lap3d_tril(6, 6, 6)
    et = elimination_tree(Lp, Li)
    out.append((et, np.diff(Lp).astype(np.float64), 8, 2, 2))
    out.append((et, np.ones(len(et)), 4, 1, 2))
    bt = postorder_binary_tree(8)
    out.append((bt, rng.integers(1, 1000, len(bt)).astype(np.float64), 16, 2, 2))
    out.append((bt, np.ones(len(bt)), 3, 1, 3))
    chain = np.arange(1, 51, dtype=np.int32)
    chain[-1] = -1
    out.append((chain, np.ones(50), 4, 1, 2))
    star = np.full(40, 39, np.int32)
    star[-1] = -1
    out.append((star, np.arange(1, 41, dtype=np.float64), 4, 1, 2))
    iso = np.full(30, -1, np.int32)
    out.append((iso, rng.permutation(30).astype(np.float64) + 1, 1, 1, 2))
    return out


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref/libsymref.so missing: run `make -C oracle ref` first")
    arrays = {}
    for c, (parent, cost, ip, mld, dr) in enumerate(cases()):
        lp, pp, part = ref.getCoarseLevelSet_6(parent, cost, ip, mld, dr)
        arrays[f"parent{c}"] = parent
        arrays[f"cost{c}"] = cost
        arrays[f"params{c}"] = np.array([ip, mld, dr], np.int32)
        arrays[f"level_ptr{c}"] = lp
        arrays[f"part_ptr{c}"] = pp
        arrays[f"partition{c}"] = part
    arrays["count"] = np.array([len(cases())], np.int32)
    np.savez_compressed(os.path.join(HERE, "tree_lbc.npz"), **arrays)


if __name__ == "__main__":
    main()
