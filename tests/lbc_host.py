"""Host restatement of the device coarsened-schedule inspector
(sympiler_amd/csrc/lbc.hip, symamd_coarsen): the same steps in numpy/scipy,
used by the tests to pin the device arrays entry for entry on small inputs.

The reference's own coarsener (get_coarse_levelSet_DAG_CSC_tree, called at
demo/sptrsv_demo_utils.h:83-88) lives in the un-vendored aggregation
submodule, so its arrays cannot be reproduced; this restates the build's
algorithm, and both are held to the reference's schedule invariants
(oracle.check_schedule) and to bit-exact LBC solves.
"""
import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

import oracle


def _fixed_cost(Lp, cost):
    c = np.diff(Lp).astype(np.float64) if cost is None else np.asarray(cost, np.float64).copy()
    c[~(c > 0)] = 0.0
    c = np.minimum(c, 2.0 ** 40)
    return np.rint(c * 1024.0).astype(np.int64)


def _roots(Lp, Li, lev, cl):
    """Smallest row of each row's component, over the dependency edges whose
    two rows lie in the same coarse level (cl: coarse level per level)."""
    n = len(Lp) - 1
    rows = np.repeat(np.arange(n), np.diff(Lp))
    keep = (Li != rows) & (cl[lev[Li]] == cl[lev[rows]])
    g = coo_matrix((np.ones(int(keep.sum())), (rows[keep], Li[keep])), shape=(n, n))
    _, lab = connected_components(g, directed=True, connection="weak")
    first = np.full(lab.max() + 1, n, np.int64)
    np.minimum.at(first, lab, np.arange(n))
    return first[lab]


def lbc_schedule(Lp, Li, cost=None, parts=6, min_level_dist=-1, div_rate=4000):
    Lp = np.asarray(Lp, np.int64)
    Li = np.asarray(Li, np.int64)
    n = len(Lp) - 1
    parts = max(parts, 1)
    if n == 0:
        return np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros(0, np.int32)
    lp, ls, _ = oracle.levelset_csr(Lp.astype(np.int32), Li.astype(np.int32))
    nlev = len(lp) - 1
    lev = np.empty(n, np.int64)
    for l in range(nlev):
        lev[ls[lp[l]:lp[l + 1]]] = l
    c64 = _fixed_cost(Lp, cost)
    if min_level_dist > 0:
        cid = np.arange(nlev) // min_level_dist
    else:
        cap = min(nlev, max(div_rate, 1))
        T = 0
        while (2 << T) <= cap:
            T += 1
        usable = {}
        for t in range(1, T + 1):
            blk = np.arange(nlev) >> t
            root = _roots(Lp, Li, lev, blk)
            nb = (nlev + (1 << t) - 1) >> t
            ccost = np.zeros(n, np.int64)
            np.add.at(ccost, root, c64)
            btot = np.zeros(nb, np.int64)
            np.add.at(btot, blk[lev], c64)
            isr = root == np.arange(n)
            bcnt = np.bincount(blk[lev[isr]], minlength=nb)
            bmax = np.zeros(nb, np.int64)
            np.maximum.at(bmax, blk[lev[isr]], ccost[isr])
            usable[t] = [bcnt[b] >= parts and int(bmax[b]) * parts <= int(btot[b])
                         for b in range(nb)]
        cid = np.empty(nlev, np.int64)
        a = c = 0
        while a < nlev:
            span = 1
            for t in range(T, 0, -1):
                if a % (1 << t) == 0 and usable[t][a >> t]:
                    span = min(1 << t, nlev - a)
                    break
            cid[a:a + span] = c
            a += span
            c += 1
    nc = int(cid[-1]) + 1
    root = _roots(Lp, Li, lev, cid)
    ccost = np.zeros(n, np.int64)
    np.add.at(ccost, root, c64)
    bin_of = np.zeros(n, np.int64)
    rts = np.nonzero(root == np.arange(n))[0]
    for c in range(nc):
        mine = rts[cid[lev[rts]] == c]
        order = mine[np.lexsort((mine, -ccost[mine]))]
        load = np.zeros(parts, np.int64)
        for r in order:
            b = int(np.argmin(load))  # first bin of least load
            load[b] += ccost[r]
            bin_of[r] = b
    key = cid[lev[ls]] * parts + bin_of[root[ls]]
    order = np.argsort(key, kind="stable")
    partition = ls[order].astype(np.int32)
    cnt = np.bincount(key, minlength=nc * parts)
    nz = cnt > 0
    par_ptr = np.concatenate([[0], np.cumsum(cnt[nz])]).astype(np.int32)
    before = np.concatenate([[0], np.cumsum(nz)])
    level_ptr = before[np.arange(nc + 1) * parts].astype(np.int32)
    return level_ptr, par_ptr, partition
