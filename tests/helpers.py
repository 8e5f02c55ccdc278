"""Shared helpers for the parity tests."""
import numpy as np

# Relative-error bar for FP64 results that are not bit-identical (G > 1
# lanes per row re-associates the row sum).
RTOL64 = 1e-12


def to_dev(*arrays, device="cuda:0"):
    import torch
    out = []
    for a in arrays:
        a = np.ascontiguousarray(a)
        out.append(torch.from_numpy(a).to(device))
    return out if len(out) > 1 else out[0]


def bitwise_equal(a, b) -> bool:
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def rel_err(x, ref) -> float:
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.max(np.abs(ref))
    return float(np.max(np.abs(x - ref)) / (den if den > 0 else 1.0))


def etree_lower(parent):
    """CSR lower triangle whose dependency DAG is the forest `parent`
    (row p depends on every child j with parent[j] == p)."""
    n = len(parent)
    rows = [[] for _ in range(n)]
    for j, p in enumerate(parent):
        if p >= 0:
            rows[p].append(j)
    Lp = np.zeros(n + 1, np.int32)
    Li, Lx = [], []
    for i in range(n):
        cols = sorted(rows[i])
        Li.extend(cols + [i])
        Lx.extend([-1.0] * len(cols) + [2.0 + len(cols)])
        Lp[i + 1] = len(Li)
    return Lp, np.asarray(Li, np.int32), np.asarray(Lx, np.float64)


# Known answer of the reference test Catch_tests/sptrsv_tests.cpp:17-44:
# a 10x10 lower triangle in CSC with all-ones values, b = L * ones.
KA_CSC_P = np.array([0, 3, 6, 9, 13, 16, 18, 20, 23, 25, 26], np.int32)
KA_CSC_I = np.array([0, 6, 9, 1, 2, 5, 2, 5, 8, 3, 4, 5, 8, 4, 5, 8, 5, 7, 6, 7,
                     7, 8, 9, 8, 9, 9], np.int32)
KA_B = np.array([1, 1, 2, 1, 2, 5, 2, 3, 5, 4], np.float64)


def known_answer_csr():
    """The same 10x10 L as a CSR lower triangle (diagonal last)."""
    import oracle
    Cx = np.ones(len(KA_CSC_I))
    # CSC of L is the CSR of L^T; transposing it gives the CSR of L
    Rp, Ri, Rx, _ = oracle.csr_to_csc(KA_CSC_P, KA_CSC_I, Cx)
    return Rp, Ri, Rx


def general_with_diags(Lp, Li, Lx, upper_per_row, seed):
    """A general CSR matrix A whose lower triangle (diagonal included) is the
    given L, plus up to `upper_per_row` random strictly-upper entries per row
    (columns ascending), and the position of each row's diagonal -- the
    layout sptrsv_csr_ilu / sptrsv_csr_ilu_lbc take
    (sparse_blas/sptrsv.cpp:28-38, :109-126)."""
    rng = np.random.default_rng(seed)
    n = len(Lp) - 1
    Ap = np.zeros(n + 1, np.int32)
    Ai, Ax, diags = [], [], np.zeros(n, np.int32)
    for i in range(n):
        lo, hi = Lp[i], Lp[i + 1]
        Ai.extend(Li[lo:hi].tolist())
        Ax.extend(Lx[lo:hi].tolist())
        diags[i] = len(Ai) - 1
        m = min(upper_per_row, n - 1 - i)
        if m > 0:
            cols = np.sort(rng.choice(np.arange(i + 1, n), size=rng.integers(0, m + 1),
                                      replace=False))
            Ai.extend(cols.tolist())
            Ax.extend(rng.uniform(-1.0, 1.0, len(cols)).tolist())
        Ap[i + 1] = len(Ai)
    return Ap, np.asarray(Ai, np.int32), np.asarray(Ax, np.float64), diags


def ilu_lower_host(Ap, Ai, Ax, diags, unit):
    """Row i of the lower factor inside A: entries [Ap[i], diags[i]) then the
    diagonal (value 1.0 if `unit`, else Ax[diags[i]])."""
    n = len(Ap) - 1
    Lp = np.zeros(n + 1, np.int32)
    Li, Lx = [], []
    for i in range(n):
        Li.extend(Ai[Ap[i]:diags[i]].tolist() + [i])
        Lx.extend(Ax[Ap[i]:diags[i]].tolist() + [1.0 if unit else Ax[diags[i]]])
        Lp[i + 1] = len(Li)
    return Lp, np.asarray(Li, np.int32), np.asarray(Lx, np.float64)


def lbc_schedules(lp, ls):
    """Two valid 3-array schedules (level_ptr, par_ptr, partition) built from
    a level set: every row its own partition, one coarse level per level;
    and three levels per coarse level, one partition each (rows in level
    order)."""
    nl = len(lp) - 1
    per_row = (np.asarray(lp, np.int32), np.arange(len(ls) + 1, dtype=np.int32),
               np.asarray(ls, np.int32))
    cl = list(range(0, nl, 3)) + [nl]
    coarse = (np.arange(len(cl), dtype=np.int32),
              np.array([lp[c] for c in cl], np.int32), np.asarray(ls, np.int32))
    return {"per_row": per_row, "coarse3": coarse}


def elimination_tree(Lp, Li):
    """Elimination tree of the symmetric matrix whose lower triangle is the
    CSR (Lp, Li): parent[j] = the smallest i > j connected to j through
    rows < i (Liu's algorithm with path compression).  Every entry L[i, j]
    (i > j) has i among j's ancestors, so a schedule valid for the tree is
    valid for the solve with L."""
    n = len(Lp) - 1
    parent = np.full(n, -1, np.int32)
    anc = np.full(n, -1, np.int64)
    for i in range(n):
        for j in Li[Lp[i]:Lp[i + 1]]:
            j = int(j)
            while j != -1 and j < i:
                nxt = int(anc[j])
                anc[j] = i
                if nxt == -1:
                    parent[j] = i
                j = nxt
    return parent


def postorder_binary_tree(levels):
    """Parent array of a complete binary tree of 2**levels - 1 nodes
    numbered in postorder (children before parents), the supernodal tree
    shape of sympiler_amd.workloads.bcsc_tree."""
    parent = []

    def build(depth):
        if depth == 1:
            parent.append(-1)
            return len(parent) - 1
        a = build(depth - 1)
        b = build(depth - 1)
        parent.append(-1)
        me = len(parent) - 1
        parent[a] = parent[b] = me
        return me

    build(levels)
    return np.asarray(parent, np.int32)
