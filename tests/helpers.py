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
