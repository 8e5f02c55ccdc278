"""Writes the golden fixtures in tests/golden/ from the REFERENCE'S OWN
compiled functions (oracle/_ref/libsymref.so, built by `make -C oracle ref`
from the sources under /root/reference).  Run in the development container:

    python tests/golden/make_golden.py

Inputs come from the deterministic generators in sympiler_amd/workloads.py,
so any machine regenerates the same matrices and checks its results against
these vectors / digests without the reference being present.

Fixtures (data only):
  c1_x.npy            reference sptrsv_csr solution of workload c1 (2D 100x100)
  digests.json        SHA-256 + norms of reference outputs on c2, c3, c4, c5:
                      IC(0) factors (spic0_csc_left), forward solves
                      (sptrsv_csr), the transposed solve (ltsolve) and SpMV
  tree_levels.npz     parsy getLevelSet (TreeUtils.cpp:89-141) level sets of
                      random forests (inputs and outputs)
  spic0_csr_4x4.npy   the reference's spic0_csr factor of the 4x4-grid 2D
                      Laplacian, which is NOT textbook IC(0) (kept to record
                      why the workloads use the spic0_csc_left order)
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from oracle import ref  # noqa: E402
from sympiler_amd import workloads as W  # noqa: E402


def digest(a):
    a = np.ascontiguousarray(a, np.float64)
    return {"sha256": hashlib.sha256(a.tobytes()).hexdigest(),
            "l2": float(np.linalg.norm(a)), "linf": float(np.max(np.abs(a))),
            "n": int(a.size)}


def ref_ic0_csr(Lp, Li, A):
    """IC(0) by the reference's spic0_csc_left, returned in CSR order."""
    Cp, Ci, Cx, mp = oracle.csr_to_csc(Lp, Li, A)
    pp, ps = oracle.prune_set(Lp, Li)
    Lc, ok = ref.spic0_csc_left(Cp, Ci, Cx, pp, ps)
    assert ok
    return Lc[mp]


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref/libsymref.so missing: run `make -C oracle ref` first")
    out = {}

    w1 = W.config1()
    x1 = ref.sptrsv_csr(w1.Lp, w1.Li, w1.Lx, w1.b)
    np.save(os.path.join(HERE, "c1_x.npy"), x1)
    out["c1_x"] = digest(x1)

    for name, g in (("c5", 64), ("c2", 128)):
        Lp, Li, A = W.lap3d_tril(g, g, g, 6.01, -1.0)
        Lx = ref_ic0_csr(Lp, Li, A)
        out[f"{name}_ic0_Lx"] = digest(Lx)
        if name == "c2":
            x_true = W.uniform(1, 0, len(Lp) - 1)
            b = ref.spmv_csr(Lp, Li, Lx, x_true)
            out["c2_b"] = digest(b)
            out["c2_x"] = digest(ref.sptrsv_csr(Lp, Li, Lx, b))
        else:
            b = W.uniform(4, 0, len(Lp) - 1)
            x = ref.sptrsv_csr(Lp, Li, Lx, b)
            out["c5_x"] = digest(x)
            Cp, Ci, Cx, _ = oracle.csr_to_csc(Lp, Li, Lx)
            out["c5_ltsolve_x"] = digest(ref.ltsolve(Cp, Ci, Cx, x))

    w3 = W.config3(k=2)
    for c in range(2):
        out[f"c3_x_col{c}"] = digest(ref.sptrsv_csr(w3.Lp, w3.Li, w3.Lx, w3.b[c]))
    out["c3_nnz"] = w3.nnz

    w4 = W.config4()
    xs = []
    offs = w4.row_offsets
    for s in range(len(offs) - 1):
        r0, r1 = offs[s], offs[s + 1]
        p0, p1 = w4.Lp[r0], w4.Lp[r1]
        xs.append(ref.sptrsv_csr(w4.Lp[r0:r1 + 1] - p0, w4.Li[p0:p1], w4.Lx[p0:p1],
                                 w4.b[r0:r1]))
    out["c4_x"] = digest(np.concatenate(xs))
    out["c4_n"] = int(offs[-1])

    rng = np.random.default_rng(12345)
    trees = {}
    for t, n in enumerate((1, 7, 50, 400)):
        parent = np.full(n, -1, np.int32)
        for j in range(n - 1):
            if rng.random() < 0.85:
                parent[j] = rng.integers(j + 1, n)
        lp, ls = ref.getLevelSet(parent)
        trees[f"parent{t}"] = parent
        trees[f"level_ptr{t}"] = lp
        trees[f"level_set{t}"] = ls
    np.savez(os.path.join(HERE, "tree_levels.npz"), **trees)

    Lp, Li, A = W.lap2d_tril(4, 4, 4.0, -1.0)
    # spic0_csr accepts the CSC lower triangle (diagonal first)
    Cp, Ci, Cx, _ = oracle.csr_to_csc(Lp, Li, A)
    np.save(os.path.join(HERE, "spic0_csr_4x4.npy"), ref.spic0_csr(Cp, Ci, Cx))

    with open(os.path.join(HERE, "digests.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out, indent=1, sort_keys=True))


if __name__ == "__main__":
    main()
