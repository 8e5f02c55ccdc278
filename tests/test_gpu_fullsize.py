"""Full-size GPU parity of the paths the bench times.

  * the Zipf(1.8) n = 1e6 SpTRSM at k = 8, 32 and 128, column-major (the
    reference's multi-RHS convention, x[idx + j*n] in
    sparse_blas/parsy/triangularSolve/Triangular_BCSC.cpp:114): one lane per
    row bit-identical to the oracle's column-by-column sptrsv_csr
    (sparse_blas/sptrsv.cpp:17-25) on the sync-free executor and on the
    one-lane level path replayed from a graph; the split level executor and
    the tuned choice within 1e-12; columns 0 and 1 against the digests
    written from the reference build;
  * the 512-system uneven batch on its timed path: the tuner's choice
    captured as a hipGraph and replayed, against the reference-build digest;
  * the 128^3 IC(0) factor on the plain level-set executor (one launch per
    level, eager and replayed from a graph), bit-identical.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
import sympiler_amd as sa
from sympiler_amd import workloads as W
from tests.helpers import RTOL64, bitwise_equal, rel_err, to_dev

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, np.float64).tobytes()).hexdigest()


def _gold():
    with open(os.path.join(GOLD, "digests.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def zipf():
    w = W.config3(k=1)
    assert w.nnz == _gold()["c3_nnz"]
    return w


@pytest.mark.parametrize("k", [8, 32, 128])
def test_c3_sptrsm_full_size(dev, zipf, k):
    import torch
    w = zipf
    gold = _gold()
    B = W.uniform(2, 0, w.n * k).reshape(k, w.n)  # column-major n x k
    X_ref = oracle.sptrsm_csr(w.Lp, w.Li, w.Lx, B)
    assert _sha(X_ref[0]) == gold["c3_x_col0"]["sha256"]
    assert _sha(X_ref[1]) == gold["c3_x_col1"]["sha256"]
    plan = sa.Plan(*to_dev(w.Lp, w.Li, w.Lx, device=dev))
    dB = torch.from_numpy(B).to(dev)
    # one lane per (row, column): the reference's order, bit for bit
    plan.set_lanes(1)
    X = dB.clone()
    plan.solve(X, layout="col", executor="syncfree")
    plan.check()
    Xh = X.cpu().numpy()
    assert bitwise_equal(Xh, X_ref), k
    # the one-lane level path replayed from a graph: bit for bit as well
    X.copy_(dB)
    plan.solve(X, layout="col", executor="graph")
    plan.check()
    assert bitwise_equal(X.cpu().numpy(), X_ref), ("graph", k)
    # the split level executor (long rows re-associated): the 1e-12 bar
    X.copy_(dB)
    plan.solve(X, layout="col", executor="split")
    plan.check()
    assert rel_err(X.cpu().numpy(), X_ref) <= RTOL64, ("split", k)
    # the tuned choice (what the bench times), within the 1e-12 bar
    plan.set_lanes(0)
    summary = plan.tune(k=k, layout="col")
    assert "best exec=" in summary
    X.copy_(dB)
    plan.solve(X, layout="col")
    plan.check()
    Xh = X.cpu().numpy()
    assert rel_err(Xh, X_ref) <= RTOL64, (k, summary.splitlines()[-1])
    res = plan.residual(X, dB, layout="col")
    assert float(res.max()) < 1e-13
    del X, dB


def test_c4_tuned_graph_replay_matches_reference_digest(dev):
    import torch
    w = W.config4()
    gold = _gold()
    plan = sa.Plan(*to_dev(w.Lp, w.Li, w.Lx, device=dev), row_offsets=w.row_offsets)
    assert plan.nsys == 512 and plan.n == gold["c4_n"]
    summary = plan.tune(k=1)
    ex, lanes = plan.choice(1)
    b = to_dev(w.b, device=dev)
    X = b.clone()
    g = plan.graph(X)  # the automatic (tuned) executor, captured
    scratch = torch.zeros(1 << 20, dtype=torch.float64, device=dev)
    for rep in range(3):
        X.copy_(b)
        scratch.add_(1.0)  # eager work between replays
        g.launch()
        torch.cuda.synchronize()
        plan.check()
        xh = X.cpu().numpy()
        if ex == "regions" or lanes <= 1:
            # one lane per row: the reference's serial solve of every system
            assert _sha(xh) == gold["c4_x"]["sha256"], (rep, summary.splitlines()[-1])
        else:
            x_ref = oracle.sptrsv_csr(w.Lp, w.meta["global_Li"], w.Lx, w.b)
            assert rel_err(xh, x_ref) <= RTOL64, rep
        assert rel_err(xh, w.x_true) < 1e-12


@pytest.mark.parametrize("executor", ["levelset", "graph"])
def test_c2_plain_levelset_full_size(dev, executor):
    w = W.config2()
    gold = _gold()
    plan = sa.Plan(*to_dev(w.Lp, w.Li, w.Lx, device=dev))
    plan.set_lanes(1)
    x = to_dev(w.b, device=dev)
    plan.solve(x, executor=executor)
    plan.check()
    assert _sha(x.cpu().numpy()) == gold["c2_x"]["sha256"]
