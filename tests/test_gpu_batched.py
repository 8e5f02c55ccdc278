"""GPU tests of the batched solve of uneven independent systems, of the
captured hipGraph replay, of the schedule tuner and of the coarsened
(level_ptr / par_ptr / partition) schedules."""
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


def test_batch_512_systems_matches_reference_digest(dev):
    w = W.config4()
    with open(os.path.join(GOLD, "digests.json")) as f:
        gold = json.load(f)
    plan = sa.Plan(*to_dev(w.Lp, w.Li, w.Lx, device=dev), row_offsets=w.row_offsets)
    assert plan.nsys == 512 and plan.n == gold["c4_n"]
    plan.set_lanes(1)
    x = to_dev(w.b, device=dev)
    plan.solve(x, executor="syncfree")
    plan.check()
    xh = x.cpu().numpy()
    # the reference's serial solve of every system (digest written from
    # the reference build), bit for bit
    assert _sha(xh) == gold["c4_x"]["sha256"]
    assert rel_err(xh, w.x_true) < 1e-12


@pytest.mark.parametrize("executor", ["syncfree", "graph"])
def test_graph_replays_match_eager(dev, executor):
    import torch
    w = W.config4(count=64, lo=16, hi=96, seed=5)
    plan = sa.Plan(*to_dev(w.Lp, w.Li, w.Lx, device=dev), row_offsets=w.row_offsets)
    plan.set_lanes(1)
    b = to_dev(w.b, device=dev)
    x_ref = oracle.sptrsv_csr(w.Lp, w.meta["global_Li"], w.Lx, w.b)
    B = b.clone()
    g = plan.graph(B, executor=executor)
    scratch = torch.zeros(1 << 20, dtype=torch.float64, device=dev)
    for rep in range(4):
        B.copy_(b)                      # eager work between replays
        scratch.add_(1.0)
        eager = b.clone()
        plan.solve(eager, executor="syncfree")
        g.launch()
        torch.cuda.synchronize()
        plan.check()
        assert bitwise_equal(B.cpu().numpy(), x_ref), rep
        assert bitwise_equal(eager.cpu().numpy(), x_ref), rep


def test_tuner_picks_a_correct_schedule(dev):
    Lp, Li, Lx = W.ic0_lap3d(32, 0.01)
    plan = sa.Plan(*to_dev(Lp, Li, Lx, device=dev))
    b = W.uniform(2, 0, len(Lp) - 1)
    x_ref = oracle.sptrsv_csr(Lp, Li, Lx, b)
    summary = plan.tune(k=1)
    assert "best exec=" in summary
    x = to_dev(b, device=dev)
    plan.solve(x)  # AUTO uses the tuned choice
    plan.check()
    assert rel_err(x.cpu().numpy(), x_ref) <= RTOL64
    # multi-RHS tuning, row-major
    k = 8
    B = W.uniform(3, 0, len(b) * k).reshape(k, -1)
    X_ref = oracle.sptrsm_csr(Lp, Li, Lx, B)
    plan.tune(k=k, layout="row")
    dB = to_dev(np.ascontiguousarray(B.T), device=dev)
    plan.solve(dB, layout="row")
    plan.check()
    assert rel_err(dB.cpu().numpy().T, X_ref) <= RTOL64
