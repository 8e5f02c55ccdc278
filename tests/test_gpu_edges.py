"""Edge cases of the device paths: an empty matrix (n = 0) and a single row,
for the solves (every executor, one and several right-hand sides), IC(0),
ILU(0) and the supernodal solve."""
import numpy as np
import pytest
import torch

import sympiler_amd as sa
from tests.helpers import bitwise_equal, to_dev

pytestmark = pytest.mark.gpu

EXECUTORS = ["syncfree", "levelset", "graph", "regions", "auto"]


def test_empty_matrix(dev):
    Lp = torch.zeros(1, dtype=torch.int32, device=dev)
    Li = torch.zeros(0, dtype=torch.int32, device=dev)
    Lx = torch.zeros(0, dtype=torch.float64, device=dev)
    plan = sa.Plan(Lp, Li, Lx)
    assert plan.n == 0 and plan.levels == 0
    for ex in EXECUTORS:
        plan.solve(torch.zeros(0, dtype=torch.float64, device=dev), executor=ex)
    plan.check()
    assert plan.ic0(Lx).numel() == 0
    assert sa.ilu0(Lp, Li, Lx, torch.zeros(0, dtype=torch.int32, device=dev)).numel() == 0
    sp = sa.SupernodalPlan(np.zeros(1, np.int32), np.zeros(0, np.int32),
                           np.zeros(1, np.int32), np.zeros(0, np.int32))
    sp.solve(Lx, torch.zeros(0, dtype=torch.float64, device=dev))
    sp.check()


def test_single_row(dev):
    Lp = np.array([0, 1], np.int32)
    Li = np.array([0], np.int32)
    Lx = np.array([4.0])
    dLp, dLi, dLx = to_dev(Lp, Li, Lx)
    plan = sa.Plan(dLp, dLi, dLx)
    b = np.array([3.0])
    for ex in EXECUTORS:
        x = to_dev(b.copy())
        plan.solve(x, executor=ex)
        plan.check()
        assert bitwise_equal(x.cpu().numpy(), b / 4.0), ex
    B = to_dev(np.array([[3.0, -1.0, 7.0]]))  # one row, three right-hand sides
    plan.solve(B, layout="row")
    plan.check()
    assert bitwise_equal(B.cpu().numpy(), np.array([[0.75, -0.25, 1.75]]))
    assert bitwise_equal(plan.ic0(dLx).cpu().numpy(), np.array([2.0]))
    l = sa.ilu0(dLp, dLi, dLx, to_dev(np.array([0], np.int32)))
    assert bitwise_equal(l.cpu().numpy(), Lx)
    sp = sa.SupernodalPlan(np.array([0, 1], np.int32), np.array([1], np.int32),
                           np.array([0, 1], np.int32), np.array([0], np.int32))
    x = to_dev(b.copy())
    sp.solve(dLx, x)
    sp.check()
    assert bitwise_equal(x.cpu().numpy(), b / 4.0)
