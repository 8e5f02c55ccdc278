"""GPU tests of the adjoint (grad_b = L^-T grad_x, grad_L on L's pattern)
through torch.autograd.Function, in FP64 and FP32.

Checked against: the oracle's ltsolve-based adjoint (bit-identical with one
lane per row), a dense FP64 reference gradient (n = 4096), central finite
differences (h = 1e-6, 64 random entries of L and of b) and torch's own
gradcheck on a small system.
"""
import numpy as np
import pytest

import oracle
import sympiler_amd as sa
from sympiler_amd import workloads as W
from sympiler_amd.autograd import TriangularSolve
from tests.helpers import bitwise_equal, to_dev

pytestmark = pytest.mark.gpu


def _setup(g, dev, dtype=None):
    import torch
    Lp, Li, Lx = W.ic0_lap3d(g, 0.01)
    dLp, dLi, dLx = to_dev(Lp, Li, Lx, device=dev)
    if dtype is not None:
        dLx = dLx.to(dtype)
    plan = sa.Plan(dLp, dLi, dLx)
    return Lp, Li, Lx, plan, dLx


def _grads(plan, dLx, b, executor="auto"):
    import torch
    Lx_t = dLx.clone().requires_grad_(True)
    b_t = b.clone().requires_grad_(True)
    x = TriangularSolve.apply(Lx_t, b_t, plan, executor)
    loss = 0.5 * (x * x).sum()
    loss.backward()
    plan.check()
    return x.detach(), Lx_t.grad, b_t.grad


@pytest.mark.parametrize("executor", ["syncfree", "levelset", "regions", "auto"])
def test_adjoint_16cubed_bitwise_and_dense(dev, executor):
    Lp, Li, Lx, plan, dLx = _setup(16, dev)
    plan.set_lanes(1)
    n = len(Lp) - 1
    b = W.uniform(4, 0, n)
    x, gL, gb = _grads(plan, dLx, to_dev(b, device=dev), executor)
    x_ref = oracle.sptrsv_csr(Lp, Li, Lx, b)
    gb_ref, gL_ref = oracle.adjoint(Lp, Li, Lx, x_ref, x_ref)
    assert bitwise_equal(x.cpu().numpy(), x_ref)
    assert bitwise_equal(gb.cpu().numpy(), gb_ref)
    assert bitwise_equal(gL.cpu().numpy(), gL_ref)
    # dense FP64 reference gradient
    D = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(Lp))
    D[rows, Li] = Lx
    xd = np.linalg.solve(D, b)
    gbd = np.linalg.solve(D.T, xd)
    gLd = -np.outer(gbd, xd)[rows, Li]
    assert np.max(np.abs(gb.cpu().numpy() - gbd)) / np.max(np.abs(gbd)) < 1e-12
    assert np.max(np.abs(gL.cpu().numpy() - gLd)) / np.max(np.abs(gLd)) < 1e-12


def test_adjoint_64cubed_finite_differences(dev):
    import torch
    Lp, Li, Lx, plan, dLx = _setup(64, dev)
    n = len(Lp) - 1
    b = to_dev(W.uniform(4, 0, n), device=dev)
    x, gL, gb = _grads(plan, dLx, b)
    h = 1e-6

    def loss(Lx_, b_):
        plan.update_values(Lx_)
        y = b_.clone()
        plan.solve(y)
        return 0.5 * float((y * y).sum())

    rng = np.random.default_rng(7)
    gLh, gbh = gL.cpu().numpy(), gb.cpu().numpy()
    loss0 = loss(dLx, b)
    # central differences: truncation O(h^2 |f'''|) is far below the
    # cancellation error of the two FP64 losses, about eps * loss / h
    cancel = 16 * np.finfo(np.float64).eps * loss0 / h
    for p in rng.choice(len(Lx), 64, replace=False):
        e = torch.zeros_like(dLx)
        e[p] = h
        fd = (loss(dLx + e, b) - loss(dLx - e, b)) / (2 * h)
        assert abs(fd - gLh[p]) <= 1e-6 * abs(gLh[p]) + cancel, (p, fd, gLh[p])
    plan.update_values(dLx)
    for i in rng.choice(n, 64, replace=False):
        e = torch.zeros_like(b)
        e[i] = h
        fd = (loss(dLx, b + e) - loss(dLx, b - e)) / (2 * h)
        assert abs(fd - gbh[i]) <= 1e-6 * abs(gbh[i]) + cancel, (i, fd, gbh[i])


def test_adjoint_fp32_against_fp64(dev):
    import torch
    Lp, Li, Lx, plan64, dLx = _setup(64, dev)
    plan32 = sa.Plan(*to_dev(Lp, Li, device=dev), dLx.to(torch.float32))
    n = len(Lp) - 1
    b = to_dev(W.uniform(4, 0, n), device=dev)
    x64, gL64, gb64 = _grads(plan64, dLx, b)
    x32, gL32, gb32 = _grads(plan32, dLx.to(torch.float32), b.to(torch.float32))
    rel = lambda a, r: float((a.double() - r).norm() / r.norm())
    # FP32 rounding through 190 dependent levels: 1e-4 relative in norm
    assert rel(x32, x64) < 1e-4
    assert rel(gb32, gb64) < 1e-4
    assert rel(gL32, gL64) < 1e-4


def test_gradcheck_small(dev):
    import torch
    Lp, Li, Lx = W.random_lower(40, 0.1, 1, seed=3)
    dLp, dLi, dLx = to_dev(Lp, Li, Lx, device=dev)
    plan = sa.Plan(dLp, dLi, dLx)
    b = to_dev(W.uniform(3, 0, 40), device=dev)
    Lx_t = dLx.clone().requires_grad_(True)
    b_t = b.clone().requires_grad_(True)
    assert torch.autograd.gradcheck(lambda a, c: TriangularSolve.apply(a, c, plan),
                                    (Lx_t, b_t), eps=1e-6, atol=1e-7, rtol=1e-5)


def test_forward_fp32_matches_fp32_oracle_order(dev):
    # FP32 forward with one lane per row repeats the row sequence in FP32;
    # compare against the FP64 solution at FP32 accuracy
    import torch
    Lp, Li, Lx = W.ic0_lap3d(32, 0.01)
    plan = sa.Plan(*to_dev(Lp, Li, device=dev), to_dev(Lx, device=dev).to(torch.float32))
    b = W.uniform(5, 0, len(Lp) - 1)
    x = to_dev(b, device=dev).to(torch.float32)
    plan.solve(x)
    plan.check()
    x_ref = oracle.sptrsv_csr(Lp, Li, Lx, b)
    err = np.max(np.abs(x.cpu().numpy().astype(np.float64) - x_ref)) / np.max(np.abs(x_ref))
    assert err < 1e-5


@pytest.mark.parametrize("executor", ["auto", "syncfree"])
def test_block_autograd_bitwise_and_dense(dev, executor):
    # b as an (n, k) block: forward = row-major SpTRSM, backward = block
    # adjoint; bitwise the oracle (loss 0.5 ||X||_F^2, so GX = X) and within
    # 1e-12 of the dense FP64 gradient
    Lp, Li, Lx, plan, dLx = _setup(12, dev)
    plan.set_lanes(1)
    n, k = len(Lp) - 1, 8
    B = W.uniform(60, 0, n * k).reshape(n, k)
    X, gL, GB = _grads(plan, dLx, to_dev(np.ascontiguousarray(B), device=dev), executor)
    X_ref = oracle.sptrsm_csr(Lp, Li, Lx, np.ascontiguousarray(B.T)).T
    assert bitwise_equal(X.cpu().numpy(), X_ref)
    GB_ref, gL_ref = oracle.adjoint_block(Lp, Li, Lx, X_ref, X_ref)
    assert bitwise_equal(GB.cpu().numpy(), GB_ref)
    assert bitwise_equal(gL.cpu().numpy(), gL_ref)
    D = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(Lp))
    D[rows, Li] = Lx
    Xd = np.linalg.solve(D, B)
    GBd = np.linalg.solve(D.T, Xd)
    gLd = -(GBd @ Xd.T)[rows, Li]
    assert np.max(np.abs(GB_ref - GBd)) / np.max(np.abs(GBd)) < 1e-12
    assert np.max(np.abs(gL_ref - gLd)) / np.max(np.abs(gLd)) < 1e-12


def test_two_forwards_one_plan_then_backward(dev):
    """Two forwards with different values on one plan, one backward through
    both: each gradient is taken against the values of its own forward."""
    import torch
    Lp, Li, Lx, plan, dLx = _setup(12, dev)
    plan.set_lanes(1)
    n = len(Lp) - 1
    Lx2 = Lx * 1.25  # a second matrix on the same pattern
    b1, b2 = W.uniform(4, 0, n), W.uniform(4, 1, n)
    L1 = dLx.clone().requires_grad_(True)
    L2 = to_dev(Lx2, device=dev).requires_grad_(True)
    t1 = to_dev(b1, device=dev).requires_grad_(True)
    t2 = to_dev(b2, device=dev).requires_grad_(True)
    x1 = TriangularSolve.apply(L1, t1, plan, "syncfree")
    x2 = TriangularSolve.apply(L2, t2, plan, "syncfree")
    (0.5 * (x1 * x1).sum() + 0.5 * (x2 * x2).sum()).backward()
    plan.check()
    for Lv, bv, xt, Lt, bt in ((Lx, b1, x1, L1, t1), (Lx2, b2, x2, L2, t2)):
        x_ref = oracle.sptrsv_csr(Lp, Li, Lv, bv)
        gb_ref, gL_ref = oracle.adjoint(Lp, Li, Lv, x_ref, x_ref)
        assert bitwise_equal(xt.detach().cpu().numpy(), x_ref)
        assert bitwise_equal(bt.grad.cpu().numpy(), gb_ref)
        assert bitwise_equal(Lt.grad.cpu().numpy(), gL_ref)
    # in-place change of the values after a forward: the plan notices
    with torch.no_grad():
        L1.mul_(0.5)
    x3 = TriangularSolve.apply(L1.detach(), to_dev(b1, device=dev), plan, "syncfree")
    plan.check()
    x3_ref = oracle.sptrsv_csr(Lp, Li, Lx * 0.5, b1)
    assert bitwise_equal(x3.cpu().numpy(), x3_ref)
