"""Region order and owner schedule of the sync-free executor: solves stay
bit-identical to the oracle whichever way units are ordered and dealt
(owner schedule, round-robin, tickets, graph replay), and the plan's level
sets stay the canonical ones (ascending rows inside a level)."""
import numpy as np
import pytest

import oracle
import sympiler_amd as sa
from sympiler_amd import workloads as W
from tests.helpers import RTOL64, bitwise_equal, rel_err, to_dev

pytestmark = pytest.mark.gpu


def _plan(Lp, Li, Lx, dev, mode=1, rows=None, handle=None):
    h = handle or sa.Handle()
    h.set_regions(mode, rows)
    return sa.Plan(*to_dev(Lp, Li, Lx, device=dev), handle=h), h


def test_grid_regions_are_cubes(dev):
    # 32^3 Laplacian pattern, regions of at most 512 rows: 8x8x8 blocks
    Lp, Li, Lx = W.ic0_lap3d(32, 0.01)
    plan, _ = _plan(Lp, Li, Lx, dev)
    count, inside = plan.regions()
    assert count == (32 // 8) ** 3
    # per axis 31 * 32^2 edges, of which the 3 * 32^2 at x (or y, z) = 8, 16,
    # 24 cross a region boundary
    assert abs(inside - (1.0 - 3.0 / 31.0)) < 1e-12


def test_level_sets_stay_canonical(dev):
    Lp, Li, Lx = W.ic0_lap3d(16, 0.01)
    plan, _ = _plan(Lp, Li, Lx, dev, mode=2, rows=64)
    assert plan.regions()[0] > 0
    lp, ls = plan.level_sets()
    lp_o, ls_o, _ = oracle.levelset_csr(Lp, Li)
    assert np.array_equal(lp, lp_o) and np.array_equal(ls, ls_o)


@pytest.mark.parametrize("owners", [True, False])
def test_c2_single_rhs_owner_schedule(dev, owners):
    w = W.config2()
    x_ref = oracle.sptrsv_csr(w.Lp, w.Li, w.Lx, w.b)
    plan, h = _plan(w.Lp, w.Li, w.Lx, dev)
    assert plan.regions()[0] == 4096
    h.set_owners(owners)
    for _ in range(2):  # the second solve runs on the next epoch of the tags
        x = to_dev(w.b, device=dev)
        plan.solve(x, executor="syncfree")
        plan.check()
        assert bitwise_equal(x.cpu().numpy(), x_ref)


def test_c2_multi_rhs_and_adjoint(dev):
    Lp, Li, Lx = W.ic0_lap3d(64, 0.01)
    n = len(Lp) - 1
    plan, h = _plan(Lp, Li, Lx, dev)
    assert plan.regions()[0] == 512
    k = 32
    B = W.uniform(5, 0, n * k).reshape(k, n)
    X_ref = oracle.sptrsm_csr(Lp, Li, Lx, B)
    for layout in ("row", "col"):
        dB = to_dev(np.ascontiguousarray(B.T) if layout == "row" else B, device=dev)
        plan.solve(dB, layout=layout, executor="syncfree")
        plan.check()
        X = dB.cpu().numpy()
        assert bitwise_equal(X.T if layout == "row" else X, X_ref), layout
    import torch
    b = W.uniform(6, 0, n)
    x = oracle.sptrsv_csr(Lp, Li, Lx, b)
    gb_ref, gL_ref = oracle.adjoint(Lp, Li, Lx, x, x)
    plan.set_lanes(1)
    dx = to_dev(x, device=dev)
    gL = torch.empty(len(Lx), dtype=torch.float64, device=dev)
    gb, _ = plan.adjoint(dx, dx.clone(), gLx=gL, executor="syncfree")
    plan.check()
    assert bitwise_equal(gb.cpu().numpy(), gb_ref)
    assert bitwise_equal(gL.cpu().numpy(), gL_ref)


@pytest.mark.parametrize("deal", ["owners", "round_robin", "tickets"])
def test_forced_regions_on_irregular_matrix(dev, deal):
    # a random lower triangle has few edges inside regions: forced regions
    # give many cross-wave dependencies, the hardest case for the schedule
    Lp, Li, Lx = W.random_lower(20000, 0.0005, 1, seed=31)
    n = len(Lp) - 1
    plan, h = _plan(Lp, Li, Lx, dev, mode=2, rows=64)
    assert plan.regions()[0] > 0
    h.set_owners(deal == "owners")
    h.set_tickets(deal == "tickets")
    b = W.uniform(31, 0, n)
    x_ref = oracle.sptrsv_csr(Lp, Li, Lx, b)
    for lanes in (1, 4):
        plan.set_lanes(lanes)
        x = to_dev(b, device=dev)
        plan.solve(x, executor="syncfree")
        plan.check()
        if lanes == 1:
            assert bitwise_equal(x.cpu().numpy(), x_ref)
        else:
            assert rel_err(x.cpu().numpy(), x_ref) <= RTOL64
    plan.set_lanes(0)
    B = W.uniform(32, 0, n * 8).reshape(8, n)
    dB = to_dev(np.ascontiguousarray(B.T), device=dev)
    plan.solve(dB, layout="row", executor="syncfree")
    plan.check()
    assert bitwise_equal(dB.cpu().numpy().T, oracle.sptrsm_csr(Lp, Li, Lx, B))


# ------------------------------------------------- region executor (LDS)
_RH = []


def _rh():
    """Handle for the region-executor tests: a 5 s watchdog, so a wait that
    never ends fails the test quickly instead of after the 20 s default."""
    if not _RH:
        _RH.append(sa.Handle(timeout_ms=5000.0))
    return _RH[0]


def _solve_regions(plan, B, dev, layout="row"):
    dB = to_dev(B, device=dev)
    plan.solve(dB, layout=layout, executor="regions")
    plan.check()
    return dB.cpu().numpy()


def test_region_executor_c2(dev):
    w = W.config2()
    plan = sa.Plan(*to_dev(w.Lp, w.Li, w.Lx, device=dev), handle=_rh())
    x_ref = oracle.sptrsv_csr(w.Lp, w.Li, w.Lx, w.b)
    for _ in range(2):  # second solve: next epoch of the region tags
        assert bitwise_equal(_solve_regions(plan, w.b, dev), x_ref)
    n = len(w.Lp) - 1
    k = 32
    B = W.uniform(40, 0, n * k).reshape(k, n)
    X_ref = oracle.sptrsm_csr(w.Lp, w.Li, w.Lx, B)
    X = _solve_regions(plan, np.ascontiguousarray(B.T), dev, "row")
    assert bitwise_equal(X.T, X_ref)


@pytest.mark.parametrize("k,layout", [(1, "row"), (3, "row"), (8, "col"), (64, "row"),
                                      (32, "col"), (5, "col-direct")])
def test_region_executor_grids(dev, k, layout):
    # column-major blocks go through the row-major workspace by default;
    # "col-direct" solves them in place (strided staging through registers)
    h = _rh()
    if layout == "col-direct":
        h = sa.Handle(timeout_ms=5000.0)
        h.set_colmajor_direct(True)
        layout = "col"
    for Lp, Li, Lx in (W.ic0_lap3d(24, 0.01), W.lap2d_tril(150, 97)):
        n = len(Lp) - 1
        plan = sa.Plan(*to_dev(Lp, Li, Lx, device=dev), handle=h)
        B = W.uniform(41 + k, 0, n * k).reshape(k, n)
        X_ref = oracle.sptrsm_csr(Lp, Li, Lx, B)
        X = _solve_regions(plan, np.ascontiguousarray(B.T) if layout == "row" else B, dev,
                           layout)
        assert bitwise_equal(X.T if layout == "row" else X, X_ref)


@pytest.mark.parametrize("waves", [2, 4])
def test_region_executor_multiwave(dev, waves):
    # tasks of eight columns solved by 2 or 4 waves (one workgroup) per task:
    # full and partial column groups, direct and staged (column-major) loads,
    # the default cap and 512
    h = sa.Handle(timeout_ms=5000.0)
    h.set_region_waves(waves)
    for Lp, Li, Lx in (W.ic0_lap3d(24, 0.01), W.lap2d_tril(150, 97)):
        n = len(Lp) - 1
        plan = sa.Plan(*to_dev(Lp, Li, Lx, device=dev), handle=h)
        for k, cap, layout in ((8, 0, "row"), (32, 512, "row"), (13, 0, "row"), (8, 0, "col")):
            h.set_region_cap(cap)
            B = W.uniform(60 + k, 0, n * k).reshape(k, n)
            X_ref = oracle.sptrsm_csr(Lp, Li, Lx, B)
            Bl = np.ascontiguousarray(B.T) if layout == "row" else B
            for _ in range(2):  # second solve: next epoch of the task tags
                X = _solve_regions(plan, Bl, dev, layout)
                assert bitwise_equal(X.T if layout == "row" else X, X_ref), (k, cap, layout)


@pytest.mark.parametrize("mat", ["lap3d", "lap2d", "random"])
def test_region_adjoint_bitwise(dev, mat):
    # L^T solved on the transposed region schedule (rows numbered backwards),
    # out of place, then grad-L: bitwise the reference's ltsolve and the
    # closed form, on two calls, and again after a value update
    import torch
    if mat == "lap3d":
        Lp, Li, Lx = W.ic0_lap3d(24, 0.01)
    elif mat == "lap2d":
        Lp, Li, Lx = W.lap2d_tril(150, 97)
    else:  # irregular: halos and overflow entries, forced onto regions
        Lp, Li, Lx = W.random_lower(20000, 0.0008, 1, seed=46)
    n = len(Lp) - 1
    plan = sa.Plan(*to_dev(Lp, Li, Lx, device=dev), handle=_rh())
    assert plan.region_schedule(1, transposed=True)[1] > 0  # runs on regions
    x, gx = W.uniform(47, 0, n), W.uniform(48, 0, n)
    dx, dgx = to_dev(x, gx, device=dev)
    gL = torch.empty(len(Lx), dtype=torch.float64, device=dev)
    for scale in (1.0, 1.0, 0.5):
        if scale != 1.0:
            Lx = Lx * scale
            plan.update_values(to_dev(Lx, device=dev))
        gb_ref, gL_ref = oracle.adjoint(Lp, Li, Lx, x, gx)
        gb, _ = plan.adjoint(dx, dgx, gLx=gL, executor="regions")
        plan.check()
        assert bitwise_equal(gb.cpu().numpy(), gb_ref)
        assert bitwise_equal(gL.cpu().numpy(), gL_ref)
        assert bitwise_equal(dgx.cpu().numpy(), gx)  # gx is not written


@pytest.mark.parametrize("k", [1, 8])
def test_region_cap_halved_to_fit(dev, k):
    # rows that read many rows outside their region: at the default cap a
    # task's halo and record overflow the workgroup's LDS, so the schedule
    # is built at a smaller cap, and the solves stay on regions and bitwise
    Lp, Li, Lx = W.random_lower(20000, 0.0008, 1, seed=46)
    n = len(Lp) - 1
    plan = sa.Plan(*to_dev(Lp, Li, Lx, device=dev), handle=_rh())
    for transposed in (False, True):
        cap, tasks = plan.region_schedule(k, transposed=transposed)
        assert 32 <= cap < 512 and tasks > 0, (transposed, cap, tasks)
    B = W.uniform(49, 0, n * k).reshape(k, n)
    X_ref = oracle.sptrsm_csr(Lp, Li, Lx, B)
    for _ in range(2):  # second solve: next epoch of the region tags
        X = _solve_regions(plan, B[0] if k == 1 else np.ascontiguousarray(B.T), dev)
        assert bitwise_equal(X if k == 1 else X.T, X_ref[0] if k == 1 else X_ref)
    if k == 1:
        x0 = W.uniform(50, 0, n)
        gb_ref, gL_ref = oracle.adjoint(Lp, Li, Lx, x0, B[0])
        gb, gL = plan.adjoint(*to_dev(x0, B[0], device=dev), executor="regions")
        plan.check()
        assert bitwise_equal(gb.cpu().numpy(), gb_ref)


@pytest.mark.parametrize("k,layout,executor", [(1, "row", "regions"), (8, "row", "regions"),
                                               (13, "row", "auto"), (32, "row", "regions"),
                                               (8, "col", "regions"), (5, "col", "syncfree"),
                                               (3, "row", "syncfree"), (2, "row", "syncfree"),
                                               (40, "row", "regions"), (64, "col", "regions"),
                                               (70, "row", "auto")])
def test_block_adjoint_bitwise(dev, k, layout, executor):
    # GB = L^-T GX for a block of k columns (region schedule in one launch,
    # or the sync-free adjoint column by column) and gL summed over the
    # columns pairwise (groups of up to 32 columns, padded past k): bitwise
    # the oracle, on two calls
    import torch
    for Lp, Li, Lx in (W.ic0_lap3d(20, 0.01), W.random_lower(6000, 0.002, 1, seed=51)):
        n = len(Lp) - 1
        plan = sa.Plan(*to_dev(Lp, Li, Lx, device=dev), handle=_rh())
        X = W.uniform(52, 0, n * k).reshape(n, k)
        GX = W.uniform(53, 0, n * k).reshape(n, k)
        GB_ref, gL_ref = oracle.adjoint_block(Lp, Li, Lx, X, GX)
        lay = (lambda a: np.ascontiguousarray(a)) if layout == "row" else \
            (lambda a: np.ascontiguousarray(a.T))
        dX, dGX = to_dev(lay(X), lay(GX), device=dev)
        gL = torch.empty(len(Lx), dtype=torch.float64, device=dev)
        for _ in range(2):
            GB, _ = plan.adjoint_block(dX, dGX, gLx=gL, layout=layout, executor=executor)
            plan.check()
            GBh = GB.cpu().numpy()
            assert bitwise_equal(GBh if layout == "row" else GBh.T, GB_ref)
            assert bitwise_equal(gL.cpu().numpy(), gL_ref)
        assert bitwise_equal(dGX.cpu().numpy(), lay(GX))  # GX is not written


def test_region_adjoint_fp32_matches_syncfree(dev):
    # FP32: the region kernel and the sync-free kernel (one lane per row)
    # perform the same float operations in the same order
    import torch
    Lp, Li, Lx = W.ic0_lap3d(20, 0.01)
    n = len(Lp) - 1
    dLp, dLi, dLx = to_dev(Lp, Li, Lx, device=dev)
    plan = sa.Plan(dLp, dLi, dLx.to(torch.float32), handle=_rh())
    plan.set_lanes(1)
    x, gx = (to_dev(W.uniform(s, 0, n), device=dev).to(torch.float32) for s in (49, 50))
    outs = []
    for ex in ("regions", "syncfree"):
        gL = torch.empty(len(Lx), dtype=torch.float32, device=dev)
        gb, _ = plan.adjoint(x, gx, gLx=gL, executor=ex)
        plan.check()
        outs.append((gb.cpu().numpy(), gL.cpu().numpy()))
    assert bitwise_equal(outs[0][0], outs[1][0])
    assert bitwise_equal(outs[0][1], outs[1][1])


def test_region_executor_irregular_falls_back(dev):
    # few entries inside regions and large halos: the region schedule does
    # not fit one wave's LDS and the solve runs on the sync-free executor
    Lp, Li, Lx = W.random_lower(30000, 0.0008, 1, seed=44)
    n = len(Lp) - 1
    plan = sa.Plan(*to_dev(Lp, Li, Lx, device=dev), handle=_rh())
    plan.set_lanes(1)  # the fallback then repeats the reference order exactly
    b = W.uniform(44, 0, n)
    assert bitwise_equal(_solve_regions(plan, b, dev), oracle.sptrsv_csr(Lp, Li, Lx, b))


def test_region_executor_fp32_and_value_update(dev):
    import torch
    Lp, Li, Lx = W.ic0_lap3d(20, 0.01)
    n = len(Lp) - 1
    dLp, dLi, dLx = to_dev(Lp, Li, Lx, device=dev)
    b = W.uniform(45, 0, n)
    # FP32: the region kernel and the sync-free kernel (one lane per row)
    # perform the same float operations in the same order
    p32 = sa.Plan(dLp, dLi, dLx.to(torch.float32), handle=_rh())
    p32.set_lanes(1)
    xr = to_dev(b, device=dev).to(torch.float32)
    p32.solve(xr, executor="regions")
    xs = to_dev(b, device=dev).to(torch.float32)
    p32.solve(xs, executor="syncfree")
    p32.check()
    assert torch.equal(xr.view(torch.int32), xs.view(torch.int32))
    # new values on the same pattern reach the region schedule
    plan = sa.Plan(dLp, dLi, dLx, handle=_rh())
    _solve_regions(plan, b, dev)
    Lx2 = Lx * 1.5
    plan.update_values(to_dev(Lx2, device=dev))
    assert bitwise_equal(_solve_regions(plan, b, dev), oracle.sptrsv_csr(Lp, Li, Lx2, b))


def test_region_executor_batched_and_captured(dev):
    w = W.config4(count=20, lo=32, hi=120, seed=6)
    plan = sa.Plan(*to_dev(w.Lp, w.Li, w.Lx, device=dev), row_offsets=w.row_offsets,
                   handle=_rh())
    offs = w.row_offsets
    ref = np.concatenate([
        oracle.sptrsv_csr(w.Lp[offs[s]:offs[s + 1] + 1] - w.Lp[offs[s]],
                          w.Li[w.Lp[offs[s]]:w.Lp[offs[s + 1]]],
                          w.Lx[w.Lp[offs[s]]:w.Lp[offs[s + 1]]], w.b[offs[s]:offs[s + 1]])
        for s in range(len(offs) - 1)])
    assert bitwise_equal(_solve_regions(plan, w.b, dev), ref)
    b = to_dev(w.b, device=dev)
    X = b.clone()
    g = plan.graph(X, executor="regions")
    for _ in range(3):  # captured launches take regions by ticket
        X.copy_(b)
        g.launch()
        plan.check()
        assert bitwise_equal(X.cpu().numpy(), ref)


def test_batched_grids_with_regions_graph_replay(dev):
    w = W.config4(count=16, lo=32, hi=96, seed=5)
    h = sa.Handle()
    h.set_regions(1, 256)
    plan = sa.Plan(*to_dev(w.Lp, w.Li, w.Lx, device=dev), row_offsets=w.row_offsets,
                   handle=h)
    assert plan.regions()[0] > 0
    b = to_dev(w.b, device=dev)
    offs = w.row_offsets
    ref = np.concatenate([
        oracle.sptrsv_csr(w.Lp[offs[s]:offs[s + 1] + 1] - w.Lp[offs[s]],
                          w.Li[w.Lp[offs[s]]:w.Lp[offs[s + 1]]],
                          w.Lx[w.Lp[offs[s]]:w.Lp[offs[s + 1]]], w.b[offs[s]:offs[s + 1]])
        for s in range(len(offs) - 1)])
    x = b.clone()
    plan.solve(x, executor="syncfree")
    plan.check()
    assert bitwise_equal(x.cpu().numpy(), ref)
    X = b.clone()
    g = plan.graph(X)
    for _ in range(3):  # replays inside capture use the ticket deal
        X.copy_(b)
        g.launch()
        plan.check()
        assert bitwise_equal(X.cpu().numpy(), ref)
