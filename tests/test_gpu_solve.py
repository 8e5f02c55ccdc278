"""GPU parity tests: the HIP path through the C ABI against the oracle.

The oracle is the CPU restatement in oracle/ (itself pinned bit for bit to
the reference's compiled sources and golden vectors by tests/test_oracle.py).
With one lane per (row, right-hand side) the GPU repeats the reference's
operation order without FMA contraction, so results must be BIT-IDENTICAL;
with several lanes per row the row sum is re-associated and the bar is a
relative error <= 1e-12 (FP64).
"""
import numpy as np
import pytest

import oracle
import sympiler_amd as sa
from sympiler_amd import workloads as W
from tests.helpers import (KA_B, RTOL64, bitwise_equal, etree_lower,
                           known_answer_csr, rel_err, to_dev)

pytestmark = pytest.mark.gpu

# (n, density) of the reference's "parallel vs serial" test,
# Catch_tests/sptrsv_tests.cpp:47-51
REF_RANDOM = [(10, 0.2), (100, 0.03), (500, 0.004), (1000, 0.0005), (10000, 0.0001)]


def _plan(Lp, Li, Lx, dev, dtype=None):
    import torch
    dLp, dLi, dLx = to_dev(Lp, Li, Lx, device=dev)
    if dtype is not None:
        dLx = dLx.to(dtype)
    return sa.Plan(dLp, dLi, dLx)


def _solve(plan, b, dev, executor, lanes):
    plan.set_lanes(lanes)
    x = to_dev(b, device=dev)
    plan.solve(x, executor=executor)
    plan.check()
    return x.cpu().numpy()


# ---------------------------------------------------------------- levels
@pytest.mark.parametrize("case", ["c1", "lap3d16", "random", "chain", "empty_rows"])
def test_levelset_matches_oracle(dev, case):
    if case == "c1":
        w = W.config1()
        Lp, Li = w.Lp, w.Li
    elif case == "lap3d16":
        Lp, Li, _ = W.lap3d_tril(16, 16, 16)
    elif case == "random":
        Lp, Li, _ = W.random_lower(3000, 0.002, 0, seed=11)
    elif case == "chain":
        Lp, Li, _ = W.lap2d_tril(1, 5000)  # bidiagonal: 5000 levels
    else:
        Lp, Li, _ = W.random_lower(2000, 0.0, 0, seed=1)  # diagonal only
    lp_o, ls_o, _ = oracle.levelset_csr(Lp, Li)
    dLp, dLi = to_dev(Lp, Li, device=dev)
    lp, ls = sa.levelset(dLp, dLi)
    assert np.array_equal(lp.cpu().numpy(), lp_o)
    assert np.array_equal(ls.cpu().numpy(), ls_o)


# ---------------------------------------------------------------- solves
@pytest.mark.parametrize("executor", ["syncfree", "levelset", "graph"])
def test_c1_bitwise(dev, executor):
    w = W.config1()
    x_ref = oracle.sptrsv_csr(w.Lp, w.Li, w.Lx, w.b)
    plan = _plan(w.Lp, w.Li, w.Lx, dev)
    assert plan.levels == 199
    x = _solve(plan, w.b, dev, executor, 1)
    assert bitwise_equal(x, x_ref)
    r = plan.residual(to_dev(x, device=dev), to_dev(w.b, device=dev))
    assert r[0] < 1e-14


@pytest.mark.parametrize("lanes", [2, 4, 8, 64])
def test_c1_multilane(dev, lanes):
    w = W.config1()
    x_ref = oracle.sptrsv_csr(w.Lp, w.Li, w.Lx, w.b)
    plan = _plan(w.Lp, w.Li, w.Lx, dev)
    for ex in ("syncfree", "levelset"):
        x = _solve(plan, w.b, dev, ex, lanes)
        assert rel_err(x, x_ref) <= RTOL64


@pytest.mark.parametrize("n,density", REF_RANDOM)
def test_reference_random_cases(dev, n, density):
    # the reference test solves with x = 1 as right-hand side
    Lp, Li, Lx = W.random_lower(n, density, 0, seed=n)
    b = np.ones(n)
    x_ref = oracle.sptrsv_csr(Lp, Li, Lx, b)
    plan = _plan(Lp, Li, Lx, dev)
    for ex in ("syncfree", "levelset"):
        assert bitwise_equal(_solve(plan, b, dev, ex, 1), x_ref)


def test_known_answer(dev):
    Rp, Ri, Rx = known_answer_csr()
    plan = _plan(Rp, Ri, Rx, dev)
    x = _solve(plan, KA_B, dev, "syncfree", 1)
    assert np.all(x == 1.0)


@pytest.mark.parametrize("n,density,min_off", [(4000, 0.003, 1), (20000, 0.0005, 3)])
def test_random_irregular(dev, n, density, min_off):
    Lp, Li, Lx = W.random_lower(n, density, min_off, seed=7)
    b = W.uniform(5, 0, n)
    x_ref = oracle.sptrsv_csr(Lp, Li, Lx, b)
    plan = _plan(Lp, Li, Lx, dev)
    assert bitwise_equal(_solve(plan, b, dev, "syncfree", 1), x_ref)
    for lanes in (4, 16):
        assert rel_err(_solve(plan, b, dev, "syncfree", lanes), x_ref) <= RTOL64


@pytest.mark.parametrize("deal,marks", [("round_robin", True), ("tickets", True),
                                        ("round_robin", False)])
def test_c2_full_size(dev, deal, marks):
    w = W.config2()
    x_ref = oracle.sptrsv_csr(w.Lp, w.Li, w.Lx, w.b)
    h = sa.Handle()
    h.set_tickets(deal == "tickets")
    h.set_marks(marks)
    plan = sa.Plan(*to_dev(w.Lp, w.Li, w.Lx, device=dev), handle=h)
    assert plan.levels == 3 * 128 - 2
    for _ in range(2):  # a second solve reuses the unit flags (next epoch)
        x = _solve(plan, w.b, dev, "syncfree", 1)
        assert bitwise_equal(x, x_ref)
    assert rel_err(x, w.x_true) < 1e-10
    r = plan.residual(to_dev(x, device=dev), to_dev(w.b, device=dev))
    assert r[0] < 1e-14


def test_c3_full_size(dev):
    w = W.config3(k=1)
    b = w.b[0]
    x_ref = oracle.sptrsv_csr(w.Lp, w.Li, w.Lx, b)
    plan = _plan(w.Lp, w.Li, w.Lx, dev)
    assert bitwise_equal(_solve(plan, b, dev, "syncfree", 1), x_ref)
    assert rel_err(_solve(plan, b, dev, "syncfree", 8), x_ref) <= RTOL64


# ---------------------------------------------------------------- SpTRSM
@pytest.mark.parametrize("k", [1, 3, 8, 32, 33, 128])
@pytest.mark.parametrize("layout", ["col", "row"])
def test_sptrsm(dev, k, layout):
    n = 3000
    Lp, Li, Lx = W.random_lower(n, 0.002, 1, seed=3)
    B = W.uniform(9, 0, n * k).reshape(k, n)  # column-major n x k
    X_ref = oracle.sptrsm_csr(Lp, Li, Lx, B)
    h = sa.Handle()
    plan = sa.Plan(*to_dev(Lp, Li, Lx, device=dev), handle=h)
    # column-major: through the level-ordered workspace and in place
    for direct in ((False, True) if layout == "col" else (False,)):
        h.set_colmajor_direct(direct)
        for ex in ("syncfree", "levelset", "graph"):
            plan.set_lanes(1)
            dB = to_dev(B if layout == "col" else np.ascontiguousarray(B.T), device=dev)
            plan.solve(dB, layout=layout, executor=ex)
            plan.check()
            X = dB.cpu().numpy()
            if layout == "row":
                X = X.T
            assert bitwise_equal(X, X_ref), (ex, k, layout, direct)


@pytest.mark.parametrize("k,lanes", [(8, 8), (4, 16)])
@pytest.mark.parametrize("layout", ["col", "row"])
def test_sptrsm_multilane_long_rows(dev, k, lanes, layout):
    """Several lanes per row with k > 1 on the unit executors: rows of up to
    257 entries (several chunks of dependency loads, dead slots in the last
    one), both layouts, at the 1e-12 bar."""
    n = 6000
    Lp, Li, Lx = W.zipf_lower(n, 1.8, 256, 3000, 9)
    assert int(np.diff(Lp).max()) > 64
    B = W.uniform(13, 0, n * k).reshape(k, n)
    X_ref = oracle.sptrsm_csr(Lp, Li, Lx, B)
    plan = _plan(Lp, Li, Lx, dev)
    for ex in ("syncfree", "levelset"):
        plan.set_lanes(lanes)
        dB = to_dev(B if layout == "col" else np.ascontiguousarray(B.T), device=dev)
        plan.solve(dB, layout=layout, executor=ex)
        plan.check()
        X = dB.cpu().numpy()
        X = X.T if layout == "row" else X
        assert rel_err(X, X_ref) <= RTOL64, (ex, k, lanes, layout)


@pytest.mark.parametrize("k", [1, 5, 64, 130])
@pytest.mark.parametrize("layout", ["col", "row"])
def test_level_one_lane_long_rows(dev, k, layout):
    """The one-lane-per-row level path on rows of up to 257 entries (many
    16-entry chunks, a partial last chunk, rows of one entry), column tiles
    side by side (k = 130: three 64-column tiles, the last one partial):
    bit-identical to the serial solve, eager and replayed from a graph."""
    n = 6000
    Lp, Li, Lx = W.zipf_lower(n, 1.8, 256, 3000, 9)
    assert int(np.diff(Lp).max()) > 64 and int(np.diff(Lp).min()) == 1
    B = W.uniform(17, 0, n * k).reshape(k, n)
    X_ref = oracle.sptrsm_csr(Lp, Li, Lx, B)
    Bl = B if layout == "col" else np.ascontiguousarray(B.T)
    plan = _plan(Lp, Li, Lx, dev)
    plan.set_lanes(1)
    for ex in ("levelset", "graph", "graph"):
        dB = to_dev(Bl.copy(), device=dev)
        print(f"[level-one-lane] k={k} layout={layout} executor={ex}", flush=True)
        plan.solve(dB, layout=layout, executor=ex)
        plan.check()
        X = dB.cpu().numpy()
        assert bitwise_equal(X.T if layout == "row" else X, X_ref), (ex, k, layout)


@pytest.mark.parametrize("k", [1, 5, 64, 130])
@pytest.mark.parametrize("layout", ["col", "row"])
def test_split_level_executor(dev, k, layout):
    """The split level executor: rows with more than 32 off-diagonal entries
    solved by a 1024-thread workgroup (sums re-associated: the 1e-12 bar),
    the others one lane per row and column.  A matrix with no long row is
    bit-identical.  The solve is captured once and replayed."""
    mats = {"zipf": W.zipf_lower(6000, 1.8, 256, 3000, 9), "grid": W.ic0_lap3d(12, 0.01)}
    for name, (Lp, Li, Lx) in mats.items():
        n = len(Lp) - 1
        has_long = bool((np.diff(Lp) - 1 > 32).any())
        assert has_long == (name == "zipf")
        B = W.uniform(19 + k, 0, n * k).reshape(k, n)
        X_ref = oracle.sptrsm_csr(Lp, Li, Lx, B)
        Bl = B if layout == "col" else np.ascontiguousarray(B.T)
        plan = _plan(Lp, Li, Lx, dev)
        for rep in range(3):
            dB = to_dev(Bl.copy(), device=dev)
            print(f"[split] {name} n={n} k={k} layout={layout} solve {rep}", flush=True)
            plan.solve(dB, layout=layout, executor="split")
            plan.check()
            X = dB.cpu().numpy()
            X = X.T if layout == "row" else X
            if has_long:
                assert rel_err(X, X_ref) <= RTOL64, (name, k, layout, rep)
            else:
                assert bitwise_equal(X, X_ref), (name, k, layout, rep)


@pytest.mark.parametrize("k", [1, 8, 40])
def test_split_level_executor_fp32(dev, k):
    """FP32: without long rows the split executor performs the one-lane
    level path's float operations in the same order (bit for bit); with
    long rows it stays within FP32 rounding of the FP64 oracle."""
    import torch
    for name, (Lp, Li, Lx) in (("grid", W.ic0_lap3d(12, 0.01)),
                               ("zipf", W.zipf_lower(6000, 1.8, 256, 3000, 9))):
        n = len(Lp) - 1
        B = W.uniform(23 + k, 0, n * k).reshape(k, n)
        plan = _plan(Lp, Li, Lx, dev, dtype=torch.float32)
        plan.set_lanes(1)
        out = {}
        for ex in ("split", "levelset"):
            dB = to_dev(B.copy(), device=dev).to(torch.float32)
            print(f"[split-fp32] {name} k={k} {ex}", flush=True)
            plan.solve(dB, layout="col", executor=ex)
            plan.check()
            out[ex] = dB
        if name == "grid":
            assert torch.equal(out["split"].view(torch.int32), out["levelset"].view(torch.int32))
        X_ref = oracle.sptrsm_csr(Lp, Li, Lx, B)
        assert rel_err(out["split"].double().cpu().numpy(), X_ref) <= 1e-4, name


def _crowded_level(roots=40, wide=400, tail=30, seed=5):
    """A lower triangle whose level 1 holds `wide` rows that each read all
    `roots` rows of level 0 (more long rows in one level than a launch's
    arguments carry), followed by `tail` rows reading a few of them."""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(roots):
        rows.append(([], 8.0))
    for i in range(roots, roots + wide):
        rows.append((list(range(roots)), 60.0))
    for i in range(roots + wide, roots + wide + tail):
        rows.append((sorted(rng.choice(np.arange(roots, roots + wide), 20, replace=False).tolist()), 30.0))
    Lp, Li, Lx = [0], [], []
    for i, (cols, dg) in enumerate(rows):
        Li.extend(cols)
        Lx.extend(rng.uniform(-1.0, 1.0, len(cols)).tolist())
        Li.append(i)
        Lx.append(dg)
        Lp.append(len(Li))
    return (np.array(Lp, np.int32), np.array(Li, np.int32), np.array(Lx, np.float64))


@pytest.mark.parametrize("layout", ["col", "row"])
def test_split_level_executor_crowded_level_and_value_update(dev, layout):
    """More long rows in one level than the launch arguments carry (the rest
    read from the plan's list), both thresholds, then new values on the same
    plan (the padded short rows are refreshed)."""
    Lp, Li, Lx = _crowded_level()
    n = len(Lp) - 1
    k = 24
    B = W.uniform(91, 0, n * k).reshape(k, n)
    Bl = B if layout == "col" else np.ascontiguousarray(B.T)
    h = sa.Handle()
    plan = sa.Plan(*to_dev(Lp, Li, Lx, device=dev), handle=h)
    for thr in (16, 32):
        h.set_split_long(thr)
        for vals in (Lx, Lx * 1.25):
            plan.update_values(to_dev(vals, device=dev))
            X_ref = oracle.sptrsm_csr(Lp, Li, vals, B)
            dB = to_dev(Bl.copy(), device=dev)
            print(f"[split-crowded] thr={thr} layout={layout} scale={vals[-1] / Lx[-1]}",
                  flush=True)
            plan.solve(dB, layout=layout, executor="split")
            plan.check()
            X = dB.cpu().numpy()
            assert rel_err(X.T if layout == "row" else X, X_ref) <= RTOL64, (thr, layout)
    h.set_split_long(0)


def test_tuner_considers_the_split_level_executor(dev):
    Lp, Li, Lx = W.zipf_lower(6000, 1.8, 256, 3000, 9)
    plan = _plan(Lp, Li, Lx, dev)
    log = plan.tune(k=64, layout="col")
    assert "exec=6" in log
    n = len(Lp) - 1
    B = W.uniform(71, 0, n * 64).reshape(64, n)
    X = to_dev(B.copy(), device=dev)
    plan.solve(X, layout="col")
    plan.check()
    assert rel_err(X.cpu().numpy(), oracle.sptrsm_csr(Lp, Li, Lx, B)) <= RTOL64


# ---------------------------------------------------------------- batched
def test_batched_small(dev):
    w = W.config4(count=24, lo=8, hi=40, seed=3)
    offs = w.row_offsets
    plan = sa.Plan(*to_dev(w.Lp, w.Li, w.Lx, device=dev), row_offsets=offs)
    x = _solve(plan, w.b, dev, "syncfree", 1)
    for s in range(len(offs) - 1):
        r0, r1 = offs[s], offs[s + 1]
        p0, p1 = w.Lp[r0], w.Lp[r1]
        xs = oracle.sptrsv_csr(w.Lp[r0:r1 + 1] - p0, w.Li[p0:p1], w.Lx[p0:p1], w.b[r0:r1])
        assert bitwise_equal(x[r0:r1], xs)


# ---------------------------------------------------------------- adjoint
@pytest.mark.parametrize("executor", ["syncfree", "levelset", "regions", "auto"])
def test_adjoint_bitwise(dev, executor):
    Lp, Li, Lx = W.ic0_lap3d(16)
    n = len(Lp) - 1
    b = W.uniform(4, 0, n)
    x = oracle.sptrsv_csr(Lp, Li, Lx, b)
    gx = x.copy()  # loss = 0.5 ||x||^2
    gb_ref, gL_ref = oracle.adjoint(Lp, Li, Lx, x, gx)
    plan = _plan(Lp, Li, Lx, dev)
    plan.set_lanes(1)
    import torch
    dx, dgx = to_dev(x, gx, device=dev)
    gL = torch.empty(len(Lx), dtype=torch.float64, device=dev)
    gb, _ = plan.adjoint(dx, dgx, gLx=gL, executor=executor)
    plan.check()
    assert bitwise_equal(gb.cpu().numpy(), gb_ref)
    assert bitwise_equal(gL.cpu().numpy(), gL_ref)


# ------------------------------------------------------------------ SpMV
def test_spmv_bitwise(dev):
    Lp, Li, Lx = W.random_lower(5000, 0.003, 1, seed=2)
    x = W.uniform(1, 0, 5000)
    y_ref = oracle.spmv_csr(Lp, Li, Lx, x)
    dLp, dLi, dLx, dx = to_dev(Lp, Li, Lx, x, device=dev)
    y = sa.spmv_csr(dLp, dLi, dLx, dx)
    assert bitwise_equal(y.cpu().numpy(), y_ref)
    yv = W.uniform(2, 0, 5000)
    z0 = W.uniform(3, 0, 5000)
    z_ref = oracle.spmv_csr_axpby(Lp, Li, Lx, x, yv, z0, 0.5, -2.0)
    dz = to_dev(z0, device=dev)
    sa.spmv_csr(dLp, dLi, dLx, dx, to_dev(yv, device=dev), dz, 0.5, -2.0)
    assert bitwise_equal(dz.cpu().numpy(), z_ref)


# ------------------------------------------------------------------ errors
def test_structure_errors(dev):
    Lp, Li, Lx = W.lap2d_tril(4, 4)
    bad = Li.copy()
    bad[Lp[5] - 1] = 3  # row 4: last entry is not the diagonal
    with pytest.raises(sa.SymAmdError) as e:
        _plan(Lp, bad, Lx, dev)
    assert e.value.code == 2
    zero = Lx.copy()
    zero[Lp[3] - 1] = 0.0
    with pytest.raises(sa.SymAmdError) as e:
        _plan(Lp, Li, zero, dev)
    assert e.value.code == 3


WIDE_DISABLED = ("the persistent workgroup-per-row executor exceeded the watchdog on its "
                 "first run (zipf n=4000, k=1, row-major, cooperative round-robin deal); "
                 "no cause was established, so the executor is disabled")


def test_wide_row_executor_is_disabled(dev):
    Lp, Li, Lx = W.lap2d_tril(1, 40)
    plan = sa.Plan(*to_dev(Lp, Li, Lx, device=dev))
    X = to_dev(np.ones(len(Lp) - 1), device=dev)
    with pytest.raises(sa.SymAmdError) as e:
        plan.solve(X, executor="wide")
    assert e.value.code == 7


@pytest.mark.skip(reason=WIDE_DISABLED)
@pytest.mark.parametrize("k", [1, 5, 32, 64, 128, 130])
@pytest.mark.parametrize("layout", ["row", "col"])
def test_wide_row_executor(dev, k, layout):
    # one workgroup per row, all k columns, the row's entries split over four
    # wavefronts (sums re-associated: the 1e-12 bar), on matrices with long,
    # uneven rows, a plain grid and a chain; two solves (next epoch of the
    # row tags) and a graph replay (ticketed deal)
    import torch
    mats = {"zipf": W.zipf_lower(4000, 1.8, 256, 1500, 7),
            "grid": W.ic0_lap3d(12, 0.01), "chain": W.lap2d_tril(1, 300)}
    for name, (Lp, Li, Lx) in mats.items():
        n = len(Lp) - 1
        plan = sa.Plan(*to_dev(Lp, Li, Lx, device=dev))
        B = W.uniform(61 + k, 0, n * k).reshape(k, n)
        X_ref = oracle.sptrsm_csr(Lp, Li, Lx, B)
        print(f"[wide] {name} n={n} k={k} layout={layout}", flush=True)
        Bl = np.ascontiguousarray(B.T) if layout == "row" else B
        for _ in range(2):
            X = to_dev(Bl.copy(), device=dev)
            plan.solve(X, layout=layout, executor="wide")
            plan.check()
            got = X.cpu().numpy()
            assert rel_err(got.T if layout == "row" else got, X_ref) <= RTOL64, name
        X = to_dev(Bl.copy(), device=dev)
        g = plan.graph(X, layout=layout, executor="wide")
        for _ in range(3):
            X.copy_(to_dev(Bl, device=dev))
            g.launch()
            plan.check()
            got = X.cpu().numpy()
            assert rel_err(got.T if layout == "row" else got, X_ref) <= RTOL64, name


@pytest.mark.skip(reason=WIDE_DISABLED)
def test_tuner_considers_the_wide_row_executor(dev):
    Lp, Li, Lx = W.zipf_lower(4000, 1.8, 256, 1500, 7)
    plan = sa.Plan(*to_dev(Lp, Li, Lx, device=dev))
    log = plan.tune(k=64, layout="col")
    assert "exec=5" in log
    n = len(Lp) - 1
    B = W.uniform(71, 0, n * 64).reshape(64, n)
    X = to_dev(B.copy(), device=dev)
    plan.solve(X, layout="col")
    plan.check()
    assert rel_err(X.cpu().numpy(), oracle.sptrsm_csr(Lp, Li, Lx, B)) <= RTOL64
