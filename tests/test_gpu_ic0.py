"""IC(0) on the device (symamd_ic0 / Plan.ic0) against the oracle's
restatements of sparse_blas/spic0.cpp: bit-identical factors for both
variants on grids, random SPD patterns and a heavy-row matrix, the square
root's rounding over many exponents, breakdown reporting, in-place use, and
the factor -> solve pipeline staying on the device."""
import numpy as np
import pytest
import torch

import oracle
import sympiler_amd as sa
from sympiler_amd import workloads as W
from tests.helpers import bitwise_equal, to_dev
from tests.test_ic0_oracle import spd_random

pytestmark = pytest.mark.gpu


def oracle_factor(Lp, Li, A, variant):
    """The oracle's factor mapped to the CSR positions, and the first failed
    column (spic0_csc_left stops there; spic0_csr computes every column)."""
    Cp, Ci, Cx, mp = oracle.csr_to_csc(Lp, Li, A)
    if variant == "left":
        pp, ps = oracle.prune_set(Lp, Li)
        Lc, ok = oracle.spic0_csc_left(Cp, Ci, Cx, pp, ps)
    else:
        Lc, ok = oracle.spic0_csr(Cp, Ci, Cx)
    return Lc[mp], ok


def heavy_rows(n, seed):
    """SPD lower triangle with a few rows of many entries (long update lists
    and long rows for one lane)."""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        k = int(rng.integers(60, 120)) if i % 97 == 96 else int(rng.integers(0, 4))
        k = min(k, i)
        cols = sorted(rng.choice(i, size=k, replace=False).tolist()) if k else []
        rows.append(cols)
    Lp = np.zeros(n + 1, np.int32)
    Li = []
    for i, cols in enumerate(rows):
        Li.extend(cols + [i])
        Lp[i + 1] = len(Li)
    Li = np.asarray(Li, np.int32)
    A = rng.uniform(-1.0, 1.0, len(Li))
    diag = Lp[1:] - 1
    r = np.repeat(np.arange(n), np.diff(Lp))
    w = np.abs(A)
    w[diag] = 0.0
    A[diag] = 1.0 + np.bincount(r, w, n) + np.bincount(Li, w, n)
    return Lp, Li, A


MATS = {
    "lap2d_40x30": lambda: W.lap2d_tril(40, 30, 4.0, -1.0),
    "lap3d_12": lambda: W.lap3d_tril(12, 12, 12, 6.01, -1.0),
    "random_3000": lambda: spd_random(3000, 0.003, 61),
    "heavy_2000": lambda: heavy_rows(2000, 62),
}


@pytest.mark.parametrize("variant", ["left", "csr"])
@pytest.mark.parametrize("name", list(MATS))
def test_ic0_bitwise(dev, name, variant):
    Lp, Li, A = MATS[name]()
    print(f"[ic0] {name} {variant} n={len(Lp) - 1} nnz={len(Li)}", flush=True)
    want, ok = oracle_factor(Lp, Li, A, variant)
    assert ok
    dLp, dLi, dA = to_dev(Lp, Li, A)
    plan = sa.Plan(dLp, dLi, dA)
    got = plan.ic0(dA, variant)
    assert bitwise_equal(got.cpu().numpy(), want)
    # the lists are cached; in place gives the same factor
    again = dA.clone()
    plan.ic0(again, variant, out=again)
    assert bitwise_equal(again.cpu().numpy(), want)
    assert bitwise_equal(dA.cpu().numpy(), A)


def test_ic0_sqrt_rounding_over_exponents(dev):
    # a diagonal matrix: L[i,i] = sqrt(A[i,i]), so the device square root is
    # compared with the correctly rounded one on a million values whose
    # exponents span [2^-600, 2^600]
    n = 1 << 20
    rng = np.random.default_rng(5)
    A = np.ldexp(rng.uniform(1.0, 2.0, n), rng.integers(-600, 600, n))
    A[:4] = [1.0, 2.0, 4.0, np.nextafter(1.0, 2.0)]
    Lp = np.arange(n + 1, dtype=np.int32)
    Li = np.arange(n, dtype=np.int32)
    dLp, dLi, dA = to_dev(Lp, Li, A)
    plan = sa.Plan(dLp, dLi, dA)
    L = plan.ic0(dA, "left")
    assert bitwise_equal(L.cpu().numpy(), np.sqrt(A))
    # a diagonal factor has one level and no dependencies: x = b / L[i,i]
    plan.update_values(L)
    b = W.uniform(11, 0, n)
    for ex in ("syncfree", "levelset"):
        x = to_dev(b.copy())
        plan.solve(x, executor=ex)
        plan.check()
        assert bitwise_equal(x.cpu().numpy(), b / np.sqrt(A))


@pytest.mark.parametrize("variant", ["left", "csr"])
def test_ic0_breakdown(dev, variant):
    Lp, Li, A = W.lap2d_tril(6, 6, 2.2, -1.0)  # not positive definite
    want, ok = oracle_factor(Lp, Li, A, variant)
    assert not ok
    n = len(Lp) - 1
    diag = np.asarray(Lp[1:]) - 1
    # the oracle's first failed column: its diagonal is unwritten (0, left)
    # or the square root of a pivot that is not positive (csr)
    jf = int(np.argmax(~(want[diag] > 0)))
    dLp, dLi, dA = to_dev(Lp, Li, A)
    plan = sa.Plan(dLp, dLi, dA)
    with pytest.raises(ValueError, match="breakdown"):
        plan.ic0(dA, variant)
    got, bad = plan.ic0(dA, variant, strict=False)
    got = got.cpu().numpy()
    assert bad == jf
    if variant == "left":
        # spic0_csc_left wrote only the columns before the failed one
        cols = np.asarray(Li) < jf
        assert bitwise_equal(got[cols], want[cols])
    else:
        # spic0_csr computes every column: equal bits, NaN where it has NaN
        nan = np.isnan(want)
        assert np.array_equal(np.isnan(got), nan)
        assert bitwise_equal(got[~nan], want[~nan])


def test_ic0_then_solve_on_device(dev):
    Lp, Li, A = W.lap3d_tril(16, 16, 16, 6.01, -1.0)
    want, ok = oracle_factor(Lp, Li, A, "left")
    assert ok
    dLp, dLi, dA = to_dev(Lp, Li, A)
    plan = sa.Plan(dLp, dLi, dA)
    L = plan.ic0(dA, "left")
    plan.update_values(L)
    n = len(Lp) - 1
    b = W.uniform(9, 0, n)
    x = to_dev(b.copy())
    plan.set_lanes(1)
    plan.solve(x, executor="syncfree")
    plan.check()
    assert bitwise_equal(x.cpu().numpy(), oracle.sptrsv_csr(Lp, Li, want, b))
