"""CPU test of the region executor's schedule and band protocol: the CPU
emulation in tests/region_emul.py runs the library's own records and must
reproduce the serial reference solve (sptrsv_csr, sparse_blas/sptrsv.cpp:17-25)
and, on the transposed schedule, ltsolve (sparse_blas/sptrsv_lt.cpp:11-20)
bit for bit, with every wait, halo load and LDS read checked."""
import numpy as np
import pytest

import oracle
from sympiler_amd import workloads as W
from tests.helpers import bitwise_equal
from tests.region_emul import emulate, schedule_host


def _batch():
    w = W.config4(count=12, lo=6, hi=30, seed=3)
    return w.Lp, w.meta["global_Li"], w.Lx


MATS = {
    "ic0_12": lambda: W.ic0_lap3d(12, 0.01),
    "random": lambda: W.random_lower(3000, 0.002, 1, seed=3),
    "zipf": lambda: W.zipf_lower(3000, 1.8, 64, 500, 5),
    "batch": _batch,
    "chain": lambda: W.lap2d_tril(1, 400),
    "lap2d": lambda: W.lap2d_tril(30, 20),
}


@pytest.mark.parametrize("name", sorted(MATS))
@pytest.mark.parametrize("cap,band", [(64, 1), (128, 2), (512, 3), (512, 7)])
def test_forward_schedule(name, cap, band):
    Lp, Li, Lx = MATS[name]()
    n = len(Lp) - 1
    sch = schedule_host(Lp, Li, Lx, cap, band)
    b = W.uniform(9, 0, n)
    x = emulate(sch, n, b)
    assert bitwise_equal(x, oracle.sptrsv_csr(Lp, Li, Lx, b))


@pytest.mark.parametrize("name", ["ic0_12", "random", "zipf", "chain"])
@pytest.mark.parametrize("cap,band", [(64, 2), (512, 2), (512, 5)])
def test_transposed_schedule(name, cap, band):
    Lp, Li, Lx = MATS[name]()
    n = len(Lp) - 1
    sch = schedule_host(Lp, Li, Lx, cap, band, transposed=True)
    g = W.uniform(10, 0, n)
    gb = emulate(sch, n, g)
    Cp, Ci, Cx, _ = oracle.csr_to_csc(Lp, Li, Lx)
    assert bitwise_equal(gb, oracle.ltsolve(Cp, Ci, Cx, g))


def test_band_bound_on_deep_regions():
    # a chain: one region of up to 512 slices has at most 31 bands, so its
    # bands widen; the grid's 8x8x8 cubes have 22 slices, 11 bands of 2
    Lp, Li, Lx = W.lap2d_tril(1, 1500)
    sch = schedule_host(Lp, Li, Lx, 512, 1)
    assert sch["slmax"] == 512
    assert max(int(h) for h in sch["rhead"][7::8]) <= 31
    Lp, Li, Lx = W.ic0_lap3d(16, 0.01)
    sch = schedule_host(Lp, Li, Lx, 512, 2)
    assert sch["slmax"] == 22 and set(int(h) for h in sch["rhead"][7::8]) == {11}
