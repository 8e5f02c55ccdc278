"""The device coarsened-schedule inspector (symamd_coarsen, and the
sym_lib::get_coarse_levelSet_DAG_CSC_tree drop-in with the demo's arguments,
demo/sptrsv_demo_utils.h:80-88, demo/sptrsv_demo.cpp:37): arrays equal to
the host restatement tests/lbc_host.py entry for entry, the reference's
schedule invariants (oracle.check_schedule), and LBC solves of the schedule
bit-identical to the serial sptrsv_csr (sparse_blas/sptrsv.cpp:17-25)."""
import ctypes
import os
from ctypes import POINTER, byref, c_double, c_int, c_size_t, c_void_p

import numpy as np
import pytest

import oracle
import sympiler_amd as sa
from sympiler_amd import workloads as W
from tests.helpers import bitwise_equal, to_dev
from tests.lbc_host import lbc_schedule
from tests.test_lbc_host import CASES

pytestmark = pytest.mark.gpu
LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "sympiler_amd", "libsympiler_amd.so")


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("params", [(6, -1, 4000), (1, -1, 4000), (4, 3, 4000),
                                    (32, -1, 8), (3, 0, 1)])
def test_device_schedule_equals_restatement(dev, case, params):
    import torch
    Lp, Li, Lx = CASES[case]()
    cost = W.uniform(7, 0, len(Lp) - 1, 0.0, 5.0) if case == "zipf" else None
    want = lbc_schedule(Lp, Li, cost, *params)
    dLp, dLi, dLx = to_dev(Lp, Li, Lx, device=dev)
    got = sa.coarsen(dLp, dLi, *params, cost=None if cost is None else to_dev(cost, device=dev))
    got = [t.cpu().numpy() for t in got]
    for g, w in zip(got, want):
        assert np.array_equal(g, w), (case, params)
    assert oracle.check_schedule(Lp, Li, *got) == 0
    b = W.uniform(5, 0, len(Lp) - 1)
    x = to_dev(b, device=dev)
    sa.solve_lbc(dLp, dLi, dLx, x, got[0], to_dev(got[1], device=dev), to_dev(got[2], device=dev))
    torch.cuda.synchronize()
    assert bitwise_equal(x.cpu().numpy(), oracle.sptrsv_csr(Lp, Li, Lx, b))


def test_full_size_ic0_schedule(dev):
    """The 128^3 IC(0) factor with the demo's parameters: valid, and the
    device LBC solve of it equals the reference digest."""
    import hashlib
    import json
    import torch
    w = W.config2()
    dLp, dLi, dLx = to_dev(w.Lp, w.Li, w.Lx, device=dev)
    lp, pp, part = [t.cpu().numpy() for t in sa.coarsen(dLp, dLi, 6, -1, 4000)]
    assert oracle.check_schedule(w.Lp, w.Li, lp, pp, part) == 0
    assert len(lp) - 1 == 382  # grid levels never merge with six parts
    x = to_dev(w.b, device=dev)
    sa.solve_lbc(dLp, dLi, dLx, x, lp, to_dev(pp, device=dev), to_dev(part, device=dev))
    torch.cuda.synchronize()
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "digests.json")))
    assert hashlib.sha256(x.cpu().numpy().tobytes()).hexdigest() == gold["c2_x"]["sha256"]


def test_dropin_with_demo_arguments(dev):
    """get_coarse_levelSet_DAG_CSC_tree(n, Lp, Li, stype, ...) with lp = 6,
    cp = -1, ic = 4000 and cost = row nnz, as the demo calls it, then the
    sptrsv_csr_lbc drop-in on its schedule."""
    L = ctypes.CDLL(LIB)
    Lp, Li, Lx = W.random_lower(5000, 0.001, 1, seed=12)
    n = len(Lp) - 1
    cost = np.diff(Lp).astype(np.float64)
    fn = L["_ZN7sym_lib32get_coarse_levelSet_DAG_CSC_treeEmPKiS1_iRiRPiS2_S4_S4_iiiPd"]
    fn.restype = c_int
    lno, pno = c_int(), c_int()
    plp, ppp, ppart = POINTER(c_int)(), POINTER(c_int)(), POINTER(c_int)()
    P = lambda a: a.ctypes.data_as(c_void_p)
    r = fn(c_size_t(n), P(Lp), P(Li), c_int(-1), byref(lno), byref(plp), byref(pno),
           byref(ppp), byref(ppart), c_int(6), c_int(-1), c_int(4000), P(cost))
    assert r == 1
    lp = np.ctypeslib.as_array(plp, (lno.value + 1,)).copy()
    pp = np.ctypeslib.as_array(ppp, (pno.value + 1,)).copy()
    part = np.ctypeslib.as_array(ppart, (n,)).copy()
    want = lbc_schedule(Lp, Li, cost, 6, -1, 4000)
    for g, w in zip((lp, pp, part), want):
        assert np.array_equal(g, w)
    b = W.uniform(12, 0, n)
    x = b.copy()
    L["_ZN7sym_lib14sptrsv_csr_lbcEiPiS0_PdS1_iS0_S0_S0_"](
        c_int(n), P(Lp), P(Li), P(Lx), P(x), lno, P(lp), P(pp), P(part))
    assert bitwise_equal(x, oracle.sptrsv_csr(Lp, Li, Lx, b))
