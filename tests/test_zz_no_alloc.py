"""symamd_levelset_noalloc (the level-set inspector with its working memory
in a fixed device array) against symamd_levelset through the C ABI: equal
arrays on two calls in a row, the second on a smaller input (the fixed
array keeps the first call's contents), the largest accepted size, and the
size limit refused before any device work."""
import ctypes
from ctypes import POINTER, byref, c_int, c_void_p

import numpy as np
import pytest

import oracle
import sympiler_amd as sa
from sympiler_amd import workloads as W
from tests.helpers import to_dev

pytestmark = pytest.mark.gpu
MAX_ROWS = 1 << 21


def _noalloc():
    fn = sa.lib().symamd_levelset_noalloc
    fn.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, POINTER(c_int)]
    fn.restype = c_int
    return fn


def _levels(fn, h, dLp, dLi):
    import torch
    n = dLp.numel() - 1
    lp = torch.full((n + 1,), -7, dtype=torch.int32, device=dLp.device)
    ls = torch.full((max(n, 1),), -7, dtype=torch.int32, device=dLp.device)
    lev = c_int(-1)
    rc = fn(h._h, n, c_void_p(dLp.data_ptr()), c_void_p(dLi.data_ptr()),
            c_void_p(lp.data_ptr()), c_void_p(ls.data_ptr()), byref(lev))
    return rc, lev.value, lp.cpu().numpy(), ls.cpu().numpy()[:n]


def test_noalloc_equals_levelset_two_calls(dev):
    fn = _noalloc()
    h = sa.Handle()
    cases = [W.ic0_lap3d(48)[:2], W.random_lower(5000, 0.002, 1, seed=11)[:2]]
    for Lp, Li in cases:  # the second call is the smaller one
        dLp, dLi = to_dev(Lp, Li, device=dev)
        rc, nl, lp, ls = _levels(fn, h, dLp, dLi)
        assert rc == 0, sa.lib().symamd_last_error()
        want_lp, want_ls = (t.cpu().numpy() for t in sa.levelset(dLp, dLi, handle=h))
        assert nl == len(want_lp) - 1
        assert np.array_equal(lp[:nl + 1], want_lp) and np.array_equal(ls, want_ls)
        o_lp, o_ls, _ = oracle.levelset_csr(Lp, Li)
        assert np.array_equal(lp[:nl + 1], o_lp) and np.array_equal(ls, o_ls)


def test_noalloc_largest_size_and_limit(dev):
    fn = _noalloc()
    h = sa.Handle()
    w = W.config2()  # n = 2^21 rows: the largest accepted size
    assert w.n == MAX_ROWS
    dLp, dLi = to_dev(w.Lp, w.Li, device=dev)
    rc, nl, lp, ls = _levels(fn, h, dLp, dLi)
    assert rc == 0, sa.lib().symamd_last_error()
    assert nl == 382
    want_lp, want_ls = (t.cpu().numpy() for t in sa.levelset(dLp, dLi, handle=h))
    assert np.array_equal(lp[:nl + 1], want_lp) and np.array_equal(ls, want_ls)
    # one row more is refused before any array is touched
    lev = c_int(-1)
    p = c_void_p(dLp.data_ptr())
    rc = fn(h._h, MAX_ROWS + 1, p, p, p, p, byref(lev))
    assert rc == 7 and b"exceed" in sa.lib().symamd_last_error()
