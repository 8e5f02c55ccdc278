"""CPU checks of the host workload generators (sympiler_amd/csrc/gen.cpp)."""
import numpy as np
import pytest

import oracle
from sympiler_amd import workloads as W
from tests.helpers import bitwise_equal


def test_uniform_stream_is_deterministic_and_in_range():
    a = W.uniform(1, 0, 100000)
    assert np.array_equal(a, W.uniform(1, 0, 100000))
    assert a.min() >= -1.0 and a.max() < 1.0
    assert abs(a.mean()) < 0.01
    assert not np.array_equal(a[:100], W.uniform(2, 0, 100))
    assert np.array_equal(W.uniform(1, 0, 50), a[:50])


@pytest.mark.parametrize("nx,ny", [(1, 1), (1, 7), (5, 3), (100, 100)])
def test_lap2d_structure(nx, ny):
    Lp, Li, Lx = W.lap2d_tril(nx, ny)
    n = nx * ny
    assert len(Lp) == n + 1 and Lp[-1] == len(Li)
    for i in range(0, n, max(1, n // 50)):
        cols = Li[Lp[i]:Lp[i + 1]]
        assert cols[-1] == i and np.all(np.diff(cols) > 0)
        x, y = i % nx, i // nx
        want = ([i - nx] if y else []) + ([i - 1] if x else []) + [i]
        assert list(cols) == want


def test_ic0_generator_is_the_reference_order():
    # product generator == oracle restatement of spic0_csc_left, bit for bit
    for g in (5, 16):
        Lp, Li, A = W.lap3d_tril(g, g, g, 6.01, -1.0)
        mine = W.ic0(Lp, Li, A)
        Cp, Ci, Cx, mp = oracle.csr_to_csc(Lp, Li, A)
        pp, ps = oracle.prune_set(Lp, Li)
        lc, ok = oracle.spic0_csc_left(Cp, Ci, Cx, pp, ps)
        assert ok and bitwise_equal(mine, lc[mp])


def test_ic0_breakdown_is_reported():
    Lp, Li, A = W.lap2d_tril(4, 4, 1.0, -1.0)  # not SPD
    with pytest.raises(ValueError, match="breakdown"):
        W.ic0(Lp, Li, A)


def test_zipf_matrix_statistics():
    Lp, Li, Lx = W.zipf_lower(200000, 1.8, 256, 50000, seed=2)
    n = len(Lp) - 1
    lens = np.diff(Lp) - 1
    assert lens.max() <= 256 and lens.min() >= 0
    assert lens[0] == 0
    # mean of min(Zipf(1.8), 256) is about 7.7
    assert 7.0 < lens[1000:].mean() < 8.5
    # P(count = 1) = 1 / zeta(1.8) = 0.531...
    assert abs(np.mean(lens[1000:] == 1) - 0.5313) < 0.01
    rows = np.repeat(np.arange(n), np.diff(Lp))
    off = Li != rows
    assert np.all(Li[off] < rows[off]) and np.all(rows[off] - Li[off] <= 50000)
    assert np.all(Lx[off] < 0) and np.all(Lx[off] >= -1)
    diag = Lx[Lp[1:] - 1]
    sums = np.bincount(rows[off], weights=-Lx[off], minlength=n)
    assert np.allclose(diag, 1 + sums, rtol=0, atol=1e-12)
    for i in range(1, n, 997):
        cols = Li[Lp[i]:Lp[i + 1]]
        assert np.all(np.diff(cols) > 0) and cols[-1] == i


def test_batch_dims():
    a, b = W.batch_dims(512, 32, 224, 3)
    assert a.min() >= 32 and a.max() <= 224 and b.min() >= 32 and b.max() <= 224
    assert len(set(a.tolist())) > 100
    w = W.config4(count=6, lo=3, hi=5, seed=3)
    assert w.row_offsets[-1] == w.n and len(w.row_offsets) == 7
    x = oracle.sptrsv_csr(w.Lp, w.meta["global_Li"], w.Lx, w.b)
    assert np.max(np.abs(x - w.x_true)) < 1e-13


def test_algorithmic_bytes_formulas():
    n, nnz = 2097152, 8339456
    assert W.algorithmic_bytes_forward(n, nnz, 1) == 12 * nnz + 4 * (n + 1) + 16 * n
    assert W.algorithmic_bytes_adjoint(n, nnz) == 20 * nnz + 28 * n + 4
