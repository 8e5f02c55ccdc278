"""Deterministic workloads for the sparse triangular solve (host numpy).

Wraps libsymamd_gen.so (sympiler_amd/csrc/gen.cpp).  Every matrix is a CSR
lower triangle with the diagonal last in each row (the layout of the
reference's sym_lib::sptrsv_csr, sparse_blas/sptrsv.cpp:17-25), int32
indices and float64 values.

Random streams: splitmix64 keyed by (seed, tag); tag 0 draws right-hand
sides / x_true, tag 1 the Zipf matrix, tag 2 the batch sizes, tag 3 random
test matrices.

The five workloads:
  c1  tril of the 2D 5-point Laplacian on 100x100 (diag 4, off -1);
      b = L x_true, x_true ~ U[-1, 1] (seed 0).
  c2  IC(0) of the 3D 7-point Laplacian on 128^3 (diag 6, off -1) plus
      0.01 I (shift applied to A before the factorisation); b = L x_true,
      x_true ~ U[-1, 1] (seed 1).
  c3  n = 1e6 irregular lower triangle: strictly-lower count per row
      min(Zipf(1.8), 256, min(i, 50000)) (clipped at 256), columns uniform
      in [max(0, i-50000), i), off-diagonals U[-1, 0), diagonal
      1 + sum|off| (seed 2); B column-major n x k ~ U[-1, 1] (seed 2).
  c4  512 independent systems tril(2D 5-point Laplacian a x b), a, b ~
      U{32..224} (seed 3); b_s = L_s x_true (x_true seed 3).
  c5  IC(0) of the 3D 7-point Laplacian on 64^3 (+0.01 I); b ~ U[-1, 1]
      (seed 4).
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_int64, c_uint64, c_void_p
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
GEN_PATH = os.path.join(_HERE, "libsymamd_gen.so")
_g = None


def gen_lib():
    global _g
    if _g is None:
        if not os.path.exists(GEN_PATH):
            raise RuntimeError(f"{GEN_PATH} is missing: run `make -C sympiler_amd/csrc`")
        g = ctypes.CDLL(GEN_PATH)
        vp = c_void_p
        g.symgen_uniform.argtypes = [c_uint64, c_uint64, c_int64, c_double, c_double, vp]
        g.symgen_uniform.restype = None
        g.symgen_lap2d_tril_nnz.argtypes = [c_int, c_int]
        g.symgen_lap2d_tril_nnz.restype = c_int64
        g.symgen_lap2d_tril.argtypes = [c_int, c_int, c_double, c_double, vp, vp, vp]
        g.symgen_lap2d_tril.restype = None
        g.symgen_lap3d_tril_nnz.argtypes = [c_int, c_int, c_int]
        g.symgen_lap3d_tril_nnz.restype = c_int64
        g.symgen_lap3d_tril.argtypes = [c_int, c_int, c_int, c_double, c_double, vp, vp, vp]
        g.symgen_lap3d_tril.restype = None
        g.symgen_ic0_csr.argtypes = [c_int, vp, vp, vp]
        g.symgen_ic0_csr.restype = c_int
        g.symgen_spmv.argtypes = [c_int, vp, vp, vp, vp, vp]
        g.symgen_spmv.restype = None
        g.symgen_zipf_lower.argtypes = [c_int, c_double, c_int, c_int, c_uint64, vp, vp, vp]
        g.symgen_zipf_lower.restype = c_int64
        g.symgen_batch_dims.argtypes = [c_int, c_int, c_int, c_uint64, vp, vp]
        g.symgen_batch_dims.restype = None
        g.symgen_random_lower.argtypes = [c_int, c_double, c_int, c_uint64, vp, vp, vp]
        g.symgen_random_lower.restype = c_int64
        g.symgen_bcsc.argtypes = [c_int, c_int, c_int, c_int, c_uint64, POINTER(c_int),
                                  vp, vp, vp, vp, vp]
        g.symgen_bcsc.restype = c_int64
        g.symgen_bcsc_tree.argtypes = [c_int, c_int, c_int, c_double, c_uint64,
                                       POINTER(c_int), POINTER(c_int), vp, vp, vp, vp, vp]
        g.symgen_bcsc_tree.restype = c_int64
        _g = g
    return _g


def _p(a):
    return a.ctypes.data_as(c_void_p) if a is not None else None


def uniform(seed: int, tag: int, count: int, lo=-1.0, hi=1.0) -> np.ndarray:
    out = np.empty(count, dtype=np.float64)
    gen_lib().symgen_uniform(seed, tag, count, lo, hi, _p(out))
    return out


def lap2d_tril(nx, ny, diag=4.0, off=-1.0):
    nnz = gen_lib().symgen_lap2d_tril_nnz(nx, ny)
    n = nx * ny
    Lp = np.empty(n + 1, np.int32)
    Li = np.empty(nnz, np.int32)
    Lx = np.empty(nnz, np.float64)
    gen_lib().symgen_lap2d_tril(nx, ny, diag, off, _p(Lp), _p(Li), _p(Lx))
    return Lp, Li, Lx


def lap3d_tril(nx, ny, nz, diag=6.0, off=-1.0):
    nnz = gen_lib().symgen_lap3d_tril_nnz(nx, ny, nz)
    n = nx * ny * nz
    Lp = np.empty(n + 1, np.int32)
    Li = np.empty(nnz, np.int32)
    Lx = np.empty(nnz, np.float64)
    gen_lib().symgen_lap3d_tril(nx, ny, nz, diag, off, _p(Lp), _p(Li), _p(Lx))
    return Lp, Li, Lx


def grid3d_general(g, seed=5):
    """A general (non-symmetric) matrix on the full 7-point pattern of a
    g^3 grid, row i = x + g*(y + g*z), columns ascending: off-diagonals
    ~ U[-1, 0) (splitmix64), each diagonal 1 + its row's and column's
    absolute sums (strictly diagonally dominant).  Returns (Ap, Ai, Ax,
    diags) with diags[i] the position of row i's diagonal -- the input of
    sym_lib::spilu0_csr."""
    n = g ** 3
    idx = np.arange(n, dtype=np.int64)
    x, y, z = idx % g, (idx // g) % g, idx // (g * g)
    offs = np.array([-g * g, -g, -1, 0, 1, g, g * g], np.int64)
    ok = np.stack([z > 0, y > 0, x > 0, np.ones(n, bool), x < g - 1, y < g - 1, z < g - 1],
                  axis=1)
    cols = idx[:, None] + offs[None, :]
    Ai = cols[ok].astype(np.int32)
    cnt = ok.sum(axis=1)
    Ap = np.zeros(n + 1, np.int32)
    Ap[1:] = np.cumsum(cnt)
    diags = (Ap[:-1] + ok[:, :3].sum(axis=1)).astype(np.int32)
    Ax = uniform(seed, 0, len(Ai), -1.0, 0.0)
    rows = np.repeat(idx, cnt)
    w = np.abs(Ax)
    w[diags] = 0.0
    Ax[diags] = 1.0 + np.bincount(rows, w, n) + np.bincount(Ai, w, n)
    return Ap, Ai, Ax, diags


def ic0(Lp, Li, A_lower):
    """IC(0) factor values on the pattern of the lower triangle."""
    Lx = np.array(A_lower, dtype=np.float64, copy=True)
    rc = gen_lib().symgen_ic0_csr(len(Lp) - 1, _p(Lp), _p(Li), _p(Lx))
    if rc:
        raise ValueError(f"IC(0) breakdown: non-positive pivot at row {rc - 1}")
    return Lx


def spmv(Lp, Li, Lx, x):
    y = np.empty(len(Lp) - 1, np.float64)
    gen_lib().symgen_spmv(len(Lp) - 1, _p(Lp), _p(Li), _p(Lx),
                          _p(np.ascontiguousarray(x, np.float64)), _p(y))
    return y


def zipf_lower(n, alpha=1.8, cap=256, window=50000, seed=2):
    Lp = np.empty(n + 1, np.int32)
    nnz = gen_lib().symgen_zipf_lower(n, alpha, cap, window, seed, _p(Lp), None, None)
    Li = np.empty(nnz, np.int32)
    Lx = np.empty(nnz, np.float64)
    gen_lib().symgen_zipf_lower(n, alpha, cap, window, seed, _p(Lp), _p(Li), _p(Lx))
    return Lp, Li, Lx


def batch_dims(count=512, lo=32, hi=224, seed=3):
    a = np.empty(count, np.int32)
    b = np.empty(count, np.int32)
    gen_lib().symgen_batch_dims(count, lo, hi, seed, _p(a), _p(b))
    return a, b


def random_lower(n, density, min_off=0, seed=0):
    Lp = np.empty(n + 1, np.int32)
    nnz = gen_lib().symgen_random_lower(n, density, min_off, seed, _p(Lp), None, None)
    Li = np.empty(max(nnz, 1), np.int32)
    Lx = np.empty(max(nnz, 1), np.float64)
    gen_lib().symgen_random_lower(n, density, min_off, seed, _p(Lp), _p(Li), _p(Lx))
    return Lp, Li[:nnz], Lx[:nnz]


def bcsc(n, wmax=32, rmax=32, window=4096, seed=6):
    """Supernodal (BCSC) lower triangle for sptrsv_csc_block (see
    symgen_bcsc): (supernodes, nrows, Lp, Li, Lx)."""
    nn = c_int()
    cnt = gen_lib().symgen_bcsc(n, wmax, rmax, window, seed, ctypes.byref(nn), None, None,
                                None, None, None)
    if cnt >= 2 ** 31:
        raise ValueError("supernodal value count exceeds int32 offsets")
    m = nn.value
    sup = np.empty(m + 1, np.int32)
    nrows = np.empty(max(m, 1), np.int32)[:m]
    Lp = np.empty(m + 1, np.int32)
    Li = np.empty(max(cnt, 1), np.int32)[:cnt]
    Lx = np.empty(max(cnt, 1), np.float64)[:cnt]
    gen_lib().symgen_bcsc(n, wmax, rmax, window, seed, ctypes.byref(nn), _p(sup), _p(nrows),
                          _p(Lp), _p(Li), _p(Lx))
    return sup, nrows, Lp, Li, Lx


def bcsc_tree(levels=12, wleaf=16, wmax=512, rowfac=2.0, seed=8):
    """Supernodal lower triangle shaped like a nested-dissection factor (see
    symgen_bcsc_tree): (supernodes, nrows, Lp, Li, Lx), n = supernodes[-1]."""
    nn, n = c_int(), c_int()
    cnt = gen_lib().symgen_bcsc_tree(levels, wleaf, wmax, rowfac, seed, ctypes.byref(nn),
                                     ctypes.byref(n), None, None, None, None, None)
    if cnt >= 2 ** 31:
        raise ValueError("supernodal value count exceeds int32 offsets")
    m = nn.value
    sup = np.empty(m + 1, np.int32)
    nrows = np.empty(m, np.int32)
    Lp = np.empty(m + 1, np.int32)
    Li = np.empty(cnt, np.int32)
    Lx = np.empty(cnt, np.float64)
    gen_lib().symgen_bcsc_tree(levels, wleaf, wmax, rowfac, seed, ctypes.byref(nn),
                               ctypes.byref(n), _p(sup), _p(nrows), _p(Lp), _p(Li), _p(Lx))
    return sup, nrows, Lp, Li, Lx


@dataclass
class Workload:
    name: str
    Lp: np.ndarray
    Li: np.ndarray
    Lx: np.ndarray
    b: np.ndarray                     # (n,) or (k, n) column-major blocks
    x_true: np.ndarray | None = None
    row_offsets: np.ndarray | None = None
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return len(self.Lp) - 1

    @property
    def nnz(self):
        return int(self.Lp[-1])


def config1(nx=100, ny=100, seed=0) -> Workload:
    Lp, Li, Lx = lap2d_tril(nx, ny, 4.0, -1.0)
    x = uniform(seed, 0, nx * ny)
    return Workload(f"lap2d_{nx}x{ny}_tril", Lp, Li, Lx, spmv(Lp, Li, Lx, x), x,
                    meta={"grid": (nx, ny)})


def ic0_lap3d(g, shift=0.01):
    Lp, Li, A = lap3d_tril(g, g, g, 6.0 + shift, -1.0)
    return Lp, Li, ic0(Lp, Li, A)


def config2(g=128, seed=1, shift=0.01) -> Workload:
    Lp, Li, Lx = ic0_lap3d(g, shift)
    x = uniform(seed, 0, g ** 3)
    return Workload(f"ic0_lap3d_{g}^3", Lp, Li, Lx, spmv(Lp, Li, Lx, x), x,
                    meta={"grid": (g, g, g), "shift": shift})


def config3(n=1_000_000, k=1, seed=2) -> Workload:
    Lp, Li, Lx = zipf_lower(n, 1.8, 256, 50000, seed)
    B = uniform(seed, 0, n * k).reshape(k, n)
    return Workload(f"zipf1.8_n{n}", Lp, Li, Lx, B, None, meta={"k": k})


def config4(count=512, lo=32, hi=224, seed=3) -> Workload:
    a, b = batch_dims(count, lo, hi, seed)
    Lps, Lis, Lxs, offs = [], [], [], [0]
    nnz_off = 0
    for s in range(count):
        Lp, Li, Lx = lap2d_tril(int(a[s]), int(b[s]), 4.0, -1.0)
        Lps.append(Lp[:-1] + nnz_off)
        Lis.append(Li)
        Lxs.append(Lx)
        nnz_off += int(Lp[-1])
        offs.append(offs[-1] + int(a[s]) * int(b[s]))
    Lp = np.concatenate(Lps + [np.array([nnz_off], np.int32)]).astype(np.int32)
    Li = np.concatenate(Lis).astype(np.int32)
    Lx = np.concatenate(Lxs)
    n = offs[-1]
    x = uniform(seed, 0, n)
    # b_s = L_s x_true,s with local column ids: shift to global for the product
    off = np.asarray(offs, np.int64)
    row_sys = np.repeat(np.arange(count), np.diff(off))
    Gi = (Li.astype(np.int64) + np.repeat(off[:-1][row_sys], np.diff(Lp))).astype(np.int32)
    bvec = spmv(Lp, Gi, Lx, x)
    return Workload(f"batch{count}_lap2d_tril", Lp, Li, Lx, bvec, x,
                    row_offsets=np.asarray(offs, np.int32),
                    meta={"dims": (a, b), "global_Li": Gi})


def config5(g=64, seed=4) -> Workload:
    Lp, Li, Lx = ic0_lap3d(g, 0.01)
    b = uniform(seed, 0, g ** 3)
    return Workload(f"ic0_lap3d_{g}^3", Lp, Li, Lx, b, None,
                    meta={"grid": (g, g, g), "shift": 0.01})


def algorithmic_bytes_forward(n, nnz, k, vbytes=8):
    """12*nnz + 4*(n+1) + 16*n*k for FP64: values + column ids once, row
    pointers once, B read and X written once."""
    return (vbytes + 4) * nnz + 4 * (n + 1) + 2 * vbytes * n * k


def algorithmic_bytes_adjoint_block(n, nnz, k, vbytes=8):
    """Block adjoint of k right-hand sides: L read once (12*nnz + 4*(n+1)
    for FP64), X and GX read (8nk each), GB written (8nk), grad-L written
    (8*nnz)."""
    return (vbytes + 4) * nnz + 4 * (n + 1) + 3 * vbytes * n * k + vbytes * nnz


def algorithmic_bytes_adjoint(n, nnz, vbytes=8):
    """20*nnz + 28*n + 4 for FP64: L read once (12*nnz + 4*(n+1)), gx and x
    read (8n each), gb written (8n), grad-L written (8*nnz)."""
    return (vbytes + 4) * nnz + 4 * (n + 1) + 3 * vbytes * n + vbytes * nnz
