"""The oracle -- TEST INFRASTRUCTURE ONLY.

CPU restatement (oracle/oracle.c -> liboracle.so) of the reference's
sparse triangular solve path, plus, when built in the development
container, the reference's own sources compiled by oracle/Makefile into
oracle/_ref/libsymref.so.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this package, and only as the
checker or the timed CPU baseline; the product (sympiler_amd/) never does.

Functions and the reference lines they follow:
  sptrsv_csr           sparse_blas/sptrsv.cpp:17-25
  sptrsv_csr_levelset  sparse_blas/sptrsv.cpp:40-54
  sptrsv_csr_lbc       sparse_blas/sptrsv.cpp:57-75
  sptrsv_csr_lbc_redundant  sparse_blas/sptrsv.cpp:78-106
  sptrsv_csr_ilu       sparse_blas/sptrsv.cpp:28-38 (unit lower, no divide)
  sptrsv_csr_ilu_lbc   sparse_blas/sptrsv.cpp:109-126 (divides)
  sptrsv_csc           sparse_blas/sptrsv.cpp:129-137
  ltsolve              sparse_blas/sptrsv_lt.cpp:11-20
  spmv_csr / _axpby    sparse_blas/spmv.cpp:12-21 / :23-34
  spmv_csr_block       sparse_blas/spmv.cpp:59-71
  spmv_csc / _axpby    sparse_blas/spmv.cpp:73-81 / :85-109
  spmv_csc_lbc         sparse_blas/spmv.cpp:122-143
  spmv_csc_block       sparse_blas/spmv.cpp:171-186
  spic0_csc_left       sparse_blas/spic0.cpp:11-46
  spic0_csr            sparse_blas/spic0.cpp:106-127
  spilu0_csr           sparse_blas/spilu0.cpp:11-51
  sptrsv_csc_block     sparse_blas/sptrsv.cpp:176-199 with dense_blas/BLAS.cpp
                       lsolve_BLAS (:3-96) and matvec_BLAS (:286-349)
  levelset_csr/_csc    build_levelSet_CSC (aggregation, absent); definition
                       of parsy/common/TreeUtils.cpp:89-141
  sptrsm_csr           k columns of sptrsv_csr
  adjoint              ltsolve + grad over L's pattern (closed form)
  check_schedule       validity of a level_ptr/par_ptr/partition schedule
"""
from __future__ import annotations

import ctypes
import os
from ctypes import c_double, c_int, c_int64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(_HERE, "liboracle.so")
REF_PATH = os.path.join(_HERE, "_ref", "libsymref.so")

_o = None
_r = None


def _p(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


def olib():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        L = ctypes.CDLL(ORACLE_PATH)
        vp, i, d, i64 = c_void_p, c_int, c_double, c_int64
        sigs = {
            "orc_sptrsv_csr": ([i, vp, vp, vp, vp], None),
            "orc_sptrsv_csr_levelset": ([i, vp, vp, vp, vp, i, vp, vp, i], None),
            "orc_sptrsv_csr_lbc": ([i, vp, vp, vp, vp, i, vp, vp, vp], None),
            "orc_sptrsv_csr_lbc_redundant": ([i, vp, vp, vp, vp, vp, i, vp, vp, vp], None),
            "orc_sptrsv_csr_ilu": ([i, vp, vp, vp, vp, vp], None),
            "orc_sptrsv_csr_ilu_lbc": ([i, vp, vp, vp, vp, vp, i, vp, vp, vp], None),
            "orc_sptrsv_csc": ([i, vp, vp, vp, vp], None),
            "orc_ltsolve": ([i, vp, vp, vp, vp], c_int),
            "orc_sptrsm_csr": ([i, vp, vp, vp, vp, i, i64, i64], None),
            "orc_spmv_csr": ([i, vp, vp, vp, vp, vp], None),
            "orc_spmv_csr_axpby": ([i, vp, vp, vp, vp, vp, vp, d, d], None),
            "orc_spmv_csr_block": ([i, vp, vp, vp, i, vp, vp, vp], None),
            "orc_spmv_csc": ([i, vp, vp, vp, vp, vp], None),
            "orc_spmv_csc_axpby": ([i, vp, vp, vp, vp, vp, vp, d, d, vp, vp], None),
            "orc_spmv_csc_lbc": ([i, vp, vp, i, vp, vp, vp, vp, vp], None),
            "orc_spmv_csc_block": ([i, vp, vp, vp, vp, vp, i, vp], None),
            "orc_spic0_csc_left": ([i, vp, vp, vp, vp, vp, vp], c_int),
            "orc_spic0_csr": ([i, vp, vp, vp], c_int),
            "orc_spilu0_csr": ([i, i, vp, vp, vp, vp, vp], c_int),
            "orc_sptrsv_csc_block": ([i, vp, vp, vp, vp, vp, i, vp], None),
            "orc_sptrsv_csc_block_t": ([i, vp, vp, vp, vp, vp, i, vp], None),
            "orc_levelset_csr": ([i, vp, vp, vp, vp, vp], c_int),
            "orc_levelset_csc": ([i, vp, vp, vp, vp], c_int),
            "orc_check_schedule": ([i, vp, vp, i, vp, vp, vp], c_int),
            "orc_adjoint": ([i, vp, vp, vp, vp, vp, vp, vp, vp, vp], None),
            "orc_adjoint_block": ([i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp], None),
            "orc_pairwise_dot": ([vp, vp, i], ctypes.c_double),
            "orc_csr_to_csc": ([i, vp, vp, vp, vp, vp, vp, vp], None),
        }
        for name, (args, res) in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _o = L
    return _o


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def rlib():
    """The reference's own compiled functions (oracle/_ref), or raise."""
    global _r
    if _r is None:
        if not ref_available():
            raise RuntimeError("oracle/_ref/libsymref.so not built (needs /root/reference)")
        L = ctypes.CDLL(REF_PATH)
        vp, i, d = c_void_p, c_int, c_double
        sigs = {
            "ref_sptrsv_csr": ([i, vp, vp, vp, vp], None),
            "ref_sptrsv_csr_levelset": ([i, vp, vp, vp, vp, i, vp, vp], None),
            "ref_sptrsv_csr_lbc": ([i, vp, vp, vp, vp, i, vp, vp, vp], None),
            "ref_sptrsv_csr_lbc_redundant": ([i, vp, vp, vp, vp, vp, i, vp, vp, vp, vp],
                                             None),
            "ref_max_threads": ([], c_int),
            "ref_sptrsv_csr_ilu": ([i, vp, vp, vp, vp, vp], None),
            "ref_sptrsv_csr_ilu_lbc": ([i, vp, vp, vp, vp, vp, i, vp, vp, vp], None),
            "ref_sptrsv_csc": ([i, vp, vp, vp, vp], None),
            "ref_sptrsv_csc_block": ([i, vp, vp, vp, vp, vp, i, vp], None),
            "ref_ltsolve": ([i, vp, vp, vp, vp], c_int),
            "ref_spmv_csr": ([i, vp, vp, vp, vp, vp], None),
            "ref_spmv_csr_axpby": ([i, vp, vp, vp, vp, vp, vp, d, d], None),
            "ref_spmv_csr_block": ([i, vp, vp, vp, i, vp, vp, vp], None),
            "ref_spmv_csc": ([i, vp, vp, vp, vp, vp], None),
            "ref_spmv_csc_axpby": ([i, vp, vp, vp, vp, vp, vp, d, d], None),
            "ref_spmv_csc_axpby_vec": ([i, vp, vp, vp, vp, vp, vp, vp, vp], None),
            "ref_spmv_csc_lbc": ([i, vp, vp, i, vp, vp, vp, vp, vp], None),
            "ref_spmv_csc_block": ([i, vp, vp, vp, vp, vp, i, vp], None),
            "ref_spic0_csc_left": ([i, vp, vp, vp, vp, vp, vp], c_int),
            "ref_spic0_csr": ([i, vp, vp, vp], None),
            "ref_getLevelSet": ([i, vp, vp, vp], c_int),
            "ref_getCoarseLevelSet_6": ([i, vp, vp, i, i, i, vp, vp, vp, vp, vp], None),
            "ref_lsolve_BLAS": ([i, i, vp, vp], None),
            "ref_matvec_BLAS": ([i, i, i, vp, vp, vp], None),
            "ref_parsy_dlsolve": ([i, i, vp, vp], None),
            "ref_parsy_dmatvec": ([i, i, i, vp, vp, vp], None),
            "ref_set_threads": ([i], None),
        }
        for name, (args, res) in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _r = L
    return _r


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ------------------------------------------------------------- restatement
def sptrsv_csr(Lp, Li, Lx, b):
    x = _f64(b).copy()
    olib().orc_sptrsv_csr(len(Lp) - 1, _p(_i32(Lp)), _p(_i32(Li)), _p(_f64(Lx)), _p(x))
    return x


def sptrsv_csr_levelset(Lp, Li, Lx, b, level_ptr, level_set, nthreads=0):
    x = _f64(b).copy()
    olib().orc_sptrsv_csr_levelset(len(Lp) - 1, _p(_i32(Lp)), _p(_i32(Li)),
                                   _p(_f64(Lx)), _p(x), len(level_ptr) - 1,
                                   _p(_i32(level_ptr)), _p(_i32(level_set)), nthreads)
    return x


def sptrsv_csr_lbc(Lp, Li, Lx, b, level_ptr, par_ptr, partition):
    x = _f64(b).copy()
    olib().orc_sptrsv_csr_lbc(len(Lp) - 1, _p(_i32(Lp)), _p(_i32(Li)), _p(_f64(Lx)),
                              _p(x), len(level_ptr) - 1, _p(_i32(level_ptr)),
                              _p(_i32(par_ptr)), _p(_i32(partition)))
    return x


def sptrsv_csr_lbc_redundant(Lp, Li, Lx, x0, b, level_ptr, par_ptr, partition):
    x = _f64(x0).copy()
    olib().orc_sptrsv_csr_lbc_redundant(len(Lp) - 1, _p(_i32(Lp)), _p(_i32(Li)),
                                        _p(_f64(Lx)), _p(x), _p(_f64(b)),
                                        len(level_ptr) - 1, _p(_i32(level_ptr)),
                                        _p(_i32(par_ptr)), _p(_i32(partition)))
    return x


def sptrsv_csr_ilu(Ap, Ai, Ax, diags, b):
    x = _f64(b).copy()
    olib().orc_sptrsv_csr_ilu(len(Ap) - 1, _p(_i32(Ap)), _p(_i32(Ai)), _p(_f64(Ax)),
                              _p(_i32(diags)), _p(x))
    return x


def sptrsv_csr_ilu_lbc(Ap, Ai, Ax, diags, b, level_ptr, par_ptr, partition):
    x = _f64(b).copy()
    olib().orc_sptrsv_csr_ilu_lbc(len(Ap) - 1, _p(_i32(Ap)), _p(_i32(Ai)), _p(_f64(Ax)),
                                  _p(_i32(diags)), _p(x), len(level_ptr) - 1,
                                  _p(_i32(level_ptr)), _p(_i32(par_ptr)),
                                  _p(_i32(partition)))
    return x


def sptrsv_csc(Cp, Ci, Cx, b):
    x = _f64(b).copy()
    olib().orc_sptrsv_csc(len(Cp) - 1, _p(_i32(Cp)), _p(_i32(Ci)), _p(_f64(Cx)), _p(x))
    return x


def ltsolve(Cp, Ci, Cx, b):
    x = _f64(b).copy()
    olib().orc_ltsolve(len(Cp) - 1, _p(_i32(Cp)), _p(_i32(Ci)), _p(_f64(Cx)), _p(x))
    return x


def sptrsm_csr(Lp, Li, Lx, B):
    """B: (k, n) array = column-major n x k block; returns X of the same shape."""
    X = _f64(B).copy()
    k, n = X.shape
    olib().orc_sptrsm_csr(n, _p(_i32(Lp)), _p(_i32(Li)), _p(_f64(Lx)), _p(X), k, 1, n)
    return X


def spmv_csr(Ap, Ai, Ax, x):
    y = np.empty(len(Ap) - 1, np.float64)
    olib().orc_spmv_csr(len(Ap) - 1, _p(_i32(Ap)), _p(_i32(Ai)), _p(_f64(Ax)),
                        _p(_f64(x)), _p(y))
    return y


def spmv_csr_axpby(Ap, Ai, Ax, x, y, z, a, b):
    z = _f64(z).copy()
    olib().orc_spmv_csr_axpby(len(Ap) - 1, _p(_i32(Ap)), _p(_i32(Ai)), _p(_f64(Ax)),
                              _p(_f64(x)), _p(_f64(y)), _p(z), a, b)
    return z


def spmv_csr_block(Ap, Ai, Ax, supernodes, x, y):
    """spmv_csr_block: y (copied) with the supernodes' rows set to A x."""
    y = _f64(y).copy()
    olib().orc_spmv_csr_block(len(Ap) - 1, _p(_i32(Ap)), _p(_i32(Ai)), _p(_f64(Ax)),
                              len(supernodes) - 1, _p(_i32(supernodes)), _p(_f64(x)), _p(y))
    return y


def spmv_csc(Cp, Ci, Cx, x, y):
    """spmv_csc: y (copied) += A x, A in CSC with len(Cp) - 1 columns."""
    y = _f64(y).copy()
    olib().orc_spmv_csc(len(Cp) - 1, _p(_i32(Cp)), _p(_i32(Ci)), _p(_f64(Cx)), _p(_f64(x)),
                        _p(y))
    return y


def spmv_csc_axpby(Cp, Ci, Cx, x, y, z, a, b):
    """spmv_csc's axpby forms: z (copied) += A x, then a z + b y; a and b
    scalars or per-row arrays."""
    z = _f64(z).copy()
    vec = not np.isscalar(a)
    av, bv = (_f64(a), _f64(b)) if vec else (None, None)
    olib().orc_spmv_csc_axpby(len(Cp) - 1, _p(_i32(Cp)), _p(_i32(Ci)), _p(_f64(Cx)),
                              _p(_f64(x)), _p(_f64(y)), _p(z), 0.0 if vec else float(a),
                              0.0 if vec else float(b), _p(av) if vec else None,
                              _p(bv) if vec else None)
    return z


def spmv_csc_lbc(sets, ids, Cp, Ci, Cx, x, y):
    y = _f64(y).copy()
    olib().orc_spmv_csc_lbc(len(sets) - 1, _p(_i32(sets)), _p(_i32(ids)), len(Cp) - 1,
                            _p(_i32(Cp)), _p(_i32(Ci)), _p(_f64(Cx)), _p(_f64(x)), _p(y))
    return y


def spmv_csc_block(Cp, Ci, Cx, x, y, supernodes):
    y = _f64(y).copy()
    olib().orc_spmv_csc_block(len(Cp) - 1, _p(_i32(Cp)), _p(_i32(Ci)), _p(_f64(Cx)),
                              _p(_f64(x)), _p(y), len(supernodes) - 1, _p(_i32(supernodes)))
    return y


def spic0_csc_left(Ap, Ai, Ax, prunePtr, pruneSet):
    Lx = np.zeros(len(Ax), np.float64)
    ok = olib().orc_spic0_csc_left(len(Ap) - 1, _p(_i32(Ap)), _p(_i32(Ai)), _p(_f64(Ax)),
                                   _p(_i32(prunePtr)), _p(_i32(pruneSet)), _p(Lx))
    return Lx, bool(ok)


def spic0_csr(Ap, Ai, Ax):
    """In place on a copy of the CSC values: (Lx, all pivots positive)."""
    Lx = _f64(Ax).copy()
    ok = olib().orc_spic0_csr(len(Ap) - 1, _p(Lx), _p(_i32(Ap)), _p(_i32(Ai)))
    return Lx, bool(ok)


def spilu0_csr(Ap, Ai, Ax, A_diag):
    """(l, first row with a zero U diagonal or -1; -2 on a structure error)."""
    nnz = int(Ap[-1])
    l = np.zeros(nnz, np.float64)
    bad = olib().orc_spilu0_csr(len(Ap) - 1, nnz, _p(_i32(Ap)), _p(_i32(Ai)), _p(_f64(Ax)),
                                _p(_i32(A_diag)), _p(l))
    return l, int(bad)


def sptrsv_csc_block(sup, nrows, Lp, Li, Lx, b):
    x = _f64(b).copy()
    olib().orc_sptrsv_csc_block(len(x), _p(_i32(Lp)), _p(_i32(Li)), _p(_i32(nrows)),
                                _p(_f64(Lx)), _p(x), len(sup) - 1, _p(_i32(sup)))
    return x


def sptrsv_csc_block_t(sup, nrows, Lp, Li, Lx, b):
    """L^T x = b on the supernodal storage (blockedLTsolve's semantics, this
    build's operation order: see orc_sptrsv_csc_block_t)."""
    x = _f64(b).copy()
    olib().orc_sptrsv_csc_block_t(len(x), _p(_i32(Lp)), _p(_i32(Li)), _p(_i32(nrows)),
                                  _p(_f64(Lx)), _p(x), len(sup) - 1, _p(_i32(sup)))
    return x


def levelset_csr(Lp, Li):
    n = len(Lp) - 1
    lp = np.zeros(n + 1, np.int32)
    ls = np.zeros(max(n, 1), np.int32)
    lev = np.zeros(max(n, 1), np.int32)
    nl = olib().orc_levelset_csr(n, _p(_i32(Lp)), _p(_i32(Li)), _p(lp), _p(ls), _p(lev))
    return lp[: nl + 1], ls[:n], lev[:n]


def levelset_csc(Cp, Ci):
    n = len(Cp) - 1
    lp = np.zeros(n + 1, np.int32)
    ls = np.zeros(max(n, 1), np.int32)
    nl = olib().orc_levelset_csc(n, _p(_i32(Cp)), _p(_i32(Ci)), _p(lp), _p(ls))
    return lp[: nl + 1], ls[:n]


def check_schedule(Lp, Li, level_ptr, par_ptr, partition) -> int:
    return olib().orc_check_schedule(len(Lp) - 1, _p(_i32(Lp)), _p(_i32(Li)),
                                     len(level_ptr) - 1, _p(_i32(level_ptr)),
                                     _p(_i32(par_ptr)), _p(_i32(partition)))


def csr_to_csc(Ap, Ai, Ax):
    n = len(Ap) - 1
    nnz = int(Ap[-1])
    Cp = np.zeros(n + 1, np.int32)
    Ci = np.zeros(max(nnz, 1), np.int32)
    Cx = np.zeros(max(nnz, 1), np.float64)
    mp = np.zeros(max(nnz, 1), np.int32)
    olib().orc_csr_to_csc(n, _p(_i32(Ap)), _p(_i32(Ai)), _p(_f64(Ax)), _p(Cp), _p(Ci),
                          _p(Cx), _p(mp))
    return Cp, Ci[:nnz], Cx[:nnz], mp[:nnz]


def adjoint(Lp, Li, Lx, x, gx):
    """(gb, gL): gb = L^-T gx (ltsolve on the CSC of L), gL[p] = -gb[i] x[j]."""
    n = len(Lp) - 1
    Cp, Ci, Cx, _ = csr_to_csc(Lp, Li, Lx)
    gb = np.empty(n, np.float64)
    gL = np.empty(len(Lx), np.float64)
    olib().orc_adjoint(n, _p(_i32(Lp)), _p(_i32(Li)), _p(Cp), _p(Ci), _p(Cx),
                       _p(_f64(x)), _p(_f64(gx)), _p(gb), _p(gL))
    return gb, gL


def adjoint_block(Lp, Li, Lx, X, GX):
    """(GB, gL) for row-major (n, k) blocks: GB = L^-T GX per column,
    gL[p] = -(pairwise sum over c of GB[i,c] X[j,c]) (pairwise_dot's order)."""
    n = len(Lp) - 1
    X = np.ascontiguousarray(X, np.float64)
    GX = np.ascontiguousarray(GX, np.float64)
    k = X.shape[1]
    Cp, Ci, Cx, _ = csr_to_csc(Lp, Li, Lx)
    GB = np.empty((n, k), np.float64)
    gL = np.empty(len(Lx), np.float64)
    olib().orc_adjoint_block(n, k, _p(_i32(Lp)), _p(_i32(Li)), _p(Cp), _p(Ci), _p(Cx),
                             _p(X), _p(GX), _p(GB), _p(gL))
    return GB, gL


def pairwise_dot(g, x):
    """The k products g[c]*x[c] added in the block adjoint's pairwise order:
    groups of G columns (pow2_ceil(k) for k <= 32, else 32) reduced by
    halving, group sums in ascending order."""
    g, x = _f64(g), _f64(x)
    return float(olib().orc_pairwise_dot(_p(g), _p(x), len(g)))


def prune_set(Lp, Li):
    """Row pattern of the strict lower part, per column c: the s < c with
    L[c, s] != 0, ascending (the prune set spic0_csc_left walks)."""
    n = len(Lp) - 1
    ptr = np.zeros(n + 1, np.int32)
    lens = np.diff(Lp) - 1
    ptr[1:] = np.cumsum(lens)
    keep = np.ones(len(Li), bool)
    keep[np.asarray(Lp[1:]) - 1] = False
    return ptr, _i32(np.asarray(Li)[keep])


# ------------------------------------------------------------ reference .so
class ref:
    """Thin ctypes calls onto the reference's own compiled functions."""

    @staticmethod
    def sptrsv_csr(Lp, Li, Lx, b):
        x = _f64(b).copy()
        rlib().ref_sptrsv_csr(len(Lp) - 1, _p(_i32(Lp).copy()), _p(_i32(Li).copy()),
                              _p(_f64(Lx).copy()), _p(x))
        return x

    @staticmethod
    def sptrsv_csr_levelset(Lp, Li, Lx, b, level_ptr, level_set):
        x = _f64(b).copy()
        rlib().ref_sptrsv_csr_levelset(len(Lp) - 1, _p(_i32(Lp)), _p(_i32(Li)),
                                       _p(_f64(Lx)), _p(x), len(level_ptr) - 1,
                                       _p(_i32(level_ptr)), _p(_i32(level_set)))
        return x

    @staticmethod
    def sptrsv_csr_lbc(Lp, Li, Lx, b, level_ptr, par_ptr, partition):
        x = _f64(b).copy()
        rlib().ref_sptrsv_csr_lbc(len(Lp) - 1, _p(_i32(Lp).copy()), _p(_i32(Li).copy()),
                                  _p(_f64(Lx).copy()), _p(x), len(level_ptr) - 1,
                                  _p(_i32(level_ptr).copy()), _p(_i32(par_ptr).copy()),
                                  _p(_i32(partition).copy()))
        return x

    @staticmethod
    def sptrsv_csr_lbc_redundant(Lp, Li, Lx, x0, b, level_ptr, par_ptr, partition):
        x = _f64(x0).copy()
        n = len(Lp) - 1
        tmp = np.zeros(max(1, rlib().ref_max_threads() * n))
        rlib().ref_sptrsv_csr_lbc_redundant(
            n, _p(_i32(Lp).copy()), _p(_i32(Li).copy()), _p(_f64(Lx).copy()), _p(x),
            _p(_f64(b).copy()), len(level_ptr) - 1, _p(_i32(level_ptr).copy()),
            _p(_i32(par_ptr).copy()), _p(_i32(partition).copy()), _p(tmp))
        return x

    @staticmethod
    def sptrsv_csr_ilu(Ap, Ai, Ax, diags, b):
        x = _f64(b).copy()
        rlib().ref_sptrsv_csr_ilu(len(Ap) - 1, _p(_i32(Ap)), _p(_i32(Ai)), _p(_f64(Ax)),
                                  _p(_i32(diags)), _p(x))
        return x

    @staticmethod
    def sptrsv_csr_ilu_lbc(Ap, Ai, Ax, diags, b, level_ptr, par_ptr, partition):
        x = _f64(b).copy()
        rlib().ref_sptrsv_csr_ilu_lbc(len(Ap) - 1, _p(_i32(Ap)), _p(_i32(Ai)),
                                      _p(_f64(Ax)), _p(_i32(diags)), _p(x),
                                      len(level_ptr) - 1, _p(_i32(level_ptr)),
                                      _p(_i32(par_ptr)), _p(_i32(partition)))
        return x

    @staticmethod
    def sptrsv_csc(Cp, Ci, Cx, b):
        x = _f64(b).copy()
        rlib().ref_sptrsv_csc(len(Cp) - 1, _p(_i32(Cp).copy()), _p(_i32(Ci).copy()),
                              _p(_f64(Cx).copy()), _p(x))
        return x

    @staticmethod
    def ltsolve(Cp, Ci, Cx, b):
        x = _f64(b).copy()
        rlib().ref_ltsolve(len(Cp) - 1, _p(_i32(Cp).copy()), _p(_i32(Ci).copy()),
                           _p(_f64(Cx).copy()), _p(x))
        return x

    @staticmethod
    def spmv_csr(Ap, Ai, Ax, x):
        y = np.empty(len(Ap) - 1, np.float64)
        rlib().ref_spmv_csr(len(Ap) - 1, _p(_i32(Ap)), _p(_i32(Ai)), _p(_f64(Ax)),
                            _p(_f64(x)), _p(y))
        return y

    @staticmethod
    def spmv_csr_axpby(Ap, Ai, Ax, x, y, z, a, b):
        z = _f64(z).copy()
        rlib().ref_spmv_csr_axpby(len(Ap) - 1, _p(_i32(Ap)), _p(_i32(Ai)), _p(_f64(Ax)),
                                  _p(_f64(x)), _p(_f64(y)), _p(z), a, b)
        return z

    @staticmethod
    def spmv_csr_block(Ap, Ai, Ax, supernodes, x, y):
        y = _f64(y).copy()
        rlib().ref_spmv_csr_block(len(Ap) - 1, _p(_i32(Ap)), _p(_i32(Ai)), _p(_f64(Ax)),
                                  len(supernodes) - 1, _p(_i32(supernodes)), _p(_f64(x)),
                                  _p(y))
        return y

    @staticmethod
    def spmv_csc(Cp, Ci, Cx, x, y):
        y = _f64(y).copy()
        rlib().ref_spmv_csc(len(Cp) - 1, _p(_i32(Cp)), _p(_i32(Ci)), _p(_f64(Cx)),
                            _p(_f64(x)), _p(y))
        return y

    @staticmethod
    def spmv_csc_axpby(Cp, Ci, Cx, x, y, z, a, b):
        z = _f64(z).copy()
        if np.isscalar(a):
            rlib().ref_spmv_csc_axpby(len(Cp) - 1, _p(_i32(Cp)), _p(_i32(Ci)), _p(_f64(Cx)),
                                      _p(_f64(x)), _p(_f64(y)), _p(z), float(a), float(b))
        else:
            rlib().ref_spmv_csc_axpby_vec(len(Cp) - 1, _p(_i32(Cp)), _p(_i32(Ci)),
                                          _p(_f64(Cx)), _p(_f64(x)), _p(_f64(y)), _p(z),
                                          _p(_f64(a).copy()), _p(_f64(b).copy()))
        return z

    @staticmethod
    def spmv_csc_lbc(sets, ids, Cp, Ci, Cx, x, y):
        y = _f64(y).copy()
        rlib().ref_spmv_csc_lbc(len(sets) - 1, _p(_i32(sets)), _p(_i32(ids)), len(Cp) - 1,
                                _p(_i32(Cp)), _p(_i32(Ci)), _p(_f64(Cx)), _p(_f64(x)), _p(y))
        return y

    @staticmethod
    def spmv_csc_block(Cp, Ci, Cx, x, y, supernodes):
        y = _f64(y).copy()
        rlib().ref_spmv_csc_block(len(Cp) - 1, _p(_i32(Cp)), _p(_i32(Ci)), _p(_f64(Cx)),
                                  _p(_f64(x)), _p(y), len(supernodes) - 1,
                                  _p(_i32(supernodes)))
        return y

    @staticmethod
    def spic0_csc_left(Ap, Ai, Ax, prunePtr, pruneSet):
        Lx = np.zeros(len(Ax), np.float64)
        ok = rlib().ref_spic0_csc_left(len(Ap) - 1, _p(_i32(Ap)), _p(_i32(Ai)),
                                       _p(_f64(Ax)), _p(_i32(prunePtr)),
                                       _p(_i32(pruneSet)), _p(Lx))
        return Lx, bool(ok)

    @staticmethod
    def sptrsv_csc_block(sup, nrows, Lp, Li, Lx, b):
        # the reference only reads Lx (lsolve_BLAS / matvec_BLAS), so it is
        # passed without a copy
        x = _f64(b).copy()
        rlib().ref_sptrsv_csc_block(len(x), _p(_i32(Lp)), _p(_i32(Li)), _p(_i32(nrows)),
                                    _p(np.ascontiguousarray(Lx, np.float64)), _p(x),
                                    len(sup) - 1, _p(_i32(sup)))
        return x

    @staticmethod
    def spic0_csr(Ap, Ai, Ax):
        Lx = _f64(Ax).copy()
        rlib().ref_spic0_csr(len(Ap) - 1, _p(Lx), _p(_i32(Ap).copy()), _p(_i32(Ai).copy()))
        return Lx

    @staticmethod
    def set_threads(n):
        rlib().ref_set_threads(int(n))

    @staticmethod
    def getLevelSet(parent):
        n = len(parent)
        lp = np.zeros(n + 1, np.int32)
        ls = np.zeros(max(n, 1), np.int32)
        nl = rlib().ref_getLevelSet(n, _p(_i32(parent)), _p(lp), _p(ls))
        return lp[: nl + 1], ls[:n]

    @staticmethod
    def dense_kernels(M, ldm, ncol, nrow, rhs, vec, out, parsy):
        """The dense kernels of the supernodal solves on one column-major
        block M (leading dimension ldm): the triangular solve of its top
        ncol x ncol part on a copy of rhs, and the product of the nrow x ncol
        part at M[ncol:] with vec into a copy of out.  parsy=False runs
        sym_lib's lsolve_BLAS/matvec_BLAS (dense_blas/BLAS.cpp:3, :286),
        parsy=True parsy's dlsolve_blas_nonUnit/dmatvec_blas
        (parsy/triangularSolve/BLAS.cpp:12, :123)."""
        M = _f64(M).copy()
        r = _f64(rhs).copy()
        o = _f64(out).copy()
        v = _f64(vec).copy()
        L = rlib()
        if parsy:
            L.ref_parsy_dlsolve(ldm, ncol, _p(M), _p(r))
            L.ref_parsy_dmatvec(ldm, nrow, ncol, _p(M[ncol:]), _p(v), _p(o))
        else:
            L.ref_lsolve_BLAS(ldm, ncol, _p(M), _p(r))
            L.ref_matvec_BLAS(ldm, nrow, ncol, _p(M[ncol:]), _p(v), _p(o))
        return r, o

    @staticmethod
    def getCoarseLevelSet_6(parent, cost, inner_parts, min_level_dist, div_rate):
        """(level_ptr, part_ptr, partition) of the reference's tree-level
        coarsening (parsy/symbolic/InspectionLevel_06.cpp:17).  The reference
        build keeps its assertions, one of which compares an integer sum of
        the costs with the float one: give integer-valued costs."""
        n = len(parent)
        lp = np.zeros(n + 1, np.int32)
        pp = np.zeros(n + 1, np.int32)
        pt = np.zeros(max(n, 1), np.int32)
        nl, npart = ctypes.c_int(0), ctypes.c_int(0)
        c = _f64(cost).copy()
        rlib().ref_getCoarseLevelSet_6(n, _p(_i32(parent)), _p(c), int(inner_parts),
                                       int(min_level_dist), int(div_rate), _p(lp),
                                       _p(pp), _p(pt), ctypes.byref(nl),
                                       ctypes.byref(npart))
        return lp[: nl.value + 1], pp[: npart.value + 1], pt[:n]
