"""sympiler_amd -- MI355X (gfx950) sparse lower-triangular solve.

Host-side Python mirror of the C ABI in include/sympiler_amd.h, loaded with
ctypes from the in-tree libsympiler_amd.so.  Device arrays are torch CUDA
tensors (ROCm); their data pointers and torch's current HIP stream are what
cross the boundary.  There is no CPU fallback: if the HIP library is missing
every entry point raises.

Reference mapping (reference include/sympiler/sparse_blas_lib.h):
  Plan.solve(k=1)        sym_lib::sptrsv_csr / sptrsv_csr_levelset
                         (sparse_blas/sptrsv.cpp:17-25, :40-54)
  Plan.solve(k>1)        k independent sptrsv_csr solves, one per column
  Plan.adjoint           sym_lib::ltsolve (sparse_blas/sptrsv_lt.cpp:11-20)
                         + grad over L's values
  levelset()             build_levelSet_CSC (aggregation submodule, absent)
  coarsen()              get_coarse_levelSet_DAG_CSC_tree (absent)
  spmv_csr()             sym_lib::spmv_csr (sparse_blas/spmv.cpp:12-34)
  solve_lbc()            sym_lib::sptrsv_csr_lbc (sparse_blas/sptrsv.cpp:57-75)
  solve_lbc_redundant()  sym_lib::sptrsv_csr_lbc_redundant (sptrsv.cpp:78-106)
  ilu_lower()            the lower factor read by sym_lib::sptrsv_csr_ilu /
                         sptrsv_csr_ilu_lbc (sptrsv.cpp:28-38, :109-126)
  Plan.ic0()             IC(0) on the plan's pattern (spic0.cpp:11-46,
                         :106-127)
  ilu0(), Ilu0Plan       sym_lib::spilu0_csr (spilu0.cpp:11-51)
  SupernodalPlan         sym_lib::sptrsv_csc_block (sptrsv.cpp:176-199)
"""
from __future__ import annotations

import atexit
import ctypes
import os
import weakref
from ctypes import (POINTER, byref, c_char_p, c_double, c_int, c_int64,
                    c_void_p)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsympiler_amd.so")
GEN_PATH = os.path.join(_HERE, "libsymamd_gen.so")

F64, F32 = 0, 1
COL_MAJOR, ROW_MAJOR = 0, 1
EXEC = {"auto": 0, "levelset": 1, "syncfree": 2, "graph": 3, "regions": 4, "wide": 5,
        "split": 6}
STATUS = {0: "OK", 1: "INVALID_ARG", 2: "STRUCTURE", 3: "ZERO_PIVOT", 4: "HIP",
          5: "ALLOC", 6: "TIMEOUT", 7: "NOT_SUPPORTED", 8: "NOT_SPD"}


class SymAmdError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None
# Native objects still alive; destroyed explicitly at interpreter exit (in
# creation-reverse order: graphs, plans, then handles), before module
# teardown, so no device memory is freed from a finalizer while the runtime
# and the profiler tools that wrap it are shutting down.
_live = {"graph": weakref.WeakSet(), "plan": weakref.WeakSet(), "handle": weakref.WeakSet()}
_closing = False


def _release_all():
    global _closing
    if _lib is not None:
        for kind in ("graph", "plan", "handle"):
            for obj in list(_live[kind]):
                try:
                    obj.close()
                except Exception:
                    pass
    _closing = True


atexit.register(_release_all)


def lib() -> ctypes.CDLL:
    """The loaded HIP library; raises if it was never built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C sympiler_amd/csrc` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        vp, i32, i64, dbl = c_void_p, c_int, c_int64, c_double
        sig = {
            "symamd_create": [POINTER(vp), vp],
            "symamd_destroy": [vp],
            "symamd_set_stream": [vp, vp],
            "symamd_set_timeout_ms": [vp, dbl],
            "symamd_set_option": [vp, i32, i64],
            "symamd_levelset": [vp, i32, vp, vp, vp, vp, POINTER(i32)],
            "symamd_coarsen": [vp, i32, vp, vp, vp, i32, i32, i32, vp, vp, vp,
                               POINTER(i32), POINTER(i32)],
            "symamd_tree_coarsen": [i32, vp, vp, i32, i32, i32, vp, vp, vp,
                                    POINTER(i32), POINTER(i32)],
            "symamd_solve_lbc": [vp, i32, vp, vp, vp, vp, i32, vp, vp, vp],
            "symamd_solve_lbc_redundant": [vp, i32, vp, vp, vp, vp, vp, i32, vp, vp, vp],
            "symamd_ilu_lower": [vp, i32, vp, vp, vp, vp, i32, vp, vp, vp,
                                 POINTER(i32)],
            "symamd_analyze": [vp, i32, vp, vp, vp, i32, POINTER(vp)],
            "symamd_analyze_batched": [vp, i32, vp, vp, vp, vp, i32, POINTER(vp)],
            "symamd_plan_update_values": [vp, vp, vp],
            "symamd_plan_info": [vp, POINTER(i32), POINTER(i64), POINTER(i32),
                                 POINTER(i32)],
            "symamd_plan_levels": [vp, vp, vp],
            "symamd_plan_regions": [vp, POINTER(i32), POINTER(dbl)],
            "symamd_region_schedule": [vp, vp, i32, i32, POINTER(i32), POINTER(i32)],
            "symamd_choice": [vp, vp, i32, i32, POINTER(i32), POINTER(i32)],
            "symamd_plan_destroy": [vp],
            "symamd_plan_set_lanes": [vp, i32],
            "symamd_solve": [vp, vp, vp, i32, i64, i32, i32],
            "symamd_plan_check": [vp, vp],
            "symamd_graph_create": [vp, vp, vp, i32, i64, i32, i32, POINTER(vp)],
            "symamd_graph_launch": [vp, vp],
            "symamd_graph_destroy": [vp],
            "symamd_adjoint": [vp, vp, vp, vp, vp, vp, i32],
            "symamd_adjoint_block": [vp, vp, vp, vp, vp, vp, i32, i64, i32, i32],
            "symamd_spmv_csr": [vp, i32, vp, vp, vp, vp, vp],
            "symamd_spmv_csr_axpby": [vp, i32, vp, vp, vp, vp, vp, vp, dbl, dbl],
            "symamd_residual": [vp, vp, vp, vp, i32, i64, i32, vp],
            "symamd_tune": [vp, vp, i32, i32, c_char_p, i32],
            "symamd_ic0": [vp, vp, i32, vp, vp, POINTER(i32)],
            "symamd_ilu0": [vp, i32, vp, vp, vp, vp, vp, POINTER(i32)],
            "symamd_ilu0_analyze": [vp, i32, vp, vp, vp, POINTER(vp)],
            "symamd_ilu0_factor": [vp, vp, vp, vp, POINTER(i32)],
            "symamd_ilu0_destroy": [vp],
            "symamd_sn_analyze": [vp, i32, i32, vp, vp, vp, vp, POINTER(vp)],
            "symamd_sn_solve": [vp, vp, vp, vp],
            "symamd_sn_solve_block": [vp, vp, vp, vp, i32, i64],
            "symamd_sn_solve_transposed": [vp, vp, vp, vp, i32, i64],
            "symamd_sn_check": [vp, vp],
            "symamd_sn_destroy": [vp],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = c_int
        L.symamd_last_error.restype = c_char_p
        L.symamd_last_error.argtypes = []
        L.symamd_version.restype = c_char_p
        L.symamd_version.argtypes = []
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        raise SymAmdError(rc, lib().symamd_last_error().decode())


def _ptr(t) -> c_void_p:
    return c_void_p(0 if t is None else t.data_ptr())


def _require_cuda(*tensors):
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise ValueError("sympiler_amd expects device (cuda) tensors")
        if t is not None and not t.is_contiguous():
            raise ValueError("sympiler_amd expects contiguous tensors")


def _dtype_code(dt) -> int:
    import torch
    if dt == torch.float64:
        return F64
    if dt == torch.float32:
        return F32
    raise ValueError(f"unsupported dtype {dt}")


class Handle:
    """A library handle bound to a HIP stream (torch's current stream when
    `stream` is None)."""

    def __init__(self, stream=None, timeout_ms: float | None = None):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream()
        self._h = c_void_p()
        _check(lib().symamd_create(byref(self._h), c_void_p(stream.cuda_stream)))
        self.stream = stream
        _live["handle"].add(self)
        if timeout_ms is not None:
            _check(lib().symamd_set_timeout_ms(self._h, float(timeout_ms)))

    def set_tickets(self, on: bool):
        """Persistent kernels take units from a ticket counter (True) or are
        dealt round-robin from a cooperative launch (False, default)."""
        _check(lib().symamd_set_option(self._h, 1, int(bool(on))))

    def set_colmajor_direct(self, on: bool):
        """Solve column-major multi-RHS blocks in place (True) instead of
        through the plan's level-ordered workspace (False, default)."""
        _check(lib().symamd_set_option(self._h, 5, int(bool(on))))

    def set_marks(self, on: bool):
        """Marked sync-free protocol (True, default): ready tags are
        published without waiting for the value stores; readers re-load
        values still carrying the mark."""
        _check(lib().symamd_set_option(self._h, 6, int(bool(on))))

    def set_max_waves(self, waves: int):
        """Cap the wavefronts of persistent launches (0: fill the GPU)."""
        _check(lib().symamd_set_option(self._h, 4, int(waves)))

    def set_regions(self, mode: int, max_rows: int | None = None):
        """Region order for plans analysed with this handle: 0 off (default),
        1 when at least half of L's entries fall inside regions, 2 always;
        `max_rows` caps the region size (default 512)."""
        _check(lib().symamd_set_option(self._h, 7, int(mode)))
        if max_rows is not None:
            _check(lib().symamd_set_option(self._h, 8, int(max_rows)))

    def set_region_cap(self, rows: int):
        """Region size cap of the region executor (0: chosen from k)."""
        _check(lib().symamd_set_option(self._h, 10, int(rows)))

    def set_region_waves(self, waves: int):
        """Wavefronts per task of the region executor: 1, 2 or 4 for tasks
        of eight columns (0: chosen from k)."""
        _check(lib().symamd_set_option(self._h, 11, int(waves)))

    def set_region_band(self, slices: int):
        """Slices per published band of the region executor (0: default)."""
        _check(lib().symamd_set_option(self._h, 12, int(slices)))

    def set_region_mw(self, on: bool):
        """Region tasks with a separate memory wavefront (True, default)."""
        _check(lib().symamd_set_option(self._h, 13, int(bool(on))))

    def set_split_long(self, entries: int):
        """Off-diagonal entries above which the split level executor gives a
        row a whole workgroup (0: by column-tile width)."""
        _check(lib().symamd_set_option(self._h, 14, int(entries)))

    def set_owners(self, on: bool):
        """Run plans with regions on the owner schedule (True, default) or
        deal their units like other plans (False)."""
        _check(lib().symamd_set_option(self._h, 9, int(bool(on))))

    def set_trace(self, buf):
        """Record per-unit clock stamps of sync-free launches into `buf`
        (a uint64/int64 cuda tensor of 3 entries per unit), or stop (None)."""
        if buf is None:
            _check(lib().symamd_set_option(self._h, 2, 0))
            _check(lib().symamd_set_option(self._h, 3, 0))
            return
        _require_cuda(buf)
        if buf.element_size() != 8:
            raise ValueError("trace buffer needs 64-bit entries")
        _check(lib().symamd_set_option(self._h, 2, buf.data_ptr()))
        _check(lib().symamd_set_option(self._h, 3, buf.numel()))

    def close(self):
        """Destroy the native handle (idempotent)."""
        if getattr(self, "_h", None) and _lib is not None:
            _lib.symamd_destroy(self._h)
        self._h = None

    def __del__(self):
        if not _closing:
            self.close()


_default_handle = None


def default_handle() -> Handle:
    global _default_handle
    if _default_handle is None:
        _default_handle = Handle()
    return _default_handle


class Plan:
    """Analysed lower-triangular L (CSR, diagonal last), resident on the GPU.

    Lp, Li: int32 cuda tensors; Lx: float64 or float32 cuda tensor.
    If `row_offsets` is given, the rows form independent systems with column
    ids local to each system (symamd_analyze_batched).
    """

    def __init__(self, Lp, Li, Lx, handle: Handle | None = None,
                 row_offsets=None):
        import torch
        _require_cuda(Lp, Li, Lx)
        if Lp.dtype != torch.int32 or Li.dtype != torch.int32:
            raise ValueError("Lp and Li must be int32")
        self.h = handle or default_handle()
        self.n = Lp.numel() - 1
        self.dtype = Lx.dtype
        self._p = c_void_p()
        if row_offsets is None:
            _check(lib().symamd_analyze(self.h._h, self.n, _ptr(Lp), _ptr(Li),
                                        _ptr(Lx), _dtype_code(Lx.dtype),
                                        byref(self._p)))
            self.nsys = 1
        else:
            off = (c_int * len(row_offsets))(*[int(v) for v in row_offsets])
            self.nsys = len(row_offsets) - 1
            _check(lib().symamd_analyze_batched(self.h._h, self.nsys, off,
                                                _ptr(Lp), _ptr(Li), _ptr(Lx),
                                                _dtype_code(Lx.dtype),
                                                byref(self._p)))
        _live["plan"].add(self)
        n, nnz, lev, mx = c_int(), c_int64(), c_int(), c_int()
        _check(lib().symamd_plan_info(self._p, byref(n), byref(nnz), byref(lev),
                                      byref(mx)))
        self.nnz, self.levels, self.max_level_rows = nnz.value, lev.value, mx.value
        self.values_generation = 0  # incremented by every update_values

    def close(self):
        """Destroy the native plan (idempotent)."""
        if getattr(self, "_p", None) and _lib is not None:
            _lib.symamd_plan_destroy(self._p)
        self._p = None

    def __del__(self):
        if not _closing:
            self.close()


    # ----------------------------------------------------------- inspection
    def level_sets(self):
        """(level_ptr, level_set) as numpy int32 arrays."""
        import numpy as np
        lp = np.zeros(self.levels + 1, dtype=np.int32)
        ls = np.zeros(max(self.n, 1), dtype=np.int32)
        _check(lib().symamd_plan_levels(self._p, lp.ctypes.data_as(c_void_p),
                                        ls.ctypes.data_as(c_void_p)))
        return lp, ls[: self.n]

    def choice(self, k=1, layout="col"):
        """(executor name, lanes per row) the automatic choice takes for a
        forward solve of k right-hand sides in `layout`."""
        ex, lanes = c_int(), c_int()
        lay = ROW_MAJOR if layout == "row" else COL_MAJOR
        _check(lib().symamd_choice(self.h._h, self._p, int(k), lay, byref(ex), byref(lanes)))
        return {v: n for n, v in EXEC.items()}[ex.value], lanes.value

    def regions(self):
        """(number of regions, fraction of L's entries inside a region);
        0 regions when the plan has no region order."""
        cnt, inside = c_int(), c_double()
        _check(lib().symamd_plan_regions(self._p, byref(cnt), byref(inside)))
        return cnt.value, inside.value

    def region_schedule(self, k=1, transposed=False):
        """(region cap, tasks) of the schedule the region executor runs for
        a forward (or, transposed, adjoint) solve of k right-hand sides;
        (0, 0) when such a solve runs sync-free."""
        cap, tasks = c_int(), c_int()
        _check(lib().symamd_region_schedule(self.h._h, self._p, int(k), int(bool(transposed)),
                                            byref(cap), byref(tasks)))
        return cap.value, tasks.value

    def set_lanes(self, lanes: int):
        _check(lib().symamd_plan_set_lanes(self._p, int(lanes)))

    def update_values(self, Lx):
        _require_cuda(Lx)
        if Lx.dtype != self.dtype:
            raise ValueError("value dtype differs from the plan's")
        _check(lib().symamd_plan_update_values(self.h._h, self._p, _ptr(Lx)))
        self.values_generation += 1

    def ic0(self, Ax, variant="left", out=None, strict=True):
        """Zero-fill incomplete Cholesky of the SPD matrix whose lower triangle
        Ax holds on this plan's pattern (FP64, the caller's CSR order); returns
        L on the same pattern, bit-identical to sym_lib::spic0_csc_left
        (variant "left", sparse_blas/spic0.cpp:11-46) or sym_lib::spic0_csr
        ("csr", :106-127).  out may be Ax (in place).  A pivot that fails the
        variant's test raises ValueError when strict, else the factor is
        returned with the failed row: (L, row or -1)."""
        import torch
        _require_cuda(Ax)
        if Ax.dtype != torch.float64 or self.dtype != torch.float64:
            raise ValueError("IC(0) needs an FP64 plan and values")
        var = {"left": 0, "csr": 1}[variant]
        if out is None:
            out = torch.empty_like(Ax)
        _require_cuda(out)
        bad = c_int(-1)
        rc = lib().symamd_ic0(self.h._h, self._p, var, _ptr(Ax), _ptr(out), byref(bad))
        if rc == 8 and not strict:
            return out, int(bad.value)
        if rc == 8:
            raise ValueError(f"IC(0) breakdown: pivot of row {bad.value} is not positive")
        _check(rc)
        return out if strict else (out, -1)

    # ---------------------------------------------------------------- solve
    @staticmethod
    def _layout(B, layout):
        if B.dim() == 1:
            return 1, B.numel(), COL_MAJOR
        if layout in ("col", COL_MAJOR):
            # torch (k, n) contiguous == column-major n x k with ld = n
            return B.shape[0], B.shape[1], COL_MAJOR
        if layout in ("row", ROW_MAJOR):
            return B.shape[1], B.shape[1], ROW_MAJOR
        raise ValueError(f"unknown layout {layout}")

    def solve(self, B, layout="col", executor="auto"):
        """Solve L X = B in place.  B: (n,) vector; or (k, n) for
        column-major right-hand sides (layout='col'); or (n, k) for
        row-major (layout='row')."""
        _require_cuda(B)
        if B.dtype != self.dtype:
            raise ValueError("right-hand side dtype differs from the plan's")
        k, ld, lay = self._layout(B, layout)
        _check(lib().symamd_solve(self.h._h, self._p, _ptr(B), k, ld, lay,
                                  EXEC[executor]))
        return B

    def check(self):
        """Synchronise and raise if a sync-free watchdog expired."""
        _check(lib().symamd_plan_check(self.h._h, self._p))

    def graph(self, B, layout="col", executor="auto") -> "Graph":
        _require_cuda(B)
        k, ld, lay = self._layout(B, layout)
        g = c_void_p()
        _check(lib().symamd_graph_create(self.h._h, self._p, _ptr(B), k, ld,
                                         lay, EXEC[executor], byref(g)))
        return Graph(self, g, B)

    def adjoint(self, x, gx, gb=None, gLx=None, executor="auto"):
        """gb = L^-T gx and gLx[p] = -gb[i] * x[j] over L's pattern."""
        import torch
        _require_cuda(x, gx, gb, gLx)
        if gb is None:
            gb = torch.empty_like(gx)
        _check(lib().symamd_adjoint(self.h._h, self._p, _ptr(x), _ptr(gx),
                                    _ptr(gb), _ptr(gLx), EXEC[executor]))
        return gb, gLx

    def adjoint_block(self, X, GX, GB=None, gLx=None, layout="row", executor="auto"):
        """Block adjoint of X = L^-1 B: GB = L^-T GX column by column and
        gLx[p] = -(sum over columns c, ascending, of GB[i,c] X[j,c]).  X, GX
        and GB are 2-D tensors in `layout` (as for solve)."""
        import torch
        _require_cuda(X, GX, GB, gLx)
        if GB is None:
            GB = torch.empty_like(GX)
        k, ld, lay = self._layout(GX, layout)
        for t in (X, GB):
            if self._layout(t, layout) != (k, ld, lay):
                raise ValueError("X, GX and GB must share shape and layout")
        _check(lib().symamd_adjoint_block(self.h._h, self._p, _ptr(X), _ptr(GX), _ptr(GB),
                                          _ptr(gLx), int(k), int(ld), lay, EXEC[executor]))
        return GB, gLx

    def residual(self, X, B, layout="col"):
        """Relative residual ||L X - B|| / ||B|| per column (list of floats)."""
        import numpy as np
        _require_cuda(X, B)
        k, ld, lay = self._layout(X, layout)
        out = np.zeros(k, dtype=np.float64)
        _check(lib().symamd_residual(self.h._h, self._p, _ptr(X), _ptr(B), k,
                                     ld, lay, out.ctypes.data_as(c_void_p)))
        return out

    def tune(self, k=1, layout="col") -> str:
        buf = ctypes.create_string_buffer(4096)
        lay = COL_MAJOR if layout in ("col", COL_MAJOR) else ROW_MAJOR
        _check(lib().symamd_tune(self.h._h, self._p, int(k), lay, buf, 4096))
        return buf.value.decode()


class Graph:
    """A captured solve (hipGraph) replayed on the handle's stream."""

    def __init__(self, plan: Plan, g, B):
        self.plan, self._g, self._B = plan, g, B
        _live["graph"].add(self)

    def launch(self):
        _check(lib().symamd_graph_launch(self.plan.h._h, self._g))

    def close(self):
        """Destroy the native graph (idempotent)."""
        if getattr(self, "_g", None) and _lib is not None:
            _lib.symamd_graph_destroy(self._g)
        self._g = None

    def __del__(self):
        if not _closing:
            self.close()


# ------------------------------------------------------------ free functions
def levelset(Lp, Li, handle: Handle | None = None):
    """GPU level sets of a device CSR lower triangle -> (level_ptr, level_set)
    int32 cuda tensors."""
    import torch
    _require_cuda(Lp, Li)
    h = handle or default_handle()
    n = Lp.numel() - 1
    lp = torch.empty(n + 1, dtype=torch.int32, device=Lp.device)
    ls = torch.empty(max(n, 1), dtype=torch.int32, device=Lp.device)
    lev = c_int()
    _check(lib().symamd_levelset(h._h, n, _ptr(Lp), _ptr(Li), _ptr(lp), _ptr(ls),
                                 byref(lev)))
    return lp[: lev.value + 1], ls[:n]


def coarsen(Lp, Li, inner_parts=6, min_level_dist=-1, div_rate=4000, cost=None,
            handle: Handle | None = None):
    """Coarsened (level_ptr, par_ptr, partition) schedule built on the GPU,
    int32 cuda tensors; the parameters of get_coarse_levelSet_DAG_CSC_tree
    (demo/sptrsv_demo_utils.h:83-88; defaults are the demo's,
    demo/sptrsv_demo.cpp:37).  cost: per-row float64 cuda tensor (None: row
    nnz)."""
    import torch
    _require_cuda(Lp, Li, cost)
    h = handle or default_handle()
    n = Lp.numel() - 1
    lp = torch.empty(n + 1, dtype=torch.int32, device=Lp.device)
    pp = torch.empty(n + 1, dtype=torch.int32, device=Lp.device)
    part = torch.empty(max(n, 1), dtype=torch.int32, device=Lp.device)
    lno, pno = c_int(), c_int()
    _check(lib().symamd_coarsen(h._h, n, _ptr(Lp), _ptr(Li), _ptr(cost), int(inner_parts),
                                int(min_level_dist), int(div_rate), _ptr(lp), _ptr(pp),
                                _ptr(part), byref(lno), byref(pno)))
    return lp[: lno.value + 1], pp[: pno.value + 1], part[:n]


def tree_coarsen(parent, node_cost, inner_parts, min_level_dist, div_rate):
    """Tree-level coarsening of an elimination tree (host arrays, no device
    work): parent[i] < 0 marks a root.  Returns numpy int32 arrays
    (level_ptr, part_ptr, partition) equal to those of
    sym_lib::parsy::getCoarseLevelSet_6 (parsy/symbolic/InspectionLevel_06.cpp:17)."""
    import numpy as np
    parent = np.ascontiguousarray(parent, dtype=np.int32)
    cost = np.ascontiguousarray(node_cost, dtype=np.float64)
    n = parent.size
    if cost.size != n:
        raise ValueError("node_cost must have one entry per node")
    lp = np.zeros(n + 1, np.int32)
    pp = np.zeros(n + 1, np.int32)
    part = np.zeros(max(n, 1), np.int32)
    lno, pno = c_int(), c_int()
    _check(lib().symamd_tree_coarsen(
        n, parent.ctypes.data_as(c_void_p), cost.ctypes.data_as(c_void_p),
        int(inner_parts), int(min_level_dist), int(div_rate),
        lp.ctypes.data_as(c_void_p), pp.ctypes.data_as(c_void_p),
        part.ctypes.data_as(c_void_p), byref(lno), byref(pno)))
    return lp[: lno.value + 1], pp[: pno.value + 1], part[:n]


def solve_lbc(Lp, Li, Lx, x, level_ptr, par_ptr, partition,
              handle: Handle | None = None):
    """Execute a caller schedule in place on x (FP64)."""
    import numpy as np
    _require_cuda(Lp, Li, Lx, x, par_ptr, partition)
    h = handle or default_handle()
    lp_host = np.ascontiguousarray(level_ptr.cpu().numpy() if hasattr(level_ptr, "cpu")
                                   else level_ptr, dtype=np.int32)
    _check(lib().symamd_solve_lbc(h._h, Lp.numel() - 1, _ptr(Lp), _ptr(Li),
                                  _ptr(Lx), _ptr(x), len(lp_host) - 1,
                                  lp_host.ctypes.data_as(c_void_p), _ptr(par_ptr),
                                  _ptr(partition)))
    return x


def solve_lbc_redundant(Lp, Li, Lx, x, b, level_ptr, par_ptr, partition,
                        handle: Handle | None = None):
    """Execute a caller schedule with right-hand side b, writing x, with the
    arithmetic of sptrsv_csr_lbc_redundant (products from zero, b added
    last); FP64."""
    import numpy as np
    _require_cuda(Lp, Li, Lx, x, b, par_ptr, partition)
    h = handle or default_handle()
    lp_host = np.ascontiguousarray(level_ptr.cpu().numpy() if hasattr(level_ptr, "cpu")
                                   else level_ptr, dtype=np.int32)
    _check(lib().symamd_solve_lbc_redundant(h._h, Lp.numel() - 1, _ptr(Lp), _ptr(Li),
                                            _ptr(Lx), _ptr(x), _ptr(b), len(lp_host) - 1,
                                            lp_host.ctypes.data_as(c_void_p),
                                            _ptr(par_ptr), _ptr(partition)))
    return x


def ilu_lower(Ap, Ai, Ax, diags, unit_diag: bool, handle: Handle | None = None):
    """The lower factor inside a general CSR matrix A with diagonal positions
    `diags`, as a CSR lower triangle (diagonal last): (Lp, Li, Lx).  With
    unit_diag the diagonal is 1.0 (sptrsv_csr_ilu does not divide), else
    A's value at diags[i] (sptrsv_csr_ilu_lbc divides by it)."""
    import torch
    _require_cuda(Ap, Ai, Ax, diags)
    h = handle or default_handle()
    n = Ap.numel() - 1
    cap = max(int(Ap[-1].item()) if n > 0 else 0, 1)
    Lp = torch.empty(n + 1, dtype=torch.int32, device=Ap.device)
    Li = torch.empty(cap, dtype=torch.int32, device=Ap.device)
    Lx = torch.empty(cap, dtype=torch.float64, device=Ap.device)
    nnz = c_int()
    _check(lib().symamd_ilu_lower(h._h, n, _ptr(Ap), _ptr(Ai), _ptr(Ax), _ptr(diags),
                                  1 if unit_diag else 0, _ptr(Lp), _ptr(Li), _ptr(Lx),
                                  byref(nnz)))
    return Lp, Li[: nnz.value], Lx[: nnz.value]


def ilu0(Ap, Ai, Ax, diags, out=None, strict=True, handle: Handle | None = None):
    """ILU(0) of a general CSR matrix A (diags[i] = position of row i's
    diagonal) on A's pattern: L strictly below the diagonal (unit diagonal
    implied), U on and above, bit-identical to sym_lib::spilu0_csr
    (sparse_blas/spilu0.cpp:11-51).  out may be Ax (in place).  A zero U
    diagonal raises ValueError when strict, else (l, row or -1) is
    returned."""
    import torch
    _require_cuda(Ap, Ai, Ax, diags)
    h = handle or default_handle()
    n = Ap.numel() - 1
    if out is None:
        out = torch.empty_like(Ax)
    _require_cuda(out)
    bad = c_int(-1)
    rc = lib().symamd_ilu0(h._h, n, _ptr(Ap), _ptr(Ai), _ptr(Ax), _ptr(diags), _ptr(out),
                           byref(bad))
    if rc == 3 and not strict:
        return out, int(bad.value)
    if rc == 3:
        raise ValueError(f"ILU(0) breakdown: U[{bad.value},{bad.value}] is zero")
    _check(rc)
    return out if strict else (out, -1)


class Ilu0Plan:
    """ILU(0) of one pattern, factored many times: the structure check, the
    schedule and the update lists are built once (symamd_ilu0_analyze), each
    factor() is the numeric phase only (symamd_ilu0_factor); results as
    ilu0()."""

    def __init__(self, Ap, Ai, diags, handle: Handle | None = None):
        _require_cuda(Ap, Ai, diags)
        self.h = handle or default_handle()
        self.n = Ap.numel() - 1
        self._p = c_void_p()
        _check(lib().symamd_ilu0_analyze(self.h._h, self.n, _ptr(Ap), _ptr(Ai), _ptr(diags),
                                         byref(self._p)))
        _live["plan"].add(self)

    def factor(self, Ax, out=None, strict=True):
        import torch
        _require_cuda(Ax)
        if out is None:
            out = torch.empty_like(Ax)
        _require_cuda(out)
        bad = c_int(-1)
        rc = lib().symamd_ilu0_factor(self.h._h, self._p, _ptr(Ax), _ptr(out), byref(bad))
        if rc == 3 and not strict:
            return out, int(bad.value)
        if rc == 3:
            raise ValueError(f"ILU(0) breakdown: U[{bad.value},{bad.value}] is zero")
        _check(rc)
        return out if strict else (out, -1)

    def close(self):
        if getattr(self, "_p", None) and _lib is not None:
            _lib.symamd_ilu0_destroy(self._p)
        self._p = None

    def __del__(self):
        if not _closing:
            self.close()


class SupernodalPlan:
    """Forward solve on a supernodal (BCSC) lower triangle, bit-identical to
    sym_lib::sptrsv_csc_block (sparse_blas/sptrsv.cpp:176-199).  The
    structure (numpy int32: supernodes, nrows, Lp, Li) is analysed once;
    solve(Lx, x) overwrites the device vector x (b in, x out)."""

    def __init__(self, supernodes, nrows, Lp, Li, handle: Handle | None = None):
        import numpy as np
        self.h = handle or default_handle()
        arrs = [np.ascontiguousarray(a, np.int32) for a in (supernodes, nrows, Lp, Li)]
        self._keep = arrs
        m = len(arrs[0]) - 1
        self.n = int(arrs[0][-1]) if m >= 0 else 0
        self._p = c_void_p()
        _check(lib().symamd_sn_analyze(self.h._h, self.n, m,
                                       *(a.ctypes.data_as(c_void_p) for a in arrs),
                                       byref(self._p)))
        _live["plan"].add(self)

    def solve(self, Lx, x):
        """x: (n,) or a column-major block (k, n) (row c = column c of X),
        overwritten with the solution."""
        _require_cuda(Lx)
        if x.dim() == 1:
            _require_cuda(x)
            _check(lib().symamd_sn_solve(self.h._h, self._p, _ptr(Lx), _ptr(x)))
        else:
            if not x.is_cuda or x.stride(1) != 1 or x.dtype != Lx.dtype:
                raise ValueError("x: a (k, n) device block with unit stride along n")
            k, n = x.shape
            _check(lib().symamd_sn_solve_block(self.h._h, self._p, _ptr(Lx), _ptr(x), k,
                                               x.stride(0)))
        return x

    def solve_transposed(self, Lx, x):
        """L^T X = B in place: x is (n,) or a column-major block (k, n)
        (parsy's blockedLTsolve semantics, Triangular_BCSC.cpp:125-186)."""
        _require_cuda(Lx)
        if not x.is_cuda or x.stride(-1) != 1 or x.dtype != Lx.dtype:
            raise ValueError("x: a device vector or (k, n) block with unit stride along n")
        k, ld = (1, x.numel()) if x.dim() == 1 else (x.shape[0], x.stride(0))
        _check(lib().symamd_sn_solve_transposed(self.h._h, self._p, _ptr(Lx), _ptr(x), k, ld))
        return x

    def check(self):
        _check(lib().symamd_sn_check(self.h._h, self._p))

    def close(self):
        if getattr(self, "_p", None) and _lib is not None:
            _lib.symamd_sn_destroy(self._p)
        self._p = None

    def __del__(self):
        if not _closing:
            self.close()


def spmv_csr(Ap, Ai, Ax, x, y=None, z=None, a=None, b=None,
             handle: Handle | None = None):
    """y = A x, or (with y, z, a, b) z = a * (z + A x) + b * y; FP64."""
    import torch
    h = handle or default_handle()
    n = Ap.numel() - 1
    if z is None:
        out = torch.empty(n, dtype=torch.float64, device=Ap.device)
        _require_cuda(Ap, Ai, Ax, x)
        _check(lib().symamd_spmv_csr(h._h, n, _ptr(Ap), _ptr(Ai), _ptr(Ax),
                                     _ptr(x), _ptr(out)))
        return out
    _require_cuda(Ap, Ai, Ax, x, y, z)
    _check(lib().symamd_spmv_csr_axpby(h._h, n, _ptr(Ap), _ptr(Ai), _ptr(Ax),
                                       _ptr(x), _ptr(y), _ptr(z), float(a),
                                       float(b)))
    return z


def version() -> str:
    return lib().symamd_version().decode()
