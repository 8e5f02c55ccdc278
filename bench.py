"""Benchmark: sparse lower-triangular solve on one MI355X.

Workload (default): the IC(0) factor of the 3D 7-point Laplacian on a 128^3
grid (+0.01 I before factorising; n = 2,097,152, nnz(L) = 8,339,456, 382
levels), FP64.  One step = one solve of L X = B for k = 32 right-hand sides
(row-major, B = L X_true with X_true ~ U[-1, 1], seed 1), inputs resident in
HBM.  Before each step B is restored from a pristine copy outside the timed
region (the solve overwrites B in place, as the reference does; its demo
re-initialises x before every timed run, demo/FusionDemo.cpp:59-62).

Reported beside the headline (in "detail"): the single-RHS plain level-set
solve (one launch per level), the same launches replayed from a hipGraph,
the single-RHS sync-free solve, the adjoint (gb = L^-T gx and grad over
L's pattern; automatic choice and sync-free), the block adjoint of the
headline block (GB = L^-T GX for all its columns, grad-L summed over them),
and the column-major 32-RHS solve.  "cpu_baseline" times the
reference's OpenMP level-set solve (oracle/_ref, the reference's own source
compiled for this host; the C restatement in oracle/ when that is absent) on
this host's cores for a bounded sample of right-hand sides.

Usage: python bench.py [--steps K] [--warmup W] [--k 32] [--layout row]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("SpTRSV RHS-solves/s + HBM GB/s, batched/adjoint vs plain GPU, "
          "rel err≤1e-12")
HBM_PEAK_GBS = 8000.0
# the kernel each executor launches once per solve
KERNELS = {"syncfree": "k_syncfree", "regions": "k_regions"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--k", type=int, default=32)
    ap.add_argument("--layout", default="row", choices=["row", "col"])
    ap.add_argument("--executor", default="auto")
    ap.add_argument("--no-detail", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--tune", action="store_true", help="run the schedule tuner first")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other workloads (Zipf SpTRSM, batch, adjoint 64^3)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a device copy of 1 GiB (x5) to scale HBM counters")
    ap.add_argument("--factors", action="store_true",
                    help="only time the device IC(0) and ILU(0) factorizations")
    ap.add_argument("--supernodal", action="store_true",
                    help="only time the supernodal (BCSC) solve")
    ap.add_argument("--leg", choices=["k1", "adjoint", "adjoint_block", "zipf", "batch"],
                    help="only run one workload (for profiling): the single right-hand "
                         "side region solve or the adjoint of the --grid factor, its "
                         "block adjoint at --k, the Zipf SpTRSM at --k (column-major, "
                         "tuned), or the 512-system batch replayed from a graph")
    return ap.parse_args()


def run_leg(args):
    """One workload alone, --warmup untimed and --steps timed repetitions,
    printed as one JSON line (HIP event time per repetition on the launch
    stream, and the algorithmic bytes of one repetition)."""
    import torch

    import sympiler_amd as sa
    from sympiler_amd import workloads as W
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    leg, k = args.leg, args.k
    out = {"leg": leg}
    up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    reset = lambda: None
    if leg in ("k1", "adjoint", "adjoint_block"):
        Lp, Li, Lx = W.ic0_lap3d(args.grid, 0.01)
        n, nnz = len(Lp) - 1, int(Lp[-1])
        plan = sa.Plan(up(Lp), up(Li), up(Lx))
        out["workload"] = f"ic0_lap3d_{args.grid}^3"
        if leg == "k1":
            b, x = up(W.uniform(1, 0, n)), up(np.zeros(n))
            fn, reset = (lambda: plan.solve(x, executor="regions")), (lambda: x.copy_(b))
            nbytes = W.algorithmic_bytes_forward(n, nnz, 1)
        elif leg == "adjoint":
            xs, gx = up(W.uniform(1, 0, n)), up(W.uniform(2, 0, n))
            gb, gL = torch.empty_like(gx), torch.empty(nnz, dtype=torch.float64, device=dev)
            fn = lambda: plan.adjoint(xs, gx, gb, gL)
            nbytes = W.algorithmic_bytes_adjoint(n, nnz)
        else:
            X, GX = up(W.uniform(1, 0, n * k).reshape(n, k)), up(W.uniform(2, 0, n * k).reshape(n, k))
            GB, gL = torch.empty_like(GX), torch.empty(nnz, dtype=torch.float64, device=dev)
            fn = lambda: plan.adjoint_block(X, GX, GB, gL, layout="row")
            nbytes = W.algorithmic_bytes_adjoint_block(n, nnz, k)
            out["k"] = k
    elif leg == "zipf":
        w3 = W.config3(k=1)
        plan = sa.Plan(up(w3.Lp), up(w3.Li), up(w3.Lx))
        B0 = up(W.uniform(2, 0, w3.n * k).reshape(k, w3.n))
        X = torch.empty_like(B0)
        out.update(workload="zipf_n1e6", k=k)
        if args.executor == "auto":
            out["tuned"] = plan.tune(k=k, layout="col").strip().splitlines()[-1]
        else:
            out["executor"] = args.executor
        fn = lambda: plan.solve(X, layout="col", executor=args.executor)
        reset = lambda: X.copy_(B0)
        nbytes = W.algorithmic_bytes_forward(w3.n, w3.nnz, k)
    else:
        w4 = W.config4()
        plan = sa.Plan(up(w4.Lp), up(w4.Li), up(w4.Lx), row_offsets=w4.row_offsets)
        b4 = up(w4.b)
        X4 = b4.clone()
        out.update(workload="batch512", tuned=plan.tune(k=1).strip().splitlines()[-1])
        g = plan.graph(X4)
        fn, reset = g.launch, (lambda: X4.copy_(b4))
        nbytes = W.algorithmic_bytes_forward(w4.n, w4.nnz, 1)
    _, evs = timed_steps(fn, reset, args.steps, args.warmup, stream, torch)
    plan.check()
    ms = float(np.mean(evs))
    out.update(steps=args.steps, warmup=args.warmup, ms=round(ms, 4),
               algorithmic_bytes=int(nbytes), gbs=round(nbytes / (ms * 1e-3) / 1e9, 1))
    print(json.dumps(out), flush=True)


def calibrate_copy():
    """Five device-to-device copies of 1 GiB: 1 GiB read + 1 GiB written per
    copy kernel.  Used to scale FETCH_SIZE / WRITE_SIZE counter readings."""
    import torch
    a = torch.empty(1 << 27, dtype=torch.float64, device="cuda:0").uniform_()
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "copy", "bytes_read_per_copy": 1 << 30,
                      "bytes_written_per_copy": 1 << 30, "copies": 5}))


def unit_executor_handoff(torch, sa, W, dev, stream, levels, ms_1):
    """Cost of one dependent cross-wave step of the unit (sync-free)
    executor: a chain (bidiagonal) system with one row per level has no
    parallelism and almost no bytes, so its solve time / levels is the
    hand-off cost t_sync, and levels * t_sync bounds any unit-executor solve
    with that many levels.  It does not bound the region executor, whose
    critical path crosses only region-to-region hand-offs."""
    out = {}
    Lp, Li, Lx = W.lap2d_tril(1, levels)  # bidiagonal: `levels` levels
    plan = sa.Plan(*(torch.from_numpy(a).to(dev) for a in (Lp, Li, Lx)))
    for k in (1, 32):
        B0 = torch.from_numpy(W.uniform(3, 0, levels * k).reshape(levels, k)).to(dev)
        X = torch.empty_like(B0)
        _, ev = timed_steps(lambda: plan.solve(X, layout="row", executor="syncfree"),
                            lambda: X.copy_(B0), 10, 2, stream, torch)
        plan.check()
        t = float(np.mean(ev))
        out[f"k{k}"] = {"chain_ms": round(t, 4), "t_sync_us": round(1e3 * t / levels, 3)}
    out["syncfree_k1_ms_over_chain"] = round(ms_1 / out["k1"]["chain_ms"], 2)
    return out


def other_configs(torch, sa, W, dev, stream):
    """The other workloads, a few timed repetitions each (seconds of GPU
    time): Zipf-irregular SpTRSM (n = 1e6, k = 1/8/32/128, column-major), the
    512-system uneven batch replayed from a hipGraph, the 64^3 IC(0)
    adjoint (forward vs forward+backward, FP64 and FP32), and the device IC(0)
    and ILU(0) factorizations on the 128^3 grid."""
    out = {}
    # Zipf(1.8) irregular lower triangle, B column-major (k, n)
    w3 = W.config3(k=1)
    plan3 = sa.Plan(*(torch.from_numpy(a).to(dev) for a in (w3.Lp, w3.Li, w3.Lx)))
    zipf = {"n": w3.n, "nnz": w3.nnz, "levels": plan3.levels}
    for k in (1, 8, 32, 128):
        B0 = torch.from_numpy(W.uniform(2, 0, w3.n * k).reshape(k, w3.n)).to(dev)
        X = torch.empty_like(B0)
        tune = plan3.tune(k=k, layout="col").strip().splitlines()[-1]
        wl, el = timed_steps(lambda: plan3.solve(X, layout="col"), lambda: X.copy_(B0),
                             5, 1, stream, torch)
        plan3.check()
        res = float(plan3.residual(X, B0, layout="col").max())
        ms = float(np.mean(el))
        zipf[f"k{k}"] = {"kernel_ms": round(ms, 4), "rhs_solves_per_s": round(k / (ms * 1e-3), 1),
                         "gbs": round(W.algorithmic_bytes_forward(w3.n, w3.nnz, k) / (ms * 1e-3) / 1e9, 1),
                         "max_residual": res, "tuned": tune}
        del B0, X
    out["zipf_sptrsm_col_major"] = zipf
    log(f"zipf: {zipf}")
    del plan3

    # 512 uneven systems, one captured launch replayed
    w4 = W.config4()
    plan4 = sa.Plan(*(torch.from_numpy(a).to(dev) for a in (w4.Lp, w4.Li, w4.Lx)),
                    row_offsets=w4.row_offsets)
    b4 = torch.from_numpy(w4.b).to(dev)
    X4 = b4.clone()
    tune = plan4.tune(k=1).strip().splitlines()[-1]
    g = plan4.graph(X4)
    wg, eg = timed_steps(g.launch, lambda: X4.copy_(b4), 10, 2, stream, torch)
    plan4.check()
    err = float(np.max(np.abs(X4.cpu().numpy() - w4.x_true)))
    ms = float(np.mean(eg))
    out["batch512"] = {"n_total": w4.n, "nnz_total": w4.nnz, "levels": plan4.levels,
                       "graph_replay_ms": round(ms, 4),
                       "system_solves_per_s": round(512 / (ms * 1e-3), 1),
                       "rows_per_s": round(w4.n / (ms * 1e-3), 1),
                       "gbs": round(W.algorithmic_bytes_forward(w4.n, w4.nnz, 1) / (ms * 1e-3) / 1e9, 1),
                       "max_abs_err_vs_x_true": err, "tuned": tune}
    log(f"batch512: {out['batch512']}")
    del g, plan4

    # adjoint through autograd on the 64^3 IC(0) factor, FP64 and FP32
    from sympiler_amd.autograd import TriangularSolve
    Lp5, Li5, Lx5 = W.ic0_lap3d(64, 0.01)
    n5 = len(Lp5) - 1
    b5 = torch.from_numpy(W.uniform(4, 0, n5)).to(dev)
    adj = {"n": n5, "nnz": int(Lp5[-1])}
    cold5 = ColdCache(torch, dev)
    for name, dt in (("fp64", torch.float64), ("fp32", torch.float32)):
        dLx = torch.from_numpy(Lx5).to(dev).to(dt)
        plan5 = sa.Plan(torch.from_numpy(Lp5).to(dev), torch.from_numpy(Li5).to(dev), dLx)
        bb = b5.to(dt)
        Lx_t = dLx.clone().requires_grad_(True)
        b_t = bb.clone().requires_grad_(True)

        def fwd():
            with torch.no_grad():
                TriangularSolve.apply(dLx, bb, plan5)

        def fwdbwd():
            x = TriangularSolve.apply(Lx_t, b_t, plan5)
            (0.5 * (x * x).sum()).backward()

        wf, _ = timed_steps(fwd, lambda: None, 10, 2, stream, torch)
        wb, _ = timed_steps(fwdbwd, lambda: None, 10, 2, stream, torch)
        wfc, _ = timed_steps(fwd, cold5(), 10, 2, stream, torch)
        wbc, _ = timed_steps(fwdbwd, cold5(), 10, 2, stream, torch)
        plan5.check()
        adj[name] = {"forward_ms": round(1e3 * float(np.mean(wf)), 4),
                     "forward_backward_ms": round(1e3 * float(np.mean(wb)), 4),
                     "ratio": round(float(np.mean(wb)) / float(np.mean(wf)), 2),
                     "cold_forward_ms": round(1e3 * float(np.mean(wfc)), 4),
                     "cold_forward_backward_ms": round(1e3 * float(np.mean(wbc)), 4)}
        del plan5
    out["adjoint_64cubed_autograd"] = adj
    log(f"adjoint 64^3: {adj}")
    del Lx_t, b_t

    out.update(factor_configs(torch, sa, W, dev, stream))
    out["supernodal_bcsc_tree13"] = supernodal_config(torch, sa, W, dev, stream)
    return out


def supernodal_config(torch, sa, W, dev, stream):
    """Supernodal (BCSC) forward solve on a nested-dissection-shaped factor
    (13 tree levels, 8,191 supernodes, 221,040 columns, 36 M block values):
    symamd_sn_solve, bit-identical to sptrsv_csc_block."""
    sup, nrows, Lp, Li, Lx = W.bcsc_tree(13)
    n = int(sup[-1])
    b = W.uniform(8, 0, n)
    plan = sa.SupernodalPlan(sup, nrows, Lp, Li)
    dLx = torch.from_numpy(Lx).to(dev)
    db = torch.from_numpy(b).to(dev)
    x = db.clone()
    wt, et = timed_steps(lambda: plan.solve(dLx, x), lambda: x.copy_(db), 10, 2, stream, torch)
    plan.check()
    ms = float(np.mean(et))
    byts = 8 * len(Lx) + 16 * n
    out = {"n": n, "supernodes": len(sup) - 1, "values": len(Lx),
           "widest": int(np.diff(sup).max()), "kernel_ms": round(ms, 4),
           "gbs": round(byts / (ms * 1e-3) / 1e9, 1)}
    # 32 right-hand sides, column-major (symamd_sn_solve_block)
    k = 32
    B0 = torch.from_numpy(W.uniform(8, 1, k * n).reshape(k, n)).to(dev)
    X = B0.clone()
    wk, ek = timed_steps(lambda: plan.solve(dLx, X), lambda: X.copy_(B0), 5, 1, stream, torch)
    plan.check()
    mk = float(np.mean(ek))
    out["k32"] = {"kernel_ms": round(mk, 4), "rhs_solves_per_s": round(k / (mk * 1e-3), 1),
                  "vs_k1_rhs_per_s": round(k * ms / mk, 2)}
    # the transposed solve L^T x = b on the same plan (blockedLTsolve), one column
    xt = db.clone()
    _, tt = timed_steps(lambda: plan.solve_transposed(dLx, xt), lambda: xt.copy_(db), 5, 1,
                        stream, torch)
    plan.check()
    mt = float(np.mean(tt))
    out["transposed"] = {"kernel_ms": round(mt, 4), "gbs": round(byts / (mt * 1e-3) / 1e9, 1)}
    log(f"supernodal tree13: {out}")
    del plan, dLx, db, x, xt, B0, X
    return out


def factor_configs(torch, sa, W, dev, stream):
    """Device IC(0) (both variants) of the 128^3 Laplacian and ILU(0) of the
    general 128^3 7-point matrix, timed on the launch stream."""
    out = {}
    # IC(0) of the 128^3 7-point Laplacian on the device (symamd_ic0), both
    # variants; checked against the package's CPU generator (spic0_csc_left
    # order)
    g6 = 128
    Lp6, Li6, A6 = W.lap3d_tril(g6, g6, g6, 6.01, -1.0)
    dA6 = torch.from_numpy(A6).to(dev)
    plan6 = sa.Plan(torch.from_numpy(Lp6).to(dev), torch.from_numpy(Li6).to(dev), dA6)
    ic0 = {"n": len(Lp6) - 1, "nnz": int(Lp6[-1]), "levels": plan6.levels}
    for variant in ("left", "csr"):
        t0 = time.perf_counter()
        L6 = plan6.ic0(dA6, variant)
        first = time.perf_counter() - t0
        wi, _ = timed_steps(lambda: plan6.ic0(dA6, variant, out=L6), lambda: None, 5, 1,
                            stream, torch)
        ic0[variant] = {"first_call_ms": round(1e3 * first, 2),
                        "factor_ms": round(1e3 * float(np.mean(wi)), 3)}
        if variant == "left":
            ic0[variant]["bitwise_vs_cpu_generator"] = bool(np.array_equal(
                L6.cpu().numpy().view(np.int64), W.ic0(Lp6, Li6, A6).view(np.int64)))
    out["ic0_lap3d_128cubed"] = ic0
    log(f"ic0 128^3: {ic0}")
    del plan6, dA6, L6

    # ILU(0) of a general matrix on the full 7-point pattern of the 128^3
    # grid: the one-shot call (schedule, pairs and factor) and the numeric
    # factorization of an analysed pattern
    Ap7, Ai7, Ax7, d7 = W.grid3d_general(128)
    dAp7, dAi7, dAx7, dd7 = (torch.from_numpy(a).to(dev) for a in (Ap7, Ai7, Ax7, d7))
    l7 = torch.empty_like(dAx7)
    wl7, _ = timed_steps(lambda: sa.ilu0(dAp7, dAi7, dAx7, dd7, out=l7), lambda: None, 3, 1,
                         stream, torch)
    t0 = time.perf_counter()
    iplan = sa.Ilu0Plan(dAp7, dAi7, dd7)
    an_ms = 1e3 * (time.perf_counter() - t0)
    wf7, _ = timed_steps(lambda: iplan.factor(dAx7, out=l7), lambda: None, 5, 1, stream, torch)
    out["ilu0_grid3d_128cubed"] = {"n": len(Ap7) - 1, "nnz": int(Ap7[-1]),
                                   "one_shot_ms": round(1e3 * float(np.mean(wl7)), 2),
                                   "analyze_ms": round(an_ms, 2),
                                   "factor_ms": round(1e3 * float(np.mean(wf7)), 3)}
    del iplan
    log(f"ilu0 128^3: {out['ilu0_grid3d_128cubed']}")
    del dAp7, dAi7, dAx7, dd7, l7
    return out


class ColdCache:
    """Evicts the caches between repetitions: writes a 1 GiB buffer (four
    times the 256 MB MALL) before each timed step, outside the timed region."""

    def __init__(self, torch, dev):
        self.buf = torch.empty(1 << 27, dtype=torch.float64, device=dev)

    def __call__(self, reset=None):
        def r():
            if reset is not None:
                reset()
            self.buf.fill_(0.5)
        return r


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def dropin_times(W, Lp, Li, Lx, b):
    """sym_lib::sptrsv_csr bound by its mangled C++ name and called on host
    arrays, as a program relinked against the library calls it (wall clock,
    PCIe copies included): the first call on a pattern (upload, analysis,
    solve, download) and later calls with new values on the same pattern
    (the kept analysis: the pattern compared, values and right-hand side
    uploaded, values gathered, solve, download).  For the 100 x 100 grid
    Laplacian (the reference demo's scale) and the headline factor."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                   "sympiler_amd", "libsympiler_amd.so"))
    fn = lib["_ZN7sym_lib10sptrsv_csrEiPiS0_PdS1_"]
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    w1 = W.config1()
    out = {}
    for name, (p, i, v, bb) in (("lap2d_100x100", (w1.Lp, w1.Li, w1.Lx, w1.b)),
                                ("ic0_lap3d_128", (Lp, Li, Lx, b))):
        n = len(p) - 1
        vals = [v, v * 1.5]               # alternate values on the same pattern
        x = bb.copy()
        t0 = time.perf_counter()
        fn(ctypes.c_int(n), P(p), P(i), P(vals[0]), P(x))
        first = time.perf_counter() - t0
        times = []
        for r in range(6):
            x = bb.copy()
            t0 = time.perf_counter()
            fn(ctypes.c_int(n), P(p), P(i), P(vals[(r + 1) % 2]), P(x))
            times.append(time.perf_counter() - t0)
        out[name] = {"n": n, "nnz": int(p[-1]), "first_call_ms": round(1e3 * first, 3),
                     "repeat_call_ms": round(1e3 * float(np.median(times)), 3)}
    return out


def timed_steps(fn, reset, steps, warmup, stream, torch):
    """Run warmup + steps; each step: reset (untimed), sync, fn, sync.
    Returns (wall seconds per step list, event ms per step list)."""
    for _ in range(warmup):
        reset()
        fn()
    torch.cuda.synchronize()
    walls, evs = [], []
    for _ in range(steps):
        reset()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - t0)
        evs.append(e0.elapsed_time(e1))
    return walls, evs


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.leg:
        run_leg(args)
        return
    if args.factors or args.supernodal:
        import torch
        import sympiler_amd as sa
        from sympiler_amd import workloads as W
        dev = torch.device("cuda:0")
        fn = factor_configs if args.factors else supernodal_config
        print(json.dumps({"factors" if args.factors else "supernodal":
                          fn(torch, sa, W, dev, torch.cuda.current_stream())}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the multi-GPU "
                          "sharding is a design note (shard right-hand sides or "
                          "systems across GPUs), no multi-GPU path is built"}))
        sys.exit(2)
    import torch

    import sympiler_amd as sa
    from sympiler_amd import workloads as W

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    g, k = args.grid, args.k
    t0 = time.time()
    Lp, Li, Lx = W.ic0_lap3d(g, 0.01)
    n, nnz = len(Lp) - 1, int(Lp[-1])
    Xt = W.uniform(1, 0, n * k).reshape(k, n)  # column c = X_true[:, c]
    Bcols = np.stack([W.spmv(Lp, Li, Lx, Xt[c]) for c in range(k)])  # (k, n)
    log(f"workload IC(0) 3D Laplacian {g}^3: n={n} nnz={nnz} k={k} "
        f"(generated in {time.time() - t0:.1f}s)")

    dLp, dLi, dLx = (torch.from_numpy(a).to(dev) for a in (Lp, Li, Lx))
    t0 = time.time()
    plan = sa.Plan(dLp, dLi, dLx)
    torch.cuda.synchronize()
    analyze_s = time.time() - t0
    log(f"analyze: {analyze_s*1e3:.1f} ms, levels={plan.levels}, "
        f"max level rows={plan.max_level_rows}")
    if args.tune:
        log("tune k=%d: %s" % (k, plan.tune(k, args.layout).replace("\n", "; ")))
    # the executor the automatic choice takes (building the region schedule
    # when it is a candidate: timed apart from the analysis)
    t0 = time.time()
    chosen, _ = plan.choice(k, args.layout)
    schedule_s = time.time() - t0
    executor = chosen if args.executor == "auto" else args.executor
    log(f"executor: {executor} (choice made in {schedule_s*1e3:.0f} ms)")

    if args.layout == "row":
        B0 = torch.from_numpy(np.ascontiguousarray(Bcols.T)).to(dev)  # (n, k)
    else:
        B0 = torch.from_numpy(Bcols).to(dev)  # (k, n) == column-major
    X = torch.empty_like(B0)

    def reset():
        X.copy_(B0)

    def step():
        plan.solve(X, layout=args.layout, executor=args.executor)

    walls, evs = timed_steps(step, reset, args.steps, args.warmup, stream, torch)
    plan.check()
    ms_step = 1e3 * float(np.mean(walls))
    ms_kernel = float(np.mean(evs))
    value = k / (ms_step * 1e-3)
    bytes_alg = W.algorithmic_bytes_forward(n, nnz, k)
    achieved = bytes_alg / (ms_kernel * 1e-3) / 1e9
    Xh = X.cpu().numpy()
    Xh = Xh.T if args.layout == "row" else Xh
    err_true = float(np.max(np.abs(Xh - Xt)) / np.max(np.abs(Xt)))
    res = plan.residual(X, B0, layout=args.layout)
    log(f"k={k} {args.layout}: {ms_step:.3f} ms/step wall, {ms_kernel:.3f} ms kernel, "
        f"{value:.0f} RHS-solves/s, {achieved:.0f} GB/s, max residual {res.max():.2e}")

    detail = {"analyze_ms": round(analyze_s * 1e3, 2),
              "schedule_ms": round(schedule_s * 1e3, 2),
              "headline_kernel_ms": round(ms_kernel, 4),
              "headline_gbs": round(achieved, 1),
              "headline_rel_err_vs_x_true": err_true,
              "headline_max_residual": float(res.max())}

    if not args.no_detail:
        b1 = torch.from_numpy(Bcols[0].copy()).to(dev)
        x1 = torch.empty_like(b1)
        bytes1 = W.algorithmic_bytes_forward(n, nnz, 1)

        def reset1():
            x1.copy_(b1)

        singles = {}
        for name, ex, lanes in (("levelset_plain", "levelset", 1),
                                ("levelset_graph", "graph", 1),
                                ("syncfree", "syncfree", 1),
                                ("regions", "regions", 1)):
            plan.set_lanes(lanes)
            w1, e1 = timed_steps(lambda: plan.solve(x1, executor=ex), reset1, 10, 2,
                                 stream, torch)
            plan.check()
            r = plan.residual(x1, b1)[0]
            singles[name] = {"ms": round(1e3 * float(np.mean(w1)), 4),
                             "kernel_ms": round(float(np.mean(e1)), 4),
                             "rhs_solves_per_s": round(1.0 / float(np.mean(w1)), 1),
                             "gbs": round(bytes1 / (np.mean(e1) * 1e-3) / 1e9, 1),
                             "residual": float(r)}
            log(f"k=1 {name}: {singles[name]}")
        # the single-RHS working set (142 MB) fits the 256 MB MALL: the same
        # solves with the caches evicted before each one
        cold = ColdCache(torch, dev)
        for name, ex in (("syncfree_cold", "syncfree"), ("regions_cold", "regions"),
                         ("levelset_graph_cold", "graph")):
            plan.set_lanes(1)
            w1, e1 = timed_steps(lambda: plan.solve(x1, executor=ex), cold(reset1), 10, 2,
                                 stream, torch)
            plan.check()
            singles[name] = {"ms": round(1e3 * float(np.mean(w1)), 4),
                             "kernel_ms": round(float(np.mean(e1)), 4),
                             "rhs_solves_per_s": round(1.0 / float(np.mean(w1)), 1),
                             "gbs": round(bytes1 / (np.mean(e1) * 1e-3) / 1e9, 1)}
            log(f"k=1 {name}: {singles[name]}")
        plan.set_lanes(0)
        detail["single_rhs"] = singles
        # the plain level-set baseline: one kernel per level, launched eagerly
        # or replayed from a graph, whichever is faster on this box
        plain = max(singles["levelset_plain"]["rhs_solves_per_s"],
                    singles["levelset_graph"]["rhs_solves_per_s"])
        detail["plain_levelset_rhs_solves_per_s"] = plain
        detail["headline_vs_plain_levelset"] = round(value / plain, 2)
        detail["headline_frac_hbm_peak"] = round(achieved / HBM_PEAK_GBS, 4)
        detail["unit_executor_handoff"] = unit_executor_handoff(
            torch, sa, W, dev, stream, plan.levels, singles["syncfree"]["kernel_ms"])
        log(f"unit executor hand-off: {detail['unit_executor_handoff']}")

        # fused adjoint: gb = L^-T gx, grad over L's pattern; loss 0.5||x||^2
        xs = x1.clone()
        gx = xs.clone()
        gb = torch.empty_like(gx)
        gL = torch.empty(nnz, dtype=torch.float64, device=dev)
        ab = W.algorithmic_bytes_adjoint(n, nnz)
        # the automatic choice (the transposed region schedule on this matrix),
        # then the sync-free adjoint for comparison; both include grad-L
        for key, ex, rs in (("adjoint", "auto", lambda: None),
                            ("adjoint_cold", "auto", cold()),
                            ("adjoint_syncfree", "syncfree", lambda: None)):
            wa, ea = timed_steps(lambda: plan.adjoint(xs, gx, gb, gL, executor=ex),
                                 rs, 10, 2, stream, torch)
            plan.check()
            detail[key] = {"executor": ex,
                           "ms": round(1e3 * float(np.mean(wa)), 4),
                           "kernel_ms": round(float(np.mean(ea)), 4),
                           "rhs_solves_per_s": round(1.0 / float(np.mean(wa)), 1),
                           "gbs": round(ab / (np.mean(ea) * 1e-3) / 1e9, 1),
                           "vs_plain_levelset": round(1.0 / float(np.mean(wa)) / plain, 2)}
            log(f"{key}: {detail[key]}")
        # the block adjoint of the headline block: GB = L^-T GX for all its
        # columns in one launch, then grad-L summed over the columns
        kb = X.shape[1] if args.layout == "row" else X.shape[0]
        GXb = X.clone()
        GBb = torch.empty_like(GXb)
        wb, eb = timed_steps(lambda: plan.adjoint_block(X, GXb, GBb, gL, layout=args.layout),
                             lambda: None, 10, 2, stream, torch)
        plan.check()
        bb = W.algorithmic_bytes_adjoint_block(n, nnz, kb)
        detail["adjoint_block"] = {"k": kb, "layout": args.layout,
                                   "ms": round(1e3 * float(np.mean(wb)), 4),
                                   "kernel_ms": round(float(np.mean(eb)), 4),
                                   "rhs_solves_per_s": round(kb / float(np.mean(wb)), 1),
                                   "gbs": round(bb / (np.mean(eb) * 1e-3) / 1e9, 1),
                                   "frac_hbm_peak": round(bb / (np.mean(eb) * 1e-3) / 1e9
                                                          / HBM_PEAK_GBS, 4),
                                   "vs_plain_levelset": round(kb / float(np.mean(wb)) / plain, 2)}
        log(f"adjoint_block: {detail['adjoint_block']}")
        del GXb, GBb

        # the headline block on the region executor
        wr, er = timed_steps(lambda: plan.solve(X, layout=args.layout, executor="regions"),
                             reset, 10, 2, stream, torch)
        plan.check()
        res_r = plan.residual(X, B0, layout=args.layout)
        detail["regions_k"] = {"ms": round(1e3 * float(np.mean(wr)), 4),
                               "kernel_ms": round(float(np.mean(er)), 4),
                               "rhs_solves_per_s": round(k / float(np.mean(wr)), 1),
                               "gbs": round(bytes_alg / (np.mean(er) * 1e-3) / 1e9, 1),
                               "max_residual": float(res_r.max())}
        log(f"regions k={k}: {detail['regions_k']}")

        if args.layout == "row":
            Bc = torch.from_numpy(Bcols).to(dev)
            Xc = torch.empty_like(Bc)
            wc, ec = timed_steps(lambda: plan.solve(Xc, layout="col"),
                                 lambda: Xc.copy_(Bc), 5, 1, stream, torch)
            plan.check()
            detail["col_major_k"] = {"ms": round(1e3 * float(np.mean(wc)), 4),
                                     "rhs_solves_per_s": round(k / float(np.mean(wc)), 1),
                                     "gbs": round(bytes_alg / (np.mean(ec) * 1e-3) / 1e9, 1)}
            log(f"col-major k={k}: {detail['col_major_k']}")
            del Bc, Xc

    if not args.no_detail:
        detail["dropin_sptrsv_csr"] = dropin_times(W, Lp, Li, Lx, Bcols[0])
        log(f"drop-in sptrsv_csr: {detail['dropin_sptrsv_csr']}")

    if not args.no_detail and not args.no_configs:
        detail["configs"] = other_configs(torch, sa, W, dev, stream)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(Lp, Li, Lx, Bcols, X, args, k)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "RHS-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": (f"synthetic: IC(0) of the 3D 7-point Laplacian {g}^3 (+0.01 I), "
                 f"B = L X_true, X_true ~ U[-1,1] (splitmix64 seed 1)"),
        "config": {"workload": f"ic0_lap3d_{g}^3 SpTRSM k={k} {args.layout}-major",
                   "n": n, "nnz": nnz, "levels": plan.levels, "k": k,
                   "layout": args.layout, "executor": executor},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(f"ic0_lap3d_{g}^3", k, args.layout,
                                                 KERNELS.get(executor, "")),
                     "traffic_source": "profiles/hbm_traffic.txt (rocprofv3 --pmc "
                                       "FETCH_SIZE, WRITE_SIZE; copy-calibrated)",
                     "kernel": KERNELS.get(executor, executor) + " (one launch per solve)",
                     "algorithmic_bytes_per_launch": int(bytes_alg)},
        "cpu_baseline": cpu,
        "detail": detail,
    }
    print(json.dumps(out), flush=True)


def measured_traffic(workload, k, layout, kernel):
    """HBM bytes per launch of the headline kernel as measured by the
    committed rocprofv3 counter passes (profiles/hbm_traffic.txt: workload,
    k, layout, kernel, bytes), or None when it was not profiled."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 5 and not line.startswith("#") and f[0] == workload \
                and f[1] == str(k) and f[2] == layout and f[3] == kernel:
            return int(f[4])
    return None


def cpu_model():
    """The host CPU's model name and counts as lscpu reports them."""
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keep = ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)")
        return "; ".join(line.strip() for line in out.splitlines()
                         if line.split(":")[0].strip() in keep)
    except Exception as e:  # lscpu missing: say so rather than guess
        return f"lscpu unavailable ({e})"


def cpu_baseline(Lp, Li, Lx, Bcols, Xgpu, args, k):
    """Reference OpenMP level-set solve on this host, one right-hand side
    at a time (the reference has no multi-RHS CSR solve), for a bounded
    sample; also checks the GPU solution of the sampled columns."""
    import oracle
    ncores = len(os.sched_getaffinity(0))
    lp, ls, _ = oracle.levelset_csr(Lp, Li)
    if oracle.ref_available():
        kind = "reference"
        oracle.ref.set_threads(ncores)
        solve = lambda b: oracle.ref.sptrsv_csr_levelset(Lp, Li, Lx, b, lp, ls)
    else:
        kind = "port"
        solve = lambda b: oracle.sptrsv_csr_levelset(Lp, Li, Lx, b, lp, ls, ncores)
    Xh = Xgpu.cpu().numpy()
    Xh = Xh.T if args.layout == "row" else Xh
    t_total, done, worst = 0.0, 0, 0.0
    while t_total < args.cpu_seconds and done < 64 * k:
        c = done % k
        b = Bcols[c]
        t0 = time.perf_counter()
        x = solve(b)
        t_total += time.perf_counter() - t0
        if done < k:
            worst = max(worst, float(np.max(np.abs(Xh[c] - x)) / np.max(np.abs(x))))
        done += 1
    rate = done / t_total
    # the C restatement (oracle.c) beside the reference build on the same
    # matrix, cores and right-hand side: their time ratio relates a port
    # timing on another host to the reference's
    ratio = None
    if kind == "reference":
        b = Bcols[0]
        tr, tp = [], []
        for _ in range(3):
            t0 = time.perf_counter()
            oracle.ref.sptrsv_csr_levelset(Lp, Li, Lx, b, lp, ls)
            tr.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            oracle.sptrsv_csr_levelset(Lp, Li, Lx, b, lp, ls, ncores)
            tp.append(time.perf_counter() - t0)
        ratio = round(min(tp) / min(tr), 3)
    # the reference's IC(0) (spic0_csc_left, serial) on the 128^3 Laplacian,
    # beside the device factor timed in the detail
    ic0_ms = None
    if kind == "reference":
        from sympiler_amd import workloads as W
        Lp6, Li6, A6 = W.lap3d_tril(128, 128, 128, 6.01, -1.0)
        Cp, Ci, Cx, _ = oracle.csr_to_csc(Lp6, Li6, A6)
        pp, ps = oracle.prune_set(Lp6, Li6)
        t0 = time.perf_counter()
        oracle.ref.spic0_csc_left(Cp, Ci, Cx, pp, ps)
        ic0_ms = round(1e3 * (time.perf_counter() - t0), 1)
    # ILU(0) of the general 128^3 7-point matrix: the oracle's serial
    # restatement of spilu0_csr (the reference's file is not buildable here)
    from sympiler_amd import workloads as W
    Ap7, Ai7, Ax7, d7 = W.grid3d_general(128)
    t0 = time.perf_counter()
    oracle.spilu0_csr(Ap7, Ai7, Ax7, d7)
    ilu_ms = round(1e3 * (time.perf_counter() - t0), 1)
    # the supernodal solve of the detail: the reference's serial
    # sptrsv_csc_block (or its restatement)
    sup8, nr8, Lp8, Li8, Lx8 = W.bcsc_tree(13)
    b8 = W.uniform(8, 0, int(sup8[-1]))
    blk = oracle.ref.sptrsv_csc_block if kind == "reference" else oracle.sptrsv_csc_block
    t0 = time.perf_counter()
    blk(sup8, nr8, Lp8, Li8, Lx8, b8)
    sn_ms = round(1e3 * (time.perf_counter() - t0), 1)
    # the reference's serial sptrsv_csr on the drop-in timing's two matrices
    # (what a relinked call replaces)
    serial = {}
    w1 = W.config1()
    ser = oracle.ref.sptrsv_csr if kind == "reference" else oracle.sptrsv_csr
    for name, (p, i, v, bb) in (("lap2d_100x100", (w1.Lp, w1.Li, w1.Lx, w1.b)),
                                ("ic0_lap3d_128", (Lp, Li, Lx, Bcols[0]))):
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            ser(p, i, v, bb)
            ts.append(time.perf_counter() - t0)
        serial[name] = round(1e3 * min(ts), 3)
    log(f"cpu baseline ({kind}, {ncores} threads): "
        f"{rate:.2f} RHS-solves/s over {done} solves; GPU vs CPU max rel err {worst:.2e}")
    return {"value": round(rate, 3), "unit": "RHS-solves/s",
            "cores": ncores, "kind": kind, "cpu": cpu_model(),
            "host_logical_cpus": os.cpu_count(),
            "port_over_reference_time": ratio,
            "sample": f"{done} single-RHS level-set solves of the same matrix "
                      f"({t_total:.1f}s)",
            "gpu_vs_cpu_max_rel_err": worst,
            "ic0_spic0_csc_left_128cubed_ms_1_thread": ic0_ms,
            "ilu0_port_spilu0_csr_grid3d_128cubed_ms_1_thread": ilu_ms,
            "sptrsv_csc_block_tree13_ms_1_thread": sn_ms,
            "sptrsv_csr_ms_1_thread": serial}


if __name__ == "__main__":
    main()
