// executors.hip -- forward and adjoint sparse triangular solve kernels for
// gfx950.
//
// Row arithmetic (reference sparse_blas/sptrsv.cpp:17-25, sptrsv_csr):
//   x[i] -= L[i,j] * x[j] for the stored j < i in ascending order,
//   then x[i] /= L[i,i]   (diagonal last in the row).
// With one lane per (row, right-hand side) the lane repeats exactly that
// sequence; this file is compiled with -ffp-contract=off, so the result is
// bit-identical to the reference's serial solve.  With G > 1 lanes per row
// the off-diagonal sum is split over the G lanes and reduced with
// cross-lane shuffles, then x = (b - sum) / L[i,i].
//
// Transposed solve (reference sparse_blas/sptrsv_lt.cpp:11-20, ltsolve):
//   gb[i] -= L[r,i] * gb[r] for the stored r > i ascending, then
//   gb[i] /= L[i,i]; rows visited in reverse level order.
//
// Lane layout inside a 64-wide wavefront: lane = (row, split, column) with
// the column index fastest, so with row-major right-hand sides the KT lanes
// of one row read and write KT consecutive values of that row.  Each lane
// carries NR rows at once and issues their loads together, so a wave keeps
// NR * (64 / (G*KT)) rows in flight.
//
// Executors
//   LEVELSET  one launch per level; plain loads and stores (kernel
//             boundaries order the levels).
//   SYNCFREE  one persistent launch over work units (runs of R rows inside
//             one level, in level order).  The inspector records, for each
//             unit, the units holding its rows' dependencies; a wave waits
//             on those units' ready flags (one coherent load per dependency
//             unit, with back-off), solves its rows and publishes its own
//             unit flag.  Units are dealt to waves round-robin from a
//             cooperative launch (all waves co-resident), or taken from a
//             ticket counter when the launch is captured in a graph; either
//             way every unit a wave waits on is held by a running wave that
//             reaches it first, so the grid cannot deadlock.  Waits are
//             bounded by a watchdog on the 100 MHz clock that records an
//             error and lets the wave proceed, so every wave exits.
#include "common.h"

namespace symamd {

namespace {

template <typename T>
struct FwdArgs {
  int n;
  int k;
  int64_t rs, cs;            // element (i, c) at X[i*rs + c*cs]
  const int *rp;             // level-ordered row pointers
  const int *ci;             // column ids (original numbering)
  const T *val;              // values
  const int *perm;           // position -> row
  T *X;
  int *state;
  const int *ub;             // unit bounds (positions)
  const int *dptr, *dlist;   // dependency units of each unit
  int *uflags;               // unit ready tags
  int U;
  int ticketed;              // 1: units from the ticket counter
  uint64_t timeout;
  int total_waves;
};

template <typename T>
struct AdjArgs {
  int n;
  const int *trp, *tci;      // level-ordered transposed copy
  const T *tval;
  const int *rp, *ci;        // level-ordered forward copy (for gL)
  const int *orp;            // caller CSR row pointers
  const int *perm;
  const T *x, *gx;
  T *gb, *gL;
  int *state;
  const int *ub;
  const int *dptr, *dlist;   // reverse dependency units
  int *uflags;
  int U;
  int ticketed;
  uint64_t timeout;
  int total_waves;
};

// Wait until every dependency unit of unit u carries `target`.  One lane per
// dependency unit, one coherent load each, sleeping between rounds.
__device__ __forceinline__ void wait_units(const int *__restrict__ dptr,
                                           const int *__restrict__ dlist,
                                           int u, const int *uflags,
                                           int target, int *state,
                                           uint64_t timeout) {
  const int lane = threadIdx.x & 63;
  const int d0 = dptr[u], d1 = dptr[u + 1];
  for (int base = d0; base < d1; base += 64) {
    const int t = base + lane;
    bool pending = t < d1;
    const int *f = pending ? &uflags[dlist[t]] : uflags;
    if (pending) pending = ld_dev(f) != target;
    if (!__any(pending)) continue;
    const uint64_t t0 = wall_ticks();
    int nap = 1, rounds = 0;
    while (__any(pending)) {
      for (int z = 0; z < nap; ++z) __builtin_amdgcn_s_sleep(1);
      if (nap < 4) nap <<= 1;
      if (pending) pending = ld_dev(f) != target;
      if ((++rounds & 31) == 0 && wall_ticks() - t0 > timeout) {
        if (pending) atomicOr(&state[ST_ERROR], 1);
        pending = false;
      }
    }
  }
  asm volatile("" ::: "memory");
}

template <typename T, int G, int KT>
__device__ __forceinline__ T lane_sum(T v) {
#pragma unroll
  for (int off = KT; off < G * KT; off <<= 1) v += __shfl_xor(v, off);
  return v;
}

template <bool SYNC, typename T>
__device__ __forceinline__ T ld_x(const T *p) {
  if constexpr (SYNC) return ld_dev(p);
  else return *p;
}
template <bool SYNC, typename T>
__device__ __forceinline__ void st_x(T *p, T v) {
  if constexpr (SYNC) st_dev(p, v);
  else *p = v;
}

// Solve the rows at positions [q0, q1) (at most NR * 64/(G*KT) of them)
// for all k right-hand sides with one wave.
template <typename T, int G, int KT, int NR, bool SYNC>
__device__ __forceinline__ void fwd_rows(const FwdArgs<T> &a, int q0, int q1) {
  constexpr int LPR = G * KT;
  constexpr int RPW = 64 / LPR;
  const int lane = threadIdx.x & 63;
  const int rr = lane / LPR;
  const int s = (lane / KT) % G;
  const int c = lane % KT;
  int i[NR], jb[NR], je[NR];
  T diag[NR];
  bool act[NR];
  int iters = 0;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int q = q0 + rr + r * RPW;
    act[r] = q < q1;
    i[r] = 0;
    jb[r] = 0;
    je[r] = 0;
    diag[r] = T(1);
    if (act[r]) {
      i[r] = a.perm[q];
      jb[r] = a.rp[q] + s;
      je[r] = a.rp[q + 1] - 1;
      diag[r] = a.val[je[r]];
      const int it = (je[r] - jb[r] + G - 1) / G;
      iters = it > iters ? it : iters;
    }
  }
  for (int c0 = 0; c0 < a.k; c0 += KT) {
    const int col = c0 + c;
    const bool colok = col < a.k;
    const int64_t coff = (int64_t)col * a.cs;
    T bv[NR], acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      bv[r] = (act[r] && colok) ? a.X[(int64_t)i[r] * a.rs + coff] : T(0);
      acc[r] = G == 1 ? bv[r] : T(0);
    }
    for (int t = 0; t < iters; ++t) {
      int d[NR];
      T v[NR], xd[NR];
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const int j = jb[r] + t * G;
        const bool live = j < je[r];
        d[r] = live ? a.ci[j] : -1;
        v[r] = live ? a.val[j] : T(0);
      }
#pragma unroll
      for (int r = 0; r < NR; ++r)
        xd[r] = (d[r] >= 0 && colok) ? ld_x<SYNC>(&a.X[(int64_t)d[r] * a.rs + coff])
                                     : T(0);
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        if (d[r] >= 0) {
          if (G == 1) acc[r] -= v[r] * xd[r];
          else acc[r] += v[r] * xd[r];
        }
      }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      T xi;
      if (G == 1) {
        xi = acc[r] / diag[r];
      } else {
        const T sum = lane_sum<T, G, KT>(acc[r]);
        xi = (bv[r] - sum) / diag[r];
      }
      if (act[r] && colok && s == 0) st_x<SYNC>(&a.X[(int64_t)i[r] * a.rs + coff], xi);
    }
  }
}

// Next unit of this wave: round-robin (co-resident grid) or ticket.
__device__ __forceinline__ int next_unit(int *state, int ticketed, int &cursor,
                                         int stride) {
  if (ticketed) {
    int u = 0;
    if ((threadIdx.x & 63) == 0) u = atomicAdd(&state[ST_TICKET], 1);
    return __shfl(u, 0);
  }
  const int u = cursor;
  cursor += stride;
  return u;
}

template <typename T, int G, int KT, int NR>
__global__ __launch_bounds__(256) void k_fwd_syncfree(FwdArgs<T> a) {
  const int lane = threadIdx.x & 63;
  const int target = ld_dev(&a.state[ST_EPOCH]) + 1;
  int cursor = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int stride = a.total_waves;
  for (;;) {
    const int u = next_unit(a.state, a.ticketed, cursor, stride);
    if (u >= a.U) break;
    wait_units(a.dptr, a.dlist, u, a.uflags, target, a.state, a.timeout);
    fwd_rows<T, G, KT, NR, true>(a, a.ub[u], a.ub[u + 1]);
    publish_fence();
    if (lane == 0) st_dev(&a.uflags[u], target);
  }
  persistent_finish(a.state, a.total_waves);
}

template <typename T, int G, int KT, int NR>
__global__ __launch_bounds__(256) void k_fwd_level(FwdArgs<T> a, int q0, int q1) {
  constexpr int R = NR * 64 / (G * KT);
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int qs = q0 + wave * R;
  if (qs >= q1) return;  // uniform across the wave
  fwd_rows<T, G, KT, NR, false>(a, qs, qs + R < q1 ? qs + R : q1);
}

// --------------------------------------------------------------- adjoint
// Rows at positions [q0, q1) of the transposed solve, G lanes per row, NR
// rows per lane; then grad over L's pattern for the same rows.
template <typename T, int G, int NR, bool SYNC>
__device__ __forceinline__ void adj_rows(const AdjArgs<T> &a, int q0, int q1) {
  constexpr int RPW = 64 / G;
  const int lane = threadIdx.x & 63;
  const int rr = lane / G;
  const int s = lane % G;
  int i[NR], tb[NR], te[NR];
  T diag[NR], g[NR], acc[NR];
  bool act[NR];
  int iters = 0;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int q = q0 + rr + r * RPW;
    act[r] = q < q1;
    i[r] = 0;
    tb[r] = 0;
    te[r] = 0;
    diag[r] = T(1);
    g[r] = T(0);
    if (act[r]) {
      i[r] = a.perm[q];
      const int b0 = a.trp[q];
      tb[r] = b0 + 1 + s;
      te[r] = a.trp[q + 1];
      diag[r] = a.tval[b0];
      g[r] = a.gx[i[r]];
      const int it = (te[r] - tb[r] + G - 1) / G;
      iters = it > iters ? it : iters;
    }
    acc[r] = G == 1 ? g[r] : T(0);
  }
  for (int t = 0; t < iters; ++t) {
    int d[NR];
    T v[NR], gr[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int j = tb[r] + t * G;
      const bool live = j < te[r];
      d[r] = live ? a.tci[j] : -1;
      v[r] = live ? a.tval[j] : T(0);
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) gr[r] = d[r] >= 0 ? ld_x<SYNC>(&a.gb[d[r]]) : T(0);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      if (d[r] >= 0) {
        if (G == 1) acc[r] -= v[r] * gr[r];
        else acc[r] += v[r] * gr[r];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    T gbi;
    if (G == 1) {
      gbi = acc[r] / diag[r];
    } else {
      const T sum = lane_sum<T, G, 1>(acc[r]);
      gbi = (g[r] - sum) / diag[r];
    }
    if (act[r] && s == 0) st_x<SYNC>(&a.gb[i[r]], gbi);
    if (a.gL && act[r]) {
      const int q = q0 + rr + r * RPW;
      const int fb = a.rp[q], len = a.rp[q + 1] - fb;
      const int ob = a.orp[i[r]];
      for (int t = s; t < len; t += G) a.gL[ob + t] = -(gbi * a.x[a.ci[fb + t]]);
    }
  }
}

template <typename T, int G, int NR>
__global__ __launch_bounds__(256) void k_adj_syncfree(AdjArgs<T> a) {
  const int lane = threadIdx.x & 63;
  const int target = ld_dev(&a.state[ST_EPOCH]) + 1;
  int cursor = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int stride = a.total_waves;
  for (;;) {
    const int t = next_unit(a.state, a.ticketed, cursor, stride);
    if (t >= a.U) break;
    const int u = a.U - 1 - t;  // reverse level order
    wait_units(a.dptr, a.dlist, u, a.uflags, target, a.state, a.timeout);
    adj_rows<T, G, NR, true>(a, a.ub[u], a.ub[u + 1]);
    publish_fence();
    if (lane == 0) st_dev(&a.uflags[u], target);
  }
  persistent_finish(a.state, a.total_waves);
}

template <typename T, int G, int NR>
__global__ __launch_bounds__(256) void k_adj_level(AdjArgs<T> a, int q0, int q1) {
  constexpr int R = NR * 64 / G;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int qs = q0 + wave * R;
  if (qs >= q1) return;
  adj_rows<T, G, NR, false>(a, qs, qs + R < q1 ? qs + R : q1);
}

// ------------------------------------------------- caller schedules
// Executes a caller-given 3-array coarsened schedule with the semantics of
// sparse_blas/sptrsv.cpp:57-75 (sptrsv_csr_lbc): one launch per coarse
// level, one lane per partition, the rows of a partition solved in their
// listed order by that lane.  Bit-identical to the serial solve.
__global__ __launch_bounds__(256) void k_lbc_level(
    const int *__restrict__ Lp, const int *__restrict__ Li,
    const double *__restrict__ Lx, double *x, const int *__restrict__ par_ptr,
    const int *__restrict__ partition, int p0, int p1) {
  const int part = p0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (part >= p1) return;
  for (int k1 = par_ptr[part]; k1 < par_ptr[part + 1]; ++k1) {
    const int i = partition[k1];
    double xi = x[i];
    const int je = Lp[i + 1] - 1;
    for (int j = Lp[i]; j < je; ++j) xi -= Lx[j] * x[Li[j]];
    x[i] = xi / Lx[je];
  }
}

// ------------------------------------------------------------ dispatch

// Co-resident grid of 256-thread blocks for a persistent kernel.
int resident_grid(symamd_handle h, const void *kernel, int units) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  int64_t cap = (int64_t)per_cu * h->num_cus;
  int64_t want = (units + 3) / 4;
  if (want < 1) want = 1;
  return (int)(want < cap ? want : cap);
}

bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
  return st != hipStreamCaptureStatusNone;
}

// Launch a persistent kernel: cooperatively (round-robin units) when
// possible, else as a plain launch taking units from the ticket counter.
template <typename Args>
int launch_persistent(symamd_handle h, const void *kfn, Args &a) {
  const int grid = resident_grid(h, kfn, a.U);
  a.total_waves = grid * 4;
  if (!capturing(h->stream) && !h->force_tickets) {
    a.ticketed = 0;
    void *args[] = {&a};
    hipError_t e = hipLaunchCooperativeKernel(kfn, dim3(grid), dim3(256), args, 0,
                                              h->stream);
    if (e == hipSuccess) return SYMAMD_OK;
    (void)hipGetLastError();
  }
  a.ticketed = 1;
  void *args[] = {&a};
  SYMAMD_HIP(hipLaunchKernel(kfn, dim3(grid), dim3(256), args, 0, h->stream));
  return SYMAMD_OK;
}

template <typename T, int G, int KT, int NR>
int run_fwd(symamd_handle h, symamd_plan p, FwdArgs<T> a, int exec) {
  constexpr int R = NR * 64 / (G * KT);
  if (exec == SYMAMD_EXEC_SYNCFREE) {
    UnitSet *us = nullptr;
    SYMAMD_TRY(plan_units(h, p, R, false, &us));
    a.ub = us->d_ub;
    a.U = us->count;
    a.dptr = us->d_dptr;
    a.dlist = us->d_dlist;
    a.uflags = us->d_uflags;
    return launch_persistent(h, (const void *)&k_fwd_syncfree<T, G, KT, NR>, a);
  }
  for (int l = 0; l < p->levels; ++l) {
    const int q0 = p->h_lvl_ptr[l], q1 = p->h_lvl_ptr[l + 1];
    const int64_t waves = (q1 - q0 + R - 1) / R;
    const int grid = (int)((waves + 3) / 4);
    k_fwd_level<T, G, KT, NR><<<grid, 256, 0, h->stream>>>(a, q0, q1);
  }
  SYMAMD_HIP(hipGetLastError());
  return SYMAMD_OK;
}

// Rows per lane: enough rows in flight per wave to hide memory latency.
template <int LPR>
constexpr int rows_per_lane() {
  return LPR >= 16 ? 8 : (LPR >= 4 ? 4 : (LPR >= 2 ? 2 : 1));
}

template <typename T, int KT>
int fwd_pick_g(symamd_handle h, symamd_plan p, FwdArgs<T> a, int exec, int g) {
#define SYMAMD_FWD_G(GV)                                                       \
  case GV:                                                                     \
    if constexpr (GV * KT <= 64)                                               \
      return run_fwd<T, GV, KT, rows_per_lane<GV * KT>()>(h, p, a, exec);      \
    break;
  switch (g) {
    SYMAMD_FWD_G(1)
    SYMAMD_FWD_G(2)
    SYMAMD_FWD_G(4)
    SYMAMD_FWD_G(8)
    SYMAMD_FWD_G(16)
    SYMAMD_FWD_G(32)
    SYMAMD_FWD_G(64)
    default: break;
  }
#undef SYMAMD_FWD_G
  return fail(SYMAMD_ERR_INVALID_ARG, "lanes per row %d not supported with a %d-column tile", g, KT);
}

template <typename T>
int fwd_dispatch(symamd_handle h, symamd_plan p, FwdArgs<T> a, int exec,
                 int g, int kt) {
  switch (kt) {
    case 1: return fwd_pick_g<T, 1>(h, p, a, exec, g);
    case 2: return fwd_pick_g<T, 2>(h, p, a, exec, g);
    case 4: return fwd_pick_g<T, 4>(h, p, a, exec, g);
    case 8: return fwd_pick_g<T, 8>(h, p, a, exec, g);
    case 16: return fwd_pick_g<T, 16>(h, p, a, exec, g);
    case 32: return fwd_pick_g<T, 32>(h, p, a, exec, g);
    case 64: return fwd_pick_g<T, 64>(h, p, a, exec, g);
    default: break;
  }
  return fail(SYMAMD_ERR_INVALID_ARG, "column tile %d not supported", kt);
}

template <typename T, int G>
int run_adj(symamd_handle h, symamd_plan p, AdjArgs<T> a, int exec) {
  constexpr int NR = 1;
  constexpr int R = NR * 64 / G;
  if (exec == SYMAMD_EXEC_SYNCFREE) {
    UnitSet *us = nullptr;
    SYMAMD_TRY(plan_units(h, p, R, true, &us));
    a.ub = us->d_ub;
    a.U = us->count;
    a.dptr = us->d_rdptr;
    a.dlist = us->d_rdlist;
    a.uflags = us->d_uflags;
    return launch_persistent(h, (const void *)&k_adj_syncfree<T, G, NR>, a);
  }
  for (int l = p->levels - 1; l >= 0; --l) {
    const int q0 = p->h_lvl_ptr[l], q1 = p->h_lvl_ptr[l + 1];
    const int64_t waves = (q1 - q0 + R - 1) / R;
    const int grid = (int)((waves + 3) / 4);
    k_adj_level<T, G, NR><<<grid, 256, 0, h->stream>>>(a, q0, q1);
  }
  SYMAMD_HIP(hipGetLastError());
  return SYMAMD_OK;
}

template <typename T>
int adj_dispatch(symamd_handle h, symamd_plan p, AdjArgs<T> a, int exec, int g) {
  switch (g) {
    case 1: return run_adj<T, 1>(h, p, a, exec);
    case 2: return run_adj<T, 2>(h, p, a, exec);
    case 4: return run_adj<T, 4>(h, p, a, exec);
    case 8: return run_adj<T, 8>(h, p, a, exec);
    case 16: return run_adj<T, 16>(h, p, a, exec);
    case 32: return run_adj<T, 32>(h, p, a, exec);
    case 64: return run_adj<T, 64>(h, p, a, exec);
    default: break;
  }
  return fail(SYMAMD_ERR_INVALID_ARG, "lanes per row %d not supported", g);
}

}  // namespace

int forward_rows_per_unit(int k, int lanes) {
  int kt = pow2_ceil(k < 64 ? (k > 0 ? k : 1) : 64);
  if (lanes * kt > 64) lanes = 64 / kt;
  const int lpr = lanes * kt;
  const int nr = lpr >= 16 ? 8 : (lpr >= 4 ? 4 : (lpr >= 2 ? 2 : 1));
  return nr * 64 / lpr;
}

int launch_lbc(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
               const double *d_Lx, double *d_x, int level_no,
               const int *h_level_ptr, const int *d_par_ptr,
               const int *d_partition) {
  (void)n;
  for (int l = 0; l < level_no; ++l) {
    const int p0 = h_level_ptr[l], p1 = h_level_ptr[l + 1];
    if (p1 <= p0) continue;
    k_lbc_level<<<(p1 - p0 + 255) / 256, 256, 0, h->stream>>>(
        d_Lp, d_Li, d_Lx, d_x, d_par_ptr, d_partition, p0, p1);
  }
  SYMAMD_HIP(hipGetLastError());
  return SYMAMD_OK;
}

int launch_forward(symamd_handle h, symamd_plan p, void *d_B, int k, int64_t rs,
                   int64_t cs, int exec, int lanes) {
  if (p->n == 0 || k == 0) return SYMAMD_OK;
  int kt = pow2_ceil(k < 64 ? k : 64);
  if (lanes * kt > 64) lanes = 64 / kt;
  if (p->dtype == SYMAMD_F64) {
    FwdArgs<double> a{};
    a.n = p->n; a.k = k; a.rs = rs; a.cs = cs;
    a.rp = p->d_rp; a.ci = p->d_ci; a.val = (const double *)p->d_val;
    a.perm = p->d_perm; a.X = (double *)d_B;
    a.state = p->d_state; a.timeout = h->timeout_ticks();
    return fwd_dispatch<double>(h, p, a, exec, lanes, kt);
  }
  FwdArgs<float> a{};
  a.n = p->n; a.k = k; a.rs = rs; a.cs = cs;
  a.rp = p->d_rp; a.ci = p->d_ci; a.val = (const float *)p->d_val;
  a.perm = p->d_perm; a.X = (float *)d_B;
  a.state = p->d_state; a.timeout = h->timeout_ticks();
  return fwd_dispatch<float>(h, p, a, exec, lanes, kt);
}

int launch_adjoint(symamd_handle h, symamd_plan p, const void *d_x,
                   const void *d_gx, void *d_gb, void *d_gL, int exec,
                   int lanes) {
  if (p->n == 0) return SYMAMD_OK;
  SYMAMD_TRY(plan_transpose(h, p));
  if (p->dtype == SYMAMD_F64) {
    AdjArgs<double> a{};
    a.n = p->n; a.trp = p->d_trp; a.tci = p->d_tci;
    a.tval = (const double *)p->d_tval; a.rp = p->d_rp; a.ci = p->d_ci;
    a.orp = p->d_orp; a.perm = p->d_perm; a.x = (const double *)d_x;
    a.gx = (const double *)d_gx; a.gb = (double *)d_gb; a.gL = (double *)d_gL;
    a.state = p->d_state; a.timeout = h->timeout_ticks();
    return adj_dispatch<double>(h, p, a, exec, lanes);
  }
  AdjArgs<float> a{};
  a.n = p->n; a.trp = p->d_trp; a.tci = p->d_tci;
  a.tval = (const float *)p->d_tval; a.rp = p->d_rp; a.ci = p->d_ci;
  a.orp = p->d_orp; a.perm = p->d_perm; a.x = (const float *)d_x;
  a.gx = (const float *)d_gx; a.gb = (float *)d_gb; a.gL = (float *)d_gL;
  a.state = p->d_state; a.timeout = h->timeout_ticks();
  return adj_dispatch<float>(h, p, a, exec, lanes);
}

}  // namespace symamd
