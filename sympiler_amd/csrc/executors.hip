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
// of one row read and write KT consecutive values of that row.
//
// Executors
//   LEVELSET  one launch per level; plain loads and stores (kernel
//             boundaries order the levels).
//   SYNCFREE  one persistent launch.  Waves take work units (a run of rows
//             inside one level) from a ticket counter in level order, wait
//             on per-row ready flags of their dependencies, and publish
//             their own rows' flags.  A unit never spans two levels, so rows
//             of one wave never wait on each other; every dependency of a
//             unit lies in an earlier unit, handed to a wave that is already
//             running, so the grid cannot deadlock.  Waits are bounded by a
//             watchdog on the 100 MHz clock that records an error and lets
//             the wave proceed, so every wave exits.
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
  int *flags;
  int *state;
  const int *ub;             // unit bounds
  int U;
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
  int *flags;
  int *state;
  const int *ub;
  int U;
  uint64_t timeout;
  int total_waves;
};

__device__ __forceinline__ void wait_ready(const int *f, int target, int *state,
                                           uint64_t timeout) {
  if (ld_dev(f) == target) return;
  const uint64_t t0 = wall_ticks();
  int spins = 0;
  while (ld_dev(f) != target) {
    __builtin_amdgcn_s_sleep(1);
    if ((++spins & 63) == 0 && wall_ticks() - t0 > timeout) {
      atomicOr(&state[ST_ERROR], 1);
      break;
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

// Solve the rows of positions [q0 + rr] for one wave.  SYNC selects the
// sync-free protocol (coherent loads, waits, flags) or plain level-set
// loads and stores.
template <typename T, int G, int KT, bool SYNC>
__device__ __forceinline__ void fwd_rows(const FwdArgs<T> &a, int q, bool active,
                                         int target) {
  const int lane = threadIdx.x & 63;
  const int s = (lane / KT) % G;
  const int c = lane % KT;
  int i = 0, jb = 0, je = 0;
  if (active) {
    i = a.perm[q];
    jb = a.rp[q];
    je = a.rp[q + 1] - 1;
  }
  const T diag = active ? a.val[je] : T(1);
  const int64_t ibase = (int64_t)i * a.rs;
  for (int c0 = 0; c0 < a.k; c0 += KT) {
    const int col = c0 + c;
    const bool act = active && col < a.k;
    const int64_t coff = (int64_t)col * a.cs;
    const T b = act ? a.X[ibase + coff] : T(0);
    T acc = G == 1 ? b : T(0);
    for (int j = jb + s; j < je; j += G) {
      const int d = a.ci[j];
      const T v = a.val[j];
      if (SYNC && c0 == 0) wait_ready(&a.flags[d], target, a.state, a.timeout);
      T xd = T(0);
      if (act) {
        const T *px = &a.X[(int64_t)d * a.rs + coff];
        xd = SYNC ? ld_dev(px) : *px;
      }
      if (G == 1) acc -= v * xd;
      else acc += v * xd;
    }
    T xi;
    if (G == 1) {
      xi = acc / diag;
    } else {
      acc = lane_sum<T, G, KT>(acc);
      xi = (b - acc) / diag;
    }
    if (act && (G == 1 || s == 0)) {
      T *po = &a.X[ibase + coff];
      if (SYNC) st_dev(po, xi);
      else *po = xi;
    }
  }
  if (SYNC) {
    publish_fence();
    if (active && s == 0 && c == 0) st_dev(&a.flags[i], target);
  }
}

template <typename T, int G, int KT>
__global__ __launch_bounds__(256) void k_fwd_syncfree(FwdArgs<T> a) {
  constexpr int RPW = 64 / (G * KT);
  const int lane = threadIdx.x & 63;
  const int rr = lane / (G * KT);
  const int target = ld_dev(&a.state[ST_EPOCH]) + 1;
  for (;;) {
    int u = 0;
    if (lane == 0) u = atomicAdd(&a.state[ST_TICKET], 1);
    u = __shfl(u, 0);
    if (u >= a.U) break;
    const int q0 = a.ub[u], q1 = a.ub[u + 1];
    const int q = q0 + rr;
    fwd_rows<T, G, KT, true>(a, q, q < q1 && rr < RPW, target);
  }
  persistent_finish(a.state, a.total_waves);
}

template <typename T, int G, int KT>
__global__ __launch_bounds__(256) void k_fwd_level(FwdArgs<T> a, int q0,
                                                   int q1) {
  const int gl = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = q0 + gl / (G * KT);
  fwd_rows<T, G, KT, false>(a, q, q < q1, 0);
}

// --------------------------------------------------------------- adjoint
template <typename T, int G, bool SYNC>
__device__ __forceinline__ void adj_row(const AdjArgs<T> &a, int q, bool active,
                                        int target) {
  const int lane = threadIdx.x & 63;
  const int s = lane % G;
  int i = 0, tb = 0, te = 0;
  if (active) {
    i = a.perm[q];
    tb = a.trp[q];
    te = a.trp[q + 1];
  }
  const T diag = active ? a.tval[tb] : T(1);
  const T g = active ? a.gx[i] : T(0);
  T acc = G == 1 ? g : T(0);
  for (int t = tb + 1 + s; t < te; t += G) {
    const int r = a.tci[t];
    const T v = a.tval[t];
    if (SYNC) wait_ready(&a.flags[r], target, a.state, a.timeout);
    const T gr = SYNC ? ld_dev(&a.gb[r]) : a.gb[r];
    if (G == 1) acc -= v * gr;
    else acc += v * gr;
  }
  T gbi;
  if (G == 1) {
    gbi = acc / diag;
  } else {
    acc = lane_sum<T, G, 1>(acc);
    gbi = (g - acc) / diag;
  }
  if (active && s == 0) {
    if (SYNC) st_dev(&a.gb[i], gbi);
    else a.gb[i] = gbi;
  }
  if (a.gL && active) {
    const int fb = a.rp[q], len = a.rp[q + 1] - fb;
    const int ob = a.orp[i];
    for (int t = s; t < len; t += G) a.gL[ob + t] = -(gbi * a.x[a.ci[fb + t]]);
  }
  if (SYNC) {
    publish_fence();
    if (active && s == 0) st_dev(&a.flags[i], target);
  }
}

template <typename T, int G>
__global__ __launch_bounds__(256) void k_adj_syncfree(AdjArgs<T> a) {
  constexpr int RPW = 64 / G;
  const int lane = threadIdx.x & 63;
  const int rr = lane / G;
  const int target = ld_dev(&a.state[ST_EPOCH]) + 1;
  for (;;) {
    int t = 0;
    if (lane == 0) t = atomicAdd(&a.state[ST_TICKET], 1);
    t = __shfl(t, 0);
    if (t >= a.U) break;
    const int u = a.U - 1 - t;  // reverse level order
    const int q0 = a.ub[u], q1 = a.ub[u + 1];
    const int q = q0 + rr;
    adj_row<T, G, true>(a, q, q < q1 && rr < RPW, target);
  }
  persistent_finish(a.state, a.total_waves);
}

template <typename T, int G>
__global__ __launch_bounds__(256) void k_adj_level(AdjArgs<T> a, int q0, int q1) {
  const int gl = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = q0 + gl / G;
  adj_row<T, G, false>(a, q, q < q1, 0);
}

// ------------------------------------------------------------ dispatch

int persistent_grid(symamd_handle h, const void *kernel, int units) {
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

template <typename T, int G, int KT>
int run_fwd(symamd_handle h, symamd_plan p, FwdArgs<T> a, int exec) {
  constexpr int LPR = G * KT;
  if (exec == SYMAMD_EXEC_SYNCFREE) {
    UnitSet *us = nullptr;
    SYMAMD_TRY(plan_units(h, p, 64 / LPR, &us));
    a.ub = us->d_ub;
    a.U = us->count;
    const void *kfn = (const void *)&k_fwd_syncfree<T, G, KT>;
    int grid = persistent_grid(h, kfn, a.U);
    a.total_waves = grid * 4;
    k_fwd_syncfree<T, G, KT><<<grid, 256, 0, h->stream>>>(a);
    SYMAMD_HIP(hipGetLastError());
  } else {
    for (int l = 0; l < p->levels; ++l) {
      const int q0 = p->h_lvl_ptr[l], q1 = p->h_lvl_ptr[l + 1];
      const int64_t lanes = (int64_t)(q1 - q0) * LPR;
      const int grid = (int)((lanes + 255) / 256);
      k_fwd_level<T, G, KT><<<grid, 256, 0, h->stream>>>(a, q0, q1);
    }
    SYMAMD_HIP(hipGetLastError());
  }
  return SYMAMD_OK;
}

template <typename T, int KT>
int fwd_pick_g(symamd_handle h, symamd_plan p, FwdArgs<T> a, int exec, int g) {
  switch (g) {
    case 1: return run_fwd<T, 1, KT>(h, p, a, exec);
    case 2: if constexpr (2 * KT <= 64) return run_fwd<T, 2, KT>(h, p, a, exec); break;
    case 4: if constexpr (4 * KT <= 64) return run_fwd<T, 4, KT>(h, p, a, exec); break;
    case 8: if constexpr (8 * KT <= 64) return run_fwd<T, 8, KT>(h, p, a, exec); break;
    case 16: if constexpr (16 * KT <= 64) return run_fwd<T, 16, KT>(h, p, a, exec); break;
    case 32: if constexpr (32 * KT <= 64) return run_fwd<T, 32, KT>(h, p, a, exec); break;
    case 64: if constexpr (64 * KT <= 64) return run_fwd<T, 64, KT>(h, p, a, exec); break;
    default: break;
  }
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
  if (exec == SYMAMD_EXEC_SYNCFREE) {
    UnitSet *us = nullptr;
    SYMAMD_TRY(plan_units(h, p, 64 / G, &us));
    a.ub = us->d_ub;
    a.U = us->count;
    const void *kfn = (const void *)&k_adj_syncfree<T, G>;
    int grid = persistent_grid(h, kfn, a.U);
    a.total_waves = grid * 4;
    k_adj_syncfree<T, G><<<grid, 256, 0, h->stream>>>(a);
    SYMAMD_HIP(hipGetLastError());
  } else {
    for (int l = p->levels - 1; l >= 0; --l) {
      const int q0 = p->h_lvl_ptr[l], q1 = p->h_lvl_ptr[l + 1];
      const int64_t lanes = (int64_t)(q1 - q0) * G;
      const int grid = (int)((lanes + 255) / 256);
      k_adj_level<T, G><<<grid, 256, 0, h->stream>>>(a, q0, q1);
    }
    SYMAMD_HIP(hipGetLastError());
  }
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

}  // namespace

int forward_rows_per_unit(int k, int lanes) {
  int kt = pow2_ceil(k < 64 ? (k > 0 ? k : 1) : 64);
  if (lanes * kt > 64) lanes = 64 / kt;
  return 64 / (lanes * kt);
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
    a.perm = p->d_perm; a.X = (double *)d_B; a.flags = p->d_flags;
    a.state = p->d_state; a.timeout = h->timeout_ticks();
    return fwd_dispatch<double>(h, p, a, exec, lanes, kt);
  }
  FwdArgs<float> a{};
  a.n = p->n; a.k = k; a.rs = rs; a.cs = cs;
  a.rp = p->d_rp; a.ci = p->d_ci; a.val = (const float *)p->d_val;
  a.perm = p->d_perm; a.X = (float *)d_B; a.flags = p->d_flags;
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
    a.flags = p->d_flags; a.state = p->d_state; a.timeout = h->timeout_ticks();
    return adj_dispatch<double>(h, p, a, exec, lanes);
  }
  AdjArgs<float> a{};
  a.n = p->n; a.trp = p->d_trp; a.tci = p->d_tci;
  a.tval = (const float *)p->d_tval; a.rp = p->d_rp; a.ci = p->d_ci;
  a.orp = p->d_orp; a.perm = p->d_perm; a.x = (const float *)d_x;
  a.gx = (const float *)d_gx; a.gb = (float *)d_gb; a.gL = (float *)d_gL;
  a.flags = p->d_flags; a.state = p->d_state; a.timeout = h->timeout_ticks();
  return adj_dispatch<float>(h, p, a, exec, lanes);
}

}  // namespace symamd
