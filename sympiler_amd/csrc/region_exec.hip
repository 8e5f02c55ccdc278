// region_exec.hip -- region-schedule forward solve for gfx950.
//
// Work = tasks.  A task is one region (a compact cluster of at most `cap`
// rows, built by regions.cpp) times one group of at most KG_MAX right-hand
// side columns; column groups are independent solves, so a task waits only
// on the same column group of the regions it reads.  One wavefront solves
// one task entirely in LDS:
//   1. stage the region's entries (LDS slot code + value), diagonals, row
//      ids, slice bounds and the group's right-hand sides -- none of it
//      depends on other tasks, so it overlaps the wait;
//   2. wait until the tasks it reads from carry the launch's tag;
//   3. load the halo (the values of those regions' rows it reads) into LDS;
//   4. solve the internal levels (slices) one after another; the lanes of a
//      slice take (row, column) pairs, column fastest, every dependency
//      read from LDS; slices are ordered by the wave's own LDS ordering,
//      with no cross-wave hand-off;
//   5. store the rows, wait for the stores, publish the task's tag.
// Every global -> LDS copy keeps STAGE_U loads per lane in flight, so a
// task pays a few memory round trips instead of one per 64 values.
//
// Row arithmetic is the reference's (sparse_blas/sptrsv.cpp:17-25): the
// right-hand side minus each product in the row's stored order, then the
// division by the diagonal, one lane per (row, column), compiled without FMA
// contraction -- bit-identical to sptrsv_csr.
//
// Tasks are ordered by region rank (topological) then column group and
// taken by ticket, so the lowest unfinished task is always held by a
// running wave with all its inputs published: the grid cannot deadlock,
// whether or not all its blocks are resident (these blocks carry up to
// 64 KiB of LDS each), and the watchdog of wait_units bounds every wait.
#include "device.h"

namespace symamd {

namespace {

constexpr int KG_MAX = 8;   // right-hand side columns per task
constexpr int STAGE_U = 8;  // loads per lane in flight while staging
constexpr int PK = RegionPlan::PACK;

template <typename T>
struct RegionArgs {
  int k;
  int64_t rs, cs;  // element (i, c) at X[i*rs + c*cs]
  const T *B;      // right-hand sides (the same buffer as X: in place)
  T *X;
  const int *rptr, *rows, *sptr, *slices, *hptr, *halo, *dptr, *dlist, *eptr, *ecode;
  const int *pcnt, *pcode;
  const T *pval, *eval, *diag;
  int *tflags;     // task tags: region rank * groups + group
  int *state;
  int R, groups, smax, hmax, emax, slmax;
  int total_waves;
  uint64_t timeout;
  // optional per-task trace (100 MHz clock): [3t] task taken, [3t+1] its
  // inputs published, [3t+2] the task published
  uint64_t *trace;
  int64_t trace_cap;
};

template <int KT>
constexpr int log2c() {
  return KT <= 1 ? 0 : 1 + log2c<KT / 2>();
}

// Columns per task for k right-hand sides.
int task_cols(int k) { return k >= KG_MAX ? KG_MAX : pow2_ceil(k < 1 ? 1 : k); }

// Byte offsets of one task's LDS buffers (each 16-byte aligned), the same on
// host (launch size) and device (carving): values first, then ints.
struct RegionLds {
  size_t xs, pval, lval, ldiag, pcode, pcnt, lcode, lep, lsl, lrow, lhalo, total;
};

__host__ __device__ inline size_t lds_align(size_t v) { return (v + 15) & ~size_t(15); }

__host__ __device__ inline RegionLds region_lds(int smax, int hmax, int emax, int slmax,
                                                int kt, size_t vs) {
  RegionLds L;
  size_t o = 0;
  L.xs = o;    o = lds_align(o + (size_t)(smax + hmax) * kt * vs);  // values, halo
  L.pval = o;  o = lds_align(o + (size_t)PK * smax * vs);            // packed values
  L.lval = o;  o = lds_align(o + (size_t)emax * vs);                 // overflow values
  L.ldiag = o; o = lds_align(o + (size_t)smax * vs);
  L.pcode = o; o = lds_align(o + (size_t)PK * smax * 4);             // packed slots
  L.pcnt = o;  o = lds_align(o + (size_t)smax * 4);
  L.lcode = o; o = lds_align(o + (size_t)emax * 4);                  // overflow slots
  L.lep = o;   o = lds_align(o + (size_t)(smax + 1) * 4);
  L.lsl = o;   o = lds_align(o + (size_t)(slmax + 1) * 4);
  L.lrow = o;  o = lds_align(o + (size_t)smax * 4);
  L.lhalo = o; o = lds_align(o + (size_t)hmax * 4);
  L.total = o;
  return L;
}

// LDS bytes of one task's buffers.
size_t region_lds_bytes(const RegionPlan &rp, int kt, size_t vs) {
  return region_lds(rp.smax, rp.hmax, rp.emax, rp.slmax, kt, vs).total;
}

// LDS writes of the block are visible to all its threads (one wave: the
// wave's own ordering suffices).
template <int NT>
__device__ __forceinline__ void block_sync() {
  if constexpr (NT == 64)
    wave_lds_sync();
  else
    __syncthreads();
}

// dst[t] = src[t] for t < count (LDS <- global), STAGE_U loads per thread
// in flight, NT threads.
template <int NT, typename V>
__device__ __forceinline__ void stage(V *dst, const V *__restrict__ src, int count) {
  const int lane = threadIdx.x;
  for (int t0 = lane; t0 < count; t0 += NT * STAGE_U) {
    V v[STAGE_U];
#pragma unroll
    for (int u = 0; u < STAGE_U; ++u) {
      const int t = t0 + NT * u;
      v[u] = t < count ? src[t] : V(0);
    }
#pragma unroll
    for (int u = 0; u < STAGE_U; ++u) {
      const int t = t0 + NT * u;
      if (t < count) dst[t] = v[u];
    }
  }
}

// xs[l*KT + c] = M[rows[l]*rs + (c0 + c)*cs] for the nr rows and the kc
// columns of the group (LDS <- global), STAGE_U loads in flight; coherent
// loads when COH (values other waves published in this launch).  `rows`
// lies in LDS: a row id read from global memory would make every value load
// wait for it (and, through the in-order load counter, for all loads
// before it).
template <int NT, typename T, int KT, bool COH>
__device__ __forceinline__ void stage_rows(T *xs, const T *M, const int *rows, int nr,
                                           int kc, int c0, int64_t rs, int64_t cs) {
  constexpr int KS = log2c<KT>();
  const int lane = threadIdx.x;
  const int count = nr * KT;
  for (int t0 = lane; t0 < count; t0 += NT * STAGE_U) {
    T v[STAGE_U];
#pragma unroll
    for (int u = 0; u < STAGE_U; ++u) {
      const int t = t0 + NT * u;
      const int c = t & (KT - 1);
      v[u] = T(0);
      if (t < count && c < kc) {
        const T *p = &M[(int64_t)rows[t >> KS] * rs + (int64_t)(c0 + c) * cs];
        v[u] = COH ? ld_dev(p) : *p;
      }
    }
#pragma unroll
    for (int u = 0; u < STAGE_U; ++u) {
      const int t = t0 + NT * u;
      if (t < count) xs[t] = v[u];
    }
  }
}

// NT threads (one or four waves) per task.
template <typename T, int KT, int NT>
__global__ __launch_bounds__(NT) void k_regions(RegionArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int KS = log2c<KT>();
  // two values per thread in flight for tasks of 8 columns (their slices
  // hold up to 2 (row, column) pairs per thread), one for narrower tasks
  constexpr int VPL = KT >= 8 ? 2 : 1;
  const RegionLds L = region_lds(a.smax, a.hmax, a.emax, a.slmax, KT, sizeof(T));
  T *xs = reinterpret_cast<T *>(smem + L.xs);        // (smax + hmax) * KT values
  T *pval = reinterpret_cast<T *>(smem + L.pval);    // PK per row: packed values
  T *lval = reinterpret_cast<T *>(smem + L.lval);    // emax overflow values
  T *ldiag = reinterpret_cast<T *>(smem + L.ldiag);  // smax diagonals
  int *pcode = reinterpret_cast<int *>(smem + L.pcode);  // PK per row: LDS slots
  int *pcnt = reinterpret_cast<int *>(smem + L.pcnt);    // entries per row
  int *lcode = reinterpret_cast<int *>(smem + L.lcode);  // emax overflow slots
  int *lep = reinterpret_cast<int *>(smem + L.lep);      // smax + 1 overflow bounds
  int *lsl = reinterpret_cast<int *>(smem + L.lsl);      // slmax + 1 slice bounds
  int *lrow = reinterpret_cast<int *>(smem + L.lrow);    // smax row ids
  int *lhalo = reinterpret_cast<int *>(smem + L.lhalo);  // hmax halo row ids
  __shared__ int s_task;
  const int lane = threadIdx.x;
  const int target = ld_dev(&a.state[ST_EPOCH]) + 1;
  const int ntasks = a.R * a.groups;
  int cursor = 0;
  for (;;) {
    int task;
    if constexpr (NT == 64) {
      task = next_unit(a.state, 1, cursor, 0);
    } else {
      // the previous task's last block_sync ordered every read of s_task
      if (lane == 0) s_task = atomicAdd(&a.state[ST_TICKET], 1);
      __syncthreads();
      task = s_task;
    }
    if (task >= ntasks) break;
    trace_mark(a.trace, a.trace_cap, task, 0);
    const int r = task / a.groups, g = task - r * a.groups;
    const int c0 = g * KT;
    const int kc = a.k - c0 < KT ? a.k - c0 : KT;
    const int base = a.rptr[r], nr = a.rptr[r + 1] - base;
    const int e0 = a.eptr[base], ne = a.eptr[base + nr] - e0;
    const int s0 = a.sptr[r], ns = a.sptr[r + 1] - s0 - 1;
    const int h0 = a.hptr[r], nh = a.hptr[r + 1] - h0;
    // 1. stage
    stage<NT>(pcode, a.pcode + (size_t)PK * base, PK * nr);
    stage<NT>(pval, a.pval + (size_t)PK * base, PK * nr);
    stage<NT>(pcnt, a.pcnt + base, nr);
    stage<NT>(lcode, a.ecode + e0, ne);
    stage<NT>(lval, a.eval + e0, ne);
    stage<NT>(lrow, a.rows + base, nr);
    stage<NT>(ldiag, a.diag + base, nr);
    stage<NT>(lsl, a.slices + s0, ns + 1);
    stage<NT>(lhalo, a.halo + h0, nh);
    for (int t = lane; t <= nr; t += NT) lep[t] = a.eptr[base + t] - e0;
    block_sync<NT>();
    stage_rows<NT, T, KT, false>(xs, a.B, lrow, nr, kc, c0, a.rs, a.cs);
    // 2. inputs published (same column group of the regions read); one wave
    // polls, the others wait at the barrier
    if (lane < 64)
      wait_units(a.dptr, a.dlist, r, a.tflags, target, a.state, a.timeout, a.groups, g);
    if constexpr (NT > 64) __syncthreads();
    trace_mark(a.trace, a.trace_cap, task, 1);
    // 3. halo
    stage_rows<NT, T, KT, true>(xs + (size_t)a.smax * KT, a.X, lhalo, nh, kc, c0, a.rs,
                                a.cs);
    block_sync<NT>();
    // 4. internal levels.  Per (row, column) pair: the row's packed slot
    // codes and values (addresses known from the row alone, one round of LDS
    // reads), then the dependency values, then the products in entry order
    // (padding slots read slot 0 and are not accumulated), then the row's
    // overflow entries in order, then the division.  VPL pairs per thread
    // are in flight together.
    for (int s = 0; s < ns; ++s) {
      const int q1 = lsl[s + 1] * KT;
      for (int t0 = lsl[s] * KT + lane; t0 < q1; t0 += NT * VPL) {
        int ll[VPL], cc[VPL], cnt[VPL], cd[VPL][PK];
        T acc[VPL], pv[VPL][PK], xv[VPL][PK];
        bool ok[VPL];
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
          const int t = t0 + NT * v;
          cc[v] = t & (KT - 1);
          ok[v] = t < q1 && cc[v] < kc;
          ll[v] = ok[v] ? t >> KS : 0;
          cnt[v] = pcnt[ll[v]];
#pragma unroll
          for (int u = 0; u < PK; ++u) {
            cd[v][u] = pcode[PK * ll[v] + u];
            pv[v][u] = pval[PK * ll[v] + u];
          }
          acc[v] = ok[v] ? xs[t] : T(0);
        }
#pragma unroll
        for (int v = 0; v < VPL; ++v)
#pragma unroll
          for (int u = 0; u < PK; ++u) xv[v][u] = xs[cd[v][u] * KT + cc[v]];
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
#pragma unroll
          for (int u = 0; u < PK; ++u)
            if (u < cnt[v]) acc[v] -= pv[v][u] * xv[v][u];
          if (cnt[v] > PK)
            for (int e = lep[ll[v]]; e < lep[ll[v] + 1]; ++e)
              acc[v] -= lval[e] * xs[lcode[e] * KT + cc[v]];
          if (ok[v]) xs[t0 + NT * v] = acc[v] / ldiag[ll[v]];
        }
      }
      block_sync<NT>();
    }
    // 5. store; every wave waits for its own stores, then one thread
    // publishes (the barrier also ends every read of this task's LDS)
    for (int t = lane; t < nr * KT; t += NT) {
      const int c = t & (KT - 1);
      if (c < kc)
        st_dev(&a.X[(int64_t)lrow[t >> KS] * a.rs + (int64_t)(c0 + c) * a.cs], xs[t]);
    }
    publish_fence();
    if constexpr (NT > 64) __syncthreads();
    if (lane == 0) st_dev(&a.tflags[task], target);
    trace_mark(a.trace, a.trace_cap, task, 2);
  }
  persistent_finish(a.state, a.total_waves);
}

template <typename T>
__global__ void k_region_gather(int64_t m, const int *__restrict__ src,
                                const T *__restrict__ from, T *to) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) to[t] = from[src[t]];
}

template <typename T>
int gather_one(symamd_handle h, symamd_plan p, RegionPlan &rp) {
  const T *val = (const T *)p->d_val;
  if (rp.entries > 0)
    k_region_gather<T><<<(unsigned)((rp.entries + 255) / 256), 256, 0, h->stream>>>(
        rp.entries, rp.d_esrc, val, (T *)rp.d_eval);
  const int64_t packed = (int64_t)PK * p->n;
  k_region_gather<T><<<(unsigned)((packed + 255) / 256), 256, 0, h->stream>>>(
      packed, rp.d_psrc, val, (T *)rp.d_pval);
  k_region_gather<T><<<(unsigned)((p->n + 255) / 256), 256, 0, h->stream>>>(
      p->n, rp.d_dsrc, val, (T *)rp.d_diag);
  SYMAMD_HIP(hipGetLastError());
  return SYMAMD_OK;
}

// Tasks are taken by ticket, so the grid may be as large as the occupancy
// estimate (a block that is not resident yet simply takes later tickets).
// One wave per task: a four-wave block per task (NT = 256) did not complete
// on the 32-column 128^3 solve and is not used until the cause is known.
template <typename T, int KT>
int region_launch(symamd_handle h, symamd_plan p, const RegionPlan &rp, void *d_B,
                  int k, int64_t rs, int64_t cs, size_t lds) {
  constexpr int NT = 64;
  const void *kfn = (const void *)&k_regions<T, KT, NT>;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, NT, lds) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  const int groups = (k + KT - 1) / KT;
  const int64_t ntasks = (int64_t)rp.R * groups;
  int64_t grid = (int64_t)per_cu * h->num_cus;
  if (h->max_waves > 0 && h->max_waves < grid) grid = h->max_waves;
  if (grid > ntasks) grid = ntasks;
  if (grid < 1) grid = 1;
  RegionArgs<T> a{};
  a.k = k;
  a.rs = rs;
  a.cs = cs;
  a.B = (const T *)d_B;
  a.X = (T *)d_B;
  a.rptr = rp.d_rptr;
  a.rows = rp.d_rows;
  a.sptr = rp.d_sptr;
  a.slices = rp.d_slices;
  a.hptr = rp.d_hptr;
  a.halo = rp.d_halo;
  a.dptr = rp.d_dptr;
  a.dlist = rp.d_dlist;
  a.eptr = rp.d_eptr;
  a.ecode = rp.d_ecode;
  a.pcnt = rp.d_pcnt;
  a.pcode = rp.d_pcode;
  a.pval = (const T *)rp.d_pval;
  a.eval = (const T *)rp.d_eval;
  a.diag = (const T *)rp.d_diag;
  a.tflags = rp.d_rflags;
  a.state = p->d_state;
  a.R = rp.R;
  a.groups = groups;
  a.smax = rp.smax;
  a.hmax = rp.hmax;
  a.emax = rp.emax;
  a.slmax = rp.slmax;
  a.total_waves = (int)grid * (NT / 64);
  a.timeout = h->timeout_ticks();
  a.trace = h->trace;
  a.trace_cap = h->trace_cap;
  void *args[] = {&a};
  SYMAMD_HIP(hipLaunchKernel(kfn, dim3((unsigned)grid), dim3(NT), args, lds, h->stream));
  return SYMAMD_OK;
}

template <typename T>
int region_dispatch(symamd_handle h, symamd_plan p, const RegionPlan &rp, void *d_B,
                    int k, int64_t rs, int64_t cs, int kt, size_t lds) {
  switch (kt) {
    case 1: return region_launch<T, 1>(h, p, rp, d_B, k, rs, cs, lds);
    case 2: return region_launch<T, 2>(h, p, rp, d_B, k, rs, cs, lds);
    case 4: return region_launch<T, 4>(h, p, rp, d_B, k, rs, cs, lds);
    default: return region_launch<T, 8>(h, p, rp, d_B, k, rs, cs, lds);
  }
}

}  // namespace

// Default cap: 512 rows for one or two columns per task, 128 for eight
// (measured on the 128^3 IC(0) factor: k=1 is fastest at 512 -- 1024 and 256
// are slower -- and k=32 at 128, where one wave's LDS still allows seven
// waves per CU).
int region_cap(symamd_handle h, int k) {
  if (h->region_cap > 0) return h->region_cap;
  const int cap = pow2_floor(1024 / task_cols(k));
  return cap > 512 ? 512 : (cap < 16 ? 16 : cap);
}

bool region_preferred(symamd_handle h, symamd_plan p, int k, int layout) {
  if (k > 1 && layout == SYMAMD_COL_MAJOR) return false;
  if (!region_usable(h, p, k)) return false;
  RegionPlan *rp = nullptr;
  if (plan_regions(h, p, region_cap(h, k), &rp) != SYMAMD_OK) return false;
  return rp->inside >= 0.6;
}

int gather_region_values(symamd_handle h, symamd_plan p) {
  for (auto &kv : p->rplans) {
    RegionPlan &rp = kv.second;
    if (rp.R == 0) continue;
    if (p->dtype == SYMAMD_F64)
      SYMAMD_TRY(gather_one<double>(h, p, rp));
    else
      SYMAMD_TRY(gather_one<float>(h, p, rp));
  }
  return SYMAMD_OK;
}

bool region_usable(symamd_handle h, symamd_plan p, int k) {
  if (k < 1 || k > KG_MAX * RegionPlan::MAX_GROUPS) return false;
  RegionPlan *rp = nullptr;
  if (plan_regions(h, p, region_cap(h, k), &rp) != SYMAMD_OK || rp->R == 0) return false;
  const size_t vs = p->dtype == SYMAMD_F64 ? 8 : 4;
  return region_lds_bytes(*rp, task_cols(k), vs) <= 64 * 1024;
}

int launch_regions(symamd_handle h, symamd_plan p, void *d_B, int k, int64_t rs,
                   int64_t cs) {
  if (!region_usable(h, p, k))
    return fail(SYMAMD_ERR_NOT_SUPPORTED, "no region schedule for this matrix and k");
  RegionPlan *rp = nullptr;
  SYMAMD_TRY(plan_regions(h, p, region_cap(h, k), &rp));
  const int kt = task_cols(k);
  const size_t vs = p->dtype == SYMAMD_F64 ? 8 : 4;
  const size_t lds = region_lds_bytes(*rp, kt, vs);
  if (p->dtype == SYMAMD_F64)
    return region_dispatch<double>(h, p, *rp, d_B, k, rs, cs, kt, lds);
  return region_dispatch<float>(h, p, *rp, d_B, k, rs, cs, kt, lds);
}

}  // namespace symamd
