// region_exec.hip -- region-schedule forward solve for gfx950.
//
// One wavefront solves one region (a compact cluster of at most `cap` rows,
// built by regions.cpp) entirely in LDS:
//   1. stage the region's entries (LDS slot code + value), diagonals, row
//      ids, slice bounds and right-hand sides -- none of it depends on other
//      regions, so it overlaps the wait;
//   2. wait until every region it reads from carries the launch's tag;
//   3. load the halo (the values of those regions' rows it reads) into LDS;
//   4. solve the internal levels (slices) one after another; the lanes of a
//      slice take (row, column) pairs, column fastest, and read every
//      dependency from LDS; slices are ordered by the wave's own LDS
//      ordering, no cross-wave hand-off;
//   5. store the rows, wait for the stores, publish the region's tag.
// Row arithmetic is the reference's (sparse_blas/sptrsv.cpp:17-25): the
// right-hand side minus each product in the row's stored order, then the
// division by the diagonal, one lane per (row, column), compiled without FMA
// contraction -- bit-identical to sptrsv_csr.
//
// Regions are ranked topologically and taken by ticket, so the lowest
// unfinished region is always held by a running wave with all its inputs
// published: the grid cannot deadlock, whether or not all its blocks are
// resident, and the watchdog of wait_units bounds every wait.
#include "device.h"

namespace symamd {

namespace {

template <typename T>
struct RegionArgs {
  int k;
  int64_t rs, cs;  // element (i, c) at X[i*rs + c*cs]
  const T *B;      // right-hand sides (the same buffer as X: in place)
  T *X;
  const int *rptr, *rows, *sptr, *slices, *hptr, *halo, *dptr, *dlist, *eptr, *ecode;
  const T *eval, *diag;
  int *rflags;
  int *state;
  int R, smax, hmax, emax, slmax;
  int ticketed;
  int total_waves;
  uint64_t timeout;
};

template <int KT>
constexpr int log2c() {
  return KT <= 1 ? 0 : 1 + log2c<KT / 2>();
}

// LDS bytes of one wave's region buffers.
size_t region_lds_bytes(const RegionPlan &rp, int kt, size_t vs) {
  return ((size_t)(rp.smax + rp.hmax) * kt + rp.emax + rp.smax) * vs +
         ((size_t)rp.emax + (rp.smax + 1) + (rp.slmax + 1) + rp.smax) * sizeof(int);
}

template <typename T, int KT>
__global__ __launch_bounds__(64) void k_regions(RegionArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int KS = log2c<KT>();
  T *xs = reinterpret_cast<T *>(smem);               // (smax + hmax) * KT values
  T *lval = xs + (size_t)(a.smax + a.hmax) * KT;     // emax entry values
  T *ldiag = lval + a.emax;                          // smax diagonals
  int *lcode = reinterpret_cast<int *>(ldiag + a.smax);  // emax LDS slots
  int *lep = lcode + a.emax;                         // smax + 1 entry bounds
  int *lsl = lep + a.smax + 1;                       // slmax + 1 slice bounds
  int *lrow = lsl + a.slmax + 1;                     // smax row ids
  const int lane = threadIdx.x;
  const int target = ld_dev(&a.state[ST_EPOCH]) + 1;
  int cursor = blockIdx.x;
  for (;;) {
    const int r = next_unit(a.state, a.ticketed, cursor, a.total_waves);
    if (r >= a.R) break;
    const int base = a.rptr[r], nr = a.rptr[r + 1] - base;
    const int e0 = a.eptr[base], ne = a.eptr[base + nr] - e0;
    const int s0 = a.sptr[r], ns = a.sptr[r + 1] - s0 - 1;
    const int h0 = a.hptr[r], nh = a.hptr[r + 1] - h0;
    // 1. stage
    for (int t = lane; t <= nr; t += 64) lep[t] = a.eptr[base + t] - e0;
    for (int t = lane; t < nr; t += 64) {
      lrow[t] = a.rows[base + t];
      ldiag[t] = a.diag[base + t];
    }
    for (int t = lane; t < ne; t += 64) {
      lcode[t] = a.ecode[e0 + t];
      lval[t] = a.eval[e0 + t];
    }
    for (int t = lane; t <= ns; t += 64) lsl[t] = a.slices[s0 + t];
    wave_lds_sync();
    for (int t = lane; t < nr * KT; t += 64) {
      const int l = t >> KS, c = t & (KT - 1);
      if (c < a.k) xs[t] = a.B[(int64_t)lrow[l] * a.rs + (int64_t)c * a.cs];
    }
    // 2. inputs published
    wait_units(a.dptr, a.dlist, r, a.rflags, target, a.state, a.timeout);
    // 3. halo
    for (int t = lane; t < nh * KT; t += 64) {
      const int hh = t >> KS, c = t & (KT - 1);
      if (c < a.k)
        xs[(a.smax + hh) * KT + c] =
            ld_dev(&a.X[(int64_t)a.halo[h0 + hh] * a.rs + (int64_t)c * a.cs]);
    }
    wave_lds_sync();
    // 4. internal levels
    for (int s = 0; s < ns; ++s) {
      const int q1 = lsl[s + 1] * KT;
      for (int t = lsl[s] * KT + lane; t < q1; t += 64) {
        const int l = t >> KS, c = t & (KT - 1);
        if (c < a.k) {
          T acc = xs[t];
          for (int e = lep[l]; e < lep[l + 1]; ++e) acc -= lval[e] * xs[lcode[e] * KT + c];
          xs[t] = acc / ldiag[l];
        }
      }
      wave_lds_sync();
    }
    // 5. store, publish
    for (int t = lane; t < nr * KT; t += 64) {
      const int l = t >> KS, c = t & (KT - 1);
      if (c < a.k) st_dev(&a.X[(int64_t)lrow[l] * a.rs + (int64_t)c * a.cs], xs[t]);
    }
    publish_fence();
    if (lane == 0) st_dev(&a.rflags[r], target);
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
  k_region_gather<T><<<(unsigned)((p->n + 255) / 256), 256, 0, h->stream>>>(
      p->n, rp.d_dsrc, val, (T *)rp.d_diag);
  SYMAMD_HIP(hipGetLastError());
  return SYMAMD_OK;
}

// Regions are always taken by ticket: a wave holds a region only while it
// runs, so the deal stays deadlock-free even when fewer blocks are resident
// than the occupancy estimate promises (these blocks carry up to 64 KiB of
// LDS each, and the estimate can sit exactly at the CU's LDS capacity).
// The grid keeps one block per CU of margin below that estimate.
template <typename T, int KT>
int region_launch(symamd_handle h, symamd_plan p, const RegionPlan &rp, void *d_B,
                  int k, int64_t rs, int64_t cs, size_t lds) {
  const void *kfn = (const void *)&k_regions<T, KT>;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, 64, lds) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  if (per_cu > 1) --per_cu;
  int64_t grid = (int64_t)per_cu * h->num_cus;
  if (h->max_waves > 0 && h->max_waves < grid) grid = h->max_waves;
  if (grid > rp.R) grid = rp.R;
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
  a.eval = (const T *)rp.d_eval;
  a.diag = (const T *)rp.d_diag;
  a.rflags = rp.d_rflags;
  a.state = p->d_state;
  a.R = rp.R;
  a.smax = rp.smax;
  a.hmax = rp.hmax;
  a.emax = rp.emax;
  a.slmax = rp.slmax;
  a.total_waves = (int)grid;
  a.timeout = h->timeout_ticks();
  a.ticketed = 1;
  void *args[] = {&a};
  SYMAMD_HIP(hipLaunchKernel(kfn, dim3((unsigned)grid), dim3(64), args, lds, h->stream));
  return SYMAMD_OK;
}

template <typename T>
int region_dispatch(symamd_handle h, symamd_plan p, const RegionPlan &rp, void *d_B,
                    int k, int64_t rs, int64_t cs, int kt, size_t lds) {
  switch (kt) {
    case 1: return region_launch<T, 1>(h, p, rp, d_B, k, rs, cs, lds);
    case 2: return region_launch<T, 2>(h, p, rp, d_B, k, rs, cs, lds);
    case 4: return region_launch<T, 4>(h, p, rp, d_B, k, rs, cs, lds);
    case 8: return region_launch<T, 8>(h, p, rp, d_B, k, rs, cs, lds);
    case 16: return region_launch<T, 16>(h, p, rp, d_B, k, rs, cs, lds);
    case 32: return region_launch<T, 32>(h, p, rp, d_B, k, rs, cs, lds);
    default: return region_launch<T, 64>(h, p, rp, d_B, k, rs, cs, lds);
  }
}

}  // namespace

int region_cap(int k) {
  const int kt = pow2_ceil(k < 1 ? 1 : k);
  const int cap = pow2_floor(2048 / kt);
  return cap > 512 ? 512 : (cap < 16 ? 16 : cap);
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
  if (k < 1 || k > 64) return false;
  RegionPlan *rp = nullptr;
  if (plan_regions(h, p, region_cap(k), &rp) != SYMAMD_OK || rp->R == 0) return false;
  const size_t vs = p->dtype == SYMAMD_F64 ? 8 : 4;
  return region_lds_bytes(*rp, pow2_ceil(k), vs) <= 64 * 1024;
}

int launch_regions(symamd_handle h, symamd_plan p, void *d_B, int k, int64_t rs,
                   int64_t cs) {
  if (k < 1 || k > 64) return fail(SYMAMD_ERR_NOT_SUPPORTED, "region schedule: 1 <= k <= 64");
  RegionPlan *rp = nullptr;
  SYMAMD_TRY(plan_regions(h, p, region_cap(k), &rp));
  if (rp->R == 0) return fail(SYMAMD_ERR_NOT_SUPPORTED, "no region schedule for this matrix");
  const int kt = pow2_ceil(k);
  const size_t vs = p->dtype == SYMAMD_F64 ? 8 : 4;
  const size_t lds = region_lds_bytes(*rp, kt, vs);
  if (lds > 64 * 1024)
    return fail(SYMAMD_ERR_NOT_SUPPORTED, "region buffers exceed 64 KiB of LDS");
  if (p->dtype == SYMAMD_F64)
    return region_dispatch<double>(h, p, *rp, d_B, k, rs, cs, kt, lds);
  return region_dispatch<float>(h, p, *rp, d_B, k, rs, cs, kt, lds);
}

}  // namespace symamd
