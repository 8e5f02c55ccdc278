// lbc.hip -- GPU inspector for the coarsened (LBC) schedule: the 3-array
// level_ptr / par_ptr / partition format that sym_lib::sptrsv_csr_lbc
// executes (sparse_blas/sptrsv.cpp:57-75).  It stands in for the
// un-vendored get_coarse_levelSet_DAG_CSC_tree (call site
// demo/sptrsv_demo_utils.h:83-88, parameters lp = threads, cp = -1,
// ic = 4000 at demo/sptrsv_demo.cpp:37) with the same inputs: the DAG of a
// lower triangle, a cost per row, the part count per coarse level
// (innerParts), a level distance (minLevelDist) and a rate (divRate).
//
// Every step runs on the device:
//   1. dependency levels (k_levels, inspector.hip) and each row's level;
//   2. coarse levels.  minLevelDist > 0 merges that many consecutive levels
//      into each coarse level.  Otherwise the merge is chosen from the DAG:
//      for every dyadic block of 2^t levels (t = 1 .. log2 of the smaller of
//      divRate and the level count), the weakly connected components of the
//      rows inside the block (edges with both ends in it) are found by a
//      lock-free union-find (hooking the larger root under the smaller with
//      compare-and-swap, path halving in the finds); a block is usable when
//      it holds at least innerParts components and its costliest component
//      is no costlier than an even share, cost / innerParts (so the
//      worst-fit packing below stays within twice the even share); from the
//      first level up, the largest usable block that starts there is taken,
//      else the single level;
//   3. components of each coarse level (one more union-find pass), sorted by
//      cost, largest first (ties by smallest row), packed worst-fit into
//      innerParts bins (the first bin of least accumulated cost), one
//      wavefront per coarse level;
//   4. rows grouped by (coarse level, bin) with a stable radix sort of the
//      level set, so inside a part rows keep (level, row id) order and every
//      dependency inside the part precedes its reader; empty bins are
//      dropped.
// Costs are summed in fixed point (1/1024 units, 64-bit integers), so the
// atomic sums, and with them the schedule, do not depend on thread timing.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>
#include <vector>

#include "common.h"

namespace symamd {

namespace {

constexpr int LBC_MAX_PARTS = 4096;  // bins per coarse level held in LDS

inline unsigned nblk(int64_t m, int bs = 256) { return (unsigned)((m + bs - 1) / bs); }

// Root of x: follow parents (which only ever decrease) to the row that is
// its own parent, halving the path on the way.
__device__ __forceinline__ int uf_root(int *comp, int x) {
  int cur = ld_dev(&comp[x]);
  if (cur != x) {
    int prev = x, next;
    while (cur > (next = ld_dev(&comp[cur]))) {
      st_dev(&comp[prev], next);
      prev = cur;
      cur = next;
    }
  }
  return cur;
}

__device__ __forceinline__ void uf_union(int *comp, int a, int b) {
  int ra = uf_root(comp, a), rb = uf_root(comp, b);
  while (ra != rb) {
    if (ra < rb) {
      const int t = ra;
      ra = rb;
      rb = t;
    }
    // hook the larger root under the smaller one
    const int old = atomicCAS(&comp[ra], ra, rb);
    if (old == ra) break;
    ra = uf_root(comp, old);
  }
}

// Coarse level of level l: cid[l] when given, else l >> shift.
__device__ __forceinline__ int coarse_of(const int *cid, int shift, int l) {
  return cid ? cid[l] : l >> shift;
}

// lev[level_set[q]] = the level of position q (binary search of level_ptr).
__global__ void k_row_level(int n, int nlev, const int *__restrict__ lp,
                            const int *__restrict__ ls, int *lev) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int lo = 0, hi = nlev - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (lp[mid] <= q) lo = mid;
    else hi = mid - 1;
  }
  lev[ls[q]] = lo;
}

// Fixed-point cost of every row: the caller's cost (negative and NaN read
// as 0, capped at 2^40), or the row's entry count when no cost is given.
__global__ void k_row_cost(int n, const int *__restrict__ Lp, const double *__restrict__ cost,
                           uint64_t *c64) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double c = cost ? cost[i] : (double)(Lp[i + 1] - Lp[i]);
  if (!(c > 0.0)) c = 0.0;
  if (c > 1099511627776.0) c = 1099511627776.0;
  c64[i] = (uint64_t)__double2ll_rn(c * 1024.0);
}

__global__ void k_uf_init(int n, int *comp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) comp[i] = i;
}

// Union every row with its dependencies in the same coarse level (the
// diagonal, last in the row, is skipped).
__global__ void k_uf_hook(int n, const int *__restrict__ Lp, const int *__restrict__ Li,
                          const int *__restrict__ lev, const int *__restrict__ cid, int shift,
                          int *comp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ci = coarse_of(cid, shift, lev[i]);
  for (int e = Lp[i], e1 = Lp[i + 1] - 1; e < e1; ++e) {
    const int d = Li[e];
    if (coarse_of(cid, shift, lev[d]) == ci) uf_union(comp, i, d);
  }
}

// After the hooks: every row points at its root.  Each thread only reads
// the (now fixed) forest and writes its own entry, so a concurrent reader
// sees either the old parent or the root -- both ancestors.
__global__ void k_uf_flatten(int n, int *comp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int r = ld_dev(&comp[i]), next;
  const int old = r;
  while (r > (next = ld_dev(&comp[r]))) r = next;
  if (r != old) st_dev(&comp[i], r);
}

// Component costs at their roots, and the cost of every coarse level.
__global__ void k_comp_cost(int n, const int *__restrict__ comp, const int *__restrict__ lev,
                            const int *__restrict__ cid, int shift,
                            const uint64_t *__restrict__ c64, unsigned long long *ccost,
                            unsigned long long *btot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long c = c64[i];
  if (c == 0) return;
  atomicAdd(&ccost[comp[i]], c);
  atomicAdd(&btot[coarse_of(cid, shift, lev[i])], c);
}

// Components per coarse level and the costliest one.
__global__ void k_block_stats(int n, const int *__restrict__ comp, const int *__restrict__ lev,
                              const int *__restrict__ cid, int shift,
                              const unsigned long long *__restrict__ ccost, int *bcnt,
                              unsigned long long *bmax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || comp[i] != i) return;
  const int b = coarse_of(cid, shift, lev[i]);
  atomicAdd(&bcnt[b], 1);
  atomicMax(&bmax[b], ccost[i]);
}

// Roots in ascending row order: flag, then scatter by the flag's scan.
__global__ void k_root_flag(int n, const int *__restrict__ comp, int *flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = comp[i] == i;
  if (i == n) flag[n] = 0;
}
__global__ void k_root_list(int n, const int *__restrict__ comp, const int *__restrict__ pos,
                            const unsigned long long *__restrict__ ccost, int *roots,
                            unsigned long long *key) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || comp[i] != i) return;
  roots[pos[i]] = i;
  key[pos[i]] = ~ccost[i];  // ascending key = descending cost
}
__global__ void k_root_coarse(int m, const int *__restrict__ roots, const int *__restrict__ lev,
                              const int *__restrict__ cid, int *key) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) key[k] = cid[lev[roots[k]]];
}

// seg[c] = first position of coarse level c in the sorted keys.
__global__ void k_segments(int m, int nc, const int *__restrict__ skey, int *seg) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const int c = skey[k], prev = k == 0 ? -1 : skey[k - 1];
  for (int t = prev + 1; t <= c; ++t) seg[t] = k;
  if (k == m - 1)
    for (int t = c + 1; t <= nc; ++t) seg[t] = m;
}

// Worst-fit packing, one wavefront per coarse level: components in order
// (largest first) each go to the first bin of least accumulated cost.
__global__ __launch_bounds__(64) void k_pack(int nc, const int *__restrict__ seg,
                                             const int *__restrict__ roots,
                                             const unsigned long long *__restrict__ ccost,
                                             int parts, int *bin_of) {
  __shared__ unsigned long long load[LBC_MAX_PARTS];
  const int lane = threadIdx.x;
  for (int c = blockIdx.x; c < nc; c += gridDim.x) {
    for (int b = lane; b < parts; b += 64) load[b] = 0;
    __syncthreads();
    for (int k = seg[c]; k < seg[c + 1]; ++k) {
      const int r = roots[k];
      // (load, bin) minimum over the lane's bins, then across the wave
      unsigned long long best = ~0ull;
      int bb = INT_MAX;
      for (int b = lane; b < parts; b += 64)
        if (load[b] < best) {
          best = load[b];
          bb = b;
        }
      for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long ob = __shfl_xor(best, off);
        const int obb = __shfl_xor(bb, off);
        if (ob < best || (ob == best && obb < bb)) {
          best = ob;
          bb = obb;
        }
      }
      if (lane == 0) {
        load[bb] += ccost[r];
        bin_of[r] = bb;
      }
      __syncthreads();
    }
    __syncthreads();
  }
}

// Part key of every level-set position: coarse level * parts + bin.
__global__ void k_part_key(int n, int parts, const int *__restrict__ ls,
                           const int *__restrict__ lev, const int *__restrict__ cid,
                           const int *__restrict__ comp, const int *__restrict__ bin_of,
                           int *key, int *cnt) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int i = ls[q];
  const int kq = cid[lev[i]] * parts + bin_of[comp[i]];
  key[q] = kq;
  atomicAdd(&cnt[kq], 1);
}

__global__ void k_nonempty(int m, const int *__restrict__ cnt, int *nz) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) nz[j] = cnt[j] > 0;
  if (j == m) nz[m] = 0;
}

// par_ptr of the nonempty bins and level_ptr of the coarse levels.
__global__ void k_lbc_ptrs(int m, int nc, int parts, int n, const int *__restrict__ cnt,
                           const int *__restrict__ off, const int *__restrict__ pidx,
                           int *par_ptr, int *level_ptr) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m && cnt[j] > 0) par_ptr[pidx[j]] = off[j];
  if (j == m) par_ptr[pidx[m]] = n;
  if (j <= nc) level_ptr[j] = pidx[j == nc ? m : j * parts];
}

template <typename T>
int ex_scan(hipStream_t s, const T *in, T *out, size_t m) {
  size_t bytes = 0;
  SYMAMD_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, T(0), m, rocprim::plus<T>(), s));
  void *tmp = nullptr;
  SYMAMD_HIP(hipMallocAsync(&tmp, bytes ? bytes : 4, s));
  const hipError_t e =
      rocprim::exclusive_scan(tmp, bytes, in, out, T(0), m, rocprim::plus<T>(), s);
  SYMAMD_HIP(hipFreeAsync(tmp, s));
  SYMAMD_HIP(e);
  return SYMAMD_OK;
}

template <typename K, typename V>
int sort_pairs(hipStream_t s, const K *kin, K *kout, const V *vin, V *vout, size_t m,
               int end_bit) {
  size_t bytes = 0;
  SYMAMD_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, m, 0, end_bit, s));
  void *tmp = nullptr;
  SYMAMD_HIP(hipMallocAsync(&tmp, bytes ? bytes : 4, s));
  const hipError_t e =
      rocprim::radix_sort_pairs(tmp, bytes, kin, kout, vin, vout, m, 0, end_bit, s);
  SYMAMD_HIP(hipFreeAsync(tmp, s));
  SYMAMD_HIP(e);
  return SYMAMD_OK;
}

int bits_for(int64_t v) {
  int b = 1;
  while (b < 63 && ((int64_t)1 << b) <= v) ++b;
  return b;
}

// Device buffers of one run, released together.
struct Bufs {
  std::vector<void *> ptrs;
  hipStream_t s;
  explicit Bufs(hipStream_t st) : s(st) {}
  template <typename T>
  T *get(int64_t count) {
    void *p = nullptr;
    if (hipMallocAsync(&p, sizeof(T) * (size_t)(count > 0 ? count : 1), s) != hipSuccess)
      return nullptr;
    ptrs.push_back(p);
    return (T *)p;
  }
  ~Bufs() {
    for (void *p : ptrs) (void)hipFreeAsync(p, s);
    (void)hipStreamSynchronize(s);
  }
};

// One union-find pass over the coarse levels given by (cid, shift); leaves
// comp flattened (every row points at its component's smallest row).
int components(hipStream_t s, int n, const int *Lp, const int *Li, const int *lev,
               const int *cid, int shift, int *comp) {
  k_uf_init<<<nblk(n), 256, 0, s>>>(n, comp);
  k_uf_hook<<<nblk(n), 256, 0, s>>>(n, Lp, Li, lev, cid, shift, comp);
  k_uf_flatten<<<nblk(n), 256, 0, s>>>(n, comp);
  SYMAMD_HIP(hipGetLastError());
  return SYMAMD_OK;
}

}  // namespace

int lbc_schedule(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
                 const double *d_cost, int parts, int min_level_dist, int div_rate,
                 int *d_level_ptr, int *d_par_ptr, int *d_partition, int *level_no,
                 int *part_no) {
  if (parts < 1) parts = 1;
  if (parts > LBC_MAX_PARTS)
    return fail(SYMAMD_ERR_NOT_SUPPORTED, "coarsened schedule: at most %d parts per coarse "
                "level (asked for %d)", LBC_MAX_PARTS, parts);
  hipStream_t s = h->stream;
  if (n == 0) {
    const int zero = 0;
    SYMAMD_HIP(hipMemcpyAsync(d_level_ptr, &zero, sizeof(int), hipMemcpyHostToDevice, s));
    SYMAMD_HIP(hipMemcpyAsync(d_par_ptr, &zero, sizeof(int), hipMemcpyHostToDevice, s));
    SYMAMD_HIP(hipStreamSynchronize(s));
    *level_no = 0;
    *part_no = 0;
    return SYMAMD_OK;
  }
  Bufs B(s);
  int *st = B.get<int>(ST_WORDS);
  int *lp = B.get<int>(n + 1);
  int *ls = B.get<int>(n);
  int *lev = B.get<int>(n);
  int *comp = B.get<int>(n);
  uint64_t *c64 = B.get<uint64_t>(n);
  unsigned long long *ccost = B.get<unsigned long long>(n);
  if (!st || !lp || !ls || !lev || !comp || !c64 || !ccost)
    return fail(SYMAMD_ERR_ALLOC, "coarsened schedule: device buffers");
  SYMAMD_TRY(check_structure(h, n, d_Lp, d_Li, st));
  int nlev = 0;
  std::vector<int> h_lp;
  SYMAMD_TRY(compute_levels(h, n, d_Lp, d_Li, st, lp, ls, &nlev, &h_lp));
  k_row_level<<<nblk(n), 256, 0, s>>>(n, nlev, lp, ls, lev);
  k_row_cost<<<nblk(n), 256, 0, s>>>(n, d_Lp, d_cost, c64);
  SYMAMD_HIP(hipGetLastError());

  // coarse level of every level
  std::vector<int> cid(nlev);
  if (min_level_dist > 0) {
    for (int l = 0; l < nlev; ++l) cid[l] = l / min_level_dist;
  } else {
    const int64_t cap = std::min<int64_t>(nlev, div_rate > 1 ? div_rate : 1);
    int T = 0;
    while (((int64_t)2 << T) <= cap) ++T;
    // usable[t][b]: block b of 2^t levels holds enough, even enough
    // components
    std::vector<std::vector<char>> usable(T + 1);
    int *bcnt = B.get<int>(nlev);
    unsigned long long *btot = B.get<unsigned long long>(nlev);
    unsigned long long *bmax = B.get<unsigned long long>(nlev);
    if (!bcnt || !btot || !bmax) return fail(SYMAMD_ERR_ALLOC, "coarsened schedule: buffers");
    std::vector<int> hcnt(nlev);
    std::vector<unsigned long long> htot(nlev), hmax(nlev);
    for (int t = 1; t <= T; ++t) {
      const int nb = (nlev + (1 << t) - 1) >> t;
      SYMAMD_TRY(components(s, n, d_Lp, d_Li, lev, nullptr, t, comp));
      SYMAMD_HIP(hipMemsetAsync(ccost, 0, sizeof(unsigned long long) * n, s));
      SYMAMD_HIP(hipMemsetAsync(bcnt, 0, sizeof(int) * nb, s));
      SYMAMD_HIP(hipMemsetAsync(btot, 0, sizeof(unsigned long long) * nb, s));
      SYMAMD_HIP(hipMemsetAsync(bmax, 0, sizeof(unsigned long long) * nb, s));
      k_comp_cost<<<nblk(n), 256, 0, s>>>(n, comp, lev, nullptr, t, c64, ccost, btot);
      k_block_stats<<<nblk(n), 256, 0, s>>>(n, comp, lev, nullptr, t, ccost, bcnt, bmax);
      SYMAMD_HIP(hipGetLastError());
      SYMAMD_HIP(hipMemcpyAsync(hcnt.data(), bcnt, sizeof(int) * nb, hipMemcpyDeviceToHost, s));
      SYMAMD_HIP(hipMemcpyAsync(htot.data(), btot, sizeof(unsigned long long) * nb,
                                hipMemcpyDeviceToHost, s));
      SYMAMD_HIP(hipMemcpyAsync(hmax.data(), bmax, sizeof(unsigned long long) * nb,
                                hipMemcpyDeviceToHost, s));
      SYMAMD_HIP(hipStreamSynchronize(s));
      usable[t].assign(nb, 0);
      for (int b = 0; b < nb; ++b)
        usable[t][b] = hcnt[b] >= parts &&
                       (unsigned __int128)hmax[b] * (unsigned)parts <= htot[b];
    }
    int c = 0;
    for (int a = 0; a < nlev; ++c) {
      int span = 1;
      for (int t = T; t >= 1; --t)
        if ((a & ((1 << t) - 1)) == 0 && usable[t][a >> t]) {
          span = std::min(1 << t, nlev - a);
          break;
        }
      for (int l = a; l < a + span; ++l) cid[l] = c;
      a += span;
    }
  }
  const int nc = cid.empty() ? 0 : cid.back() + 1;

  // components of the chosen coarse levels, packed into bins
  int *d_cid = B.get<int>(nlev);
  int *flag = B.get<int>(n + 1);
  int *pos = B.get<int>(n + 1);
  if (!d_cid || !flag || !pos) return fail(SYMAMD_ERR_ALLOC, "coarsened schedule: buffers");
  SYMAMD_HIP(hipMemcpyAsync(d_cid, cid.data(), sizeof(int) * nlev, hipMemcpyHostToDevice, s));
  SYMAMD_TRY(components(s, n, d_Lp, d_Li, lev, d_cid, 0, comp));
  SYMAMD_HIP(hipMemsetAsync(ccost, 0, sizeof(unsigned long long) * n, s));
  {
    unsigned long long *btot = B.get<unsigned long long>(nc);
    if (!btot) return fail(SYMAMD_ERR_ALLOC, "coarsened schedule: buffers");
    SYMAMD_HIP(hipMemsetAsync(btot, 0, sizeof(unsigned long long) * nc, s));
    k_comp_cost<<<nblk(n), 256, 0, s>>>(n, comp, lev, d_cid, 0, c64, ccost, btot);
  }
  k_root_flag<<<nblk(n + 1), 256, 0, s>>>(n, comp, flag);
  SYMAMD_HIP(hipGetLastError());
  SYMAMD_TRY(ex_scan<int>(s, flag, pos, (size_t)n + 1));
  int m = 0;
  SYMAMD_HIP(hipMemcpyAsync(&m, pos + n, sizeof(int), hipMemcpyDeviceToHost, s));
  SYMAMD_HIP(hipStreamSynchronize(s));
  int *roots = B.get<int>(m), *roots2 = B.get<int>(m), *ckey = B.get<int>(m),
      *ckey2 = B.get<int>(m), *seg = B.get<int>(nc + 1), *bin_of = B.get<int>(n);
  unsigned long long *wkey = B.get<unsigned long long>(m), *wkey2 = B.get<unsigned long long>(m);
  if (!roots || !roots2 || !ckey || !ckey2 || !seg || !bin_of || !wkey || !wkey2)
    return fail(SYMAMD_ERR_ALLOC, "coarsened schedule: buffers");
  k_root_list<<<nblk(n), 256, 0, s>>>(n, comp, pos, ccost, roots, wkey);
  SYMAMD_HIP(hipGetLastError());
  // largest cost first, then (stable) by coarse level: ties keep row order
  SYMAMD_TRY((sort_pairs<unsigned long long, int>(s, wkey, wkey2, roots, roots2, (size_t)m, 64)));
  k_root_coarse<<<nblk(m), 256, 0, s>>>(m, roots2, lev, d_cid, ckey);
  SYMAMD_HIP(hipGetLastError());
  SYMAMD_TRY((sort_pairs<int, int>(s, ckey, ckey2, roots2, roots, (size_t)m, bits_for(nc))));
  k_segments<<<nblk(m), 256, 0, s>>>(m, nc, ckey2, seg);
  k_pack<<<(unsigned)std::min(nc, 65536), 64, 0, s>>>(nc, seg, roots, ccost, parts, bin_of);
  SYMAMD_HIP(hipGetLastError());

  // rows grouped by (coarse level, bin), level-set order inside
  const int64_t mb = (int64_t)nc * parts;
  if (mb >= INT_MAX) return fail(SYMAMD_ERR_NOT_SUPPORTED, "coarsened schedule: too many bins");
  int *key = B.get<int>(n), *key2 = B.get<int>(n), *cnt = B.get<int>(mb + 1),
      *off = B.get<int>(mb + 1), *nz = B.get<int>(mb + 1), *pidx = B.get<int>(mb + 1);
  if (!key || !key2 || !cnt || !off || !nz || !pidx)
    return fail(SYMAMD_ERR_ALLOC, "coarsened schedule: buffers");
  SYMAMD_HIP(hipMemsetAsync(cnt, 0, sizeof(int) * (mb + 1), s));
  k_part_key<<<nblk(n), 256, 0, s>>>(n, parts, ls, lev, d_cid, comp, bin_of, key, cnt);
  SYMAMD_HIP(hipGetLastError());
  SYMAMD_TRY((sort_pairs<int, int>(s, key, key2, ls, d_partition, (size_t)n, bits_for(mb))));
  SYMAMD_TRY(ex_scan<int>(s, cnt, off, (size_t)mb + 1));
  k_nonempty<<<nblk(mb + 1), 256, 0, s>>>((int)mb, cnt, nz);
  SYMAMD_TRY(ex_scan<int>(s, nz, pidx, (size_t)mb + 1));
  k_lbc_ptrs<<<nblk(std::max<int64_t>(mb, nc) + 1), 256, 0, s>>>((int)mb, nc, parts, n, cnt,
                                                                off, pidx, d_par_ptr,
                                                                d_level_ptr);
  SYMAMD_HIP(hipGetLastError());
  int np = 0;
  SYMAMD_HIP(hipMemcpyAsync(&np, pidx + mb, sizeof(int), hipMemcpyDeviceToHost, s));
  SYMAMD_HIP(hipStreamSynchronize(s));
  *level_no = nc;
  *part_no = np;
  return SYMAMD_OK;
}

}  // namespace symamd
