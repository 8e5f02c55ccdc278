// ic0.hip -- zero-fill incomplete Cholesky, IC(0), on the plan's pattern for
// gfx950: the factor the solves take, computed on the device.
//
// Reference: sparse_blas/spic0.cpp:11-46 (spic0_csc_left: left-looking on the
// CSC lower triangle of A, diagonal first; :48-99 the same column routine
// under an LBC schedule) and :106-127 (spic0_csr: right-looking, in place on
// the same arrays; :130-160 under an LBC schedule with atomic updates).
//
// Both are restated per entry (r, c) of L, in the CSR of L the plan holds
// (diagonal last):
//   v = A[r,c];  v -= L[r,s] * L[c,s]  for s in S(r,c), ascending;
//   L[c,c] = sqrt(v)                    (r == c)
//   L[r,c] = v / L[c,c]                 (r > c)
// spic0_csc_left: S(r,c) = { s < c : L[r,s] != 0 and L[c,s] != 0 }.  Column c
//   walks its prune set (the row pattern of c, ascending) and subtracts
//   L[r,s] * L[c,s] from f[r] for every row r of column s at or below c.
// spic0_csr: the same set restricted to the s whose last stored row is r.
//   Its update loop (spic0.cpp:117-123) walks l from m while the row of the
//   NEXT entry is <= the target's row, so it reaches a match only at the last
//   entry of column s, whose successor is the diagonal of column s + 1
//   (row s + 1 <= r).  spic0_csr_lbc applies the same updates with atomics
//   in an order that depends on the threads; this file applies them in the
//   serial routine's order.
// In every case the updates of one entry come in ascending s (the outer
// column loop), so a lane that applies S(r,c) in ascending order repeats the
// reference's operations on that entry.  Products, subtractions, divisions
// and square roots are IEEE operations compiled without contraction, so the
// factor is bit-identical to the reference's.
//
// Inspector: for every entry the (position of L[r,s], position of L[c,s])
// pairs of S(r,c), found by merging the column ids of row r below c with
// those of row c below its diagonal; counted, scanned, then written.
// Executor: the sync-free unit schedule of the forward solve (row r waits
// for the rows c of its pattern, which are exactly the rows IC(0) reads);
// one lane per row computes the row's entries in order, then its diagonal.
#include <rocprim/device/device_scan.hpp>

#include <climits>

#include "device.h"

namespace symamd {

namespace {

inline int blocks_of(int64_t m, int bs = 256) { return (int)((m + bs - 1) / bs); }

// last[s] = the largest row r with L[r,s] stored (-1 on entry).
__global__ void k_last_row(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                           int *last) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (int p = rp[r]; p < rp[r + 1]; ++p) atomicMax(&last[ci[p]], r);
}

// emit(a, b) for every s of S(r, c) in ascending order, where p is the
// position of (r, c): a runs over row r below column c, b over row c below
// its diagonal (for the diagonal, c == r, both over row r: pairs (a, a)).
template <int VAR, typename F>
__device__ __forceinline__ void pairs_of(const int *__restrict__ rp, const int *__restrict__ ci,
                                         const int *__restrict__ last, int r, int p, F emit) {
  const int c = ci[p];
  int a = rp[r], b = rp[c];
  const int ae = p, be = rp[c + 1] - 1;
  while (a < ae && b < be) {
    const int sa = ci[a], sb = ci[b];
    if (sa == sb) {
      if (VAR == SYMAMD_IC0_LEFT || last[sa] == r) emit(a, b);
      ++a;
      ++b;
    } else if (sa < sb) {
      ++a;
    } else {
      ++b;
    }
  }
}

// cnt[p] = |S(r, c)|; dpos[p] = position of L[c,c] (the divisor of p).
template <int VAR>
__global__ void k_pair_count(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                             const int *__restrict__ last, int64_t nnz, int64_t *cnt,
                             int *dpos) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r > n) return;
  if (r == n) {
    cnt[nnz] = 0;
    return;
  }
  for (int p = rp[r]; p < rp[r + 1]; ++p) {
    int64_t k = 0;
    pairs_of<VAR>(rp, ci, last, r, p, [&](int, int) { ++k; });
    cnt[p] = k;
    dpos[p] = rp[ci[p] + 1] - 1;
  }
}

template <int VAR>
__global__ void k_pair_fill(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                            const int *__restrict__ last, const int64_t *__restrict__ pp,
                            int *pa, int *pb) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (int p = rp[r]; p < rp[r + 1]; ++p) {
    int64_t t = pp[p];
    pairs_of<VAR>(rp, ci, last, r, p, [&](int a, int b) {
      pa[t] = a;
      pb[t] = b;
      ++t;
    });
  }
}

__global__ void k_set_int(int *p, int v) { *p = v; }

struct Ic0Args {
  const int *rp, *ci;        // CSR of L in the caller's order (diagonal last)
  const double *Ax;          // A's values on that pattern
  double *Lx;                // the factor (may alias Ax)
  const int64_t *pp;         // pairs of position p: [pp[p], pp[p+1])
  const int *pa, *pb;        //   positions of L[r,s] and L[c,s]
  const int *dpos;           // position of L[c,c] for position p = (r, c)
  const int *perm;           // level position -> row
  const int *ub, *dptr, *dlist;  // units and their dependency units
  int *uflags, *state;
  int *bad;                  // smallest row whose pivot failed (INT_MAX: none)
  int U, ticketed, total_waves, variant;
  uint64_t timeout;
};

// Rows of at most IC0_RB entries keep their values in registers: the loads
// that do not depend on the row's own results (A, the pair ranges, the
// divisors) are issued first, and the row is stored once at the end, so no
// load waits behind a store.  (Variants that also held the pairs' indices
// and values in registers measured slower on the 128^3 Laplacian: 5.0 and
// 6.4 ms against 4.0 ms.)
constexpr int IC0_RB = 8;

__device__ __forceinline__ double pick(const double (&v)[IC0_RB], int i) {
  double r = v[0];
#pragma unroll
  for (int j = 1; j < IC0_RB; ++j) r = i == j ? v[j] : r;
  return r;
}

__device__ __forceinline__ bool pivot_fails(int variant, double v) {
  // spic0_csc_left stops at f <= 0 (a NaN pivot passes);
  // spic0_csr asserts pivot > 0 (a NaN pivot fails)
  return variant == SYMAMD_IC0_LEFT ? v <= 0.0 : !(v > 0.0);
}

__device__ __forceinline__ void finish_entry(const Ic0Args &a, int r, bool diag, double dg,
                                             double &v) {
  if (!diag) {
    v = v / dg;
  } else {
    if (pivot_fails(a.variant, v)) atomicMin(a.bad, r);
    v = __builtin_sqrt(v);
  }
}

__device__ __forceinline__ void ic0_short_row(const Ic0Args &a, int r, int b0, int m) {
  double ax[IC0_RB], dg[IC0_RB], own[IC0_RB];
  int64_t pq[IC0_RB + 1];
#pragma unroll
  for (int j = 0; j < IC0_RB; ++j) {
    ax[j] = j < m ? a.Ax[b0 + j] : 0.0;
    own[j] = 0.0;
  }
#pragma unroll
  for (int j = 0; j <= IC0_RB; ++j) pq[j] = j <= m ? a.pp[b0 + j] : 0;
#pragma unroll
  for (int j = 0; j < IC0_RB; ++j) dg[j] = j < m - 1 ? ld_dev(&a.Lx[a.dpos[b0 + j]]) : 1.0;
#pragma unroll
  for (int j = 0; j < IC0_RB; ++j) {
    if (j < m) {
      double v = ax[j];
      const bool diag = j == m - 1;
      for (int64_t t = pq[j]; t < pq[j + 1]; ++t) {
        const int ib = a.pb[t];
        // L[r,s] is this row's; L[c,s] another row's, or this row's on the
        // diagonal (c == r)
        const double x = pick(own, a.pa[t] - b0);
        const double y = diag ? pick(own, ib - b0) : ld_dev(&a.Lx[ib]);
        v -= x * y;
      }
      finish_entry(a, r, diag, dg[j], v);
      own[j] = v;
    }
  }
#pragma unroll
  for (int j = 0; j < IC0_RB; ++j)
    if (j < m) st_dev(&a.Lx[b0 + j], own[j]);
}

// Longer rows: entry by entry through memory (a lane reads back its own
// row's entries, written earlier by the same lane, with device-scope loads).
__device__ __forceinline__ void ic0_long_row(const Ic0Args &a, int r, int b0, int d) {
  for (int p = b0; p <= d; ++p) {
    double v = a.Ax[p];
    for (int64_t t = a.pp[p], te = a.pp[p + 1]; t < te; ++t)
      v -= ld_dev(&a.Lx[a.pa[t]]) * ld_dev(&a.Lx[a.pb[t]]);
    finish_entry(a, r, p == d, p < d ? ld_dev(&a.Lx[a.dpos[p]]) : 1.0, v);
    st_dev(&a.Lx[p], v);
  }
}

// One wave per unit (a run of at most 64 rows of one level), one lane per
// row.  Values of other rows are read with device-scope loads once their
// units carry this launch's tag.
__global__ __launch_bounds__(256) void k_ic0(Ic0Args a) {
  const int lane = threadIdx.x & 63;
  const int target = ld_dev(&a.state[ST_EPOCH]) + 1;
  int cursor = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  for (;;) {
    const int u = next_unit(a.state, a.ticketed, cursor, a.total_waves);
    if (u >= a.U) break;
    const int q = a.ub[u] + lane;
    const bool act = q < a.ub[u + 1];
    wait_units(a.dptr, a.dlist, u, a.uflags, target, a.state, a.timeout);
    if (act) {
      const int r = a.perm[q];
      const int b0 = a.rp[r], m = a.rp[r + 1] - b0;
      if (m <= IC0_RB)
        ic0_short_row(a, r, b0, m);
      else
        ic0_long_row(a, r, b0, b0 + m - 1);
    }
    publish_fence();
    if (lane == 0) st_dev(&a.uflags[u], target);
  }
  persistent_finish(a.state, a.total_waves);
}

int ic0_grid(symamd_handle h, int units) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)&k_ic0, 256, 0) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  int64_t cap = (int64_t)per_cu * h->num_cus;
  if (h->max_waves > 0 && (h->max_waves + 3) / 4 < cap) cap = (h->max_waves + 3) / 4;
  int64_t want = (units + 3) / 4;
  if (want < 1) want = 1;
  return (int)(want < cap ? want : cap);
}

bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
  return st != hipStreamCaptureStatusNone;
}

// The pair lists of a variant, built on first use.
int plan_ic0_pairs(symamd_handle h, symamd_plan p, int variant, Ic0Pairs **out) {
  auto it = p->ic0.find(variant);
  if (it != p->ic0.end()) {
    *out = &it->second;
    return SYMAMD_OK;
  }
  SYMAMD_TRY(plan_oci(h, p));
  const int n = p->n;
  const int64_t nnz = p->nnz;
  Ic0Pairs pr;
  int *d_last = nullptr;
  int64_t *d_cnt = nullptr;
  void *tmp = nullptr;
  int rc = SYMAMD_OK;
  auto done = [&](int code) {
    if (d_last) (void)hipFree(d_last);
    if (d_cnt) (void)hipFree(d_cnt);
    if (tmp) (void)hipFree(tmp);
    if (code != SYMAMD_OK) {
      if (pr.d_pp) (void)hipFree(pr.d_pp);
      if (pr.d_pa) (void)hipFree(pr.d_pa);
      if (pr.d_pb) (void)hipFree(pr.d_pb);
      if (pr.d_dpos) (void)hipFree(pr.d_dpos);
    }
    return code;
  };
#define IC0_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return done(fail(SYMAMD_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_))); \
  } while (0)
  IC0_HIP(hipMalloc(&d_last, sizeof(int) * (n > 0 ? n : 1)));
  IC0_HIP(hipMalloc(&d_cnt, sizeof(int64_t) * (nnz + 1)));
  IC0_HIP(hipMalloc(&pr.d_pp, sizeof(int64_t) * (nnz + 1)));
  IC0_HIP(hipMalloc(&pr.d_dpos, sizeof(int) * (nnz > 0 ? nnz : 1)));
  IC0_HIP(hipMemsetAsync(d_last, 0xff, sizeof(int) * (n > 0 ? n : 1), h->stream));
  if (n > 0) {
    k_last_row<<<blocks_of(n), 256, 0, h->stream>>>(n, p->d_orp, p->d_oci, d_last);
    IC0_HIP(hipGetLastError());
  }
  if (variant == SYMAMD_IC0_LEFT)
    k_pair_count<SYMAMD_IC0_LEFT><<<blocks_of(n + 1), 256, 0, h->stream>>>(
        n, p->d_orp, p->d_oci, d_last, nnz, d_cnt, pr.d_dpos);
  else
    k_pair_count<SYMAMD_IC0_CSR><<<blocks_of(n + 1), 256, 0, h->stream>>>(
        n, p->d_orp, p->d_oci, d_last, nnz, d_cnt, pr.d_dpos);
  IC0_HIP(hipGetLastError());
  size_t bytes = 0;
  IC0_HIP(rocprim::exclusive_scan(nullptr, bytes, d_cnt, pr.d_pp, (int64_t)0,
                                  (size_t)(nnz + 1), rocprim::plus<int64_t>(), h->stream));
  IC0_HIP(hipMalloc(&tmp, bytes ? bytes : 4));
  IC0_HIP(rocprim::exclusive_scan(tmp, bytes, d_cnt, pr.d_pp, (int64_t)0, (size_t)(nnz + 1),
                                  rocprim::plus<int64_t>(), h->stream));
  IC0_HIP(hipMemcpyAsync(&pr.pairs, pr.d_pp + nnz, sizeof(int64_t), hipMemcpyDeviceToHost,
                         h->stream));
  IC0_HIP(hipStreamSynchronize(h->stream));
  IC0_HIP(hipMalloc(&pr.d_pa, sizeof(int) * (pr.pairs > 0 ? pr.pairs : 1)));
  IC0_HIP(hipMalloc(&pr.d_pb, sizeof(int) * (pr.pairs > 0 ? pr.pairs : 1)));
  if (n > 0) {
    if (variant == SYMAMD_IC0_LEFT)
      k_pair_fill<SYMAMD_IC0_LEFT><<<blocks_of(n), 256, 0, h->stream>>>(
          n, p->d_orp, p->d_oci, d_last, pr.d_pp, pr.d_pa, pr.d_pb);
    else
      k_pair_fill<SYMAMD_IC0_CSR><<<blocks_of(n), 256, 0, h->stream>>>(
          n, p->d_orp, p->d_oci, d_last, pr.d_pp, pr.d_pa, pr.d_pb);
    IC0_HIP(hipGetLastError());
  }
  IC0_HIP(hipStreamSynchronize(h->stream));
#undef IC0_HIP
  rc = done(SYMAMD_OK);
  if (rc != SYMAMD_OK) return rc;
  auto ins = p->ic0.emplace(variant, pr);
  *out = &ins.first->second;
  return SYMAMD_OK;
}

}  // namespace

int ic0_factor(symamd_handle h, symamd_plan p, int variant, const double *d_Ax, double *d_Lx,
               int *bad_row) {
  Ic0Pairs *pr = nullptr;
  SYMAMD_TRY(plan_ic0_pairs(h, p, variant, &pr));
  UnitSet *us = nullptr;
  SYMAMD_TRY(plan_units(h, p, 64, false, &us));
  if (!p->d_ic0_bad) SYMAMD_HIP(hipMalloc(&p->d_ic0_bad, sizeof(int)));
  k_set_int<<<1, 1, 0, h->stream>>>(p->d_ic0_bad, INT_MAX);
  SYMAMD_HIP(hipGetLastError());
  Ic0Args a;
  a.rp = p->d_orp;
  a.ci = p->d_oci;
  a.Ax = d_Ax;
  a.Lx = d_Lx;
  a.pp = pr->d_pp;
  a.pa = pr->d_pa;
  a.pb = pr->d_pb;
  a.dpos = pr->d_dpos;
  a.perm = p->d_perm;
  a.ub = us->d_ub;
  a.dptr = us->d_dptr;
  a.dlist = us->d_dlist;
  a.uflags = us->d_uflags;
  a.state = p->d_state;
  a.bad = p->d_ic0_bad;
  a.U = us->count;
  a.variant = variant;
  a.timeout = h->timeout_ticks();
  const int grid = ic0_grid(h, a.U);
  a.total_waves = grid * 4;
  bool launched = false;
  if (!stream_capturing(h->stream) && !h->force_tickets) {
    a.ticketed = 0;
    void *args[] = {&a};
    if (hipLaunchCooperativeKernel((const void *)&k_ic0, dim3(grid), dim3(256), args, 0,
                                   h->stream) == hipSuccess)
      launched = true;
    else
      (void)hipGetLastError();
  }
  if (!launched) {
    a.ticketed = 1;
    void *args[] = {&a};
    SYMAMD_HIP(hipLaunchKernel((const void *)&k_ic0, dim3(grid), dim3(256), args, 0,
                               h->stream));
  }
  int bad = INT_MAX;
  SYMAMD_HIP(hipMemcpyAsync(&bad, p->d_ic0_bad, sizeof(int), hipMemcpyDeviceToHost,
                            h->stream));
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  *bad_row = bad == INT_MAX ? -1 : bad;
  return SYMAMD_OK;
}

void free_ic0_pairs(symamd_plan p) {
  for (auto &kv : p->ic0) {
    void *q[] = {kv.second.d_pp, kv.second.d_pa, kv.second.d_pb, kv.second.d_dpos};
    for (void *x : q)
      if (x) (void)hipFree(x);
  }
  p->ic0.clear();
  if (p->d_ic0_bad) (void)hipFree(p->d_ic0_bad);
  p->d_ic0_bad = nullptr;
}

}  // namespace symamd
