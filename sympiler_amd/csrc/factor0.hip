// factor0.hip -- zero-fill incomplete factorizations on the device for
// gfx950: IC(0) on a plan's pattern (the factor the solves take) and ILU(0)
// of a general CSR matrix.
//
// IC(0).  Reference: sparse_blas/spic0.cpp:11-46 (spic0_csc_left:
// left-looking on the CSC lower triangle of A, diagonal first; :48-99 the
// same column routine under an LBC schedule) and :106-127 (spic0_csr:
// right-looking, in place on the same arrays; :130-160 under an LBC schedule
// with atomic updates).  Both are restated per entry (r, c) of L, in the CSR
// of L the plan holds (diagonal last):
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
//
// ILU(0).  Reference: sparse_blas/spilu0.cpp:11-51 (spilu0_csr, row by row,
// in place on a copy of A; :54-110 the same row routine under an LBC
// schedule).  Per entry (i, c) of A:
//   v = A[i,c];  v -= L[i,k] * U[k,c]  for k in K(i,c), ascending;
//   L[i,c] = v / U[c,c]  (c < i);   U[i,c] = v  (c >= i)
// with K(i,c) = { k < min(i, c) : A stores (i, k) and (k, c) }: row i visits
// its k < i in ascending order, divides L[i,k] by U[k,k] and subtracts
// L[i,k] * U[k,c] from every entry of row i whose column c appears in the
// upper part of row k.
//
// In every case the updates of one entry come in ascending order of the
// eliminated index (the outer loop of the reference), so a lane that applies
// them in that order repeats the reference's operations on that entry.
// Products, subtractions, divisions and square roots are IEEE operations
// compiled without contraction, so the factors are bit-identical.
//
// Inspector: for every entry the (position in its own row, position in the
// other row) pairs, counted, scanned, then written.  Executor: the sync-free
// unit schedule of the forward solve over the lower pattern (row i waits for
// the rows k of its lower part, which are exactly the rows it reads); one
// lane per row computes the row's entries in order.
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

// ILU(0) structure: every row ascending, columns in [0, n), the diagonal at
// adiag[i].  err = 1 otherwise.
__global__ void k_ilu_structure(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                const int *__restrict__ adiag, int *err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b0 = rp[i], b1 = rp[i + 1], d = adiag[i];
  bool ok = b0 <= d && d < b1 && ci[d] == i;
  for (int p = b0; ok && p < b1; ++p)
    ok = ci[p] >= 0 && ci[p] < n && (p == b0 || ci[p - 1] < ci[p]);
  if (!ok) atomicOr(err, 1);
}

// ILU(0) pairs of row i: for every k = ci[a] < i in ascending order, the
// columns c that both the upper part of row k (positions b) and row i after
// a (positions t) store: entry t subtracts l[a] * l[b].  COUNT: cnt[t] += 1
// (and the divisor position of every entry); else written at cur[t]++.
template <bool COUNT>
__global__ void k_ilu_pairs(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                            const int *__restrict__ adiag, int64_t nnz, int64_t *cnt,
                            int *dpos, int *pa, int *pb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  if (i == n) {
    if (COUNT) cnt[nnz] = 0;
    return;
  }
  const int b0 = rp[i], b1 = rp[i + 1];
  if (COUNT)
    for (int t = b0; t < b1; ++t) {
      cnt[t] = 0;
      dpos[t] = ci[t] < i ? adiag[ci[t]] : -1;
    }
  for (int a = b0; a < adiag[i]; ++a) {
    const int k = ci[a];
    int b = adiag[k] + 1, t = a + 1;
    const int be = rp[k + 1];
    while (b < be && t < b1) {
      const int cb = ci[b], ct = ci[t];
      if (cb == ct) {
        if (COUNT) {
          cnt[t] += 1;
        } else {
          const int64_t w = cnt[t]++;
          pa[w] = a;
          pb[w] = b;
        }
        ++b;
        ++t;
      } else if (cb < ct) {
        ++b;
      } else {
        ++t;
      }
    }
  }
}

constexpr int MODE_IC0 = 0, MODE_ILU0 = 1;

struct FactorArgs {
  const int *rp, *ci;        // the factored pattern in the caller's order
  const double *Ax;          // A's values on it
  double *Lx;                // the factor (may alias Ax)
  const int64_t *pp;         // pairs of position p: [pp[p], pp[p+1])
  const int *pa, *pb;        //   own-row and other-row positions
  const int *dpos;           // divisor position of p (IC(0): off-diagonals;
                             // ILU(0): -1 for no division)
  const int *adiag;          // ILU(0): position of each row's diagonal
  const int *perm;           // level position -> row
  const int *ub, *dptr, *dlist;  // units and their dependency units
  int *uflags, *state;
  int *bad;                  // smallest row whose pivot failed (INT_MAX: none)
  int U, ticketed, total_waves, variant;
  uint64_t timeout;
};

// Rows of at most SHORT_RB entries keep their values in registers and the
// lane's LDS slots, and are stored once at the end, so no load waits behind
// a store.
constexpr int SHORT_RB = 8;

// The pivot test of each routine: spic0_csc_left stops at f <= 0 (a NaN
// pivot passes); spic0_csr asserts pivot > 0 (a NaN pivot fails);
// spilu0_csr asserts U[i,i] != 0.
template <int MODE>
__device__ __forceinline__ bool pivot_fails(int variant, double v) {
  if (MODE == MODE_ILU0) return v == 0.0;
  return variant == SYMAMD_IC0_LEFT ? v <= 0.0 : !(v > 0.0);
}

// Finish entry `off` of row r (offset `dr` is the diagonal): divide by the
// divisor (when it has one), or take the IC(0) square root on the diagonal.
template <int MODE>
__device__ __forceinline__ void finish_entry(const FactorArgs &a, int r, int off, int dr,
                                             bool has_div, double dg, double &v) {
  if (has_div) v = v / dg;
  if (off == dr) {
    if (pivot_fails<MODE>(a.variant, v)) atomicMin(a.bad, r);
    if (MODE == MODE_IC0) v = __builtin_sqrt(v);
  }
}

// A short row's first SHORT_PB pairs, staged in the lane's own LDS slots
// (read back by pair index, which registers cannot be without scratch
// memory): the offset in this row of the first operand, the offset of the
// second when it is this row's too (the IC(0) diagonal) or -1, and the
// other row's value otherwise.
constexpr int SHORT_PB = 8;

struct PairSlots {
  double own[SHORT_RB];  // this row's finished entries, read by offset
  double y[SHORT_PB];    // other rows' operands
  int yp[SHORT_PB];      // their positions
  int16_t ia[SHORT_PB];
  int16_t ib[SHORT_PB];
};

// A short row in two phases.  prep: everything that depends only on the
// structure and on A (values of A, pair ranges, divisor positions, the pairs
// staged in the slots) -- issued before the wait for the rows it reads, so
// those loads overlap the wait.  fin: one round of loads of the other rows'
// values (divisors and pair operands), then the arithmetic in entry order,
// then one store per entry.
struct ShortRow {
  double ax[SHORT_RB];
  int64_t pq[SHORT_RB + 1];
  int dp[SHORT_RB];
  int64_t t0;
  bool staged;
};

template <int MODE>
__device__ __forceinline__ void short_prep(const FactorArgs &a, int b0, int m, int dr,
                                           ShortRow &w, PairSlots &ps) {
#pragma unroll
  for (int j = 0; j < SHORT_RB; ++j) w.ax[j] = j < m ? a.Ax[b0 + j] : 0.0;
#pragma unroll
  for (int j = 0; j <= SHORT_RB; ++j) w.pq[j] = j <= m ? a.pp[b0 + j] : 0;
#pragma unroll
  for (int j = 0; j < SHORT_RB; ++j)
    w.dp[j] = j < m && (MODE == MODE_ILU0 || j != dr) ? a.dpos[b0 + j] : -1;
  w.t0 = w.pq[0];
  int64_t t1 = 0;
#pragma unroll
  for (int j = 1; j <= SHORT_RB; ++j) t1 = j == m ? w.pq[j] : t1;
  w.staged = t1 - w.t0 <= SHORT_PB;
  if (w.staged) {
    const int np = (int)(t1 - w.t0);
#pragma unroll
    for (int k = 0; k < SHORT_PB; ++k) {
      const int ia = k < np ? a.pa[w.t0 + k] : b0;
      const int ib = k < np ? a.pb[w.t0 + k] : b0;
      ps.ia[k] = (int16_t)(ia - b0);
      // another row's operand: its position, loaded after the wait
      ps.ib[k] = (int16_t)(ib < b0 ? -1 : ib - b0);
      ps.yp[k] = ib;
    }
  }
}

template <int MODE>
__device__ __forceinline__ void short_fin(const FactorArgs &a, int r, int b0, int m, int dr,
                                          const ShortRow &w, PairSlots &ps) {
  double dg[SHORT_RB], own[SHORT_RB];
#pragma unroll
  for (int j = 0; j < SHORT_RB; ++j) dg[j] = w.dp[j] >= 0 ? ld_dev(&a.Lx[w.dp[j]]) : 1.0;
  if (w.staged) {
    double y[SHORT_PB];
#pragma unroll
    for (int k = 0; k < SHORT_PB; ++k) y[k] = ps.ib[k] < 0 ? ld_dev(&a.Lx[ps.yp[k]]) : 0.0;
#pragma unroll
    for (int k = 0; k < SHORT_PB; ++k) ps.y[k] = y[k];
  }
#pragma unroll
  for (int j = 0; j < SHORT_RB; ++j) {
    if (j < m) {
      double v = w.ax[j];
      if (w.staged) {
        for (int k = (int)(w.pq[j] - w.t0), ke = (int)(w.pq[j + 1] - w.t0); k < ke; ++k) {
          const int kb = ps.ib[k];
          const double y = kb >= 0 ? ps.own[kb] : ps.y[k];
          v -= ps.own[ps.ia[k]] * y;
        }
      } else {
        // the pair's first operand is this row's; the second another row's,
        // or this row's on the IC(0) diagonal (c == r)
        for (int64_t t = w.pq[j]; t < w.pq[j + 1]; ++t) {
          const int ib = a.pb[t];
          const double x = ps.own[a.pa[t] - b0];
          const double y = ib >= b0 ? ps.own[ib - b0] : ld_dev(&a.Lx[ib]);
          v -= x * y;
        }
      }
      finish_entry<MODE>(a, r, j, dr, w.dp[j] >= 0, dg[j], v);
      own[j] = v;
      ps.own[j] = v;
    }
  }
#pragma unroll
  for (int j = 0; j < SHORT_RB; ++j)
    if (j < m) st_dev(&a.Lx[b0 + j], own[j]);
}

// Longer rows: entry by entry through memory (a lane reads back its own
// row's entries, written earlier by the same lane, with device-scope loads).
template <int MODE>
__device__ __forceinline__ void long_row(const FactorArgs &a, int r, int b0, int m, int dr) {
  for (int j = 0; j < m; ++j) {
    const int p = b0 + j;
    double v = a.Ax[p];
    for (int64_t t = a.pp[p], te = a.pp[p + 1]; t < te; ++t)
      v -= ld_dev(&a.Lx[a.pa[t]]) * ld_dev(&a.Lx[a.pb[t]]);
    const int dp = MODE == MODE_ILU0 || j != dr ? a.dpos[p] : -1;
    finish_entry<MODE>(a, r, j, dr, dp >= 0, dp >= 0 ? ld_dev(&a.Lx[dp]) : 1.0, v);
    st_dev(&a.Lx[p], v);
  }
}

// One wave per unit (a run of at most 64 rows of one level), one lane per
// row.  Values of other rows are read with device-scope loads once their
// units carry this launch's tag.
template <int MODE>
__global__ __launch_bounds__(256) void k_factor0(FactorArgs a) {
  __shared__ PairSlots slots[256];
  const int lane = threadIdx.x & 63;
  const int target = ld_dev(&a.state[ST_EPOCH]) + 1;
  int cursor = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  for (;;) {
    const int u = next_unit(a.state, a.ticketed, cursor, a.total_waves);
    if (u >= a.U) break;
    const int q = a.ub[u] + lane;
    const bool act = q < a.ub[u + 1];
    int r = 0, b0 = 0, m = 0, dr = 0;
    ShortRow w;
    if (act) {
      r = a.perm[q];
      b0 = a.rp[r];
      m = a.rp[r + 1] - b0;
      dr = MODE == MODE_IC0 ? m - 1 : a.adiag[r] - b0;
      if (m <= SHORT_RB) short_prep<MODE>(a, b0, m, dr, w, slots[threadIdx.x]);
    }
    wait_units(a.dptr, a.dlist, u, a.uflags, target, a.state, a.timeout);
    if (act) {
      if (m <= SHORT_RB)
        short_fin<MODE>(a, r, b0, m, dr, w, slots[threadIdx.x]);
      else
        long_row<MODE>(a, r, b0, m, dr);
    }
    publish_fence();
    if (lane == 0) st_dev(&a.uflags[u], target);
  }
  persistent_finish(a.state, a.total_waves);
}

bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
  return st != hipStreamCaptureStatusNone;
}

// Persistent launch over the plan's forward units (co-resident grid,
// cooperative unless capturing; else units from the ticket counter), then
// wait for it and return the failed-pivot row (-1: none).
template <int MODE>
int launch_factor0(symamd_handle h, symamd_plan p, FactorArgs a, int *bad_row) {
  UnitSet *us = nullptr;
  SYMAMD_TRY(plan_units(h, p, 64, false, &us));
  if (!p->d_ic0_bad) SYMAMD_HIP(hipMalloc(&p->d_ic0_bad, sizeof(int)));
  k_set_int<<<1, 1, 0, h->stream>>>(p->d_ic0_bad, INT_MAX);
  SYMAMD_HIP(hipGetLastError());
  a.perm = p->d_perm;
  a.ub = us->d_ub;
  a.dptr = us->d_dptr;
  a.dlist = us->d_dlist;
  a.uflags = us->d_uflags;
  a.state = p->d_state;
  a.bad = p->d_ic0_bad;
  a.U = us->count;
  a.timeout = h->timeout_ticks();
  const void *kfn = (const void *)&k_factor0<MODE>;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, 256, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  int64_t cap = (int64_t)per_cu * h->num_cus;
  if (h->max_waves > 0 && (h->max_waves + 3) / 4 < cap) cap = (h->max_waves + 3) / 4;
  int64_t want = (a.U + 3) / 4;
  if (want < 1) want = 1;
  const int grid = (int)(want < cap ? want : cap);
  a.total_waves = grid * 4;
  bool launched = false;
  if (!stream_capturing(h->stream) && !h->force_tickets) {
    a.ticketed = 0;
    void *args[] = {&a};
    if (hipLaunchCooperativeKernel(kfn, dim3(grid), dim3(256), args, 0, h->stream) ==
        hipSuccess)
      launched = true;
    else
      (void)hipGetLastError();
  }
  if (!launched) {
    a.ticketed = 1;
    void *args[] = {&a};
    SYMAMD_HIP(hipLaunchKernel(kfn, dim3(grid), dim3(256), args, 0, h->stream));
  }
  int bad = INT_MAX;
  SYMAMD_HIP(hipMemcpyAsync(&bad, p->d_ic0_bad, sizeof(int), hipMemcpyDeviceToHost,
                            h->stream));
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  *bad_row = bad == INT_MAX ? -1 : bad;
  return SYMAMD_OK;
}

// Device buffers released on every exit path.
struct Scratch {
  std::vector<void *> ptrs;
  ~Scratch() {
    for (void *q : ptrs)
      if (q) (void)hipFree(q);
  }
  template <typename T>
  int alloc(T **out, size_t count) {
    *out = nullptr;
    SYMAMD_HIP(hipMalloc((void **)out, sizeof(T) * (count ? count : 1)));
    ptrs.push_back(*out);
    return SYMAMD_OK;
  }
};

int scan64(symamd_handle h, const int64_t *in, int64_t *out, int64_t count, Scratch &sc) {
  size_t bytes = 0;
  SYMAMD_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, (int64_t)0, (size_t)count,
                                     rocprim::plus<int64_t>(), h->stream));
  uint8_t *tmp = nullptr;
  SYMAMD_TRY(sc.alloc(&tmp, bytes));
  SYMAMD_HIP(rocprim::exclusive_scan(tmp, bytes, in, out, (int64_t)0, (size_t)count,
                                     rocprim::plus<int64_t>(), h->stream));
  return SYMAMD_OK;
}

// The IC(0) pair lists of a variant, built on first use and kept by the plan.
int plan_ic0_pairs(symamd_handle h, symamd_plan p, int variant, Ic0Pairs **out) {
  auto it = p->ic0.find(variant);
  if (it != p->ic0.end()) {
    *out = &it->second;
    return SYMAMD_OK;
  }
  SYMAMD_TRY(plan_oci(h, p));
  const int n = p->n;
  const int64_t nnz = p->nnz;
  Scratch sc, keep;
  Ic0Pairs pr;
  int *d_last = nullptr;
  int64_t *d_cnt = nullptr;
  SYMAMD_TRY(sc.alloc(&d_last, n));
  SYMAMD_TRY(sc.alloc(&d_cnt, nnz + 1));
  SYMAMD_TRY(keep.alloc(&pr.d_pp, nnz + 1));
  SYMAMD_TRY(keep.alloc(&pr.d_dpos, nnz));
  SYMAMD_HIP(hipMemsetAsync(d_last, 0xff, sizeof(int) * (n > 0 ? n : 1), h->stream));
  if (n > 0) {
    k_last_row<<<blocks_of(n), 256, 0, h->stream>>>(n, p->d_orp, p->d_oci, d_last);
    SYMAMD_HIP(hipGetLastError());
  }
  if (variant == SYMAMD_IC0_LEFT)
    k_pair_count<SYMAMD_IC0_LEFT><<<blocks_of(n + 1), 256, 0, h->stream>>>(
        n, p->d_orp, p->d_oci, d_last, nnz, d_cnt, pr.d_dpos);
  else
    k_pair_count<SYMAMD_IC0_CSR><<<blocks_of(n + 1), 256, 0, h->stream>>>(
        n, p->d_orp, p->d_oci, d_last, nnz, d_cnt, pr.d_dpos);
  SYMAMD_HIP(hipGetLastError());
  SYMAMD_TRY(scan64(h, d_cnt, pr.d_pp, nnz + 1, sc));
  SYMAMD_HIP(hipMemcpyAsync(&pr.pairs, pr.d_pp + nnz, sizeof(int64_t), hipMemcpyDeviceToHost,
                            h->stream));
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  SYMAMD_TRY(keep.alloc(&pr.d_pa, pr.pairs));
  SYMAMD_TRY(keep.alloc(&pr.d_pb, pr.pairs));
  if (n > 0) {
    if (variant == SYMAMD_IC0_LEFT)
      k_pair_fill<SYMAMD_IC0_LEFT><<<blocks_of(n), 256, 0, h->stream>>>(
          n, p->d_orp, p->d_oci, d_last, pr.d_pp, pr.d_pa, pr.d_pb);
    else
      k_pair_fill<SYMAMD_IC0_CSR><<<blocks_of(n), 256, 0, h->stream>>>(
          n, p->d_orp, p->d_oci, d_last, pr.d_pp, pr.d_pa, pr.d_pb);
    SYMAMD_HIP(hipGetLastError());
  }
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  keep.ptrs.clear();  // owned by the plan from here on
  auto ins = p->ic0.emplace(variant, pr);
  *out = &ins.first->second;
  return SYMAMD_OK;
}

}  // namespace

int ic0_factor(symamd_handle h, symamd_plan p, int variant, const double *d_Ax, double *d_Lx,
               int *bad_row) {
  Ic0Pairs *pr = nullptr;
  SYMAMD_TRY(plan_ic0_pairs(h, p, variant, &pr));
  FactorArgs a{};
  a.rp = p->d_orp;
  a.ci = p->d_oci;
  a.Ax = d_Ax;
  a.Lx = d_Lx;
  a.pp = pr->d_pp;
  a.pa = pr->d_pa;
  a.pb = pr->d_pb;
  a.dpos = pr->d_dpos;
  a.adiag = nullptr;
  a.variant = variant;
  return launch_factor0<MODE_IC0>(h, p, a, bad_row);
}

void ilu0_destroy(symamd_ilu0_s *q) {
  if (!q) return;
  void *ptrs[] = {q->d_Ap, q->d_Ai, q->d_Adiag, q->d_pp, q->d_pa, q->d_pb, q->d_dpos};
  for (void *x : ptrs)
    if (x) (void)hipFree(x);
  if (q->sched) symamd_plan_destroy(q->sched);
  delete q;
}

namespace {

int ilu0_analyze_into(symamd_handle h, symamd_ilu0_s *q, const int *d_Ap, const int *d_Ai,
                      const int *d_Adiag) {
  const int n = q->n;
  int nnz32 = 0;
  SYMAMD_HIP(hipMemcpyAsync(&nnz32, d_Ap + n, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  const int64_t nnz = nnz32;
  q->nnz = nnz;
  // the plan keeps its own copy of A's pattern
  SYMAMD_HIP(hipMalloc(&q->d_Ap, sizeof(int) * (n + 1)));
  SYMAMD_HIP(hipMalloc(&q->d_Ai, sizeof(int) * (nnz > 0 ? nnz : 1)));
  SYMAMD_HIP(hipMalloc(&q->d_Adiag, sizeof(int) * (n > 0 ? n : 1)));
  SYMAMD_HIP(hipMemcpyAsync(q->d_Ap, d_Ap, sizeof(int) * (n + 1), hipMemcpyDeviceToDevice,
                            h->stream));
  if (nnz > 0)
    SYMAMD_HIP(hipMemcpyAsync(q->d_Ai, d_Ai, sizeof(int) * nnz, hipMemcpyDeviceToDevice,
                              h->stream));
  if (n > 0)
    SYMAMD_HIP(hipMemcpyAsync(q->d_Adiag, d_Adiag, sizeof(int) * n, hipMemcpyDeviceToDevice,
                              h->stream));
  if (n == 0) return SYMAMD_OK;
  Scratch sc;
  int *d_err = nullptr;
  SYMAMD_TRY(sc.alloc(&d_err, 1));
  SYMAMD_HIP(hipMemsetAsync(d_err, 0, sizeof(int), h->stream));
  k_ilu_structure<<<blocks_of(n), 256, 0, h->stream>>>(n, q->d_Ap, q->d_Ai, q->d_Adiag, d_err);
  SYMAMD_HIP(hipGetLastError());
  int err = 0;
  SYMAMD_HIP(hipMemcpyAsync(&err, d_err, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  if (err)
    return fail(SYMAMD_ERR_STRUCTURE,
                "ILU(0): rows must be ascending with the diagonal at A_diag[i]");
  // the schedule: the plan of A's lower part with a unit diagonal (row i
  // waits for the rows k < i it stores)
  int *d_Lp = nullptr, *d_Li = nullptr;
  double *d_Lx = nullptr, *d_zero = nullptr;
  SYMAMD_TRY(sc.alloc(&d_Lp, n + 1));
  SYMAMD_TRY(sc.alloc(&d_Li, nnz));
  SYMAMD_TRY(sc.alloc(&d_Lx, nnz));
  SYMAMD_TRY(sc.alloc(&d_zero, nnz));
  SYMAMD_HIP(hipMemsetAsync(d_zero, 0, sizeof(double) * nnz, h->stream));
  int lnnz = 0;
  SYMAMD_TRY(ilu_lower(h, n, q->d_Ap, q->d_Ai, d_zero, q->d_Adiag, 1, d_Lp, d_Li, d_Lx, &lnnz));
  SYMAMD_TRY(symamd_analyze(h, n, d_Lp, d_Li, d_Lx, SYMAMD_F64, &q->sched));
  // update pairs and divisors on A's pattern
  int64_t *d_cnt = nullptr;
  SYMAMD_TRY(sc.alloc(&d_cnt, nnz + 1));
  SYMAMD_HIP(hipMalloc(&q->d_pp, sizeof(int64_t) * (nnz + 1)));
  SYMAMD_HIP(hipMalloc(&q->d_dpos, sizeof(int) * nnz));
  k_ilu_pairs<true><<<blocks_of(n + 1), 256, 0, h->stream>>>(
      n, q->d_Ap, q->d_Ai, q->d_Adiag, nnz, d_cnt, q->d_dpos, nullptr, nullptr);
  SYMAMD_HIP(hipGetLastError());
  SYMAMD_TRY(scan64(h, d_cnt, q->d_pp, nnz + 1, sc));
  SYMAMD_HIP(hipMemcpyAsync(&q->pairs, q->d_pp + nnz, sizeof(int64_t), hipMemcpyDeviceToHost,
                            h->stream));
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  SYMAMD_HIP(hipMalloc(&q->d_pa, sizeof(int) * (q->pairs > 0 ? q->pairs : 1)));
  SYMAMD_HIP(hipMalloc(&q->d_pb, sizeof(int) * (q->pairs > 0 ? q->pairs : 1)));
  // write cursors start at each entry's first pair
  SYMAMD_HIP(hipMemcpyAsync(d_cnt, q->d_pp, sizeof(int64_t) * (nnz + 1),
                            hipMemcpyDeviceToDevice, h->stream));
  k_ilu_pairs<false><<<blocks_of(n), 256, 0, h->stream>>>(
      n, q->d_Ap, q->d_Ai, q->d_Adiag, nnz, d_cnt, nullptr, q->d_pa, q->d_pb);
  SYMAMD_HIP(hipGetLastError());
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  return SYMAMD_OK;
}

}  // namespace

int ilu0_analyze(symamd_handle h, int n, const int *d_Ap, const int *d_Ai, const int *d_Adiag,
                 symamd_ilu0_s **out) {
  auto *q = new symamd_ilu0_s();
  q->n = n;
  const int rc = ilu0_analyze_into(h, q, d_Ap, d_Ai, d_Adiag);
  if (rc != SYMAMD_OK) {
    ilu0_destroy(q);
    return rc;
  }
  *out = q;
  return SYMAMD_OK;
}

int ilu0_numeric(symamd_handle h, symamd_ilu0_s *q, const double *d_Ax, double *d_l,
                 int *bad_row) {
  *bad_row = -1;
  if (q->n == 0) return SYMAMD_OK;
  FactorArgs a{};
  a.rp = q->d_Ap;
  a.ci = q->d_Ai;
  a.Ax = d_Ax;
  a.Lx = d_l;
  a.pp = q->d_pp;
  a.pa = q->d_pa;
  a.pb = q->d_pb;
  a.dpos = q->d_dpos;
  a.adiag = q->d_Adiag;
  a.variant = 0;
  SYMAMD_TRY(launch_factor0<MODE_ILU0>(h, q->sched, a, bad_row));
  return symamd_plan_check(h, q->sched);
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
