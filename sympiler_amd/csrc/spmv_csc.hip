// spmv_csc.hip -- the column-oriented SpMV forms of sparse_blas/spmv.cpp as
// row pulls, bit-identical to the reference's serial loops.
//
// The reference scatters y[Ai[j]] += Ax[j] * x[c] column by column
// (spmv_csc, spmv.cpp:73-81; the axpby forms :85-109; spmv_csc_block
// :171-186), so row r receives its products in the order its columns are
// visited; spmv_csc_lbc (:122-167) sums each part's products into a zeroed
// buffer in the order of `ids`, then adds the parts' buffers to y in part
// order.  Here an inspector lists every (row, visit position, entry) of the
// visited columns and sorts the list by (row, visit position); one lane per
// row then adds in exactly that order -- starting from y[r] (direct forms) or
// forming each part's sum from +0.0 and adding it to y[r] (LBC form).  The
// file is compiled with -ffp-contract=off.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "common.h"

namespace symamd {

namespace {

inline unsigned grid_for(int64_t m) { return (unsigned)((m + 255) / 256); }

// Column visited at position k: ids[k0 + k], or c0 + k without a list.
__device__ __forceinline__ int visit_col(const int *ids, int k0, int c0, int k) {
  return ids ? ids[k0 + k] : c0 + k;
}

// len[k] = entries of the column visited at position k (0 for a column
// outside [0, n)); len[V] = 0 closes the scan.
__global__ void k_visit_len(int V, int n, const int *__restrict__ ids, int k0, int c0,
                            const int *__restrict__ Ap, int *len, int *bad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > V) return;
  if (k == V) {
    len[V] = 0;
    return;
  }
  const int c = visit_col(ids, k0, c0, k);
  if (c < 0 || c >= n) {
    atomicOr(bad, 1);
    len[k] = 0;
    return;
  }
  len[k] = Ap[c + 1] - Ap[c];
}

// The entries of position k at off[k]: key (row << 32 | k), value the entry;
// rows outside [0, m) are flagged and sorted past every row.
__global__ void k_visit_fill(int V, int n, int m, const int *__restrict__ ids, int k0, int c0,
                             const int *__restrict__ Ap, const int *__restrict__ Ai,
                             const int *__restrict__ off, uint64_t *key, int *ent, int *bad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= V) return;
  const int c = visit_col(ids, k0, c0, k);
  if (c < 0 || c >= n) return;
  int o = off[k];
  for (int p = Ap[c]; p < Ap[c + 1]; ++p, ++o) {
    int r = Ai[p];
    if (r < 0 || r >= m) {
      atomicOr(bad, 2);
      r = m;
    }
    key[o] = ((uint64_t)(unsigned)r << 32) | (unsigned)k;
    ent[o] = p;
  }
}

// rp[r] = first sorted position of row r (rows 0..m; row m holds the
// flagged entries and everything after them).
__global__ void k_row_ptr(int64_t total, int m, const uint64_t *__restrict__ key, int *rp) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int r = (int)(key[t] >> 32);
  const int prev = t == 0 ? -1 : (int)(key[t - 1] >> 32);
  for (int q = prev + 1; q <= r && q <= m; ++q) rp[q] = (int)t;
  if (t == total - 1)
    for (int q = r + 1; q <= m; ++q) rp[q] = (int)total;
}

__global__ void k_fill_rows(int m, int v, int *rp) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r <= m) rp[r] = v;
}

// part[k] = the part of visit position k (sets[l] <= k0 + k < sets[l+1]),
// positions clamped to [k0, k0 + V).
__global__ void k_visit_part(int nparts, const int *__restrict__ sets, int k0, int V,
                             int *part) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nparts) return;
  const int kb = max(sets[l], k0), ke = min(sets[l + 1], k0 + V);
  for (int k = kb; k < ke; ++k) part[k - k0] = l;
}

__global__ void k_zero_int(int *v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *v = 0;
}

// One lane per row: its products in visit order.  Direct forms: y[r] plus
// each product in turn.  LBC form: each part's products summed from +0.0,
// each nonempty part's sum added to y[r] in part order.  The reference also
// adds the +0.0 of every part without products of this row; that changes y
// only from -0.0 to +0.0, and only before any other addition (no part sum is
// -0.0: a sum started from +0.0 never is), so -0.0 is replaced up front when
// there is at least one part.
__global__ void k_pull(int m, const int *__restrict__ rp, const uint64_t *__restrict__ key,
                       const int *__restrict__ ent, const int *__restrict__ ids, int k0,
                       int c0, const double *__restrict__ Ax, const double *__restrict__ x,
                       const int *__restrict__ part, int nparts, double *y) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  double acc = y[r];
  const int q0 = rp[r], q1 = rp[r + 1];
  if (!part) {
    for (int q = q0; q < q1; ++q) {
      const int k = (int)(unsigned)key[q];
      acc = acc + Ax[ent[q]] * x[visit_col(ids, k0, c0, k)];
    }
  } else {
    if (nparts > 0 && __double_as_longlong(acc) == (long long)0x8000000000000000ULL) acc = 0.0;
    int cur = -1;
    double t = 0.0;
    for (int q = q0; q < q1; ++q) {
      const int k = (int)(unsigned)key[q];
      const int l = part[k];
      if (l != cur) {
        if (cur >= 0) acc = acc + t;
        cur = l;
        t = 0.0;
      }
      t = t + Ax[ent[q]] * x[visit_col(ids, k0, c0, k)];
    }
    if (cur >= 0) acc = acc + t;
  }
  y[r] = acc;
}

__global__ void k_axpby(int n, const double *__restrict__ av, const double *__restrict__ bv,
                        double a, double b, const double *__restrict__ y, double *z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double ai = av ? av[i] : a, bi = bv ? bv[i] : b;
  z[i] = ai * z[i] + bi * y[i];
}

}  // namespace

// y[r] (+)= the products of the V visited columns (ids[k0 .. k0+V) or
// c0 .. c0+V-1), in the reference's order; with d_sets, the LBC form.
int spmv_pull(symamd_handle h, int n, int m, const int *d_Ap, const int *d_Ai,
              const double *d_Ax, const double *d_x, double *d_y, const int *d_ids, int k0,
              int c0, int V, int nparts, const int *d_sets) {
  if (m == 0) return SYMAMD_OK;
  int *len = nullptr, *off = nullptr, *bad = nullptr, *rp = nullptr, *ent = nullptr,
      *sent = nullptr, *part = nullptr;
  uint64_t *key = nullptr, *skey = nullptr;
  void *tmp = nullptr;
  int rc = SYMAMD_OK;
  auto alloc = [&](void **p, size_t bytes) {
    if (rc == SYMAMD_OK && hipMallocAsync(p, bytes ? bytes : 4, h->stream) != hipSuccess)
      rc = fail(SYMAMD_ERR_ALLOC, "spmv (CSC): scratch allocation failed");
  };
  alloc((void **)&len, sizeof(int) * ((size_t)V + 1));
  alloc((void **)&off, sizeof(int) * ((size_t)V + 1));
  alloc((void **)&bad, sizeof(int));
  alloc((void **)&rp, sizeof(int) * ((size_t)m + 1));
  int h_bad = 0, total = 0;
  if (rc == SYMAMD_OK) {
    k_zero_int<<<1, 64, 0, h->stream>>>(bad);
    k_visit_len<<<grid_for((int64_t)V + 1), 256, 0, h->stream>>>(V, n, d_ids, k0, c0, d_Ap,
                                                                 len, bad);
    size_t bytes = 0;
    if (rocprim::exclusive_scan(nullptr, bytes, len, off, 0, (size_t)V + 1,
                                rocprim::plus<int>(), h->stream) != hipSuccess)
      rc = fail(SYMAMD_ERR_HIP, "spmv (CSC): scan sizing failed");
    else {
      alloc(&tmp, bytes);
      if (rc == SYMAMD_OK &&
          rocprim::exclusive_scan(tmp, bytes, len, off, 0, (size_t)V + 1,
                                  rocprim::plus<int>(), h->stream) != hipSuccess)
        rc = fail(SYMAMD_ERR_HIP, "spmv (CSC): scan failed");
    }
    if (rc == SYMAMD_OK &&
        (hipMemcpyAsync(&total, off + V, sizeof(int), hipMemcpyDeviceToHost, h->stream) !=
             hipSuccess ||
         hipStreamSynchronize(h->stream) != hipSuccess))
      rc = fail(SYMAMD_ERR_HIP, "spmv (CSC): reading the entry count failed");
  }
  alloc((void **)&key, sizeof(uint64_t) * (size_t)total);
  alloc((void **)&skey, sizeof(uint64_t) * (size_t)total);
  alloc((void **)&ent, sizeof(int) * (size_t)total);
  alloc((void **)&sent, sizeof(int) * (size_t)total);
  if (rc == SYMAMD_OK && total > 0) {
    k_visit_fill<<<grid_for(V), 256, 0, h->stream>>>(V, n, m, d_ids, k0, c0, d_Ap, d_Ai, off,
                                                      key, ent, bad);
    int end_bit = 32;
    while (end_bit < 64 && (1ull << (end_bit - 32)) <= (unsigned long long)m) ++end_bit;
    size_t bytes = 0;
    void *stmp = nullptr;
    if (rocprim::radix_sort_pairs(nullptr, bytes, key, skey, ent, sent, (size_t)total, 0,
                                  end_bit, h->stream) != hipSuccess)
      rc = fail(SYMAMD_ERR_HIP, "spmv (CSC): sort sizing failed");
    alloc(&stmp, bytes);
    if (rc == SYMAMD_OK &&
        rocprim::radix_sort_pairs(stmp, bytes, key, skey, ent, sent, (size_t)total, 0,
                                  end_bit, h->stream) != hipSuccess)
      rc = fail(SYMAMD_ERR_HIP, "spmv (CSC): sort failed");
    if (stmp) (void)hipFreeAsync(stmp, h->stream);
    if (rc == SYMAMD_OK) k_row_ptr<<<grid_for(total), 256, 0, h->stream>>>(total, m, skey, rp);
  } else if (rc == SYMAMD_OK) {
    k_fill_rows<<<grid_for((int64_t)m + 1), 256, 0, h->stream>>>(m, 0, rp);
  }
  if (rc == SYMAMD_OK && d_sets) {
    alloc((void **)&part, sizeof(int) * (size_t)V);
    if (rc == SYMAMD_OK && nparts > 0)
      k_visit_part<<<grid_for(nparts), 256, 0, h->stream>>>(nparts, d_sets, k0, V, part);
  }
  if (rc == SYMAMD_OK)
    k_pull<<<grid_for(m), 256, 0, h->stream>>>(m, rp, skey, sent, d_ids, k0, c0, d_Ax, d_x,
                                                d_sets ? part : nullptr, nparts, d_y);
  if (rc == SYMAMD_OK &&
      (hipGetLastError() != hipSuccess ||
       hipMemcpyAsync(&h_bad, bad, sizeof(int), hipMemcpyDeviceToHost, h->stream) !=
           hipSuccess ||
       hipStreamSynchronize(h->stream) != hipSuccess))
    rc = fail(SYMAMD_ERR_HIP, "spmv (CSC): launch failed");
  for (void *p : {(void *)len, (void *)off, (void *)bad, (void *)rp, (void *)ent,
                  (void *)sent, (void *)part, (void *)key, (void *)skey, tmp})
    if (p) (void)hipFreeAsync(p, h->stream);
  if (rc == SYMAMD_OK && h_bad)
    rc = fail(SYMAMD_ERR_STRUCTURE, h_bad & 1 ? "spmv (CSC): a visited column lies outside [0, n)"
                                              : "spmv (CSC): a row index lies outside [0, m)");
  return rc;
}

int launch_axpby(symamd_handle h, int n, const double *d_a, const double *d_b, double a,
                 double b, const double *d_y, double *d_z) {
  if (n == 0) return SYMAMD_OK;
  k_axpby<<<grid_for(n), 256, 0, h->stream>>>(n, d_a, d_b, a, b, d_y, d_z);
  SYMAMD_HIP(hipGetLastError());
  return SYMAMD_OK;
}

}  // namespace symamd
