// inspector.hip -- GPU inspector for the sparse lower-triangular solve:
// structure validation, dependency levels, level sets, the level-ordered
// copy of L that the executors stream, and the transposed copy the adjoint
// streams.
//
// Reference: the level-set inspector build_levelSet_CSC lives in the
// un-vendored aggregation submodule (call sites Catch_tests/
// sptrsv_tests.cpp:59, demo/sptrsv_demo_utils.h:44); the definition
// implemented here is the one the in-tree analog
// sparse_blas/parsy/common/TreeUtils.cpp:89-141 (getLevelSet) computes:
// ASAP levels, rows ascending inside a level.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <map>
#include <mutex>

#include "common.h"

namespace symamd {

namespace {

// ------------------------------------------------------ structure check
// Bit 0: a row with no entries; bit 1: last entry of a row is not its
// diagonal; bit 2: columns not strictly ascending or out of range.
__global__ void k_check_structure(int n, const int *__restrict__ Lp,
                                  const int *__restrict__ Li, int *flag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int b = Lp[i], e = Lp[i + 1];
  int bad = 0;
  if (e <= b) {
    bad |= 1;
  } else {
    if (Li[e - 1] != i) bad |= 2;
    int prev = -1;
    for (int j = b; j < e; ++j) {
      int c = Li[j];
      if (c <= prev || c < 0 || c > i) { bad |= 4; break; }
      prev = c;
    }
  }
  if (bad) atomicOr(flag, bad);
}

// ------------------------------------------------------------- levels
// Sync-free longest path over the DAG of L in natural row order.  Waves
// take chunks of 64 consecutive rows from a ticket counter, so every row a
// wave waits on is held by a wave that is already running.  lev[] starts at
// -1 and is its own ready flag: lev[j] >= 0 means row j is done.  A lane
// whose dependency is not ready yet keeps its place and retries on the next
// pass of the wave-wide loop, so lanes of one wave never block each other
// (row i may depend on row i-1 in the same chunk).
__global__ void k_levels(int n, const int *__restrict__ Lp,
                         const int *__restrict__ Li, int *lev, int *state,
                         uint64_t timeout) {
  const int lane = threadIdx.x & 63;
  int local_max = -1;
  for (;;) {
    int chunk = 0;
    if (lane == 0) chunk = atomicAdd(&state[ST_TICKET], 1);
    chunk = __shfl(chunk, 0);
    const int64_t row0 = (int64_t)chunk * 64;
    if (row0 >= n) break;
    const int i = (int)row0 + lane;
    bool done = i >= n;
    int j = done ? 0 : Lp[i];
    const int je = done ? 0 : Lp[i + 1] - 1;
    int m = -1;
    uint64_t t0 = 0;
    int spins = 0;
    for (;;) {
      if (!done) {
        while (j < je) {
          int v = ld_dev(&lev[Li[j]]);
          if (v < 0) break;
          m = v > m ? v : m;
          ++j;
        }
        if (j >= je) {
          st_dev(&lev[i], m + 1);
          local_max = m + 1 > local_max ? m + 1 : local_max;
          done = true;
        }
      }
      if (__all(done)) break;
      if ((++spins & 255) == 0) {
        uint64_t now = wall_ticks();
        if (t0 == 0) t0 = now;
        else if (now - t0 > timeout) {
          // give up: publish something so dependants drain, report error
          if (!done) {
            atomicOr(&state[ST_ERROR], 1);
            st_dev(&lev[i], m + 1);
            done = true;
          }
        }
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  // wave max -> global max
  for (int off = 32; off > 0; off >>= 1) {
    int o = __shfl_xor(local_max, off);
    local_max = o > local_max ? o : local_max;
  }
  if (lane == 0) atomicMax(&state[ST_SCRATCH], local_max);
}

// level_ptr from level-sorted keys: position p starts level key[p] when it
// differs from key[p-1] (levels are contiguous 0..nlev-1).
__global__ void k_level_ptr(int n, int nlev, const int *__restrict__ skey,
                            int *level_ptr) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int l = skey[p];
  int prev = p == 0 ? -1 : skey[p - 1];
  for (int t = prev + 1; t <= l; ++t) level_ptr[t] = p;
  if (p == n - 1) level_ptr[nlev] = n;
}

// ----------------------------------------------------- plan arrays
__global__ void k_perm_rowlen(int n, const int *__restrict__ perm,
                              const int *__restrict__ Lp, int *len) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) {
    int i = perm[p];
    len[p] = Lp[i + 1] - Lp[i];
  }
  if (p == n) len[p] = 0;
}

// One wave per 64 consecutive level-ordered rows; each lane copies one
// row's entries (rows are short, 1-8 entries for most workloads).
template <typename T>
__global__ void k_perm_copy(int n, const int *__restrict__ perm,
                            const int *__restrict__ Lp,
                            const int *__restrict__ Li,
                            const T *__restrict__ Lx,
                            const int *__restrict__ rp, int *ci, T *val,
                            int *fmap) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int i = perm[p];
  int src = Lp[i], len = Lp[i + 1] - src, dst = rp[p];
  for (int t = 0; t < len; ++t) {
    ci[dst + t] = Li[src + t];
    if (Lx) val[dst + t] = Lx[src + t];
    fmap[dst + t] = src + t;
  }
}

template <typename T>
__global__ void k_gather(int64_t m, const int *__restrict__ map,
                         const T *__restrict__ src, T *dst) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) dst[t] = src[map[t]];
}

template <typename T>
__global__ void k_diag_zero(int n, const int *__restrict__ rp,
                            const T *__restrict__ val, int *flag) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n && val[rp[p + 1] - 1] == T(0)) atomicOr(flag, 8);
}

// Transpose keys: entry t of the level-ordered copy is (row perm[p],
// column ci[t]); key = column * 2^32 + row orders it column-major with rows
// ascending (diagonal first in each column of a lower triangle).
__global__ void k_transpose_keys(int n, const int *__restrict__ perm,
                                 const int *__restrict__ rp,
                                 const int *__restrict__ ci, uint64_t *key,
                                 int *idx) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint64_t row = (uint32_t)perm[p];
  for (int t = rp[p]; t < rp[p + 1]; ++t) {
    key[t] = ((uint64_t)(uint32_t)ci[t] << 32) | row;
    idx[t] = t;
  }
}

// Column pointers of the natural-order CSC from column-sorted keys.
__global__ void k_col_ptr(int64_t nnz, int n, const uint64_t *__restrict__ key,
                          int *cp) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnz) return;
  int c = (int)(key[t] >> 32);
  int prev = t == 0 ? -1 : (int)(key[t - 1] >> 32);
  for (int q = prev + 1; q <= c; ++q) cp[q] = (int)t;
  if (t == nnz - 1)
    for (int q = c + 1; q <= n; ++q) cp[q] = (int)nnz;
}

// Level-ordered transposed copy: position q holds column perm[q] of L.
template <typename T>
__global__ void k_transpose_copy(int n, const int *__restrict__ perm,
                                 const int *__restrict__ cp,
                                 const uint64_t *__restrict__ skey,
                                 const int *__restrict__ sidx,
                                 const int *__restrict__ trp,
                                 const int *__restrict__ fmap,
                                 const T *__restrict__ val, int *tci, T *tval,
                                 int *tmap) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int c = perm[q];
  int src = cp[c], len = cp[c + 1] - src, dst = trp[q];
  for (int t = 0; t < len; ++t) {
    uint64_t k = skey[src + t];
    int e = sidx[src + t];
    tci[dst + t] = (int)(k & 0xffffffffu);
    tval[dst + t] = val[e];
    tmap[dst + t] = fmap[e];
  }
}

__global__ void k_collen(int n, const int *__restrict__ perm,
                         const int *__restrict__ cp, int *len) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) {
    int c = perm[q];
    len[q] = cp[c + 1] - cp[c];
  }
  if (q == n) len[q] = 0;
}

// Work units: level l splits into ceil(size/rows) units of at most `rows`
// consecutive positions.  upl = units per level, scanned into uoff.
// Work units: each level is cut, in order, into runs of at most `rows`
// rows holding at most `emax` entries of the row copy `rp` (a single row
// longer than emax forms a unit of its own).  One thread per level scans
// its rows twice: first to count the units (upl), then, after the scan of
// the counts into uoff, to write the unit starts (ub).
// With a region order (preg != null) a unit also ends where the region of
// the positions changes, so every unit lies inside one region.
// One wavefront scans a level: the lengths (and regions) of 64 positions are
// loaded together, then the greedy cut runs over them on uniform values.
__device__ __forceinline__ int unit_scan(const int *__restrict__ rp,
                                         const int *__restrict__ preg, int q0,
                                         int q1, int rows, int emax,
                                         int *ub_out) {
  const int lane = threadIdx.x & 63;
  int cnt = 0, in_rows = 0, ents = 0, prev = 0;
  for (int b = q0; b < q1; b += 64) {
    const int q = b + lane;
    int len = 0, reg = 0;
    if (q < q1) {
      len = rp[q + 1] - rp[q];
      reg = preg ? preg[q] : 0;
    }
    const int m = q1 - b < 64 ? q1 - b : 64;
    for (int j = 0; j < m; ++j) {
      const int lj = __builtin_amdgcn_readlane(len, j);
      const int rj = __builtin_amdgcn_readlane(reg, j);
      if (in_rows == rows || (in_rows > 0 && ents + lj > emax) ||
          (in_rows > 0 && preg && rj != prev)) {
        in_rows = 0;
        ents = 0;
      }
      if (in_rows == 0) {
        if (ub_out && lane == 0) ub_out[cnt] = b + j;
        ++cnt;
      }
      ++in_rows;
      ents += lj;
      prev = rj;
    }
  }
  return cnt;
}

// One wavefront per level (wave nlev writes the trailing zero).
__global__ void k_units_per_level(int nlev, const int *__restrict__ lp,
                                  const int *__restrict__ rp,
                                  const int *__restrict__ preg, int rows,
                                  int emax, int *upl) {
  const int l = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (l < nlev) {
    const int c = unit_scan(rp, preg, lp[l], lp[l + 1], rows, emax, nullptr);
    if ((threadIdx.x & 63) == 0) upl[l] = c;
  }
  if (l == nlev && (threadIdx.x & 63) == 0) upl[l] = 0;
}

__global__ void k_unit_bounds(int nlev, int U, int n, const int *__restrict__ lp,
                              const int *__restrict__ rp,
                              const int *__restrict__ preg,
                              const int *__restrict__ uoff, int rows, int emax,
                              int *ub) {
  const int l = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (l < nlev) unit_scan(rp, preg, lp[l], lp[l + 1], rows, emax, ub + uoff[l]);
  if (l == nlev && (threadIdx.x & 63) == 0) ub[U] = n;
}

inline int blocks_for(int64_t m, int bs = 256) {
  return (int)((m + bs - 1) / bs);
}

// exclusive scan of len[0..m] (m+1 entries, len[m] = 0) into out[0..m]
int scan_exclusive(symamd_handle h, const int *len, int *out, int64_t m1) {
  size_t bytes = 0;
  SYMAMD_HIP(rocprim::exclusive_scan(nullptr, bytes, len, out, 0, (size_t)m1,
                                     rocprim::plus<int>(), h->stream));
  void *tmp = nullptr;
  SYMAMD_HIP(hipMallocAsync(&tmp, bytes ? bytes : 4, h->stream));
  hipError_t e = rocprim::exclusive_scan(tmp, bytes, len, out, 0, (size_t)m1,
                                         rocprim::plus<int>(), h->stream);
  SYMAMD_HIP(hipFreeAsync(tmp, h->stream));
  SYMAMD_HIP(e);
  return SYMAMD_OK;
}

// Lower factor stored inside a general CSR matrix A with diagonal positions
// diags[i] (the L of ILU(0), as sparse_blas/sptrsv.cpp:28-38 and :109-126
// read it: entries [Ap[i], diags[i]) and the value at diags[i]).
// len[i] = row length of the extracted factor; bad |= 1 when diags[i] lies
// outside row i.
__global__ void k_ilu_rowlen(int n, const int *__restrict__ Ap,
                             const int *__restrict__ diags, int *len, int *bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  if (i == n) {
    len[n] = 0;
    return;
  }
  const int d = diags[i];
  const bool ok = d >= Ap[i] && d < Ap[i + 1];
  if (!ok) atomicOr(bad, 1);
  len[i] = ok ? d - Ap[i] + 1 : 1;
}

__global__ void k_ilu_copy(int n, const int *__restrict__ Ap, const int *__restrict__ Ai,
                           const double *__restrict__ Ax, const int *__restrict__ diags,
                           int unit, const int *__restrict__ Lp, int *Li, double *Lx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int a0 = Ap[i], d = diags[i], o = Lp[i];
  for (int j = a0; j < d; ++j) {
    Li[o + j - a0] = Ai[j];
    Lx[o + j - a0] = Ax[j];
  }
  Li[o + d - a0] = i;
  Lx[o + d - a0] = unit ? 1.0 : Ax[d];
}

}  // namespace

int ilu_lower(symamd_handle h, int n, const int *d_Ap, const int *d_Ai,
              const double *d_Ax, const int *d_diags, int unit, int *d_Lp,
              int *d_Li, double *d_Lx, int *nnz_out) {
  if (n == 0) {
    SYMAMD_HIP(hipMemsetAsync(d_Lp, 0, sizeof(int), h->stream));
    SYMAMD_HIP(hipStreamSynchronize(h->stream));
    *nnz_out = 0;
    return SYMAMD_OK;
  }
  int *len = nullptr;  // n+1 row lengths, then the error flag
  SYMAMD_HIP(hipMallocAsync(&len, sizeof(int) * (n + 2), h->stream));
  int *bad = len + n + 1;
  int rc = SYMAMD_OK;
  int hb[2] = {0, 0};
  if (hipMemsetAsync(bad, 0, sizeof(int), h->stream) != hipSuccess) {
    rc = fail(SYMAMD_ERR_HIP, "hipMemsetAsync failed");
  } else {
    k_ilu_rowlen<<<blocks_for(n + 1), 256, 0, h->stream>>>(n, d_Ap, d_diags, len, bad);
    rc = scan_exclusive(h, len, d_Lp, n + 1);
  }
  if (rc == SYMAMD_OK &&
      (hipMemcpyAsync(&hb[0], bad, sizeof(int), hipMemcpyDeviceToHost, h->stream) !=
           hipSuccess ||
       hipMemcpyAsync(&hb[1], d_Lp + n, sizeof(int), hipMemcpyDeviceToHost, h->stream) !=
           hipSuccess ||
       hipStreamSynchronize(h->stream) != hipSuccess))
    rc = fail(SYMAMD_ERR_HIP, "reading the factor size failed");
  if (rc == SYMAMD_OK && hb[0])
    rc = fail(SYMAMD_ERR_INVALID_ARG, "diags[i] does not lie inside row i");
  if (rc == SYMAMD_OK) {
    k_ilu_copy<<<blocks_for(n), 256, 0, h->stream>>>(n, d_Ap, d_Ai, d_Ax, d_diags, unit,
                                                     d_Lp, d_Li, d_Lx);
    if (hipGetLastError() != hipSuccess) rc = fail(SYMAMD_ERR_HIP, "k_ilu_copy launch failed");
  }
  (void)hipFreeAsync(len, h->stream);
  if (rc == SYMAMD_OK) {
    SYMAMD_HIP(hipStreamSynchronize(h->stream));
    *nnz_out = hb[1];
  }
  return rc;
}

namespace {

// v[0 .. m) = value (a kernel, not a memset node, so callers may capture it).
__global__ void k_fill_int(int64_t m, int value, int *v) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) v[t] = value;
}

// Levels and level sets with the caller's buffers: lev and skey of n ints,
// the sort's temporary storage tmp of tmp_bytes (nullptr: allocated here).
int levels_core(symamd_handle h, int n, const int *d_Lp, const int *d_Li, int *d_state,
                int *d_lev, int *d_skey, void *tmp, size_t tmp_bytes, int *d_level_ptr,
                int *d_level_set, int *levels, std::vector<int> *h_level_ptr) {
  *levels = 0;
  if (n == 0) {
    k_fill_int<<<1, 64, 0, h->stream>>>(1, 0, d_level_ptr);
    SYMAMD_HIP(hipGetLastError());
    SYMAMD_HIP(hipStreamSynchronize(h->stream));
    if (h_level_ptr) h_level_ptr->assign(1, 0);
    return SYMAMD_OK;
  }
  k_fill_int<<<blocks_for(n), 256, 0, h->stream>>>(n, -1, d_lev);
  k_fill_int<<<1, 64, 0, h->stream>>>(ST_WORDS, 0, d_state);
  int waves_needed = (n + 63) / 64;
  int grid = (waves_needed + 3) / 4;
  int cap = h->num_cus * 8;
  if (grid > cap) grid = cap;
  k_levels<<<grid, 256, 0, h->stream>>>(n, d_Lp, d_Li, d_lev, d_state,
                                        h->timeout_ticks());
  SYMAMD_HIP(hipGetLastError());
  int st[ST_WORDS];
  SYMAMD_HIP(hipMemcpyAsync(st, d_state, sizeof(st), hipMemcpyDeviceToHost,
                            h->stream));
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  if (st[ST_ERROR]) return fail(SYMAMD_ERR_TIMEOUT, "level inspector watchdog expired");
  const int nlev = st[ST_SCRATCH] + 1;
  unsigned end_bit = 1;
  while ((1u << end_bit) <= (unsigned)(nlev - 1) && end_bit < 31) ++end_bit;
  size_t bytes = 0;
  rocprim::counting_iterator<int> rows(0);
  SYMAMD_HIP(rocprim::radix_sort_pairs(nullptr, bytes, (unsigned *)d_lev,
                                       (unsigned *)d_skey, rows, d_level_set,
                                       (size_t)n, 0, end_bit, h->stream));
  void *own = nullptr;
  if (!tmp) {
    SYMAMD_HIP(hipMallocAsync(&own, bytes ? bytes : 4, h->stream));
    tmp = own;
  } else if (bytes > tmp_bytes) {
    return fail(SYMAMD_ERR_NOT_SUPPORTED,
                "level sets of %d rows need %zu bytes of sort storage, more than the %zu "
                "of the fixed buffer", n, bytes, tmp_bytes);
  }
  SYMAMD_HIP(rocprim::radix_sort_pairs(tmp, bytes, (unsigned *)d_lev,
                                       (unsigned *)d_skey, rows, d_level_set,
                                       (size_t)n, 0, end_bit, h->stream));
  k_level_ptr<<<blocks_for(n), 256, 0, h->stream>>>(n, nlev, d_skey,
                                                    d_level_ptr);
  SYMAMD_HIP(hipGetLastError());
  if (own) SYMAMD_HIP(hipFreeAsync(own, h->stream));
  if (h_level_ptr) {
    h_level_ptr->resize(nlev + 1);
    SYMAMD_HIP(hipMemcpyAsync(h_level_ptr->data(), d_level_ptr,
                              sizeof(int) * (nlev + 1), hipMemcpyDeviceToHost,
                              h->stream));
  }
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  *levels = nlev;
  return SYMAMD_OK;
}

}  // namespace

int check_structure(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
                    int *d_state) {
  if (n == 0) return SYMAMD_OK;
  k_fill_int<<<1, 64, 0, h->stream>>>(1, 0, d_state + ST_SCRATCH);
  k_check_structure<<<blocks_for(n), 256, 0, h->stream>>>(n, d_Lp, d_Li,
                                                          d_state + ST_SCRATCH);
  SYMAMD_HIP(hipGetLastError());
  int flag = 0;
  SYMAMD_HIP(hipMemcpyAsync(&flag, d_state + ST_SCRATCH, sizeof(int),
                            hipMemcpyDeviceToHost, h->stream));
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  if (flag & 1) return fail(SYMAMD_ERR_STRUCTURE, "a row has no entries (no diagonal)");
  if (flag & 2) return fail(SYMAMD_ERR_STRUCTURE, "a row's last entry is not its diagonal (CSR lower triangle needs the diagonal last)");
  if (flag & 4) return fail(SYMAMD_ERR_STRUCTURE, "column indices not strictly ascending, negative, or above the diagonal");
  return SYMAMD_OK;
}

int compute_levels(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
                   int *d_state, int *d_level_ptr, int *d_level_set,
                   int *levels, std::vector<int> *h_level_ptr) {
  int *d_lev = nullptr, *d_skey = nullptr;
  if (n > 0) {
    SYMAMD_HIP(hipMallocAsync(&d_lev, sizeof(int) * n, h->stream));
    SYMAMD_HIP(hipMallocAsync(&d_skey, sizeof(int) * n, h->stream));
  }
  const int rc = levels_core(h, n, d_Lp, d_Li, d_state, d_lev, d_skey, nullptr, 0, d_level_ptr,
                             d_level_set, levels, h_level_ptr);
  if (d_lev) (void)hipFreeAsync(d_lev, h->stream);
  if (d_skey) (void)hipFreeAsync(d_skey, h->stream);
  if (n > 0) SYMAMD_HIP(hipStreamSynchronize(h->stream));
  return rc;
}

// ------------------------------------------------ allocation-free level sets
// The level-set inspector for callers that cannot allocate device memory:
// its state words, per-row levels, sort keys and sort storage live in one
// fixed device array (each GPU holds its own copy for the life of the
// process).  Host code reaches the array only through the address
// hipGetSymbolAddress gives for the current device, looked up once per
// device and kept.
constexpr int NOALLOC_MAX_ROWS = 1 << 21;
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)48 << 20;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
std::map<int, uint8_t *> g_scratch_addr;  // device -> its scratch address

int scratch_address(int device, uint8_t **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_addr.find(device);
  if (it != g_scratch_addr.end()) {
    *out = it->second;
    return SYMAMD_OK;
  }
  int cur = -1;
  SYMAMD_HIP(hipGetDevice(&cur));
  if (cur != device)
    return fail(SYMAMD_ERR_INVALID_ARG,
                "symamd_levelset_noalloc: the handle's device (%d) is not the current one (%d)",
                device, cur);
  uint8_t *p = nullptr;
  const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(SYMAMD_ERR_HIP, "symamd_levelset_noalloc: hipGetSymbolAddress failed: %s",
                hipGetErrorString(e));
  g_scratch_addr[device] = p;
  *out = p;
  return SYMAMD_OK;
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

int levelset_noalloc(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
                     int *d_level_ptr, int *d_level_set, int *levels) {
  if (n > NOALLOC_MAX_ROWS)
    return fail(SYMAMD_ERR_NOT_SUPPORTED,
                "symamd_levelset_noalloc: %d rows exceed the fixed buffer's limit of %d "
                "(use symamd_levelset)", n, NOALLOC_MAX_ROWS);
  uint8_t *base = nullptr;
  SYMAMD_TRY(scratch_address(h->device, &base));
  int *d_state = reinterpret_cast<int *>(base);
  const size_t o_lev = 256;
  const size_t o_skey = o_lev + align256(sizeof(int) * (size_t)n);
  const size_t o_tmp = o_skey + align256(sizeof(int) * (size_t)n);
  int *d_lev = reinterpret_cast<int *>(base + o_lev);
  int *d_skey = reinterpret_cast<int *>(base + o_skey);
  SYMAMD_TRY(check_structure(h, n, d_Lp, d_Li, d_state));
  return levels_core(h, n, d_Lp, d_Li, d_state, d_lev, d_skey, base + o_tmp,
                     NOALLOC_SCRATCH_BYTES - o_tmp, d_level_ptr, d_level_set, levels, nullptr);
}

int build_plan_arrays(symamd_handle h, symamd_plan p, const int *d_Lp,
                      const int *d_Li, const void *d_Lx) {
  const int n = p->n;
  int *d_len = nullptr;
  SYMAMD_HIP(hipMallocAsync(&d_len, sizeof(int) * (n + 1), h->stream));
  k_perm_rowlen<<<blocks_for(n + 1), 256, 0, h->stream>>>(n, p->d_perm, d_Lp,
                                                          d_len);
  SYMAMD_HIP(hipGetLastError());
  SYMAMD_TRY(scan_exclusive(h, d_len, p->d_rp, n + 1));
  SYMAMD_HIP(hipFreeAsync(d_len, h->stream));
  SYMAMD_HIP(hipMemcpyAsync(p->d_orp, d_Lp, sizeof(int) * (n + 1),
                            hipMemcpyDeviceToDevice, h->stream));
  if (n > 0) {
    if (p->dtype == SYMAMD_F64)
      k_perm_copy<double><<<blocks_for(n), 256, 0, h->stream>>>(
          n, p->d_perm, d_Lp, d_Li, (const double *)d_Lx, p->d_rp, p->d_ci,
          (double *)p->d_val, p->d_fmap);
    else
      k_perm_copy<float><<<blocks_for(n), 256, 0, h->stream>>>(
          n, p->d_perm, d_Lp, d_Li, (const float *)d_Lx, p->d_rp, p->d_ci,
          (float *)p->d_val, p->d_fmap);
    SYMAMD_HIP(hipGetLastError());
  }
  return SYMAMD_OK;
}

int gather_values(symamd_handle h, symamd_plan p, const void *d_Lx) {
  const int64_t m = p->nnz;
  if (m == 0) return SYMAMD_OK;
  if (p->dtype == SYMAMD_F64) {
    k_gather<double><<<blocks_for(m), 256, 0, h->stream>>>(
        m, p->d_fmap, (const double *)d_Lx, (double *)p->d_val);
    if (p->d_tval)
      k_gather<double><<<blocks_for(m), 256, 0, h->stream>>>(
          m, p->d_tmap, (const double *)d_Lx, (double *)p->d_tval);
  } else {
    k_gather<float><<<blocks_for(m), 256, 0, h->stream>>>(
        m, p->d_fmap, (const float *)d_Lx, (float *)p->d_val);
    if (p->d_tval)
      k_gather<float><<<blocks_for(m), 256, 0, h->stream>>>(
          m, p->d_tmap, (const float *)d_Lx, (float *)p->d_tval);
  }
  SYMAMD_HIP(hipGetLastError());
  SYMAMD_TRY(gather_region_values(h, p));
  // zero pivots
  SYMAMD_HIP(hipMemsetAsync(p->d_state + ST_SCRATCH, 0, sizeof(int), h->stream));
  if (p->dtype == SYMAMD_F64)
    k_diag_zero<double><<<blocks_for(p->n), 256, 0, h->stream>>>(
        p->n, p->d_rp, (const double *)p->d_val, p->d_state + ST_SCRATCH);
  else
    k_diag_zero<float><<<blocks_for(p->n), 256, 0, h->stream>>>(
        p->n, p->d_rp, (const float *)p->d_val, p->d_state + ST_SCRATCH);
  SYMAMD_HIP(hipGetLastError());
  int flag = 0;
  SYMAMD_HIP(hipMemcpyAsync(&flag, p->d_state + ST_SCRATCH, sizeof(int),
                            hipMemcpyDeviceToHost, h->stream));
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  if (flag & 8) return fail(SYMAMD_ERR_ZERO_PIVOT, "L has a zero diagonal entry");
  return SYMAMD_OK;
}

namespace {

// unit of every position (thread per unit; units hold at most a few hundred
// rows)
__global__ void k_unit_of_pos(int U, const int *__restrict__ ub, int *uop) {
  int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= U) return;
  for (int q = ub[u]; q < ub[u + 1]; ++q) uop[q] = u;
}

// Caller-order column ids from the level-ordered copy: entry t of the copy
// sits at caller position fmap[t].
__global__ void k_scatter_cols(int64_t nnz, const int *__restrict__ fmap,
                               const int *__restrict__ ci, int *oci) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nnz) oci[fmap[t]] = ci[t];
}

__global__ void k_inverse_perm(int n, const int *__restrict__ perm, int *ipos) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) ipos[perm[q]] = q;
}

// Dependency entries of a unit: the off-diagonal entries of its rows.  In
// the forward copy the diagonal is the last entry of a row, in the
// transposed copy the first; either way a row contributes len-1 entries.
__global__ void k_unit_dep_count(int U, const int *__restrict__ ub,
                                 const int *__restrict__ rp, int *cnt) {
  int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u > U) return;
  if (u == U) { cnt[U] = 0; return; }
  int c = 0;
  for (int q = ub[u]; q < ub[u + 1]; ++q) c += rp[q + 1] - rp[q] - 1;
  cnt[u] = c;
}

__global__ void k_unit_dep_keys(int U, const int *__restrict__ ub,
                                const int *__restrict__ rp,
                                const int *__restrict__ ci, int skip_first,
                                const int *__restrict__ ipos,
                                const int *__restrict__ uop,
                                const int *__restrict__ off, uint64_t *keys) {
  int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= U) return;
  int o = off[u];
  for (int q = ub[u]; q < ub[u + 1]; ++q) {
    const int b = rp[q] + (skip_first ? 1 : 0);
    const int e = rp[q + 1] - (skip_first ? 0 : 1);
    for (int j = b; j < e; ++j)
      keys[o++] = ((uint64_t)(uint32_t)u << 32) | (uint32_t)uop[ipos[ci[j]]];
  }
}

__global__ void k_mark_unique(int64_t m, const uint64_t *__restrict__ keys,
                              int *flag) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) flag[t] = (t == 0 || keys[t] != keys[t - 1]) ? 1 : 0;
  if (t == m) flag[m] = 0;
}

__global__ void k_compact_deps(int64_t m, const uint64_t *__restrict__ keys,
                               const int *__restrict__ flag,
                               const int *__restrict__ pos, int *dlist,
                               int *dcount) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m || !flag[t]) return;
  dlist[pos[t]] = (int)(keys[t] & 0xffffffffu);
  atomicAdd(&dcount[(int)(keys[t] >> 32)], 1);
}

// Deduplicated dependency-unit lists of every unit, from the row lists
// (rp, ci) of the level-ordered copy (forward) or of the transposed copy
// (reverse, skip_first = 1).
int build_unit_deps(symamd_handle h, symamd_plan p, const UnitSet &us,
                    const int *rp, const int *ci, int skip_first, int *uop,
                    int **d_dptr, int **d_dlist, int64_t *total) {
  const int U = us.count;
  int *d_cnt = nullptr, *d_off = nullptr;
  SYMAMD_HIP(hipMalloc(d_dptr, sizeof(int) * (U + 1)));
  SYMAMD_HIP(hipMallocAsync(&d_cnt, sizeof(int) * (U + 1), h->stream));
  SYMAMD_HIP(hipMallocAsync(&d_off, sizeof(int) * (U + 1), h->stream));
  k_unit_dep_count<<<blocks_for(U + 1), 256, 0, h->stream>>>(U, us.d_ub, rp, d_cnt);
  SYMAMD_HIP(hipGetLastError());
  SYMAMD_TRY(scan_exclusive(h, d_cnt, d_off, U + 1));
  int m = 0;
  SYMAMD_HIP(hipMemcpyAsync(&m, d_off + U, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  uint64_t *d_keys = nullptr, *d_skeys = nullptr;
  int *d_flag = nullptr, *d_pos = nullptr;
  const int64_t mm = m > 0 ? m : 1;
  SYMAMD_HIP(hipMallocAsync(&d_keys, 8 * mm, h->stream));
  SYMAMD_HIP(hipMallocAsync(&d_skeys, 8 * mm, h->stream));
  SYMAMD_HIP(hipMallocAsync(&d_flag, 4 * (mm + 1), h->stream));
  SYMAMD_HIP(hipMallocAsync(&d_pos, 4 * (mm + 1), h->stream));
  if (U > 0) {
    k_unit_dep_keys<<<blocks_for(U), 256, 0, h->stream>>>(U, us.d_ub, rp, ci, skip_first,
                                                          p->d_ipos, uop, d_off, d_keys);
    SYMAMD_HIP(hipGetLastError());
  }
  unsigned ubits = 1;
  while ((1u << ubits) < (unsigned)U && ubits < 31) ++ubits;
  if (m > 0) {
    size_t bytes = 0;
    SYMAMD_HIP(rocprim::radix_sort_keys(nullptr, bytes, d_keys, d_skeys, (size_t)m, 0,
                                        32 + ubits, h->stream));
    void *tmp = nullptr;
    SYMAMD_HIP(hipMallocAsync(&tmp, bytes ? bytes : 4, h->stream));
    SYMAMD_HIP(rocprim::radix_sort_keys(tmp, bytes, d_keys, d_skeys, (size_t)m, 0,
                                        32 + ubits, h->stream));
    SYMAMD_HIP(hipFreeAsync(tmp, h->stream));
  }
  k_mark_unique<<<blocks_for(m + 1), 256, 0, h->stream>>>(m, d_skeys, d_flag);
  SYMAMD_HIP(hipGetLastError());
  SYMAMD_TRY(scan_exclusive(h, d_flag, d_pos, m + 1));
  int uniq = 0;
  SYMAMD_HIP(hipMemcpyAsync(&uniq, d_pos + m, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  SYMAMD_HIP(hipMalloc(d_dlist, sizeof(int) * (uniq > 0 ? uniq : 1)));
  SYMAMD_HIP(hipMemsetAsync(d_cnt, 0, sizeof(int) * (U + 1), h->stream));
  if (m > 0) {  // a diagonal matrix has no dependencies
    k_compact_deps<<<blocks_for(m), 256, 0, h->stream>>>(m, d_skeys, d_flag, d_pos,
                                                          *d_dlist, d_cnt);
    SYMAMD_HIP(hipGetLastError());
  }
  SYMAMD_TRY(scan_exclusive(h, d_cnt, *d_dptr, U + 1));
  SYMAMD_HIP(hipFreeAsync(d_keys, h->stream));
  SYMAMD_HIP(hipFreeAsync(d_skeys, h->stream));
  SYMAMD_HIP(hipFreeAsync(d_flag, h->stream));
  SYMAMD_HIP(hipFreeAsync(d_pos, h->stream));
  SYMAMD_HIP(hipFreeAsync(d_cnt, h->stream));
  SYMAMD_HIP(hipFreeAsync(d_off, h->stream));
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  *total = uniq;
  return SYMAMD_OK;
}

}  // namespace

int plan_units(symamd_handle h, symamd_plan p, int rows, bool adjoint,
               UnitSet **out) {
  const int key = rows * 2 + (adjoint ? 1 : 0);
  auto it = p->units.find(key);
  if (it != p->units.end()) {
    *out = &it->second;
    return SYMAMD_OK;
  }
  if (adjoint) SYMAMD_TRY(plan_transpose(h, p));
  // units of the forward solve are cut by the entries of the forward copy,
  // those of the transposed solve by the entries of the transposed copy
  const int *rp = adjoint ? p->d_trp : p->d_rp;
  const int *ci = adjoint ? p->d_tci : p->d_ci;
  UnitSet us;
  us.rows = rows;
  const int nlev = p->levels;
  us.h_uoff.assign(nlev + 1, 0);
  if (nlev > 0) {
    int *d_upl = nullptr, *d_uoff = nullptr;
    SYMAMD_HIP(hipMallocAsync(&d_upl, sizeof(int) * (nlev + 1), h->stream));
    SYMAMD_HIP(hipMallocAsync(&d_uoff, sizeof(int) * (nlev + 1), h->stream));
    k_units_per_level<<<blocks_for(nlev + 1, 4), 256, 0, h->stream>>>(
        nlev, p->d_lvl_ptr, rp, p->d_preg, rows, UNIT_EMAX, d_upl);
    SYMAMD_HIP(hipGetLastError());
    SYMAMD_TRY(scan_exclusive(h, d_upl, d_uoff, nlev + 1));
    SYMAMD_HIP(hipMemcpyAsync(us.h_uoff.data(), d_uoff, sizeof(int) * (nlev + 1),
                              hipMemcpyDeviceToHost, h->stream));
    SYMAMD_HIP(hipStreamSynchronize(h->stream));
    const int U = us.h_uoff[nlev];
    us.count = U;
    SYMAMD_HIP(hipMalloc(&us.d_ub, sizeof(int) * (U + 1)));
    k_unit_bounds<<<blocks_for(nlev + 1, 4), 256, 0, h->stream>>>(
        nlev, U, p->n, p->d_lvl_ptr, rp, p->d_preg, d_uoff, rows, UNIT_EMAX, us.d_ub);
    SYMAMD_HIP(hipGetLastError());
    SYMAMD_HIP(hipFreeAsync(d_upl, h->stream));
    SYMAMD_HIP(hipFreeAsync(d_uoff, h->stream));
  } else {
    us.count = 0;
    SYMAMD_HIP(hipMalloc(&us.d_ub, sizeof(int)));
    SYMAMD_HIP(hipMemsetAsync(us.d_ub, 0, sizeof(int), h->stream));
  }
  const int U = us.count;
  SYMAMD_HIP(hipMalloc(&us.d_uflags, sizeof(int) * (U > 0 ? U : 1)));
  SYMAMD_HIP(hipMemsetAsync(us.d_uflags, 0, sizeof(int) * (U > 0 ? U : 1), h->stream));
  SYMAMD_TRY(plan_ipos(h, p));
  int *d_uop = nullptr;
  SYMAMD_HIP(hipMallocAsync(&d_uop, sizeof(int) * (p->n > 0 ? p->n : 1), h->stream));
  if (U > 0) {
    k_unit_of_pos<<<blocks_for(U), 256, 0, h->stream>>>(U, us.d_ub, d_uop);
    SYMAMD_HIP(hipGetLastError());
  }
  int rc = build_unit_deps(h, p, us, rp, ci, adjoint ? 1 : 0, d_uop, &us.d_dptr,
                           &us.d_dlist, &us.deps);
  (void)hipFreeAsync(d_uop, h->stream);
  (void)hipStreamSynchronize(h->stream);
  if (rc != SYMAMD_OK) {
    (void)hipFree(us.d_ub);
    (void)hipFree(us.d_uflags);
    return rc;
  }
  *out = &p->units.emplace(key, us).first->second;
  return SYMAMD_OK;
}

int plan_oci(symamd_handle h, symamd_plan p) {
  if (p->d_oci) return SYMAMD_OK;
  SYMAMD_HIP(hipMalloc(&p->d_oci, sizeof(int) * (p->nnz > 0 ? p->nnz : 1)));
  if (p->nnz > 0) {
    k_scatter_cols<<<blocks_for(p->nnz), 256, 0, h->stream>>>(p->nnz, p->d_fmap, p->d_ci,
                                                              p->d_oci);
    SYMAMD_HIP(hipGetLastError());
  }
  return SYMAMD_OK;
}

int plan_ipos(symamd_handle h, symamd_plan p) {
  if (p->d_ipos) return SYMAMD_OK;
  SYMAMD_HIP(hipMalloc(&p->d_ipos, sizeof(int) * (p->n > 0 ? p->n : 1)));
  if (p->n > 0) {
    k_inverse_perm<<<blocks_for(p->n), 256, 0, h->stream>>>(p->n, p->d_perm, p->d_ipos);
    SYMAMD_HIP(hipGetLastError());
  }
  return SYMAMD_OK;
}

int plan_transpose(symamd_handle h, symamd_plan p) {
  if (p->d_trp) return SYMAMD_OK;
  const int n = p->n;
  const int64_t nnz = p->nnz;
  const size_t vs = p->dtype == SYMAMD_F64 ? 8 : 4;
  SYMAMD_HIP(hipMalloc(&p->d_trp, sizeof(int) * (n + 1)));
  SYMAMD_HIP(hipMalloc(&p->d_tci, sizeof(int) * (nnz ? nnz : 1)));
  SYMAMD_HIP(hipMalloc(&p->d_tval, vs * (nnz ? nnz : 1)));
  SYMAMD_HIP(hipMalloc(&p->d_tmap, sizeof(int) * (nnz ? nnz : 1)));
  if (n == 0) {
    SYMAMD_HIP(hipMemsetAsync(p->d_trp, 0, sizeof(int), h->stream));
    SYMAMD_HIP(hipStreamSynchronize(h->stream));
    return SYMAMD_OK;
  }
  uint64_t *d_key = nullptr, *d_skey = nullptr;
  int *d_idx = nullptr, *d_sidx = nullptr, *d_cp = nullptr, *d_len = nullptr;
  SYMAMD_HIP(hipMallocAsync(&d_key, 8 * nnz, h->stream));
  SYMAMD_HIP(hipMallocAsync(&d_skey, 8 * nnz, h->stream));
  SYMAMD_HIP(hipMallocAsync(&d_idx, 4 * nnz, h->stream));
  SYMAMD_HIP(hipMallocAsync(&d_sidx, 4 * nnz, h->stream));
  SYMAMD_HIP(hipMallocAsync(&d_cp, 4 * (n + 1), h->stream));
  SYMAMD_HIP(hipMallocAsync(&d_len, 4 * (n + 1), h->stream));
  k_transpose_keys<<<blocks_for(n), 256, 0, h->stream>>>(n, p->d_perm, p->d_rp,
                                                         p->d_ci, d_key, d_idx);
  SYMAMD_HIP(hipGetLastError());
  size_t bytes = 0;
  SYMAMD_HIP(rocprim::radix_sort_pairs(nullptr, bytes, d_key, d_skey, d_idx,
                                       d_sidx, (size_t)nnz, 0, 64, h->stream));
  void *tmp = nullptr;
  SYMAMD_HIP(hipMallocAsync(&tmp, bytes ? bytes : 4, h->stream));
  SYMAMD_HIP(rocprim::radix_sort_pairs(tmp, bytes, d_key, d_skey, d_idx,
                                       d_sidx, (size_t)nnz, 0, 64, h->stream));
  SYMAMD_HIP(hipFreeAsync(tmp, h->stream));
  k_col_ptr<<<blocks_for(nnz), 256, 0, h->stream>>>(nnz, n, d_skey, d_cp);
  k_collen<<<blocks_for(n + 1), 256, 0, h->stream>>>(n, p->d_perm, d_cp, d_len);
  SYMAMD_HIP(hipGetLastError());
  SYMAMD_TRY(scan_exclusive(h, d_len, p->d_trp, n + 1));
  if (p->dtype == SYMAMD_F64)
    k_transpose_copy<double><<<blocks_for(n), 256, 0, h->stream>>>(
        n, p->d_perm, d_cp, d_skey, d_sidx, p->d_trp, p->d_fmap,
        (const double *)p->d_val, p->d_tci, (double *)p->d_tval, p->d_tmap);
  else
    k_transpose_copy<float><<<blocks_for(n), 256, 0, h->stream>>>(
        n, p->d_perm, d_cp, d_skey, d_sidx, p->d_trp, p->d_fmap,
        (const float *)p->d_val, p->d_tci, (float *)p->d_tval, p->d_tmap);
  SYMAMD_HIP(hipGetLastError());
  SYMAMD_HIP(hipFreeAsync(d_key, h->stream));
  SYMAMD_HIP(hipFreeAsync(d_skey, h->stream));
  SYMAMD_HIP(hipFreeAsync(d_idx, h->stream));
  SYMAMD_HIP(hipFreeAsync(d_sidx, h->stream));
  SYMAMD_HIP(hipFreeAsync(d_cp, h->stream));
  SYMAMD_HIP(hipFreeAsync(d_len, h->stream));
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  return SYMAMD_OK;
}

}  // namespace symamd
