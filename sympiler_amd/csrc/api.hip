// api.hip -- the extern "C" ABI of libsympiler_amd.so: handles, plans,
// solve/adjoint/SpMV dispatch, captured graphs, the schedule tuner and the
// coarsened-schedule inspector.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <numeric>

#include "common.h"

namespace symamd {

namespace {
thread_local std::string g_last_error;
std::map<uint64_t, std::map<int64_t, Choice>> g_tune_cache;  // fingerprint ->

uint64_t fnv1a(uint64_t h, const void *data, size_t len) {
  const unsigned char *b = (const unsigned char *)data;
  for (size_t i = 0; i < len; ++i) {
    h ^= b[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

// Element strides of an n-by-k block in `layout` with leading dimension ld.
int strides(int n, int k, int64_t ld, int layout, int64_t *rs, int64_t *cs) {
  if (layout == SYMAMD_COL_MAJOR) {
    if (k > 1 && ld < n) return fail(SYMAMD_ERR_INVALID_ARG, "column-major ld %lld < n %d", (long long)ld, n);
    *rs = 1;
    *cs = ld;
  } else if (layout == SYMAMD_ROW_MAJOR) {
    if (n > 1 && ld < k) return fail(SYMAMD_ERR_INVALID_ARG, "row-major ld %lld < k %d", (long long)ld, k);
    *rs = ld;
    *cs = 1;
  } else {
    return fail(SYMAMD_ERR_INVALID_ARG, "unknown layout %d", layout);
  }
  return SYMAMD_OK;
}

void free_plan_arrays(symamd_plan p) {
  void *ptrs[] = {p->d_lvl_ptr, p->d_perm, p->d_rp,   p->d_ci,   p->d_val,
                  p->d_fmap,    p->d_orp,  p->d_trp,  p->d_tci,  p->d_tval,
                  p->d_tmap,    p->d_state, p->d_ipos, p->d_cpos, p->d_ws,
                  p->d_lvl_set, p->d_preg,   p->d_oci};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  for (auto &kv : p->units) {
    UnitSet &u = kv.second;
    void *up[] = {u.d_ub,   u.d_dptr,  u.d_dlist,  u.d_uflags,
                  u.d_wptr, u.d_wlist, u.d_xdptr, u.d_xdlist};
    for (void *q : up)
      if (q) (void)hipFree(q);
  }
  for (auto &kv : p->split)
    if (kv.second.d_rec) (void)hipFree(kv.second.d_rec);
  for (auto &kv : p->rplans) free_region_plan(kv.second);
  for (auto &kv : p->rplans_t) free_region_plan(kv.second);
  free_ic0_pairs(p);
  for (auto &kv : p->gcache) {
    if (kv.second.second) (void)hipGraphExecDestroy(kv.second.second);
    if (kv.second.first) (void)hipGraphDestroy(kv.second.first);
  }
}

// Default schedule when nothing was tuned.
// One lane per row (bit-identical to the reference) unless rows are long on
// average; then enough lanes that a typical row takes two steps and the
// longest about eight (launches clip lanes x column tile to 64).  Measured on
// the Zipf(1.8) n = 1e6 matrix (average 8.7 entries, longest 257) with the
// sync-free executor: lanes 4 / 32 at k = 1: 11.5 / 4.3 ms; lanes 1 / 8 at
// k = 8: 30.8 / 5.7 ms; lanes 1 / 2 at k = 32: 33.2 / 16.6 ms.
Choice heuristic(symamd_plan p, int k) {
  (void)k;
  Choice c;
  c.exec = SYMAMD_EXEC_SYNCFREE;
  c.lanes = 1;
  if (p->avg_row_nnz >= 8.0)
    c.lanes = std::min(32, std::max(pow2_floor((int)(p->avg_row_nnz / 2)),
                                    pow2_ceil((p->max_row_nnz + 7) / 8)));
  return c;
}

// Longest row of L (zero on entry): a grid-stride maximum per lane, reduced
// over the wave and then the block's four waves, one atomic per block (one
// per wave made 32k atomics on one word for 2M rows, 0.38 ms).
__global__ __launch_bounds__(256) void k_max_row(int n, const int *__restrict__ Lp, int *out) {
  __shared__ int wave_max[4];
  int v = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    v = max(v, Lp[i + 1] - Lp[i]);
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
  if ((threadIdx.x & 63) == 0) wave_max[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    atomicMax(out, max(max(wave_max[0], wave_max[1]), max(wave_max[2], wave_max[3])));
}

// forward: a forward solve (the region executor is a candidate)
Choice choose(symamd_handle h, symamd_plan p, int k, int layout, int exec, bool forward) {
  Choice c;
  auto it = p->tuned.find(tune_key(k, layout));
  if (exec == SYMAMD_EXEC_AUTO) {
    if (it != p->tuned.end()) {
      c = it->second;
    } else {
      c = heuristic(p, k);
      if (forward && region_preferred(h, p, k, layout)) c.exec = SYMAMD_EXEC_REGIONS;
    }
  } else {
    c = heuristic(p, k);
    c.exec = exec;
    if (it != p->tuned.end() && it->second.exec == exec) {
      c.lanes = it->second.lanes;
      c.tile = it->second.tile;
    }
  }
  if (p->lanes_override > 0) c.lanes = p->lanes_override;
  return c;
}

}  // namespace

int tuned_region_cap(symamd_plan p, int k) {
  for (int layout : {SYMAMD_ROW_MAJOR, SYMAMD_COL_MAJOR}) {
    auto it = p->tuned.find(tune_key(k, layout));
    if (it != p->tuned.end() && it->second.exec == SYMAMD_EXEC_REGIONS && it->second.cap > 0)
      return it->second.cap;
  }
  return 0;
}

namespace {

int check_handle(symamd_handle h) {
  if (!h) return fail(SYMAMD_ERR_INVALID_ARG, "null handle");
  return SYMAMD_OK;
}

int check_plan(symamd_plan p) {
  if (!p) return fail(SYMAMD_ERR_INVALID_ARG, "null plan");
  return SYMAMD_OK;
}

int check_exec(int exec) {
  if (exec == SYMAMD_EXEC_WIDE)
    return fail(SYMAMD_ERR_NOT_SUPPORTED,
                "the persistent workgroup-per-row executor is disabled (its first run "
                "exceeded the watchdog; cause not established)");
  if (exec < SYMAMD_EXEC_AUTO || exec > SYMAMD_EXEC_LEVELSET_SPLIT)
    return fail(SYMAMD_ERR_INVALID_ARG, "unknown executor %d", exec);
  return SYMAMD_OK;
}

int analyze_common(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
                   const void *d_Lx, symamd_dtype dtype, symamd_plan *out) {
  if (dtype != SYMAMD_F64 && dtype != SYMAMD_F32)
    return fail(SYMAMD_ERR_INVALID_ARG, "unknown dtype %d", (int)dtype);
  auto *p = new symamd_plan_s();
  p->n = n;
  p->dtype = dtype;
  p->device = h->device;
  int rc = SYMAMD_OK;
  auto cleanup = [&](int code) {
    free_plan_arrays(p);
    delete p;
    return code;
  };
  int nnz_i = 0;
  int lp0 = 0;
  if (hipMemcpy(&nnz_i, d_Lp + n, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(&lp0, d_Lp, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    return cleanup(fail(SYMAMD_ERR_HIP, "cannot read Lp from the device"));
  if (lp0 != 0 || nnz_i < n) return cleanup(fail(SYMAMD_ERR_STRUCTURE, "Lp[0] must be 0 and Lp[n] >= n"));
  p->nnz = nnz_i;
  const size_t vs = dtype == SYMAMD_F64 ? 8 : 4;
  const size_t nn = (size_t)(n > 0 ? n : 1);
  const size_t mm = (size_t)(p->nnz > 0 ? p->nnz : 1);
  if (hipMalloc(&p->d_state, sizeof(int) * ST_WORDS) != hipSuccess ||
      hipMalloc(&p->d_perm, sizeof(int) * nn) != hipSuccess ||
      hipMalloc(&p->d_lvl_ptr, sizeof(int) * (nn + 1)) != hipSuccess ||
      hipMalloc(&p->d_rp, sizeof(int) * (nn + 1)) != hipSuccess ||
      hipMalloc(&p->d_orp, sizeof(int) * (nn + 1)) != hipSuccess ||
      hipMalloc(&p->d_ci, sizeof(int) * mm) != hipSuccess ||
      hipMalloc(&p->d_fmap, sizeof(int) * mm) != hipSuccess ||
      hipMalloc(&p->d_val, vs * mm) != hipSuccess)
    return cleanup(fail(SYMAMD_ERR_ALLOC, "device allocation for the plan failed (n=%d nnz=%lld)", n, (long long)p->nnz));
  if ((rc = check_structure(h, n, d_Lp, d_Li, p->d_state)) != SYMAMD_OK) return cleanup(rc);
  if ((rc = compute_levels(h, n, d_Lp, d_Li, p->d_state, p->d_lvl_ptr, p->d_perm,
                           &p->levels, &p->h_lvl_ptr)) != SYMAMD_OK)
    return cleanup(rc);
  if ((rc = order_by_regions(h, p, d_Lp, d_Li)) != SYMAMD_OK) return cleanup(rc);
  if ((rc = build_plan_arrays(h, p, d_Lp, d_Li, d_Lx)) != SYMAMD_OK) return cleanup(rc);
  if ((rc = gather_values(h, p, d_Lx)) != SYMAMD_OK) return cleanup(rc);
  if (n > 0) {
    // the state words are cleared below; ST_SCRATCH takes the longest row
    if (hipMemsetAsync(p->d_state + ST_SCRATCH, 0, sizeof(int), h->stream) != hipSuccess)
      return cleanup(fail(SYMAMD_ERR_HIP, "plan state initialisation failed"));
    k_max_row<<<std::min((n + 255) / 256, 1024), 256, 0, h->stream>>>(n, d_Lp,
                                                                    p->d_state + ST_SCRATCH);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&p->max_row_nnz, p->d_state + ST_SCRATCH, sizeof(int),
                       hipMemcpyDeviceToHost, h->stream) != hipSuccess)
      return cleanup(fail(SYMAMD_ERR_HIP, "row statistics failed"));
  }
  if (hipMemsetAsync(p->d_state, 0, sizeof(int) * ST_WORDS, h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess)
    return cleanup(fail(SYMAMD_ERR_HIP, "plan state initialisation failed"));
  for (int l = 0; l < p->levels; ++l)
    p->max_level_rows = std::max(p->max_level_rows, p->h_lvl_ptr[l + 1] - p->h_lvl_ptr[l]);
  p->avg_row_nnz = n > 0 ? (double)p->nnz / n : 1.0;
  uint64_t fp = 0xcbf29ce484222325ull;
  fp = fnv1a(fp, &n, sizeof(n));
  fp = fnv1a(fp, &p->nnz, sizeof(p->nnz));
  fp = fnv1a(fp, &dtype, sizeof(dtype));
  fp = fnv1a(fp, p->h_lvl_ptr.data(), sizeof(int) * p->h_lvl_ptr.size());
  p->fingerprint = fp;
  auto it = g_tune_cache.find(fp);
  if (it != g_tune_cache.end()) p->tuned = it->second;
  *out = p;
  return SYMAMD_OK;
}

// Adds each row's system offset to its local column ids.
__global__ void k_globalise(int n, int nsys, const int *__restrict__ off,
                            const int *__restrict__ Lp,
                            const int *__restrict__ Li, int *Gi) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = 0, hi = nsys - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= i) lo = mid; else hi = mid - 1;
  }
  const int base = off[lo];
  for (int j = Lp[i]; j < Lp[i + 1]; ++j) Gi[j] = Li[j] + base;
}

__global__ void k_fill(int64_t m, double v, void *dst, int is64) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  if (is64) ((double *)dst)[t] = v;
  else ((float *)dst)[t] = (float)v;
}

int solve_impl(symamd_handle h, symamd_plan p, void *d_B, int k, int64_t ld,
               int layout, int exec, bool graph_ok);

template <typename W>
__global__ void k_copy_strided(int n, const W *src, int64_t ss, W *dst, int64_t ds) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[(int64_t)i * ds] = src[(int64_t)i * ss];
}

// dst[i*ds] = src[i*ss] for n elements of vs (4 or 8) bytes.
int copy_strided(symamd_handle h, int n, size_t vs, const void *src, int64_t ss, void *dst,
                 int64_t ds) {
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (vs == 8)
    k_copy_strided<uint64_t><<<blocks, 256, 0, h->stream>>>(n, (const uint64_t *)src, ss,
                                                            (uint64_t *)dst, ds);
  else
    k_copy_strided<uint32_t><<<blocks, 256, 0, h->stream>>>(n, (const uint32_t *)src, ss,
                                                            (uint32_t *)dst, ds);
  SYMAMD_HIP(hipGetLastError());
  return SYMAMD_OK;
}

}  // namespace

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

int fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

namespace {

std::string graph_key(void *d_B, int k, int64_t rs, int64_t cs, const Choice &c) {
  char buf[160];
  snprintf(buf, sizeof buf, "%p/%d/%lld/%lld/%d/%d/%d/%d", d_B, k, (long long)rs,
           (long long)cs, c.exec, c.lanes, c.tile, c.cap);
  return buf;
}

// Record the per-level launches of one solve into a graph on a private
// capture stream.  Units for the sync-free path are prepared beforehand so
// nothing allocates while capturing.
// Column-major blocks of several right-hand sides are solved through the
// level-ordered workspace (coalesced); everything else in place.
bool through_workspace(symamd_handle h, int k, int layout) {
  return layout == SYMAMD_COL_MAJOR && k > 1 && !h->colmajor_direct;
}

// Everything a solve allocates or builds, done before any launch so that a
// solve can be captured into a graph.
int prepare_solve(symamd_handle h, symamd_plan p, int k, int layout, const Choice &c) {
  if (c.exec == SYMAMD_EXEC_REGIONS && region_usable(h, p, k))
    return through_workspace(h, k, layout) ? prepare_workspace(h, p, k) : SYMAMD_OK;
  if (c.exec == SYMAMD_EXEC_LEVELSET_SPLIT) {
    SplitLists *sl = nullptr;
    SYMAMD_TRY(split_lists(h, p, split_threshold(h, tile_cols(k, c.tile)), &sl));
    return through_workspace(h, k, layout) ? prepare_workspace(h, p, k) : SYMAMD_OK;
  }
  UnitSet *us = nullptr;
  SYMAMD_TRY(plan_units(h, p, forward_rows_per_unit(k, c.lanes, c.tile), false, &us));
  const int kt = tile_cols(k, c.tile);
  SYMAMD_TRY(unit_flags(h, us, (k + kt - 1) / kt));
  if (through_workspace(h, k, layout)) SYMAMD_TRY(prepare_workspace(h, p, k));
  return SYMAMD_OK;
}

int run_solve(symamd_handle h, symamd_plan p, void *d_B, int k, int64_t ld,
              int64_t rs, int64_t cs, int layout, int exec, int lanes, int tile) {
  if (exec == SYMAMD_EXEC_LEVELSET_GRAPH) exec = SYMAMD_EXEC_LEVELSET;
  if (exec == SYMAMD_EXEC_LEVELSET_SPLIT) lanes = 1;
  if (exec == SYMAMD_EXEC_REGIONS) {
    // column-major blocks go through the row-major workspace, the rest in
    // place (schedule and workspace built beforehand by prepare_solve, so
    // nothing is built here while capturing)
    if (region_usable(h, p, k))
      return through_workspace(h, k, layout) ? launch_regions_ws(h, p, d_B, d_B, k, ld, false)
                                             : launch_regions(h, p, d_B, k, rs, cs);
    exec = SYMAMD_EXEC_SYNCFREE;
  }
  if (through_workspace(h, k, layout))
    return launch_forward_ws(h, p, d_B, k, ld, exec, lanes, tile);
  return launch_forward(h, p, d_B, k, rs, cs, exec, lanes, tile);
}

int capture_solve(symamd_handle h, symamd_plan p, void *d_B, int k, int64_t ld,
                  int64_t rs, int64_t cs, int layout, const Choice &c, hipGraph_t *g,
                  hipGraphExec_t *ge) {
  SYMAMD_TRY(prepare_solve(h, p, k, layout, c));
  hipStream_t cap = nullptr;
  SYMAMD_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
  hipStream_t saved = h->stream;
  h->stream = cap;
  int rc = SYMAMD_OK;
  if (hipStreamBeginCapture(cap, hipStreamCaptureModeRelaxed) != hipSuccess) {
    rc = fail(SYMAMD_ERR_HIP, "hipStreamBeginCapture failed");
  } else {
    rc = run_solve(h, p, d_B, k, ld, rs, cs, layout, c.exec, c.lanes, c.tile);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(cap, &graph);
    if (rc == SYMAMD_OK && e != hipSuccess)
      rc = fail(SYMAMD_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    if (rc == SYMAMD_OK) {
      e = hipGraphInstantiate(ge, graph, nullptr, nullptr, 0);
      if (e != hipSuccess) rc = fail(SYMAMD_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
      *g = graph;
    } else if (graph) {
      (void)hipGraphDestroy(graph);
    }
  }
  h->stream = saved;
  (void)hipStreamDestroy(cap);
  return rc;
}

int solve_impl(symamd_handle h, symamd_plan p, void *d_B, int k, int64_t ld,
               int layout, int exec, bool graph_ok) {
  int64_t rs = 0, cs = 0;
  SYMAMD_TRY(strides(p->n, k, ld, layout, &rs, &cs));
  if (k == 0 || p->n == 0) return SYMAMD_OK;
  Choice c = choose(h, p, k, layout, exec, true);
  if ((c.exec == SYMAMD_EXEC_LEVELSET_GRAPH || c.exec == SYMAMD_EXEC_LEVELSET_SPLIT) &&
      graph_ok) {
    std::string key = graph_key(d_B, k, rs, cs, c) + (through_workspace(h, k, layout) ? "/ws" : "");
    if (c.exec == SYMAMD_EXEC_LEVELSET_SPLIT)
      key += "/split" + std::to_string(split_threshold(h, tile_cols(k, c.tile)));
    auto it = p->gcache.find(key);
    if (it == p->gcache.end()) {
      hipGraph_t g = nullptr;
      hipGraphExec_t ge = nullptr;
      SYMAMD_TRY(capture_solve(h, p, d_B, k, ld, rs, cs, layout, c, &g, &ge));
      it = p->gcache.emplace(key, std::make_pair(g, ge)).first;
    }
    SYMAMD_HIP(hipGraphLaunch(it->second.second, h->stream));
    return SYMAMD_OK;
  }
  SYMAMD_TRY(prepare_solve(h, p, k, layout, c));
  return run_solve(h, p, d_B, k, ld, rs, cs, layout, c.exec, c.lanes, c.tile);
}

}  // namespace
}  // namespace symamd

using namespace symamd;

extern "C" {

const char *symamd_last_error(void) { return g_last_error.c_str(); }

const char *symamd_version(void) { return "sympiler_amd 0.1 (gfx950)"; }

int symamd_create(symamd_handle *h, void *stream) {
  if (!h) return fail(SYMAMD_ERR_INVALID_ARG, "null handle pointer");
  auto *x = new symamd_handle_s();
  x->stream = (hipStream_t)stream;
  if (hipGetDevice(&x->device) != hipSuccess) {
    delete x;
    return fail(SYMAMD_ERR_HIP, "no HIP device");
  }
  if (hipDeviceGetAttribute(&x->num_cus, hipDeviceAttributeMultiprocessorCount,
                            x->device) != hipSuccess ||
      x->num_cus < 1)
    x->num_cus = 256;
  if (hipDeviceGetAttribute(&x->lds_per_block, hipDeviceAttributeMaxSharedMemoryPerBlock,
                            x->device) != hipSuccess ||
      x->lds_per_block < 65536)
    x->lds_per_block = 65536;
  *h = x;
  return SYMAMD_OK;
}

int symamd_destroy(symamd_handle h) {
  delete h;
  return SYMAMD_OK;
}

int symamd_set_stream(symamd_handle h, void *stream) {
  SYMAMD_TRY(check_handle(h));
  h->stream = (hipStream_t)stream;
  return SYMAMD_OK;
}

int symamd_set_option(symamd_handle h, int option, int64_t value) {
  SYMAMD_TRY(check_handle(h));
  switch (option) {
    case SYMAMD_OPT_TICKETS:
      h->force_tickets = value != 0;
      return SYMAMD_OK;
    case SYMAMD_OPT_TRACE_PTR:
      h->trace = (uint64_t *)(intptr_t)value;
      return SYMAMD_OK;
    case SYMAMD_OPT_MARKS:
      h->unmarked = value == 0;
      return SYMAMD_OK;
    case SYMAMD_OPT_COLMAJOR_DIRECT:
      h->colmajor_direct = value != 0;
      return SYMAMD_OK;
    case SYMAMD_OPT_REGIONS:
      if (value < 0 || value > 2) return fail(SYMAMD_ERR_INVALID_ARG, "regions: 0, 1 or 2");
      h->regions = (int)value;
      return SYMAMD_OK;
    case SYMAMD_OPT_OWNERS:
      h->owners = value != 0;
      return SYMAMD_OK;
    case SYMAMD_OPT_REGION_MW:
      h->region_mw = value != 0;
      return SYMAMD_OK;
    case SYMAMD_OPT_SPLIT_LONG:
      if (value < 0 || value > (1 << 30))
        return fail(SYMAMD_ERR_INVALID_ARG, "split threshold out of range");
      h->split_long = (int)value;
      return SYMAMD_OK;
    case SYMAMD_OPT_REGION_BAND:
      if (value < 0 || value > 255)
        return fail(SYMAMD_ERR_INVALID_ARG, "region band: 0..255 slices");
      h->region_band = (int)value;
      return SYMAMD_OK;
    case SYMAMD_OPT_REGION_WAVES:
      if (value != 0 && value != 1 && value != 2 && value != 4)
        return fail(SYMAMD_ERR_INVALID_ARG, "region waves: 0, 1, 2 or 4");
      h->region_waves = (int)value;
      return SYMAMD_OK;
    case SYMAMD_OPT_REGION_CAP:
      if (value != 0 && (value < 2 || value > 4096))
        return fail(SYMAMD_ERR_INVALID_ARG, "region cap: 0 or 2..4096");
      h->region_cap = (int)value;
      return SYMAMD_OK;
    case SYMAMD_OPT_REGION_ROWS:
      if (value < 2 || value > (1 << 20))
        return fail(SYMAMD_ERR_INVALID_ARG, "region rows out of range");
      h->region_rows = (int)value;
      return SYMAMD_OK;
    case SYMAMD_OPT_MAX_WAVES:
      if (value < 0 || value > (1 << 24))
        return fail(SYMAMD_ERR_INVALID_ARG, "max waves out of range");
      h->max_waves = (int)value;
      return SYMAMD_OK;
    case SYMAMD_OPT_TRACE_CAP:
      if (value < 0) return fail(SYMAMD_ERR_INVALID_ARG, "trace capacity must be >= 0");
      h->trace_cap = value;
      return SYMAMD_OK;
    default:
      return fail(SYMAMD_ERR_INVALID_ARG, "unknown option %d", option);
  }
}

int symamd_set_timeout_ms(symamd_handle h, double ms) {
  SYMAMD_TRY(check_handle(h));
  if (!(ms > 0)) return fail(SYMAMD_ERR_INVALID_ARG, "timeout must be positive");
  h->timeout_ms = ms;
  return SYMAMD_OK;
}

int symamd_levelset(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
                    int *d_level_ptr, int *d_level_set, int *levels) {
  SYMAMD_TRY(check_handle(h));
  if (n < 0 || !levels || (n > 0 && (!d_Lp || !d_Li || !d_level_ptr || !d_level_set)))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_levelset: bad arguments");
  int *d_state = nullptr;
  SYMAMD_HIP(hipMalloc(&d_state, sizeof(int) * ST_WORDS));
  int rc = check_structure(h, n, d_Lp, d_Li, d_state);
  if (rc == SYMAMD_OK)
    rc = compute_levels(h, n, d_Lp, d_Li, d_state, d_level_ptr, d_level_set,
                        levels, nullptr);
  (void)hipFree(d_state);
  return rc;
}

int symamd_levelset_noalloc(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
                            int *d_level_ptr, int *d_level_set, int *levels) {
  SYMAMD_TRY(check_handle(h));
  if (n < 0 || !levels || (n > 0 && (!d_Lp || !d_Li || !d_level_ptr || !d_level_set)))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_levelset_noalloc: bad arguments");
  return levelset_noalloc(h, n, d_Lp, d_Li, d_level_ptr, d_level_set, levels);
}

int symamd_spmv_csc(symamd_handle h, int n, int m, const int *d_Ap, const int *d_Ai,
                    const double *d_Ax, const double *d_x, double *d_y, int c0, int c1) {
  SYMAMD_TRY(check_handle(h));
  if (n < 0 || m < 0 || c0 < 0 || c1 < c0 || c1 > n ||
      (c1 > c0 && (!d_Ap || !d_Ai || !d_Ax || !d_x)) || (m > 0 && !d_y))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_spmv_csc: bad arguments");
  if (c1 == c0) return SYMAMD_OK;
  return spmv_pull(h, n, m, d_Ap, d_Ai, d_Ax, d_x, d_y, nullptr, 0, c0, c1 - c0, 0, nullptr);
}

int symamd_spmv_csc_parts(symamd_handle h, int n, int m, const int *d_Ap, const int *d_Ai,
                          const double *d_Ax, const double *d_x, double *d_y, int nparts,
                          const int *d_sets, const int *d_ids) {
  SYMAMD_TRY(check_handle(h));
  if (n < 0 || m < 0 || nparts < 0 || (nparts > 0 && !d_sets) || (m > 0 && !d_y))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_spmv_csc_parts: bad arguments");
  if (nparts == 0) return SYMAMD_OK;
  int ends[2] = {0, 0};
  SYMAMD_HIP(hipMemcpyAsync(&ends[0], d_sets, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  SYMAMD_HIP(hipMemcpyAsync(&ends[1], d_sets + nparts, sizeof(int), hipMemcpyDeviceToHost,
                            h->stream));
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  const int k0 = ends[0], V = ends[1] - ends[0];
  if (k0 < 0 || V < 0 || (V > 0 && (!d_ids || !d_Ap || !d_Ai || !d_Ax || !d_x)))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_spmv_csc_parts: bad part bounds or arrays");
  return spmv_pull(h, n, m, d_Ap, d_Ai, d_Ax, d_x, d_y, d_ids, k0, 0, V, nparts, d_sets);
}

int symamd_axpby(symamd_handle h, int n, const double *d_a, const double *d_b, double a,
                 double b, const double *d_y, double *d_z) {
  SYMAMD_TRY(check_handle(h));
  if (n < 0 || (n > 0 && (!d_y || !d_z)))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_axpby: bad arguments");
  return launch_axpby(h, n, d_a, d_b, a, b, d_y, d_z);
}

int symamd_coarsen(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
                   const double *d_cost, int inner_parts, int min_level_dist,
                   int div_rate, int *d_level_ptr, int *d_par_ptr, int *d_partition,
                   int *level_no, int *part_no) {
  SYMAMD_TRY(check_handle(h));
  if (n < 0 || !level_no || !part_no || !d_Lp || !d_level_ptr || !d_par_ptr ||
      (n > 0 && (!d_Li || !d_partition)))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_coarsen: bad arguments");
  return lbc_schedule(h, n, d_Lp, d_Li, d_cost, inner_parts, min_level_dist, div_rate,
                      d_level_ptr, d_par_ptr, d_partition, level_no, part_no);
}

int symamd_analyze(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
                   const void *d_Lx, symamd_dtype dtype, symamd_plan *plan) {
  SYMAMD_TRY(check_handle(h));
  if (n < 0 || !plan || !d_Lp || (n > 0 && (!d_Li || !d_Lx)))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_analyze: bad arguments");
  return analyze_common(h, n, d_Lp, d_Li, d_Lx, dtype, plan);
}

int symamd_analyze_batched(symamd_handle h, int nsys, const int *h_row_off,
                           const int *d_Lp, const int *d_Li, const void *d_Lx,
                           symamd_dtype dtype, symamd_plan *plan) {
  SYMAMD_TRY(check_handle(h));
  if (nsys < 1 || !h_row_off || !plan || !d_Lp)
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_analyze_batched: bad arguments");
  for (int s = 0; s < nsys; ++s)
    if (h_row_off[s + 1] < h_row_off[s])
      return fail(SYMAMD_ERR_INVALID_ARG, "row offsets must be non-decreasing");
  const int n = h_row_off[nsys];
  int nnz = 0;
  SYMAMD_HIP(hipMemcpy(&nnz, d_Lp + n, sizeof(int), hipMemcpyDeviceToHost));
  int *d_off = nullptr, *d_Gi = nullptr;
  SYMAMD_HIP(hipMalloc(&d_off, sizeof(int) * (nsys + 1)));
  SYMAMD_HIP(hipMalloc(&d_Gi, sizeof(int) * (nnz > 0 ? nnz : 1)));
  SYMAMD_HIP(hipMemcpy(d_off, h_row_off, sizeof(int) * (nsys + 1), hipMemcpyHostToDevice));
  if (n > 0) {
    k_globalise<<<(n + 255) / 256, 256, 0, h->stream>>>(n, nsys, d_off, d_Lp, d_Li, d_Gi);
    SYMAMD_HIP(hipGetLastError());
  }
  int rc = analyze_common(h, n, d_Lp, d_Gi, d_Lx, dtype, plan);
  (void)hipStreamSynchronize(h->stream);
  (void)hipFree(d_off);
  (void)hipFree(d_Gi);
  if (rc == SYMAMD_OK) {
    (*plan)->nsys = nsys;
    (*plan)->h_row_off.assign(h_row_off, h_row_off + nsys + 1);
  }
  return rc;
}

int symamd_plan_update_values(symamd_handle h, symamd_plan plan, const void *d_Lx) {
  SYMAMD_TRY(check_handle(h));
  SYMAMD_TRY(check_plan(plan));
  if (plan->nnz > 0 && !d_Lx) return fail(SYMAMD_ERR_INVALID_ARG, "null values");
  return gather_values(h, plan, d_Lx);
}

int symamd_plan_info(symamd_plan plan, int *n, int64_t *nnz, int *levels,
                     int *max_level_rows) {
  SYMAMD_TRY(check_plan(plan));
  if (n) *n = plan->n;
  if (nnz) *nnz = plan->nnz;
  if (levels) *levels = plan->levels;
  if (max_level_rows) *max_level_rows = plan->max_level_rows;
  return SYMAMD_OK;
}

int symamd_plan_levels(symamd_plan plan, int *h_level_ptr, int *h_level_set) {
  SYMAMD_TRY(check_plan(plan));
  if (h_level_ptr)
    std::copy(plan->h_lvl_ptr.begin(), plan->h_lvl_ptr.end(), h_level_ptr);
  if (h_level_set && plan->n > 0)
    SYMAMD_HIP(hipMemcpy(h_level_set, plan->d_lvl_set ? plan->d_lvl_set : plan->d_perm,
                         sizeof(int) * plan->n, hipMemcpyDeviceToHost));
  return SYMAMD_OK;
}

int symamd_choice(symamd_handle h, symamd_plan plan, int k, symamd_layout layout,
                  int *exec, int *lanes) {
  SYMAMD_TRY(check_handle(h));
  SYMAMD_TRY(check_plan(plan));
  if (k < 1) return fail(SYMAMD_ERR_INVALID_ARG, "k must be >= 1");
  Choice c = choose(h, plan, k, layout, SYMAMD_EXEC_AUTO, true);
  if (exec) *exec = c.exec;
  if (lanes) *lanes = c.lanes;
  return SYMAMD_OK;
}

int symamd_plan_regions(symamd_plan plan, int *count, double *inside) {
  SYMAMD_TRY(check_plan(plan));
  if (count) *count = plan->d_preg ? plan->nregions : 0;
  if (inside) *inside = plan->region_intra;
  return SYMAMD_OK;
}

int symamd_region_schedule(symamd_handle h, symamd_plan plan, int k, int transposed,
                           int *cap, int *tasks) {
  SYMAMD_TRY(check_handle(h));
  SYMAMD_TRY(check_plan(plan));
  if (k < 1) return fail(SYMAMD_ERR_INVALID_ARG, "k must be >= 1");
  int c = 0, r = 0;
  region_schedule_info(h, plan, k, transposed != 0, &c, &r);
  if (cap) *cap = c;
  if (tasks) *tasks = r;
  return SYMAMD_OK;
}

int symamd_plan_destroy(symamd_plan plan) {
  if (!plan) return SYMAMD_OK;
  (void)hipDeviceSynchronize();
  free_plan_arrays(plan);
  delete plan;
  return SYMAMD_OK;
}

int symamd_plan_set_lanes(symamd_plan plan, int lanes_per_row) {
  SYMAMD_TRY(check_plan(plan));
  if (lanes_per_row < 0 || lanes_per_row > 64 ||
      (lanes_per_row & (lanes_per_row - 1)) != 0)
    return fail(SYMAMD_ERR_INVALID_ARG, "lanes per row must be 0 or a power of two <= 64");
  plan->lanes_override = lanes_per_row;
  return SYMAMD_OK;
}

int symamd_solve(symamd_handle h, symamd_plan plan, void *d_B, int k, int64_t ld,
                 symamd_layout layout, symamd_executor exec) {
  SYMAMD_TRY(check_handle(h));
  SYMAMD_TRY(check_plan(plan));
  if (k < 0 || (k > 0 && plan->n > 0 && !d_B))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_solve: bad arguments");
  SYMAMD_TRY(check_exec(exec));
  return solve_impl(h, plan, d_B, k, ld, layout, exec, true);
}

int symamd_plan_check(symamd_handle h, symamd_plan plan) {
  SYMAMD_TRY(check_handle(h));
  SYMAMD_TRY(check_plan(plan));
  int st[ST_WORDS];
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  SYMAMD_HIP(hipMemcpy(st, plan->d_state, sizeof(st), hipMemcpyDeviceToHost));
  if (st[ST_ERROR]) return fail(SYMAMD_ERR_TIMEOUT, "a sync-free wait exceeded the watchdog");
  return SYMAMD_OK;
}

int symamd_graph_create(symamd_handle h, symamd_plan plan, void *d_B, int k,
                        int64_t ld, symamd_layout layout, symamd_executor exec,
                        symamd_graph *g) {
  SYMAMD_TRY(check_handle(h));
  SYMAMD_TRY(check_plan(plan));
  if (!g || k < 1 || !d_B) return fail(SYMAMD_ERR_INVALID_ARG, "symamd_graph_create: bad arguments");
  SYMAMD_TRY(check_exec(exec));
  int64_t rs = 0, cs = 0;
  SYMAMD_TRY(strides(plan->n, k, ld, layout, &rs, &cs));
  Choice c = choose(h, plan, k, layout, exec, true);
  auto *x = new symamd_graph_s();
  x->plan = plan;
  int rc = capture_solve(h, plan, d_B, k, ld, rs, cs, layout, c, &x->graph, &x->exec);
  if (rc != SYMAMD_OK) {
    delete x;
    return rc;
  }
  *g = x;
  return SYMAMD_OK;
}

int symamd_graph_launch(symamd_handle h, symamd_graph g) {
  SYMAMD_TRY(check_handle(h));
  if (!g) return fail(SYMAMD_ERR_INVALID_ARG, "null graph");
  SYMAMD_HIP(hipGraphLaunch(g->exec, h->stream));
  return SYMAMD_OK;
}

int symamd_graph_destroy(symamd_graph g) {
  if (!g) return SYMAMD_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return SYMAMD_OK;
}

int symamd_adjoint(symamd_handle h, symamd_plan plan, const void *d_x,
                   const void *d_gx, void *d_gb, void *d_gLx,
                   symamd_executor exec) {
  SYMAMD_TRY(check_handle(h));
  SYMAMD_TRY(check_plan(plan));
  if (plan->n > 0 && (!d_x || !d_gx || !d_gb))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_adjoint: bad arguments");
  SYMAMD_TRY(check_exec(exec));
  // the region executor solves L^T on the transposed region schedule; a
  // forced choice without a usable schedule runs sync-free, like forward
  if (exec == SYMAMD_EXEC_REGIONS ? region_adjoint_usable(h, plan, 1)
                                  : exec == SYMAMD_EXEC_AUTO && region_adjoint_preferred(h, plan, 1)) {
    // grad-L is written by the solve itself where its kernel supports it
    bool fused = false;
    SYMAMD_TRY(launch_regions_plan(h, plan, d_gx, d_gb, 1, 1, 0, true, d_x, d_gLx, &fused));
    return d_gLx && !fused ? launch_grad_l(h, plan, d_x, d_gb, d_gLx, 1, 1, 0) : SYMAMD_OK;
  }
  Choice c = choose(h, plan, 1, SYMAMD_COL_MAJOR, exec, false);
  int e = c.exec == SYMAMD_EXEC_LEVELSET_GRAPH || c.exec == SYMAMD_EXEC_LEVELSET_SPLIT
              ? SYMAMD_EXEC_LEVELSET : c.exec;
  if (e == SYMAMD_EXEC_REGIONS) e = SYMAMD_EXEC_SYNCFREE;
  return launch_adjoint(h, plan, d_x, d_gx, d_gb, d_gLx, e, c.lanes);
}

int symamd_adjoint_block(symamd_handle h, symamd_plan plan, const void *d_X,
                         const void *d_GX, void *d_GB, void *d_gLx, int k, int64_t ld,
                         symamd_layout layout, symamd_executor exec) {
  SYMAMD_TRY(check_handle(h));
  SYMAMD_TRY(check_plan(plan));
  if (k < 0 || (k > 0 && plan->n > 0 && (!d_X || !d_GX || !d_GB)))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_adjoint_block: bad arguments");
  SYMAMD_TRY(check_exec(exec));
  int64_t rs = 0, cs = 0;
  SYMAMD_TRY(strides(plan->n, k, ld, layout, &rs, &cs));
  if (k == 0 || plan->n == 0) return SYMAMD_OK;
  const bool regions = exec == SYMAMD_EXEC_REGIONS ? region_adjoint_usable(h, plan, k)
                       : exec == SYMAMD_EXEC_AUTO && region_adjoint_preferred(h, plan, k);
  if (regions) {
    if (k > 1 && layout == SYMAMD_COL_MAJOR && !h->colmajor_direct) {
      SYMAMD_TRY(prepare_workspace(h, plan, k));
      SYMAMD_TRY(launch_regions_ws(h, plan, d_GX, d_GB, k, ld, true));
    } else {
      SYMAMD_TRY(launch_regions_plan(h, plan, d_GX, d_GB, k, rs, cs, true));
    }
  } else {
    // one column at a time on the transposed unit executors, through
    // contiguous vectors (column c of the block at offset c*cs, stride rs)
    Choice c = choose(h, plan, 1, SYMAMD_COL_MAJOR, exec == SYMAMD_EXEC_REGIONS
                                                        ? SYMAMD_EXEC_SYNCFREE : exec, false);
    int e = c.exec == SYMAMD_EXEC_LEVELSET_GRAPH || c.exec == SYMAMD_EXEC_LEVELSET_SPLIT
                ? SYMAMD_EXEC_LEVELSET : c.exec;
    if (e == SYMAMD_EXEC_REGIONS) e = SYMAMD_EXEC_SYNCFREE;
    const size_t vs = plan->dtype == SYMAMD_F64 ? 8 : 4;
    void *tin = nullptr, *tout = nullptr;
    SYMAMD_HIP(hipMallocAsync(&tin, vs * plan->n, h->stream));
    SYMAMD_HIP(hipMallocAsync(&tout, vs * plan->n, h->stream));
    int rc = SYMAMD_OK;
    for (int col = 0; col < k && rc == SYMAMD_OK; ++col) {
      const int64_t off = (int64_t)col * cs;
      rc = copy_strided(h, plan->n, vs, (const char *)d_GX + vs * off, rs, tin, 1);
      if (rc == SYMAMD_OK) rc = launch_adjoint(h, plan, nullptr, tin, tout, nullptr, e, c.lanes);
      if (rc == SYMAMD_OK) rc = copy_strided(h, plan->n, vs, tout, 1, (char *)d_GB + vs * off, rs);
    }
    (void)hipFreeAsync(tin, h->stream);
    (void)hipFreeAsync(tout, h->stream);
    SYMAMD_TRY(rc);
  }
  return d_gLx ? launch_grad_l(h, plan, d_X, d_GB, d_gLx, k, rs, cs) : SYMAMD_OK;
}

int symamd_spmv_csr(symamd_handle h, int n, const int *d_Ap, const int *d_Ai,
                    const double *d_Ax, const double *d_x, double *d_y) {
  SYMAMD_TRY(check_handle(h));
  if (n < 0 || (n > 0 && (!d_Ap || !d_Ai || !d_Ax || !d_x || !d_y)))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_spmv_csr: bad arguments");
  return launch_spmv(h, n, d_Ap, d_Ai, d_Ax, d_x, nullptr, d_y, 0, 0, 0);
}

int symamd_spmv_csr_axpby(symamd_handle h, int n, const int *d_Ap,
                          const int *d_Ai, const double *d_Ax,
                          const double *d_x, const double *d_y, double *d_z,
                          double a, double b) {
  SYMAMD_TRY(check_handle(h));
  if (n < 0 || (n > 0 && (!d_Ap || !d_Ai || !d_Ax || !d_x || !d_y || !d_z)))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_spmv_csr_axpby: bad arguments");
  return launch_spmv(h, n, d_Ap, d_Ai, d_Ax, d_x, d_y, d_z, a, b, 1);
}

int symamd_residual(symamd_handle h, symamd_plan plan, const void *d_X,
                    const void *d_B, int k, int64_t ld, symamd_layout layout,
                    double *h_rel) {
  SYMAMD_TRY(check_handle(h));
  SYMAMD_TRY(check_plan(plan));
  if (k < 0 || !h_rel || (k > 0 && plan->n > 0 && (!d_X || !d_B)))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_residual: bad arguments");
  int64_t rs = 0, cs = 0;
  SYMAMD_TRY(strides(plan->n, k, ld, layout, &rs, &cs));
  return plan_residual(h, plan, d_X, d_B, k, rs, cs, h_rel);
}

int symamd_tune(symamd_handle h, symamd_plan plan, int k, symamd_layout layout,
                char *summary, int summary_len) {
  SYMAMD_TRY(check_handle(h));
  SYMAMD_TRY(check_plan(plan));
  if (k < 1) return fail(SYMAMD_ERR_INVALID_ARG, "k must be >= 1");
  if (plan->n == 0) return SYMAMD_OK;
  const int64_t ld = layout == SYMAMD_COL_MAJOR ? plan->n : k;
  const size_t vs = plan->dtype == SYMAMD_F64 ? 8 : 4;
  const int64_t m = (int64_t)plan->n * k;
  void *d_B = nullptr;
  SYMAMD_HIP(hipMalloc(&d_B, vs * m));
  k_fill<<<(unsigned)((m + 255) / 256), 256, 0, h->stream>>>(m, 1.0, d_B, vs == 8);
  hipEvent_t e0, e1;
  SYMAMD_HIP(hipEventCreate(&e0));
  SYMAMD_HIP(hipEventCreate(&e1));
  const int gmax_rows = std::max(1, pow2_ceil((int)std::ceil(plan->avg_row_nnz)) * 2);
  std::string log;
  Choice best;
  float best_ms = 1e30f;
  const int saved_override = plan->lanes_override;
  plan->lanes_override = 0;
  const int execs[] = {SYMAMD_EXEC_SYNCFREE, SYMAMD_EXEC_LEVELSET_GRAPH, SYMAMD_EXEC_REGIONS,
                       SYMAMD_EXEC_LEVELSET_SPLIT};
  int rc = SYMAMD_OK;
  // candidates: the executor; for the unit executors the column tile (all
  // of k up to 64 columns, or narrower tiles of 8 to 32 columns solved as
  // separate work items) and the lanes per row that share a row's entries
  for (int ex : execs) {
    // the region executor has no lanes-per-row or tile choice and needs a
    // schedule; the split level executor has one lane per row (its long rows
    // are split by the schedule itself) and no tile choice
    if (ex == SYMAMD_EXEC_REGIONS && !region_usable(h, plan, k)) continue;
    std::vector<int> tiles = {0};
    if (ex == SYMAMD_EXEC_SYNCFREE)
      for (int t = 8; t < tile_cols(k, 0); t *= 2) tiles.push_back(t);
    // region caps: for single right-hand sides 512 and 1024 rows (whichever
    // schedule fits; the cap of the handle, when set, alone)
    std::vector<int> caps = {0};
    if (ex == SYMAMD_EXEC_REGIONS && k == 1 && h->region_cap == 0) caps = {512, 1024};
    for (int tile : tiles)
    for (int cap : caps) {
      const int kt = tile_cols(k, tile);
      const int gmax = ex == SYMAMD_EXEC_REGIONS || ex == SYMAMD_EXEC_LEVELSET_SPLIT
                           ? 1 : std::min(64, gmax_rows);
      for (int g = 1; g * kt <= 64 && g <= gmax; g *= 2) {
        Choice c;
        c.exec = ex;
        c.lanes = g;
        c.tile = tile;
        c.cap = cap;
        plan->tuned[tune_key(k, layout)] = c;
        rc = solve_impl(h, plan, d_B, k, ld, layout, SYMAMD_EXEC_AUTO, true);  // warm / capture
        if (rc != SYMAMD_OK) break;
        float ms = 1e30f;
        for (int rep = 0; rep < 3 && rc == SYMAMD_OK; ++rep) {
          (void)hipEventRecord(e0, h->stream);
          rc = solve_impl(h, plan, d_B, k, ld, layout, SYMAMD_EXEC_AUTO, true);
          (void)hipEventRecord(e1, h->stream);
          (void)hipEventSynchronize(e1);
          float t = 0;
          (void)hipEventElapsedTime(&t, e0, e1);
          ms = std::min(ms, t);
        }
        if (rc != SYMAMD_OK) break;
        char line[128];
        snprintf(line, sizeof line, "exec=%d lanes=%d tile=%d cap=%d ms=%.4f\n", ex, g, kt,
                 cap, ms);
        log += line;
        if (ms < best_ms * 0.97f) {  // prefer earlier (fewer lanes, sync-free) on ties
          best_ms = ms;
          best = c;
        }
      }
      if (rc != SYMAMD_OK) break;
    }
    if (rc != SYMAMD_OK) break;
  }
  plan->lanes_override = saved_override;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(h->stream);
  // drop the graphs captured on the scratch block
  char prefix[32];
  snprintf(prefix, sizeof prefix, "%p/", d_B);
  for (auto it = plan->gcache.begin(); it != plan->gcache.end();) {
    if (it->first.compare(0, strlen(prefix), prefix) == 0) {
      (void)hipGraphExecDestroy(it->second.second);
      (void)hipGraphDestroy(it->second.first);
      it = plan->gcache.erase(it);
    } else {
      ++it;
    }
  }
  (void)hipFree(d_B);
  if (rc != SYMAMD_OK) {
    plan->tuned.erase(tune_key(k, layout));
    return rc;
  }
  plan->tuned[tune_key(k, layout)] = best;
  g_tune_cache[plan->fingerprint][tune_key(k, layout)] = best;
  char line[128];
  snprintf(line, sizeof line, "best exec=%d lanes=%d tile=%d cap=%d ms=%.4f\n", best.exec,
           best.lanes, tile_cols(k, best.tile), best.cap, best_ms);
  log += line;
  if (summary && summary_len > 0) {
    strncpy(summary, log.c_str(), summary_len - 1);
    summary[summary_len - 1] = 0;
  }
  SYMAMD_TRY(symamd_plan_check(h, plan));
  return SYMAMD_OK;
}

int symamd_solve_lbc(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
                     const double *d_Lx, double *d_x, int level_no,
                     const int *h_level_ptr, const int *d_par_ptr,
                     const int *d_partition) {
  SYMAMD_TRY(check_handle(h));
  if (n < 0 || level_no < 0 || (level_no > 0 && (!h_level_ptr || !d_par_ptr || !d_partition)) ||
      (n > 0 && (!d_Lp || !d_Li || !d_Lx || !d_x)))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_solve_lbc: bad arguments");
  return launch_lbc(h, n, d_Lp, d_Li, d_Lx, d_x, nullptr, level_no, h_level_ptr,
                    d_par_ptr, d_partition);
}

int symamd_solve_lbc_redundant(symamd_handle h, int n, const int *d_Lp,
                               const int *d_Li, const double *d_Lx, double *d_x,
                               const double *d_b, int level_no,
                               const int *h_level_ptr, const int *d_par_ptr,
                               const int *d_partition) {
  SYMAMD_TRY(check_handle(h));
  if (n < 0 || level_no < 0 || (level_no > 0 && (!h_level_ptr || !d_par_ptr || !d_partition)) ||
      (n > 0 && (!d_Lp || !d_Li || !d_Lx || !d_x || !d_b)))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_solve_lbc_redundant: bad arguments");
  return launch_lbc(h, n, d_Lp, d_Li, d_Lx, d_x, d_b, level_no, h_level_ptr,
                    d_par_ptr, d_partition);
}

int symamd_ilu_lower(symamd_handle h, int n, const int *d_Ap, const int *d_Ai,
                     const double *d_Ax, const int *d_diags, int unit_diag,
                     int *d_Lp, int *d_Li, double *d_Lx, int *nnz) {
  SYMAMD_TRY(check_handle(h));
  if (n < 0 || !nnz || !d_Lp || (n > 0 && (!d_Ap || !d_Ai || !d_Ax || !d_diags ||
                                          !d_Li || !d_Lx)))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_ilu_lower: bad arguments");
  return ilu_lower(h, n, d_Ap, d_Ai, d_Ax, d_diags, unit_diag ? 1 : 0, d_Lp, d_Li, d_Lx,
                   nnz);
}

int symamd_ic0(symamd_handle h, symamd_plan plan, symamd_ic0_variant variant,
               const double *d_Ax, double *d_Lx, int *bad_row) {
  SYMAMD_TRY(check_handle(h));
  SYMAMD_TRY(check_plan(plan));
  if (!bad_row || (variant != SYMAMD_IC0_LEFT && variant != SYMAMD_IC0_CSR) ||
      (plan->nnz > 0 && (!d_Ax || !d_Lx)))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_ic0: bad arguments");
  if (plan->dtype != SYMAMD_F64)
    return fail(SYMAMD_ERR_NOT_SUPPORTED, "symamd_ic0: FP64 plans only");
  *bad_row = -1;
  if (plan->n == 0) return SYMAMD_OK;
  SYMAMD_TRY(ic0_factor(h, plan, (int)variant, d_Ax, d_Lx, bad_row));
  SYMAMD_TRY(symamd_plan_check(h, plan));
  if (*bad_row >= 0)
    return fail(SYMAMD_ERR_NOT_SPD, "symamd_ic0: pivot of row %d is not positive", *bad_row);
  return SYMAMD_OK;
}

int symamd_ilu0_analyze(symamd_handle h, int n, const int *d_Ap, const int *d_Ai,
                        const int *d_Adiag, symamd_ilu0_plan *plan) {
  SYMAMD_TRY(check_handle(h));
  if (n < 0 || !plan || !d_Ap || (n > 0 && (!d_Ai || !d_Adiag)))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_ilu0_analyze: bad arguments");
  *plan = nullptr;
  return ilu0_analyze(h, n, d_Ap, d_Ai, d_Adiag, plan);
}

int symamd_ilu0_factor(symamd_handle h, symamd_ilu0_plan plan, const double *d_Ax,
                       double *d_l, int *bad_row) {
  SYMAMD_TRY(check_handle(h));
  if (!plan || !bad_row || (plan->nnz > 0 && (!d_Ax || !d_l)))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_ilu0_factor: bad arguments");
  SYMAMD_TRY(ilu0_numeric(h, plan, d_Ax, d_l, bad_row));
  if (*bad_row >= 0)
    return fail(SYMAMD_ERR_ZERO_PIVOT, "ILU(0): U[%d,%d] is zero", *bad_row, *bad_row);
  return SYMAMD_OK;
}

int symamd_ilu0_destroy(symamd_ilu0_plan plan) {
  ilu0_destroy(plan);
  return SYMAMD_OK;
}

int symamd_ilu0(symamd_handle h, int n, const int *d_Ap, const int *d_Ai, const double *d_Ax,
                const int *d_Adiag, double *d_l, int *bad_row) {
  SYMAMD_TRY(check_handle(h));
  if (n < 0 || !bad_row || !d_Ap || (n > 0 && (!d_Ai || !d_Ax || !d_Adiag || !d_l)))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_ilu0: bad arguments");
  symamd_ilu0_plan plan = nullptr;
  SYMAMD_TRY(ilu0_analyze(h, n, d_Ap, d_Ai, d_Adiag, &plan));
  const int rc = symamd_ilu0_factor(h, plan, d_Ax, d_l, bad_row);
  ilu0_destroy(plan);
  return rc;
}

int symamd_sn_analyze(symamd_handle h, int n, int m, const int *h_sup, const int *h_nrows,
                      const int *h_Lp, const int *h_Li, symamd_snplan *plan) {
  SYMAMD_TRY(check_handle(h));
  if (!plan) return fail(SYMAMD_ERR_INVALID_ARG, "symamd_sn_analyze: null plan");
  *plan = nullptr;
  return sn_analyze(h, n, m, h_sup, h_nrows, h_Lp, h_Li, plan);
}

int symamd_sn_solve(symamd_handle h, symamd_snplan plan, const double *d_Lx, double *d_x) {
  SYMAMD_TRY(check_handle(h));
  if (!plan || (plan->m > 0 && (!d_Lx || !d_x)))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_sn_solve: bad arguments");
  return sn_solve(h, plan, d_Lx, d_x, 1, 0);
}

int symamd_sn_solve_block(symamd_handle h, symamd_snplan plan, const double *d_Lx,
                          double *d_X, int k, int64_t ld) {
  SYMAMD_TRY(check_handle(h));
  if (!plan || k < 0 || (k > 1 && ld < plan->n) || (plan->m > 0 && k > 0 && (!d_Lx || !d_X)))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_sn_solve_block: bad arguments");
  return sn_solve(h, plan, d_Lx, d_X, k, ld);
}

int symamd_sn_solve_transposed(symamd_handle h, symamd_snplan plan, const double *d_Lx,
                               double *d_X, int k, int64_t ld) {
  SYMAMD_TRY(check_handle(h));
  if (!plan || k < 0 || (k > 1 && ld < plan->n) || (plan->m > 0 && k > 0 && (!d_Lx || !d_X)))
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_sn_solve_transposed: bad arguments");
  return sn_solve(h, plan, d_Lx, d_X, k, ld, true);
}

int symamd_sn_check(symamd_handle h, symamd_snplan plan) {
  SYMAMD_TRY(check_handle(h));
  if (!plan) return fail(SYMAMD_ERR_INVALID_ARG, "symamd_sn_check: null plan");
  return sn_check(h, plan);
}

int symamd_sn_destroy(symamd_snplan plan) {
  sn_destroy(plan);
  return SYMAMD_OK;
}

}  // extern "C"
