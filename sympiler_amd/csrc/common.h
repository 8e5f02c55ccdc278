// common.h -- internal definitions shared by the HIP translation units of
// libsympiler_amd.so: status plumbing, device-coherent memory helpers, the
// plan structure and kernel-launch declarations.
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/sympiler_amd.h"

namespace symamd {

// ------------------------------------------------------------ error state
void set_error(const char *fmt, ...);
int fail(int code, const char *fmt, ...);

#define SYMAMD_HIP(call)                                                   \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess)                                                  \
      return ::symamd::fail(SYMAMD_ERR_HIP, "%s failed: %s (%s:%d)", #call, \
                            hipGetErrorString(e_), __FILE__, __LINE__);    \
  } while (0)

#define SYMAMD_TRY(expr)         \
  do {                           \
    int s_ = (expr);             \
    if (s_ != SYMAMD_OK) return s_; \
  } while (0)

// ------------------------------------------------- device memory helpers
//
// Cross-workgroup communication inside one kernel (ready flags and the
// values they guard) uses agent-scope relaxed atomics: on gfx950 these are
// loads/stores with device scope that bypass the non-coherent per-XCD
// caches, so a value written by a wave on one XCD is seen by a wave on
// another without cache maintenance.  The producer orders "values, then
// flag" with a wait for its outstanding stores (publish_fence).

template <typename T>
__device__ __forceinline__ T ld_dev(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_dev(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// All stores this wave issued have completed (and, being device-scope, are
// visible device-wide) before any later memory instruction issues.
__device__ __forceinline__ void publish_fence() {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// Non-temporal streaming load for read-once matrix data.
template <typename T>
__device__ __forceinline__ T ld_stream(const T *p) {
  return __builtin_nontemporal_load(p);
}

// 100 MHz constant clock, used by the sync-free watchdog.
__device__ __forceinline__ uint64_t wall_ticks() {
  return __builtin_amdgcn_s_memrealtime();
}

// Shared state of one plan's persistent kernels (device memory, int32):
//   [0] epoch   -- launches completed; flags equal to epoch+1 are "ready"
//   [1] ticket  -- next work unit to hand out
//   [2] done    -- waves that finished this launch
//   [3] error   -- sticky watchdog flag
//   [4] nlev    -- scratch for the inspector
enum { ST_EPOCH = 0, ST_TICKET = 1, ST_DONE = 2, ST_ERROR = 3, ST_SCRATCH = 4,
       ST_WORDS = 8 };

// Called by every wave at the end of a persistent launch: the last wave
// resets the ticket and done counters and advances the epoch, so the next
// launch (stream-ordered) starts clean without any host-side memset.  This
// keeps the kernels replayable from a captured hipGraph.
__device__ __forceinline__ void persistent_finish(int *state, int total_waves) {
  int lane = threadIdx.x & 63;
  if (lane == 0) {
    int old = atomicAdd(&state[ST_DONE], 1);
    if (old == total_waves - 1) {
      state[ST_TICKET] = 0;
      state[ST_DONE] = 0;
      __hip_atomic_store(&state[ST_EPOCH], state[ST_EPOCH] + 1,
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// ------------------------------------------------------------------ plan

// Entries of one work unit staged in a wave's LDS slab (a single row with
// more entries forms a unit of its own and is read from global memory).
constexpr int UNIT_EMAX = 512;

// Work units: runs of at most `rows` consecutive level-ordered rows inside
// one level (and at most UNIT_EMAX entries), with the units their rows
// depend on.  Units of the forward solve depend on units of earlier levels;
// units of the transposed solve (cut by the transposed copy) on units of
// later levels.
struct UnitSet {
  int rows = 0;             // rows per unit (cap)
  int count = 0;            // number of units
  std::vector<int> h_uoff;  // first unit of each level (levels+1)
  int *d_ub = nullptr;      // unit u covers positions [ub[u], ub[u+1])
  int *d_dptr = nullptr;    // dependency units of u:
  int *d_dlist = nullptr;   //   dlist[dptr[u] .. dptr[u+1])
  int *d_uflags = nullptr;  // ready tags, one per unit and column tile
  int flag_tiles = 1;       // column tiles the tags have room for
  int64_t deps = 0;
  // owner schedule (plans with regions): unit u belongs to wave
  // owner(region of u) of `nw` co-resident waves; wave w solves
  // wlist[wptr[w] .. wptr[w+1]) in that order and waits only on the
  // dependencies owned by other waves, xdlist[xdptr[u] .. xdptr[u+1])
  int nw = 0;
  int *d_wptr = nullptr;
  int *d_wlist = nullptr;
  int *d_xdptr = nullptr;
  int *d_xdlist = nullptr;
  int64_t xdeps = 0;
};

// IC(0) update pairs of one variant (factor0.hip): entry p of the caller's CSR
// subtracts Lx[pa[t]] * Lx[pb[t]] for t in [pp[p], pp[p+1]), in order.
struct Ic0Pairs {
  int64_t pairs = 0;
  int64_t *d_pp = nullptr;  // nnz + 1
  int *d_pa = nullptr;
  int *d_pb = nullptr;
  int *d_dpos = nullptr;    // position of L[c,c] for each entry (r, c)
};

// Region schedule: each region (a compact cluster of rows, at most `cap`
// rows) is solved by one wavefront in LDS, its internal levels in order;
// regions wait for the regions they read from (a DAG, ranked
// topologically).  All arrays are region-major in rank order.
struct RegionPlan {
  static constexpr int MAX_GROUPS = 8;  // column groups per region (tags)
  static constexpr int PACK = 4;        // entries per row kept in the packed arrays
  int cap = 0;        // region size cap the clustering used
  int R = 0;          // regions (0: no region schedule for this cap)
  int smax = 0;       // max rows in a region
  int hmax = 0;       // max halo rows (rows of other regions read)
  int emax = 0;       // max overflow entries (beyond PACK per row) in a region
  int slmax = 0;      // max internal levels (slices) of a region
  int width = 0;      // most regions at one depth (longest-path level) of the region graph
  double inside = 0;  // fraction of off-diagonal entries inside regions
  // One record per region (byte layout rec_layout(), also its LDS image),
  // described by rhead[8r ..] = {rows, halo rows, overflow entries, slices,
  // dependencies, first 16-byte unit of the record in `rec`, its units, 0}.
  void *d_rec = nullptr;
  int *d_rhead = nullptr;
  int recmax = 0;  // largest record, bytes
  // matrix values of the records: rec (as values) [vdst[t]] = value at
  // position vsrc[t] of the plan's level-ordered copy
  int *d_vsrc = nullptr;
  int *d_vdst = nullptr;
  int64_t nvals = 0;
  // task progress tags, MAX_GROUPS per region: (launch target << 8) | bands
  // published, so a value from an earlier launch never satisfies a wait
  int64_t *d_rflags = nullptr;
  int band = 0;             // slices per band the schedule was built with
  int gp = 0;               // records carry grad-L positions (rec_layout's gp)
  bool transposed = false;  // schedule of L^T x = g (values from the transposed copy)
};

// Sections of a region record, in bytes from its start, each 16-byte
// aligned: the packed values (PACK per row; padding slots hold +0.0), the
// diagonals, the overflow values (entries beyond PACK, row
// by row in order), the packed LDS slot codes (uint16), the
// per-row count of off-diagonal entries (uint8, capped at PACK + 1), the
// row ids, the halo row ids, the slice bounds (ns + 1 local row offsets),
// the ranks of the regions it waits for, the overflow bounds per row (nr + 1
// ints, only when ne > 0) and the overflow slot codes (uint16).  A slot
// code is the column's local row in the region, or smax + its index in the
// halo; padding slots name slot smax + hmax, which holds zeros, so a
// padding product is +0.0 * +0.0 and subtracting it leaves any value
// unchanged, bit for bit.
struct RecLayout {
  int pval, diag, lval, code, cnt, rows, halo, slices, bands, hband, wptr, deps, lep, lcode,
      gpos, bytes;
};
// gp: the record also carries, for the transposed (adjoint) solve, the
// caller's CSR position of every entry of L it holds (PACK + 1 per row --
// the packed entries, padding 0, then the diagonal -- then one per overflow
// entry), so the solve writes grad-L itself.
__host__ __device__ inline RecLayout rec_layout(int nr, int nh, int ne, int ns, int nd,
                                                int nb, int vs, int gp = 0) {
  constexpr int PK = RegionPlan::PACK;
  RecLayout L;
  int o = 0;
  auto put = [&o](int &field, int bytes) {
    field = o;
    o = (o + bytes + 15) & ~15;
  };
  put(L.pval, PK * nr * vs);
  put(L.diag, nr * vs);
  put(L.lval, ne * vs);
  put(L.code, PK * nr * 2);
  put(L.cnt, nr);
  put(L.rows, nr * 4);
  put(L.halo, nh * 4);
  put(L.slices, (ns + 1) * 4);
  put(L.bands, (nb + 1) * 4);
  put(L.hband, (nb + 1) * 4);
  put(L.wptr, (nb + 1) * 4);
  put(L.deps, nd * 4);
  put(L.lep, ne > 0 ? (nr + 1) * 4 : 0);
  put(L.lcode, ne * 2);
  put(L.gpos, gp ? ((PK + 1) * nr + ne) * 4 : 0);
  L.bytes = o;
  return L;
}
// Slot codes are uint16: a region with its halo and the zero slot may use
// at most this many LDS slots.
constexpr int REGION_SLOTS_MAX = 65535;
// Smallest region cap tried when a schedule's task does not fit a
// workgroup's LDS at the default cap.
constexpr int REGION_CAP_MIN = 32;
// A region publishes its progress band by band (a band is a few consecutive
// slices, at most REGION_BANDS_MAX per region).  A wait entry packs the rank
// of the region read (low REGION_RANK_BITS bits) with the number of its
// bands that must be published (the high bits).
constexpr int REGION_BANDS_MAX = 31;
constexpr int REGION_RANK_BITS = 26;

// Host copy of a region schedule, before packing into records.  Rows are
// region-major in rank order; entries of every row: the first PACK packed
// per row (pcnt = all its off-diagonal entries, pcode/psrc; padding slots
// hold code smax + hmax and the diagonal's position), the rest in order in the
// overflow list (eptr over the region-major rows, ecode, esrc); psrc, esrc
// and dsrc are positions in the level-ordered copy of the matrix.
struct HostRegions {
  int R = 0, smax = 0, hmax = 0, emax = 0, slmax = 0;
  int width = 0;  // most regions at one depth of the region graph
  double inside = 0;
  std::vector<int> rptr, rows, sptr, slices, hptr, halo, dptr, dlist, eptr, ecode, esrc,
      dsrc, pcnt, pcode, psrc;
  // bands: per region nb+1 entries at bptr[r]: the first slice of each band
  // (bands), its first halo row (hband, halo sorted by the band that first
  // reads it) and its first wait entry (wptr, into the region's dlist,
  // whose entries are rank | bands_needed << REGION_RANK_BITS)
  std::vector<int> bptr, bands, hband, wptr;
};
bool build_region_schedule(int n, int64_t nnz, int levels, const int *lvl_ptr,
                           const int *perm, const int *crp, const int *cci, int cap,
                           int band_slices, HostRegions &out);
// Builder inputs of the transposed solve from the plan's level sets (perm,
// lvl_ptr) and transposed copy (trp, tci: column perm[q] of L at position q,
// diagonal first): rows and positions numbered backwards, entries in
// ascending row order of L with the diagonal last, levels reversed; vmap
// maps each entry to its index in the transposed copy.
void reversed_transpose(int n, int levels, const std::vector<int> &lvl_ptr,
                        const std::vector<int> &perm, const std::vector<int> &trp,
                        const std::vector<int> &tci, std::vector<int> &rlvl,
                        std::vector<int> &rperm, std::vector<int> &rrp, std::vector<int> &rci,
                        std::vector<int> &vmap);
// Records of a schedule for values of vs bytes: the static bytes (values
// zero), headers (8 ints per region, as RegionPlan::d_rhead) and the value
// gather list.  False when the records would exceed the 32-bit offsets.
struct HostRecords {
  std::vector<uint8_t> bytes;
  std::vector<int> rhead, vsrc, vdst;
  int recmax = 0;
};
// gmap (may be null): maps the builder's entry positions (psrc, esrc,
// dsrc) to the caller's CSR positions; when given, the records carry them
// (rec_layout's gp section).
bool pack_region_records(const HostRegions &o, int vs, HostRecords &out,
                         const std::vector<int> *gmap = nullptr);

struct Choice {
  int exec = SYMAMD_EXEC_SYNCFREE;
  int lanes = 1;  // lanes cooperating on one row's nonzeros
  int tile = 0;   // columns per tile of the unit executors (0: from k)
  int cap = 0;    // region cap of the region executor (0: the default)
};

}  // namespace symamd

struct symamd_handle_s {
  hipStream_t stream = nullptr;
  int device = 0;
  int num_cus = 0;
  int lds_per_block = 65536;  // LDS bytes one workgroup may allocate
  double timeout_ms = 20000.0;
  int force_tickets = 0;  // persistent kernels take units from a ticket counter
  uint64_t *trace = nullptr;  // per-unit timestamps of sync-free launches
  int64_t trace_cap = 0;
  int max_waves = 0;  // cap on the waves of a persistent launch (0: occupancy)
  int colmajor_direct = 0;  // solve column-major blocks in place, no workspace
  int unmarked = 0;         // sync-free tags wait for the value stores
  int regions = 0;          // 0 off, 1 when enough edges are inside, 2 always
  int region_rows = 512;    // region size cap
  int owners = 1;           // solve plans with regions on the owner schedule
  int region_cap = 0;       // region executor: region size cap (0: by k)
  int region_waves = 0;     // waves per region task (0: chosen from k)
  int region_band = 0;      // slices per published band (0: the default)
  int region_mw = 1;        // region executor: memory-wave kernel where it applies
  int split_long = 0;       // split level executor: long-row threshold (0: by tile width)
  uint64_t timeout_ticks() const { return (uint64_t)(timeout_ms * 1.0e5); }
};

// Long rows of the split level executor for one threshold: records {q,
// first entry, diagonal entry, row} grouped by level (device) and their
// levels+1 offsets (host).
struct SplitLists {
  int4 *d_rec = nullptr;
  std::vector<int> ptr;
};

struct symamd_plan_s {
  int n = 0;
  int64_t nnz = 0;
  int levels = 0;
  int max_level_rows = 0;
  symamd_dtype dtype = SYMAMD_F64;
  int device = 0;
  std::vector<int> h_lvl_ptr;          // levels+1
  int *d_lvl_ptr = nullptr;            // levels+1
  int *d_perm = nullptr;               // position -> row (level order)
  // with regions: positions of a level are ordered by (region rank, row);
  // d_lvl_set keeps the canonical level set (ascending rows in a level) and
  // d_preg the region rank of each position.  Both null without regions.
  int *d_lvl_set = nullptr;
  int *d_preg = nullptr;
  int nregions = 0;
  double region_intra = 0.0;           // fraction of entries inside regions
  int *d_rp = nullptr;                 // level-ordered row pointers, n+1
  int *d_ci = nullptr;                 // level-ordered column ids (row ids), nnz
  void *d_val = nullptr;               // level-ordered values, nnz
  int *d_fmap = nullptr;               // caller CSR position of each entry
  int *d_ipos = nullptr;               // row -> position
  int *d_cpos = nullptr;               // column ids of the copy, in positions
  void *d_ws = nullptr;                // level-ordered row-major workspace
  // split level executor, per long-row threshold
  std::map<int, SplitLists> split;
  size_t ws_bytes = 0;
  int *d_orp = nullptr;                // caller CSR row pointers, n+1
  int *d_oci = nullptr;                // caller CSR column ids, nnz (built on first use)
  // transposed copy (CSC of L, level order), built on first adjoint
  int *d_trp = nullptr;
  int *d_tci = nullptr;
  void *d_tval = nullptr;
  int *d_tmap = nullptr;
  // persistent-kernel state
  int *d_state = nullptr;              // ST_WORDS ints
  std::map<int, symamd::UnitSet> units;  // rows-per-unit -> units
  // region schedules keyed by cap * 256 + slices per band
  std::map<int64_t, symamd::RegionPlan> rplans;
  std::map<int64_t, symamd::RegionPlan> rplans_t;  // same, transposed solve
  // IC(0): update pairs per variant, and the failed-pivot word
  std::map<int, symamd::Ic0Pairs> ic0;
  int *d_ic0_bad = nullptr;
  // schedule choices
  int lanes_override = 0;
  std::map<int64_t, symamd::Choice> tuned;  // key(k, layout) -> choice
  uint64_t fingerprint = 0;
  double avg_row_nnz = 1.0;
  int max_row_nnz = 1;
  // batch bookkeeping
  int nsys = 1;
  std::vector<int> h_row_off;
  // captured per-level launch sequences, keyed by (B, k, strides, lanes)
  std::map<std::string, std::pair<hipGraph_t, hipGraphExec_t>> gcache;
};

// ILU(0) plan: A's pattern (owned copies), the schedule of its lower part
// and the update pairs of every entry (factor0.hip).
struct symamd_ilu0_s {
  int n = 0;
  int64_t nnz = 0;
  int *d_Ap = nullptr, *d_Ai = nullptr, *d_Adiag = nullptr;
  symamd_plan sched = nullptr;
  int64_t pairs = 0;
  int64_t *d_pp = nullptr;
  int *d_pa = nullptr, *d_pb = nullptr, *d_dpos = nullptr;
};

// Supernodal (BCSC) solve plan (supernodal.hip): the reference's supernode
// arrays on the device, the temporaries' offsets, the per-row lists of the
// temporaries each row subtracts, the dependency lists, tags and state.
constexpr int SN_WMAX = 1024;  // widest supernode (LDS of one wave: 8 KiB)
constexpr int SN_CHUNK = 128;  // rows below a top per task
struct symamd_snplan_s {
  int n = 0, m = 0, ntasks = 0, kcap = 1;  // kcap: columns d_T and d_flags hold
  int *d_task = nullptr;       // (supernode, chunk) per task
  int64_t tcount = 0, deps = 0;
  int *d_sup = nullptr, *d_nrows = nullptr, *d_Lp = nullptr, *d_tb = nullptr;
  int *d_cptr = nullptr, *d_cidx = nullptr, *d_dptr = nullptr, *d_dlist = nullptr;
  double *d_T = nullptr;
  int *d_flags = nullptr, *d_state = nullptr;
  // transposed solve: rows below (temporaries order), the supernodes owning
  // them per supernode
  int *d_rowb = nullptr, *d_tdptr = nullptr, *d_tdlist = nullptr;
  int *d_torder = nullptr;  // transposed solve: supernode of each ticket
};

struct symamd_graph_s {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  symamd_plan plan = nullptr;
};

namespace symamd {

// Ensures the unit partition with `rows` rows per unit exists with its
// dependency lists, for the forward solve or (adjoint) the transposed one.
int plan_units(symamd_handle h, symamd_plan p, int rows, bool adjoint,
               UnitSet **out);
// Ensures the transposed (CSC) copy exists.
int plan_transpose(symamd_handle h, symamd_plan p);
// Ensures the inverse of the level order (row -> position, d_ipos) exists.
int plan_ipos(symamd_handle h, symamd_plan p);
// Ensures the caller-order column ids (d_oci) exist.
int plan_oci(symamd_handle h, symamd_plan p);

// Kernel launchers (executors.hip / adjoint.hip / spmv.hip / inspector.hip)
int launch_forward(symamd_handle h, symamd_plan p, void *d_B, int k,
                   int64_t rs, int64_t cs, int exec, int lanes, int tile);
int launch_adjoint(symamd_handle h, symamd_plan p, const void *d_x,
                   const void *d_gx, void *d_gb, void *d_gL, int exec,
                   int lanes);
int compute_levels(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
                   int *d_state, int *d_level_ptr, int *d_level_set,
                   int *levels, std::vector<int> *h_level_ptr);
int build_plan_arrays(symamd_handle h, symamd_plan p, const int *d_Lp,
                      const int *d_Li, const void *d_Lx);
int gather_values(symamd_handle h, symamd_plan p, const void *d_Lx);
int check_structure(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
                    int *d_state);
// Level sets with every buffer in a fixed device array (inspector.hip).
int levelset_noalloc(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
                     int *d_level_ptr, int *d_level_set, int *levels);
// The column-oriented SpMV forms as row pulls (spmv_csc.hip): y[r] (+)= the
// products of the V visited columns, ids[k0 .. k0+V) or c0 .. c0+V-1, in
// visit order; with d_sets, part by part (sums from +0.0 added in order).
int spmv_pull(symamd_handle h, int n, int m, const int *d_Ap, const int *d_Ai,
              const double *d_Ax, const double *d_x, double *d_y, const int *d_ids, int k0,
              int c0, int V, int nparts, const int *d_sets);
// z[i] = a_i * z[i] + b_i * y[i] (a_i = d_a[i] when d_a is given, else a).
int launch_axpby(symamd_handle h, int n, const double *d_a, const double *d_b, double a,
                 double b, const double *d_y, double *d_z);
int launch_spmv(symamd_handle h, int n, const int *d_Ap, const int *d_Ai,
                const double *d_Ax, const double *d_x, const double *d_y,
                double *d_z, double a, double b, int mode);
int plan_residual(symamd_handle h, symamd_plan p, const void *d_X,
                  const void *d_B, int k, int64_t rs, int64_t cs,
                  double *h_rel);
int forward_rows_per_unit(int k, int lanes, int tile);
int tile_cols(int k, int tile);
// Ready tags of a unit set for `tiles` column tiles (grown on demand).
int unit_flags(symamd_handle h, UnitSet *us, int tiles);
// Column-major multi-RHS solves through a level-ordered workspace.
int prepare_workspace(symamd_handle h, symamd_plan p, int k);
// The split level executor: its long-row threshold for a column tile of kt
// columns, and the long-row lists of a threshold (built once per plan and
// threshold, outside graph capture).
int split_threshold(symamd_handle h, int kt);
int split_lists(symamd_handle h, symamd_plan p, int thr, SplitLists **out);
int launch_forward_ws(symamd_handle h, symamd_plan p, void *d_B, int k, int64_t ld,
                      int exec, int lanes, int tile);
// Coarsened 3-array schedule on the device (lbc.hip).
int lbc_schedule(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
                 const double *d_cost, int parts, int min_level_dist, int div_rate,
                 int *d_level_ptr, int *d_par_ptr, int *d_partition, int *level_no,
                 int *part_no);
// d_b == nullptr: sptrsv_csr_lbc arithmetic in place on d_x; else the
// sptrsv_csr_lbc_redundant arithmetic with right-hand side d_b.
int launch_lbc(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
               const double *d_Lx, double *d_x, const double *d_b, int level_no,
               const int *h_level_ptr, const int *d_par_ptr,
               const int *d_partition);
// Regions (regions.cpp): clustering of the rows, the region-ordered level
// permutation of a plan, and the owner schedule of a unit set for nw waves.
int cluster_regions(int n, const int *Lp, const int *Li, int max_rows,
                    std::vector<int> &region);
int order_by_regions(symamd_handle h, symamd_plan p, const int *d_Lp, const int *d_Li);
int plan_owners(symamd_handle h, symamd_plan p, UnitSet *us, int nw);
// Region schedule of the plan for region cap `cap` and the band of k
// right-hand sides (built on first use; rp->R == 0 when the region graph
// has a cycle or there is nothing to gain), of the forward solve or,
// transposed, of L^T x = g.
int plan_regions(symamd_handle h, symamd_plan p, int cap, int k, RegionPlan **out,
                 bool transposed = false);
void free_region_plan(RegionPlan &rp);
// Re-gathers the region schedules' values from the plan's level-ordered
// copies (the transposed copy for transposed schedules).
int gather_region_values(symamd_handle h, symamd_plan p);
// Region-schedule forward solve (region_exec.hip); SYMAMD_ERR_NOT_SUPPORTED
// when the plan has no usable region schedule for k right-hand sides.
int launch_regions(symamd_handle h, symamd_plan p, void *d_B, int k, int64_t rs,
                   int64_t cs);
// The forward (or transposed) region solve of a column-major block from
// d_in into d_out (which may be the same buffer) through the workspace in
// natural row order; prepare_workspace must have run.
int launch_regions_ws(symamd_handle h, symamd_plan p, const void *d_in, void *d_out, int k,
                      int64_t ld, bool transposed);
// The forward (or, transposed, L^T X = B) region solve from d_B into d_X
// (which may be the same buffer); SYMAMD_ERR_NOT_SUPPORTED when the
// schedule is not usable.
int launch_regions_plan(symamd_handle h, symamd_plan p, const void *d_B, void *d_X, int k,
                        int64_t rs, int64_t cs, bool transposed, const void *d_xf = nullptr,
                        void *d_gL = nullptr, bool *fused = nullptr);
// Whether the transposed region schedule for k right-hand sides exists
// (built on first use) and fits the LDS of a workgroup; whether the
// automatic choice takes it (at least 60 % of the entries inside regions).
bool region_adjoint_usable(symamd_handle h, symamd_plan p, int k);
bool region_adjoint_preferred(symamd_handle h, symamd_plan p, int k);
// gL[e] = -(sum over c ascending of GB[i,c] * X[j,c]) for every entry
// e = (i, j) of L in the caller's CSR order (element (r, c) at r*rs + c*cs).
int launch_grad_l(symamd_handle h, symamd_plan p, const void *d_X, const void *d_GB,
                  void *d_gL, int k, int64_t rs, int64_t cs);
// Region size cap used for k right-hand sides (the LDS of one wave holds
// the region's and its halo's values of one column group per row), unless
// the handle sets one.
int region_cap(symamd_handle h, symamd_plan p, int k);
// The region cap the tuner chose for k right-hand sides (either layout), or 0.
int tuned_region_cap(symamd_plan p, int k);
int region_band(symamd_handle h, int k);
// True when a region schedule for k right-hand sides exists (building it on
// first use) and its buffers fit one wave's LDS.
bool region_usable(symamd_handle h, symamd_plan p, int k);
// Cap and task count (regions times column groups) of the schedule a region
// solve of k right-hand sides runs; 0 and 0 when it runs sync-free.
void region_schedule_info(symamd_handle h, symamd_plan p, int k, bool transposed, int *cap,
                          int *tasks);
// Whether the automatic choice takes the region executor: a usable region
// schedule with at least 60 % of L's off-diagonal entries inside regions
// (column-major blocks only through the workspace).
bool region_preferred(symamd_handle h, symamd_plan p, int k, int layout);
// IC(0) of the plan's pattern (factor0.hip); *bad_row = smallest failed pivot
// row or -1.  Blocks.
int ic0_factor(symamd_handle h, symamd_plan p, int variant, const double *d_Ax,
               double *d_Lx, int *bad_row);
void free_ic0_pairs(symamd_plan p);
// ILU(0) of a general CSR matrix (factor0.hip): the schedule and update
// pairs once, then numeric factorizations (*bad_row = smallest row with a
// zero U diagonal or -1; blocks).
int ilu0_analyze(symamd_handle h, int n, const int *d_Ap, const int *d_Ai, const int *d_Adiag,
                 symamd_ilu0_s **out);
int ilu0_numeric(symamd_handle h, symamd_ilu0_s *q, const double *d_Ax, double *d_l,
                 int *bad_row);
void ilu0_destroy(symamd_ilu0_s *q);
// Supernodal solve (supernodal.hip): analysis on host structure arrays,
// asynchronous solve, watchdog check, destroy.
int sn_analyze(symamd_handle h, int n, int m, const int *h_sup, const int *h_nrows,
               const int *h_Lp, const int *h_Li, symamd_snplan_s **out);
int sn_solve(symamd_handle h, symamd_snplan_s *q, const double *d_Lx, double *d_x, int k,
             int64_t ldx, bool transposed = false);
int sn_check(symamd_handle h, symamd_snplan_s *q);
void sn_destroy(symamd_snplan_s *q);
int ilu_lower(symamd_handle h, int n, const int *d_Ap, const int *d_Ai,
              const double *d_Ax, const int *d_diags, int unit, int *d_Lp,
              int *d_Li, double *d_Lx, int *nnz_out);

inline int64_t tune_key(int k, int layout) { return ((int64_t)k << 2) | layout; }

inline int pow2_floor(int v) {
  int r = 1;
  while (r * 2 <= v) r *= 2;
  return r;
}
inline int pow2_ceil(int v) {
  int r = 1;
  while (r < v) r *= 2;
  return r;
}

}  // namespace symamd
