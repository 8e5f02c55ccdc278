// region_exec.hip -- region-schedule forward solve for gfx950.
//
// Work = tasks.  A task is one region (a compact cluster of at most `cap`
// rows, built by regions.cpp) times one group of at most KG_MAX right-hand
// side columns; column groups are independent solves, so a task waits only
// on the same column group of the regions it reads.  One wavefront solves
// one task entirely in LDS:
//   1. copy the region's record -- packed entries (LDS slot codes and
//      values), diagonals, row ids, halo row ids, slice bounds and the
//      regions it waits for, one contiguous block laid out as its LDS image
//      -- straight into LDS (direct global -> LDS loads, all in flight at
//      once), then the group's right-hand sides of its rows the same way;
//      none of it depends on other tasks, so it overlaps the wait;
//   2. wait until the tasks it reads from carry the launch's tag;
//   3. load the halo (the values of those regions' rows it reads) into LDS;
//   4. solve the internal levels (slices) one after another; the lanes of a
//      slice take (row, column) pairs, column fastest, every dependency
//      read from LDS; slices are ordered by the wave's own LDS ordering,
//      with no cross-wave hand-off;
//   5. store the rows, wait for the stores, publish the task's tag.
// The next task's ticket and header are fetched while this one runs, so a
// task pays about four memory round trips (record, halo, stores, and the
// wait itself).
//
// Row arithmetic is the reference's (sparse_blas/sptrsv.cpp:17-25): the
// right-hand side minus each product in the row's stored order, then the
// division by the diagonal, one lane per (row, column), compiled without FMA
// contraction -- bit-identical to sptrsv_csr.
//
// Tasks are ordered by region rank (topological) then column group and
// taken by ticket, so the lowest unfinished task is always held by a
// running wave with all its inputs published (a wave holds at most one
// further ticket, always above the one it is solving): the grid cannot
// deadlock, whether or not all its blocks are resident (these blocks carry
// up to the device's LDS per workgroup, 160 KiB on gfx950), and the watchdog
// of wait_list bounds every wait.
#include "device.h"

namespace symamd {

namespace {

constexpr int KG_MAX = 8;   // right-hand side columns per task
constexpr int STAGE_U = 8;  // loads per lane in flight while staging (strided path)
constexpr int STORE_U = 8;  // result stores per lane issued after one LDS wait
constexpr int PK = RegionPlan::PACK;
static_assert(PK == 4, "the slice loop unpacks four uint16 slot codes per row");
// cache policy bits of a direct global -> LDS load: 16 = sc1, the bits of an
// agent-scope atomic load (values other waves published in this launch)
constexpr int AUX_PLAIN = 0;
constexpr int AUX_COH = 16;

typedef __attribute__((address_space(1))) void gmem_void;
typedef __attribute__((address_space(3))) void lds_void;

template <typename T>
struct RegionArgs {
  int k;
  int64_t rs, cs;  // element (i, c) at X[i*rs + c*cs]
  const T *B;      // right-hand sides (read only for the task's own rows)
  T *X;            // solution (halo rows are read from it); may equal B
  int n;           // rows: row ids read from LDS are clamped below n
  const uint8_t *rec;  // region records
  const int *head;     // 8 ints per region (RegionPlan::d_rhead)
  int64_t *tflags;     // task progress: region rank * groups + group
  int *state;
  int R, groups, smax, hmax;
  int xsb;  // LDS bytes of the value slots; the record image follows
  // transposed single right-hand side with grad-L (records with positions):
  // the forward solution of the task's rows is loaded to LDS at xfb and
  // gL[p] = -(gb[i] * x[j]) is written for every entry the task holds
  int gp, xfb;
  const T *xf;
  T *gL;
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

// LDS of one task: (smax + hmax + 1) * KT value slots (the region's rows,
// its halo, the zero slot), then the record image.
int xs_bytes(const RegionPlan &rp, int kt, int vs) {
  return (int)((((int64_t)rp.smax + rp.hmax + 1) * kt * vs + 15) & ~int64_t(15));
}
size_t region_image_bytes(const RegionPlan &rp, int kt, int vs) {
  return (size_t)xs_bytes(rp, kt, vs) + (size_t)rp.recmax;
}
// Everything a task may hold: the slots and record image and, for records
// that carry grad-L positions (single right-hand sides), the forward
// solution of the region's rows after them.
size_t region_lds_bytes(const RegionPlan &rp, int kt, int vs) {
  const size_t base = region_image_bytes(rp, kt, vs);
  return rp.gp && kt == 1 ? ((base + 15) & ~(size_t)15) + (size_t)rp.smax * vs : base;
}

__device__ __forceinline__ void vm_wait() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// LDS dst[0 .. bytes) = global src[0 .. bytes) (both 16-byte aligned, bytes
// a multiple of 16): 16 bytes per lane per instruction, 1 KiB per
// instruction, every instruction in flight together; wave `wid` of NW takes
// every NW-th KiB.  vm_wait() completes them.
template <int NW>
__device__ __forceinline__ void dma_bytes(uint8_t *dst, const uint8_t *src, int bytes,
                                          int wid) {
  const int lane = threadIdx.x & 63;
  for (int o = wid * 1024; o < bytes; o += NW * 1024)
    if (o + 16 * lane < bytes)
      __builtin_amdgcn_global_load_lds((gmem_void *)(src + o + 16 * lane),
                                       (lds_void *)(dst + o), 16, 0, 0);
}

// xs[l*KT + c] = M[rows[l]*rs + c0 + c] for l < nr, c < kc (unit column
// stride): each row's KT slots are KT*sizeof(T) contiguous LDS bytes, moved
// ES bytes per lane by direct global -> LDS loads, 64/LPR rows per
// instruction.  The launch picks ES so that kc*sizeof(T), rs*sizeof(T) and
// the buffer are ES-aligned.  `rows` lies in LDS (ids clamped below n).
// Wave `wid` of NW takes every NW-th batch of UB instructions.
template <typename T, int KT, int ES, int AUX, int NW>
__device__ __forceinline__ void dma_rows(T *xs, const T *M, const int *rows, int nr, int kc,
                                         int c0, int64_t rs, int n, int wid) {
  constexpr int RB = KT * (int)sizeof(T);
  static_assert((ES == 16 || ES == 4) && RB % ES == 0, "row bytes: a multiple of the load size");
  constexpr int LPR = RB / ES;   // lanes per row
  constexpr int RPI = 64 / LPR;  // rows per instruction
  constexpr int UB = 8;          // instructions per batch: row ids read first
  const int lane = threadIdx.x & 63;
  const int lr = lane / LPR, part = lane - lr * LPR;
  const bool live = (part + 1) * ES <= kc * (int)sizeof(T);
  for (int b0 = wid * RPI * UB; b0 < nr; b0 += NW * RPI * UB) {
    int rid[UB];
#pragma unroll
    for (int u = 0; u < UB; ++u) {
      const int l = b0 + RPI * u + lr;
      const int id = live && l < nr ? rows[l] : -1;
      rid[u] = id < n ? id : 0;
    }
#pragma unroll
    for (int u = 0; u < UB; ++u) {
      if (rid[u] < 0) continue;
      const uint8_t *src =
          reinterpret_cast<const uint8_t *>(M + (int64_t)rid[u] * rs + c0) + part * ES;
      lds_void *dst = (lds_void *)(xs + (size_t)(b0 + RPI * u) * KT);
      // the builtin takes its size as a literal
      if constexpr (ES == 16)
        __builtin_amdgcn_global_load_lds((gmem_void *)src, dst, 16, 0, AUX);
      else
        __builtin_amdgcn_global_load_lds((gmem_void *)src, dst, 4, 0, AUX);
    }
  }
}

// Strided columns (column-major blocks): xs[l*KT + c] = M[rows[l]*rs +
// (c0 + c)*cs] through registers, STAGE_U loads in flight; coherent loads
// when COH; NT threads.
template <typename T, int KT, bool COH, int NT>
__device__ __forceinline__ void stage_rows(T *xs, const T *M, const int *rows, int nr,
                                           int kc, int c0, int64_t rs, int64_t cs, int n) {
  constexpr int KS = log2c<KT>();
  const int tid = threadIdx.x;
  const int count = nr * KT;
  for (int t0 = tid; t0 < count; t0 += NT * STAGE_U) {
    T v[STAGE_U];
#pragma unroll
    for (int u = 0; u < STAGE_U; ++u) {
      const int t = t0 + NT * u;
      const int c = t & (KT - 1);
      v[u] = T(0);
      if (t < count && c < kc) {
        const int id = rows[t >> KS];
        const T *p = &M[(int64_t)((unsigned)id < (unsigned)n ? id : 0) * rs +
                        (int64_t)(c0 + c) * cs];
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

template <typename T, int KT, int ES, bool COH, int NW>
__device__ __forceinline__ void load_rows(T *xs, const T *M, const int *rows, int nr, int kc,
                                          int c0, int64_t rs, int64_t cs, int n, int wid) {
  if constexpr (ES > 0)
    dma_rows<T, KT, ES, COH ? AUX_COH : AUX_PLAIN, NW>(xs, M, rows, nr, kc, c0, rs, n, wid);
  else
    stage_rows<T, KT, COH, 64 * NW>(xs, M, rows, nr, kc, c0, rs, cs, n);
}

// LDS writes of the task's waves are visible to all of them: one wave keeps
// its LDS operations in order (only the compiler must not reorder); several
// meet at a workgroup barrier.
template <int NW>
__device__ __forceinline__ void task_sync() {
  if constexpr (NW == 1) {
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
  } else {
    __syncthreads();
  }
}

// CPL consecutive columns of one row, one LDS access (16-byte aligned for
// CPL = 2: columns are taken in even pairs of a row of KT slots).
template <typename T, int CPL>
__device__ __forceinline__ void ld_cols(const T *p, T (&out)[CPL]) {
  if constexpr (CPL == 1) {
    out[0] = p[0];
  } else {
    static_assert(CPL == 2, "one or two columns per lane");
    if constexpr (sizeof(T) == 8) {
      const double2 v = *reinterpret_cast<const double2 *>(p);
      out[0] = v.x;
      out[1] = v.y;
    } else {
      const float2 v = *reinterpret_cast<const float2 *>(p);
      out[0] = v.x;
      out[1] = v.y;
    }
  }
}

// A lane's share of a slice: CPL consecutive columns of one row.  Operands
// that do not depend on the slice being solved: position, entry count, slot
// codes and values and the diagonal (shared by the columns), and the
// right-hand sides (a row's slots are written only when its own slice is
// solved).  A lane outside the slice reads row 0 and stores nothing.
template <typename T, int CPL>
struct Pairs {
  int t, cnt, ll;
  uint2 cw;  // four uint16 slot codes, unpacked where they are used
  bool ok;
  T pv[PK], dg, acc[CPL];
};

template <typename T, int KT, int CPL>
__device__ __forceinline__ void load_pairs(Pairs<T, CPL> &m, int t, int q1, int kc,
                                           const T *xs, const uint8_t *cnt8,
                                           const uint16_t *code, const T *pval,
                                           const T *ldiag) {
  constexpr int KS = log2c<KT>();
  m.t = t;
  m.ok = t < q1 && (t & (KT - 1)) < kc;
  const int ll = m.ok ? t >> KS : 0;
  m.ll = ll;
  m.cnt = cnt8[ll];
  m.cw = *reinterpret_cast<const uint2 *>(code + PK * ll);
#pragma unroll
  for (int u = 0; u < PK; ++u) m.pv[u] = pval[PK * ll + u];
  m.dg = ldiag[ll];
  ld_cols<T, CPL>(xs + (m.ok ? t : 0), m.acc);
}

// The dependency values of the pairs (rows of earlier slices or the halo,
// final when the slice starts): the four packed slots of each column.
template <typename T, int KT, int CPL>
__device__ __forceinline__ void read_deps(const Pairs<T, CPL> &m, const T *xs,
                                          T (&xv)[PK][CPL]) {
  const int cc = m.t & (KT - 1);
  const int cd[PK] = {(int)(m.cw.x & 0xffffu), (int)(m.cw.x >> 16), (int)(m.cw.y & 0xffffu),
                      (int)(m.cw.y >> 16)};
#pragma unroll
  for (int u = 0; u < PK; ++u) ld_cols<T, CPL>(xs + cd[u] * KT + cc, xv[u]);
}

// The row arithmetic, per column: the four packed products in entry order
// (padding products are +0.0 * +0.0 from the zero slot, which leave the
// value unchanged, so no test is needed), then the row's overflow entries in
// order (bounds clamped to [0, ne]), then the divisions, one per column,
// computed together and stored in one access.  A pair's second column may
// lie beyond kc: its slot is padding that only the same padding column of
// later rows reads and that no store to X takes, so it is written without a
// test (and the two divisions share one branch-free block).
template <typename T, int KT, int CPL>
__device__ __forceinline__ void finish_pairs(Pairs<T, CPL> &m, const T (&xv)[PK][CPL], T *xs,
                                             const int *lep, const T *lval,
                                             const uint16_t *lcode, int ne) {
  const int cc = m.t & (KT - 1);
#pragma unroll
  for (int u = 0; u < PK; ++u)
#pragma unroll
    for (int j = 0; j < CPL; ++j) m.acc[j] -= m.pv[u] * xv[u][j];
  if (m.cnt > PK)
    for (int e = max(lep[m.ll], 0), e1 = min(lep[m.ll + 1], ne); e < e1; ++e) {
      T xe[CPL];
      ld_cols<T, CPL>(xs + lcode[e] * KT + cc, xe);
#pragma unroll
      for (int j = 0; j < CPL; ++j) m.acc[j] -= lval[e] * xe[j];
    }
  if (m.ok) {
    T q[CPL];
#pragma unroll
    for (int j = 0; j < CPL; ++j) q[j] = m.acc[j] / m.dg;
    if constexpr (CPL == 1) {
      xs[m.t] = q[0];
    } else if constexpr (sizeof(T) == 8) {
      *reinterpret_cast<double2 *>(xs + m.t) = make_double2(q[0], q[1]);
    } else {
      *reinterpret_cast<float2 *>(xs + m.t) = make_float2(q[0], q[1]);
    }
  }
}

// P groups of pairs of one slice: every dependency read first (rows of one
// slice never read each other), then the arithmetic.
template <typename T, int KT, int CPL, int P>
__device__ __forceinline__ void solve_pairs(Pairs<T, CPL> (&m)[P], T *xs, const int *lep,
                                            const T *lval, const uint16_t *lcode, int ne) {
  T xv[P][PK][CPL];
#pragma unroll
  for (int p = 0; p < P; ++p) read_deps<T, KT, CPL>(m[p], xs, xv[p]);
#pragma unroll
  for (int p = 0; p < P; ++p) finish_pairs<T, KT, CPL>(m[p], xv[p], xs, lep, lval, lcode, ne);
}

// The internal levels of a task, slice after slice.  Each lane solves CPL
// columns of one row at a time, P such groups of the next slice loaded while
// this slice is solved, so only the dependency values sit on the chain (a
// slice with more groups than one pass of the lanes loads the rest as it
// goes).  Within a slice the dependency reads are issued before the next
// slice's loads, and the loop is unrolled by two with the two sets of pairs
// trading roles, so no register copies (and no wait for every LDS operation
// they would need) close a slice.  One wave's LDS operations complete in
// issue order, so a wave's next slice sees its own writes of this slice
// without a wait.  Slice bounds: lane j holds
// lsl[j]; with SHORT (ns <= 62) every bound is read from that register,
// otherwise those beyond lane 63 from LDS.  Bounds are clamped to [0, nr].
template <typename T, int KT, int CPL, int P, bool SHORT, int NW>
__device__ __forceinline__ void solve_slices(T *xs, const uint8_t *rb, const RecLayout &L,
                                             int nr, int ns, int sb, int se, int ne, int kc) {
  const int lane = threadIdx.x & 63;
  const T *pval = reinterpret_cast<const T *>(rb + L.pval);
  const T *ldiag = reinterpret_cast<const T *>(rb + L.diag);
  const T *lval = reinterpret_cast<const T *>(rb + L.lval);
  const uint16_t *code = reinterpret_cast<const uint16_t *>(rb + L.code);
  const uint8_t *cnt8 = rb + L.cnt;
  const int *lsl = reinterpret_cast<const int *>(rb + L.slices);
  const int *lep = reinterpret_cast<const int *>(rb + L.lep);
  const uint16_t *lcode = reinterpret_cast<const uint16_t *>(rb + L.lcode);
  const int myb = lane <= ns ? lsl[lane] : 0;
  auto bound = [&](int s) {
    int v;
    if constexpr (SHORT)
      v = __builtin_amdgcn_readlane(myb, s);
    else
      v = s < 64 ? __builtin_amdgcn_readlane(myb, s) : lsl[s];
    return min(max(v, 0), nr);
  };
  // Column split (NW * CPL == KT): wave w owns columns w*CPL .. w*CPL+CPL-1
  // of every row and its lanes take rows; a column depends only on the same
  // column, so a wave reads only slots it wrote itself and the waves need no
  // barrier between slices.  Otherwise the lanes of all NW waves share each
  // slice's (row, column-pair) items and meet at a barrier after it.
  constexpr bool SPLIT = NW * CPL == KT;
  const int off = SPLIT ? lane * KT + ((int)threadIdx.x >> 6) * CPL : (int)threadIdx.x * CPL;
  constexpr int STEP = SPLIT ? 64 * KT : 64 * NW * CPL;  // slots per pass of the task
  Pairs<T, CPL> cur[P], nxt[P];
  int b0 = bound(sb), b1 = se > sb ? bound(sb + 1) : 0;
  if (se > sb) {
#pragma unroll
    for (int p = 0; p < P; ++p)
      load_pairs<T, KT, CPL>(cur[p], b0 * KT + off + p * STEP, b1 * KT, kc, xs, cnt8, code,
                             pval, ldiag);
  }
  // slice s: solve c (rows [b0, b1)) while nx takes slice s + 1
  auto slice = [&](Pairs<T, CPL>(&c)[P], Pairs<T, CPL>(&nx)[P], int s) {
    // after the last slice b2 == b1: the loads are harmless and unused
    const int b2 = s + 1 < se ? bound(s + 2) : b1;
    T xv[P][PK][CPL];
#pragma unroll
    for (int p = 0; p < P; ++p) read_deps<T, KT, CPL>(c[p], xs, xv[p]);
#pragma unroll
    for (int p = 0; p < P; ++p)
      load_pairs<T, KT, CPL>(nx[p], b1 * KT + off + p * STEP, b2 * KT, kc, xs, cnt8, code,
                             pval, ldiag);
#pragma unroll
    for (int p = 0; p < P; ++p) finish_pairs<T, KT, CPL>(c[p], xv[p], xs, lep, lval, lcode, ne);
    for (int t = b0 * KT + off + P * STEP; t < b1 * KT; t += STEP) {
      Pairs<T, CPL> more[1];
      load_pairs<T, KT, CPL>(more[0], t, b1 * KT, kc, xs, cnt8, code, pval, ldiag);
      solve_pairs<T, KT, CPL, 1>(more, xs, lep, lval, lcode, ne);
    }
    task_sync<SPLIT ? 1 : NW>();
    b0 = b1;
    b1 = b2;
  };
  for (int s = sb; s < se; s += 2) {
    slice(cur, nxt, s);
    if (s + 1 >= se) break;
    slice(nxt, cur, s + 1);
  }
}

// Wait until every entry of list[0 .. count) is satisfied: entry e names
// the region of rank e & (2^REGION_RANK_BITS - 1), whose task of column
// group g must have published e >> REGION_RANK_BITS bands in this launch
// (flags[rank * groups + g] >= base + bands).  One lane per entry, one
// coherent load each; a wait longer than `timeout` ticks records an error
// and gives up.  Ranks outside [0, R) are read as rank 0.
__device__ __forceinline__ void wait_bands(const int *list, int count, const int64_t *flags,
                                           int64_t base, int *state, uint64_t timeout,
                                           int groups, int g, int R) {
  const int lane = threadIdx.x & 63;
  for (int b0 = 0; b0 < count; b0 += 64) {
    const int t = b0 + lane;
    bool pending = t < count;
    const int e = pending ? list[t] : 0;
    int rank = e & ((1 << REGION_RANK_BITS) - 1);
    rank = rank < R ? rank : 0;
    const int64_t want = base + (e >> REGION_RANK_BITS);
    const int64_t *f = &flags[(int64_t)rank * groups + g];
    if (pending) pending = ld_dev(f) < want;
    if (!__any(pending)) continue;
    const uint64_t t0 = wall_ticks();
    int rounds = 0;
    while (__any(pending)) {
      __builtin_amdgcn_s_sleep(1);
      if (pending) pending = ld_dev(f) < want;
      if ((++rounds & 63) == 0 && wall_ticks() - t0 > timeout) {
        if (pending) atomicOr(&state[ST_ERROR], 1);
        pending = false;
      }
    }
  }
  asm volatile("" ::: "memory");
}

__device__ __forceinline__ void load_head(const int *head, int r, int4 &h0, int4 &h1) {
  const int4 *p = reinterpret_cast<const int4 *>(head) + 2 * (int64_t)r;
  h0 = p[0];
  h1 = p[1];
}

// NW waves (one workgroup) per task.  ES: bytes per direct global -> LDS
// load of the row values (16 or 4), or 0 for strided columns.  With several
// waves every branch that reaches a barrier depends only on the task, which
// all of them share; wave 0 alone waits for the inputs (its wait is bounded
// by the watchdog) while the others wait at the next barrier.
template <typename T, int KT, int ES, int NW>
__global__ __launch_bounds__(64 * NW) void k_regions(RegionArgs<T> a) {
  extern __shared__ __align__(16) uint8_t smem[];
  constexpr int NT = 64 * NW;
  constexpr int KS = log2c<KT>();
  constexpr int VS = (int)sizeof(T);
  // columns per lane in the slices: two when a task has two or more (the
  // row's operands are then read once per two columns)
  constexpr int CPL = KT >= 2 ? 2 : 1;
  // groups of pairs per lane prefetched per slice (two, for the eight-column
  // tasks of the 128^3 IC(0) solve, measured 10% slower: the second group's
  // loads are wasted on most slices and load the LDS of the CU)
  constexpr int P = 1;
  __shared__ int s_next[9];             // next task and its header (NW > 1)
  T *xs = reinterpret_cast<T *>(smem);  // (smax + hmax + 1) * KT values
  uint8_t *rb = smem + a.xsb;           // record image
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int target = ld_dev(&a.state[ST_EPOCH]) + 1;
  const int64_t base = (int64_t)target << 8;  // progress of this launch
  const int ntasks = a.R * a.groups;
  // the zero slot, never written afterwards (rows and halo lie below it)
  if (tid < KT) xs[(size_t)(a.smax + a.hmax) * KT + tid] = T(0);
  int ticket = 0;
  if (tid == 0) ticket = atomicAdd(&a.state[ST_TICKET], 1);
  int task;
  if constexpr (NW == 1) {
    task = __builtin_amdgcn_readlane(ticket, 0);
  } else {
    if (tid == 0) s_next[0] = ticket;
    __syncthreads();
    task = s_next[0];
  }
  int4 h0 = make_int4(0, 0, 0, 0), h1 = h0;
  if (task < ntasks) load_head(a.head, task / a.groups, h0, h1);
  while (task < ntasks) {
    if (wid == 0) trace_mark(a.trace, a.trace_cap, task, 0);
    // the next ticket, in flight while this task runs
    if (tid == 0) ticket = atomicAdd(&a.state[ST_TICKET], 1);
    const int g = task % a.groups;
    const int c0 = g * KT;
    const int kc = a.k - c0 < KT ? a.k - c0 : KT;
    const int nr = __builtin_amdgcn_readfirstlane(h0.x);
    const int nh = __builtin_amdgcn_readfirstlane(h0.y);
    const int ne = __builtin_amdgcn_readfirstlane(h0.z);
    const int ns = __builtin_amdgcn_readfirstlane(h0.w);
    const int nd = __builtin_amdgcn_readfirstlane(h1.x);
    const int nb = __builtin_amdgcn_readfirstlane(h1.w);
    const RecLayout L = rec_layout(nr, nh, ne, ns, nd, nb, VS, a.gp);
    const int *lrow = reinterpret_cast<const int *>(rb + L.rows);
    const int *lhalo = reinterpret_cast<const int *>(rb + L.halo);
    const int *ldep = reinterpret_cast<const int *>(rb + L.deps);
    const int *lsl = reinterpret_cast<const int *>(rb + L.slices);
    const int *lband = reinterpret_cast<const int *>(rb + L.bands);
    const int *lhb = reinterpret_cast<const int *>(rb + L.hband);
    const int *lwp = reinterpret_cast<const int *>(rb + L.wptr);
    int64_t *myflag = &a.tflags[task];
    // 1. record, then the right-hand sides of its rows
    dma_bytes<NW>(rb, a.rec + (size_t)__builtin_amdgcn_readfirstlane(h1.y) * 16,
                  __builtin_amdgcn_readfirstlane(h1.z) * 16, wid);
    vm_wait();
    if constexpr (NW > 1) __syncthreads();
    load_rows<T, KT, ES, false, NW>(xs, a.B, lrow, nr, kc, c0, a.rs, a.cs, a.n, wid);
    // band by band: wait for the bands of other regions it reads, load its
    // halo rows, publish the previous band (its stores are complete once the
    // halo loads are), solve its slices, store its rows
    for (int b = 0; b < nb; ++b) {
      const int sb = __builtin_amdgcn_readfirstlane(lband[b]);
      const int se = __builtin_amdgcn_readfirstlane(lband[b + 1]);
      const int hb0 = __builtin_amdgcn_readfirstlane(lhb[b]);
      const int hb1 = __builtin_amdgcn_readfirstlane(lhb[b + 1]);
      const int wb0 = __builtin_amdgcn_readfirstlane(lwp[b]);
      const int wb1 = __builtin_amdgcn_readfirstlane(lwp[b + 1]);
      // 2. inputs published (same column group of the regions read)
      if (wid == 0 && wb1 > wb0)
        wait_bands(ldep + wb0, wb1 - wb0, a.tflags, base, a.state, a.timeout, a.groups, g, a.R);
      if constexpr (NW > 1) __syncthreads();
      if (b == 0 && wid == 0) trace_mark(a.trace, a.trace_cap, task, 1);
      // 3. this band's halo rows
      if (hb1 > hb0)
        load_rows<T, KT, ES, true, NW>(xs + (size_t)(a.smax + hb0) * KT, a.X, lhalo + hb0,
                                       hb1 - hb0, kc, c0, a.rs, a.cs, a.n, wid);
      vm_wait();
      if constexpr (NW == 1)
        wave_lds_sync();
      else
        __syncthreads();
      if (b > 0 && tid == 0) st_dev(myflag, base + b);
      // 4. this band's slices; then every wave's results are visible to all
      // (the stores read every column)
      if (ns <= 62)
        solve_slices<T, KT, CPL, P, true, NW>(xs, rb, L, nr, ns, sb, se, ne, kc);
      else
        solve_slices<T, KT, CPL, P, false, NW>(xs, rb, L, nr, ns, sb, se, ne, kc);
      if constexpr (NW == 1)
        wave_lds_sync();
      else
        __syncthreads();
      // 5. store this band's rows (row ids and values of STORE_U stores read
      // first, one wait)
      const int r0 = min(max(lsl[min(max(sb, 0), ns)], 0), nr);
      const int r1 = min(max(lsl[min(max(se, 0), ns)], 0), nr);
      for (int t0 = r0 * KT + tid; t0 < r1 * KT; t0 += NT * STORE_U) {
        int id[STORE_U];
        T v[STORE_U];
#pragma unroll
        for (int u = 0; u < STORE_U; ++u) {
          const int t = t0 + NT * u;
          const bool in = t < r1 * KT && (t & (KT - 1)) < kc;
          id[u] = in ? lrow[t >> KS] : -1;
          v[u] = xs[in ? t : 0];
        }
#pragma unroll
        for (int u = 0; u < STORE_U; ++u) {
          const int c = (t0 + NT * u) & (KT - 1);
          if ((unsigned)id[u] < (unsigned)a.n)
            st_dev(&a.X[(int64_t)id[u] * a.rs + (int64_t)(c0 + c) * a.cs], v[u]);
        }
      }
    }
    int next;
    if constexpr (NW == 1) {
      next = __builtin_amdgcn_readlane(ticket, 0);
      if (next < ntasks) load_head(a.head, next / a.groups, h0, h1);
      publish_fence();
    } else {
      // wave 0 fetches the next task and its header into LDS; every wave
      // waits for its stores; then one thread publishes
      if (wid == 0) {
        const int nx = __builtin_amdgcn_readlane(ticket, 0);
        if (lane < 8) s_next[1 + lane] = nx < ntasks ? a.head[8 * (nx / a.groups) + lane] : 0;
        if (lane == 0) s_next[0] = nx;
      }
      publish_fence();
      __syncthreads();
      next = s_next[0];
      h0 = make_int4(s_next[1], s_next[2], s_next[3], s_next[4]);
      h1 = make_int4(s_next[5], s_next[6], s_next[7], s_next[8]);
    }
    if (tid == 0) st_dev(myflag, base + nb);
    if (wid == 0) trace_mark(a.trace, a.trace_cap, task, 2);
    task = next;
  }
  persistent_finish(a.state, a.total_waves);
}

// ----------------------------------------------- memory-wave specialization
// The same tasks with the memory traffic taken off the compute waves: a
// workgroup holds NC compute waves and one memory wave.  The memory wave
// takes the tickets, copies the record, loads the right-hand sides, and band
// by band waits for the bands of the regions the band reads, loads its halo
// rows, and -- as the compute waves finish a band -- stores that band's rows,
// waits for the stores and publishes the band.  The compute waves only solve
// slices from LDS: they wait for an LDS counter of loaded bands and count the
// bands they finished in another, so a compute wave never waits for a
// global store, and a band is published as soon as its stores land while
// the next band is being solved.  Every wait is bounded by the watchdog.
template <typename T>
__device__ __forceinline__ int lds_get(const int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// One round of the waits of bands b0 .. b1-1 (at most 64 entries read):
// the first of these bands that is not ready yet (b1 when all are ready).
// Entry e of the region's list belongs to band b when wptr[b] <= e <
// wptr[b+1]; it is satisfied when flags[rank * groups + g] >= base + bands.
__device__ __forceinline__ int bands_ready(const int *list, const int *wptr, int b0, int b1,
                                           const int64_t *flags, int64_t base, int groups,
                                           int g, int R, bool marked_only) {
  const int lane = threadIdx.x & 63;
  const int e0 = wptr[b0];
  const int e1 = min(wptr[b1], e0 + 64);
  // bands whose entries are not all read in this round count as not ready
  int first = b1;
  if (wptr[b1] > e1) {
    first = b0;
    while (first < b1 && wptr[first + 1] <= e1) ++first;
  }
  const int e = e0 + lane;
  int bad = INT_MAX;
  if (e < e1) {
    const int v = list[e];
    int rank = v & ((1 << REGION_RANK_BITS) - 1);
    rank = rank < R ? rank : 0;
    const int64_t want = base + (marked_only ? 1 : (v >> REGION_RANK_BITS));
    if (ld_dev(&flags[(int64_t)rank * groups + g]) < want) {
      int b = b0;
      while (wptr[b + 1] <= e) ++b;
      bad = b;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) bad = min(bad, __shfl_xor(bad, off));
  return min(first, bad);
}

// One round over every entry of list[0 .. count): true when each region it
// names has published that its rows are marked (flag >= base + 1).
__device__ __forceinline__ bool all_marked(const int *list, int count, const int64_t *flags,
                                           int64_t base, int groups, int g, int R) {
  const int lane = threadIdx.x & 63;
  bool pending = false;
  for (int e = lane; e < count; e += 64) {
    int rank = list[e] & ((1 << REGION_RANK_BITS) - 1);
    rank = rank < R ? rank : 0;
    pending |= ld_dev(&flags[(int64_t)rank * groups + g]) < base + 1;
  }
  return !__any(pending);
}

// grad-L of the rows of bands [b0, b1) of a task of the transposed solve:
// every entry (i, j) of L the task holds gets -(gb[i] * x[j]), gb[i] final
// in LDS (the task's own rows, or its halo, loaded and no longer marked
// before the band that reads it was solved) and x[j] the row's forward
// solution (in LDS at xfl); positions from the record's gp section.
template <typename T>
__device__ __forceinline__ void grad_bands(T *gL, const T *xs, const T *xfl, const uint8_t *rb,
                                           const RecLayout &L, const int *lsl, const int *lband,
                                           int b0, int b1, int nr, int ns, int ne) {
  constexpr int PK = RegionPlan::PACK;
  const int lane = threadIdx.x & 63;
  const int *gpos = reinterpret_cast<const int *>(rb + L.gpos);
  const uint16_t *code = reinterpret_cast<const uint16_t *>(rb + L.code);
  const uint8_t *cnt8 = rb + L.cnt;
  const int *lep = reinterpret_cast<const int *>(rb + L.lep);
  const uint16_t *lcode = reinterpret_cast<const uint16_t *>(rb + L.lcode);
  const int r0 = min(max(lsl[min(lband[b0], ns)], 0), nr);
  const int r1 = min(max(lsl[min(lband[b1], ns)], 0), nr);
  for (int l = r0 + lane; l < r1; l += 64) {
    const T xj = xfl[l];
    const int *gl = gpos + (PK + 1) * l;
    const int c = cnt8[l];
#pragma unroll
    for (int u = 0; u < PK; ++u)
      if (u < c) gL[gl[u]] = -(xs[code[PK * l + u]] * xj);
    if (c > PK)
      for (int e = max(lep[l], 0), e1 = min(lep[l + 1], ne); e < e1; ++e)
        gL[gpos[(PK + 1) * nr + e]] = -(xs[lcode[e]] * xj);
    gL[gl[PK]] = -(xs[l] * xj);
  }
}

// MK (single right-hand sides): the marked protocol.  A task marks its rows
// in X (after reading their right-hand sides) and publishes "marked" once;
// a task that reads them waits only for that, loads its halo values as soon
// as it reaches their band and loads again any value still carrying the
// mark.  Stores are never waited for, and no band is published: a value is
// ready when it is no longer the mark.
template <typename T, int KT, int ES, int NC, bool MK>
__global__ __launch_bounds__(64 * (NC + 1)) void k_regions_mw(RegionArgs<T> a) {
  extern __shared__ __align__(16) uint8_t smem[];
  constexpr int KS = log2c<KT>();
  constexpr int VS = (int)sizeof(T);
  constexpr int CPL = KT >= 2 ? 2 : 1;
  constexpr int P = 1;
  constexpr int LOOK = 8;  // bands whose waits one round polls
  static_assert(NC == 1 || NC * CPL == KT, "compute waves split the columns");
  static_assert(!MK || KT == 1, "the marked protocol is for single right-hand sides");
  // [0] task, [1..8] its header, [9] bands loaded, [10 + w] bands compute
  // wave w finished
  __shared__ int s_ctl[10 + NC];
  T *xs = reinterpret_cast<T *>(smem);
  uint8_t *rb = smem + a.xsb;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
  const bool memw = wid == NC;
  const int target = ld_dev(&a.state[ST_EPOCH]) + 1;
  const int64_t base = (int64_t)target << 8;
  const int ntasks = a.R * a.groups;
  if (tid < KT) xs[(size_t)(a.smax + a.hmax) * KT + tid] = T(0);
  if (memw) {
    int t = 0;
    if (lane == 0) t = atomicAdd(&a.state[ST_TICKET], 1);
    t = __builtin_amdgcn_readlane(t, 0);
    if (lane < 8) s_ctl[1 + lane] = t < ntasks ? a.head[8 * (t / a.groups) + lane] : 0;
    if (lane == 0) s_ctl[0] = t;
  }
  __syncthreads();
  int task = s_ctl[0];
  while (task < ntasks) {
    const int nr = s_ctl[1], nh = s_ctl[2], ne = s_ctl[3], ns = s_ctl[4], nd = s_ctl[5];
    const int ru = s_ctl[6], rn = s_ctl[7], nb = s_ctl[8];
    const int g = task % a.groups;
    const int c0 = g * KT;
    const int kc = a.k - c0 < KT ? a.k - c0 : KT;
    const RecLayout L = rec_layout(nr, nh, ne, ns, nd, nb, VS, a.gp);
    const int *lrow = reinterpret_cast<const int *>(rb + L.rows);
    const int *lhalo = reinterpret_cast<const int *>(rb + L.halo);
    const int *ldep = reinterpret_cast<const int *>(rb + L.deps);
    const int *lsl = reinterpret_cast<const int *>(rb + L.slices);
    const int *lband = reinterpret_cast<const int *>(rb + L.bands);
    const int *lhb = reinterpret_cast<const int *>(rb + L.hband);
    const int *lwp = reinterpret_cast<const int *>(rb + L.wptr);
    if (wid == 0) trace_mark(a.trace, a.trace_cap, task, 0);
    // 1. the record, every wave a share of it
    dma_bytes<NC + 1>(rb, a.rec + (size_t)ru * 16, rn * 16, wid);
    vm_wait();
    if (tid <= NC) s_ctl[9 + tid] = 0;
    __syncthreads();
    if (memw) {
      int ticket = 0;
      if (lane == 0) ticket = atomicAdd(&a.state[ST_TICKET], 1);
      int64_t *myflag = &a.tflags[task];
      // 2. the right-hand sides of the rows (waited for with the first halo);
      // marked protocol: then mark the rows and publish that they are marked
      load_rows<T, KT, ES, false, 1>(xs, a.B, lrow, nr, kc, c0, a.rs, a.cs, a.n, 0);
      if constexpr (MK) {
        // grad-L: the forward solution of the rows, beside the record
        if (a.gL)
          load_rows<T, KT, ES, false, 1>(reinterpret_cast<T *>(smem + a.xfb), a.xf, lrow, nr, kc,
                                         c0, a.rs, a.cs, a.n, 0);
        vm_wait();
        for (int t = lane; t < nr; t += 64) {
          const int id = lrow[t];
          if ((unsigned)id < (unsigned)a.n)
            st_dev(&a.X[(int64_t)id * a.rs + (int64_t)c0 * a.cs], mark_value<T>());
        }
        vm_wait();
        if (lane == 0) st_dev(myflag, base + 1);
      }
      // 3. rounds: issue the stores of every band all compute waves have
      // finished and the halo loads of every band whose waits are met, wait
      // once for all of them, then publish and signal
      int loaded = 0, issued = 0, stored = 0;
      bool halo_in = false;  // marked protocol: band `loaded`'s halo is in LDS
      int graded = 0;        // marked protocol with grad-L: bands whose grad-L is written
      uint64_t t0 = wall_ticks();
      bool gave_up = false;
      while (stored < nb) {
        int done = nb;
#pragma unroll
        for (int w = 0; w < NC; ++w) done = min(done, lds_get<int>(&s_ctl[10 + w]));
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        const bool stores = done > issued;
        if (stores) {
          const int r0 = min(max(lsl[min(lband[issued], ns)], 0), nr);
          const int r1 = min(max(lsl[min(lband[done], ns)], 0), nr);
          for (int t = r0 * KT + lane; t < r1 * KT; t += 64) {
            const int c = t & (KT - 1);
            const int id = lrow[t >> KS];
            if (c < kc && (unsigned)id < (unsigned)a.n)
              st_dev(&a.X[(int64_t)id * a.rs + (int64_t)(c0 + c) * a.cs], xs[t]);
          }
          issued = done;
        }
        if constexpr (MK) {
          // stores are not waited for: a reader re-loads what is still marked
          stored = issued;
          bool moved = stores;
          if (loaded < nb) {
            if (!halo_in) {
              // once every region read has marked its rows (a region stays
              // marked for the rest of the launch), every halo value of the
              // task is loaded at once; values still marked are loaded again
              // band by band below
              if (gave_up || all_marked(ldep, nd, a.tflags, base, a.groups, g, a.R)) {
                if (nh > 0)
                  load_rows<T, KT, ES, true, 1>(xs + (size_t)a.smax * KT, a.X, lhalo, nh, kc, c0,
                                                a.rs, a.cs, a.n, 0);
                vm_wait();
                wave_lds_sync();
                halo_in = true;
                moved = true;
              }
            }
            if (halo_in) {
              const int hb0 = lhb[loaded], hb1 = lhb[loaded + 1];
              bool marked = false;
              for (int t = hb0 + lane; t < hb1; t += 64) {
                T *slot = &xs[(size_t)(a.smax + t) * KT];
                if (is_mark(*slot)) {
                  const int id = lhalo[t];
                  const T v = (unsigned)id < (unsigned)a.n
                                  ? ld_dev(&a.X[(int64_t)id * a.rs + (int64_t)c0 * a.cs])
                                  : T(0);
                  *slot = v;
                  marked |= gave_up ? false : is_mark(v);
                }
              }
              wave_lds_sync();
              if (!__any(marked)) {
                if (loaded == 0 && lane == 0) trace_mark(a.trace, a.trace_cap, task, 1);
                ++loaded;
                if (lane == 0)
                  __hip_atomic_store(&s_ctl[9], loaded, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_WORKGROUP);
                moved = true;
              }
            }
          }
          if (moved) {
            t0 = wall_ticks();
          } else if (a.gL && graded < stored) {
            // grad-L only in rounds that would otherwise wait: a band a
            // round, off the path of the halo values and the stores
            grad_bands<T>(a.gL, xs, reinterpret_cast<const T *>(smem + a.xfb), rb, L, lsl, lband,
                          graded, graded + 1, nr, ns, ne);
            ++graded;
          } else {
            __builtin_amdgcn_s_sleep(1);
            if (!gave_up && wall_ticks() - t0 > a.timeout) {
              if (lane == 0) atomicOr(&a.state[ST_ERROR], 1);
              gave_up = true;
            }
          }
          continue;
        }
        int ready = loaded;
        if (loaded < nb) {
          const int b1 = min(nb, loaded + LOOK);
          ready = gave_up ? b1 : bands_ready(ldep, lwp, loaded, b1, a.tflags, base, a.groups,
                                              g, a.R, false);
          if (ready > loaded) {
            const int hb0 = lhb[loaded], hb1 = lhb[ready];
            if (hb1 > hb0)
              load_rows<T, KT, ES, true, 1>(xs + (size_t)(a.smax + hb0) * KT, a.X, lhalo + hb0,
                                            hb1 - hb0, kc, c0, a.rs, a.cs, a.n, 0);
          }
        }
        if (stores || ready > loaded) {
          vm_wait();
          if (issued > stored) {
            stored = issued;
            if (lane == 0) st_dev(myflag, base + stored);
          }
          if (ready > loaded) {
            if (loaded == 0 && lane == 0) trace_mark(a.trace, a.trace_cap, task, 1);
            loaded = ready;
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            if (lane == 0)
              __hip_atomic_store(&s_ctl[9], loaded, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
          }
          t0 = wall_ticks();
        } else {
          __builtin_amdgcn_s_sleep(1);
          if (!gave_up && wall_ticks() - t0 > a.timeout) {
            // a wait that does not end: record it and stop waiting (the
            // compute waves then finish the task with what is loaded)
            if (lane == 0) atomicOr(&a.state[ST_ERROR], 1);
            gave_up = true;
          }
        }
      }
      if constexpr (MK)
        if (a.gL && graded < nb)
          grad_bands<T>(a.gL, xs, reinterpret_cast<const T *>(smem + a.xfb), rb, L, lsl, lband,
                        graded, nb, nr, ns, ne);
      // the next task and its header
      const int nx = __builtin_amdgcn_readlane(ticket, 0);
      if (lane < 8) s_ctl[1 + lane] = nx < ntasks ? a.head[8 * (nx / a.groups) + lane] : 0;
      if (lane == 0) s_ctl[0] = nx;
      if (lane == 0) trace_mark(a.trace, a.trace_cap, task, 2);
    } else {
      // compute wave `wid`: columns wid*CPL .. (column split) of every row
      for (int b = 0; b < nb; ++b) {
        if (lds_get<int>(&s_ctl[9]) < b + 1) {
          const uint64_t t0 = wall_ticks();
          int rounds = 0;
          while (lds_get<int>(&s_ctl[9]) < b + 1) {
            __builtin_amdgcn_s_sleep(1);
            if ((++rounds & 255) == 0 && wall_ticks() - t0 > 4 * a.timeout) {
              if (lane == 0) atomicOr(&a.state[ST_ERROR], 4);
              break;
            }
          }
        }
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        const int sb = __builtin_amdgcn_readfirstlane(lband[b]);
        const int se = __builtin_amdgcn_readfirstlane(lband[b + 1]);
        if (ns <= 62)
          solve_slices<T, KT, CPL, P, true, NC>(xs, rb, L, nr, ns, sb, se, ne, kc);
        else
          solve_slices<T, KT, CPL, P, false, NC>(xs, rb, L, nr, ns, sb, se, ne, kc);
        wave_lds_sync();
        if (lane == 0)
          __hip_atomic_store(&s_ctl[10 + wid], b + 1, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
    __syncthreads();
    task = s_ctl[0];
  }
  persistent_finish(a.state, a.total_waves);
}

template <typename T>
__global__ void k_region_gather(int64_t m, const int *__restrict__ src,
                                const int *__restrict__ dst, const T *__restrict__ from,
                                T *to) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) to[dst[t]] = from[src[t]];
}

template <typename T>
int gather_one(symamd_handle h, symamd_plan p, RegionPlan &rp) {
  const T *src = (const T *)(rp.transposed ? p->d_tval : p->d_val);
  if (rp.nvals > 0)
    k_region_gather<T><<<(unsigned)((rp.nvals + 255) / 256), 256, 0, h->stream>>>(
        rp.nvals, rp.d_vsrc, rp.d_vdst, src, (T *)rp.d_rec);
  SYMAMD_HIP(hipGetLastError());
  return SYMAMD_OK;
}

// Bytes per direct global -> LDS load of a task's row values: 16 when every
// row chunk (KT and, for the last group, k mod KT columns), the row stride
// and the buffer are 16-byte aligned, else 4; 0 (staging through registers)
// for strided columns.
template <typename T, int KT>
int load_bytes(const void *d_B, int k, int64_t rs, int64_t cs) {
  if (k > 1 && cs != 1) return 0;
  constexpr int vs = (int)sizeof(T);
  const bool a16 = (KT * vs) % 16 == 0 && ((k % KT) * vs) % 16 == 0 && (rs * vs) % 16 == 0 &&
                   reinterpret_cast<uintptr_t>(d_B) % 16 == 0;
  return a16 ? 16 : 4;
}

// Tasks are taken by ticket, so any grid size is safe (a block that is not
// resident yet simply takes later tickets).  The grid is the occupancy
// estimate, but at most a third more blocks than the tasks of the widest
// depth of the region graph: more blocks would only wait, and their polls
// slow the rest (the single-column 128^3 IC(0) solve: 0.415 ms with 256
// blocks, 0.432 ms with the 1024 that fit).
// With the memory-wave kernel (MW) a block has NW compute waves and one
// memory wave; its tasks overlap across depths of the region graph (a
// region starts once the first bands it reads are published), so the grid
// may hold up to four times the widest depth's tasks.
template <typename T, int KT, int ES, int NW, bool MW>
int region_launch(symamd_handle h, symamd_plan p, const RegionPlan &rp, const void *d_B,
                  void *d_X, int k, int64_t rs, int64_t cs, const void *d_xf = nullptr,
                  void *d_gL = nullptr) {
  const void *kfn;
  if constexpr (MW)
    kfn = (const void *)&k_regions_mw<T, KT, ES, NW, KT == 1>;
  else
    kfn = (const void *)&k_regions<T, KT, ES, NW>;
  constexpr int WAVES = NW + (MW ? 1 : 0);
  // grad-L from the solve: marked single right-hand side kernel only
  const bool grad = MW && KT == 1 && d_gL && rp.gp;
  const size_t lds_base = region_image_bytes(rp, KT, (int)sizeof(T));
  const size_t xfb = (lds_base + 15) & ~(size_t)15;
  const size_t lds = grad ? xfb + (size_t)rp.smax * sizeof(T) : lds_base;
  if (lds > 65536)
    SYMAMD_HIP(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, 64 * WAVES, lds) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  const int groups = (k + KT - 1) / KT;
  const int64_t ntasks = (int64_t)rp.R * groups;
  int64_t grid = (int64_t)per_cu * h->num_cus;
  const int64_t wide = (int64_t)rp.width * groups;
  const int64_t wcap = MW ? 4 * wide : wide + wide / 3;
  if (wide > 0 && grid > wcap) grid = wcap;
  if (h->max_waves > 0 && h->max_waves < grid) grid = h->max_waves;
  if (grid > ntasks) grid = ntasks;
  if (grid < 1) grid = 1;
  RegionArgs<T> a{};
  a.k = k;
  a.rs = rs;
  a.cs = cs;
  a.B = (const T *)d_B;
  a.X = (T *)d_X;
  a.n = p->n;
  a.rec = (const uint8_t *)rp.d_rec;
  a.head = rp.d_rhead;
  a.tflags = rp.d_rflags;
  a.state = p->d_state;
  a.R = rp.R;
  a.groups = groups;
  a.smax = rp.smax;
  a.hmax = rp.hmax;
  a.xsb = xs_bytes(rp, KT, (int)sizeof(T));
  a.gp = rp.gp;
  a.xfb = (int)xfb;
  a.xf = grad ? (const T *)d_xf : nullptr;
  a.gL = grad ? (T *)d_gL : nullptr;
  a.total_waves = (int)grid * WAVES;
  a.timeout = h->timeout_ticks();
  a.trace = h->trace;
  a.trace_cap = h->trace_cap;
  void *args[] = {&a};
  SYMAMD_HIP(hipLaunchKernel(kfn, dim3((unsigned)grid), dim3(64 * WAVES), args, lds,
                            h->stream));
  return SYMAMD_OK;
}

template <typename T, int KT, int NW>
int region_launch_es(symamd_handle h, symamd_plan p, const RegionPlan &rp, const void *d_B,
                     void *d_X, int k, int64_t rs, int64_t cs, const void *d_xf, void *d_gL,
                     bool *fused) {
  int es = load_bytes<T, KT>(d_B, k, rs, cs);
  if (d_X != d_B) es = es < load_bytes<T, KT>(d_X, k, rs, cs) ? es : load_bytes<T, KT>(d_X, k, rs, cs);
  // the memory-wave kernel: direct loads only, one compute wave or the
  // column split of eight-column tasks
  // (measured on the 128^3 IC(0) factor: single right-hand sides gain from
  // the marked protocol; eight-column tasks were slower with a memory wave,
  // 1.1 vs 0.85 ms at k = 32)
  constexpr bool MWOK = KT == 1;
  if (es == 0) return region_launch<T, KT, 0, NW, false>(h, p, rp, d_B, d_X, k, rs, cs);
  if constexpr ((KT * sizeof(T)) % 16 == 0)
    if (es == 16) {
      if constexpr (MWOK)
        if (h->region_mw) {
          *fused = d_gL && rp.gp;
          return region_launch<T, KT, 16, NW, true>(h, p, rp, d_B, d_X, k, rs, cs, d_xf, d_gL);
        }
      return region_launch<T, KT, 16, NW, false>(h, p, rp, d_B, d_X, k, rs, cs);
    }
  if constexpr (MWOK)
    if (h->region_mw) {
      *fused = d_gL && rp.gp;
      return region_launch<T, KT, 4, NW, true>(h, p, rp, d_B, d_X, k, rs, cs, d_xf, d_gL);
    }
  return region_launch<T, KT, 4, NW, false>(h, p, rp, d_B, d_X, k, rs, cs);
}

template <typename T>
int region_dispatch(symamd_handle h, symamd_plan p, const RegionPlan &rp, const void *d_B,
                    void *d_X, int k, int64_t rs, int64_t cs, int kt, int nw,
                    const void *d_xf, void *d_gL, bool *fused) {
  switch (kt) {
    case 1: return region_launch_es<T, 1, 1>(h, p, rp, d_B, d_X, k, rs, cs, d_xf, d_gL, fused);
    case 2: return region_launch_es<T, 2, 1>(h, p, rp, d_B, d_X, k, rs, cs, d_xf, d_gL, fused);
    case 4: return region_launch_es<T, 4, 1>(h, p, rp, d_B, d_X, k, rs, cs, d_xf, d_gL, fused);
    default:
      if (nw == 4)
        return region_launch_es<T, 8, 4>(h, p, rp, d_B, d_X, k, rs, cs, d_xf, d_gL, fused);
      if (nw == 2)
        return region_launch_es<T, 8, 2>(h, p, rp, d_B, d_X, k, rs, cs, d_xf, d_gL, fused);
      return region_launch_es<T, 8, 1>(h, p, rp, d_B, d_X, k, rs, cs, d_xf, d_gL, fused);
  }
}

// Waves per task: several only for tasks of eight columns (whose slices
// need several passes of one wave); four by default (measured on the 128^3
// IC(0) factor at k = 32: 0.84 ms with four waves at cap 512, 0.87 with two
// at cap 256, 0.99 with one at cap 256).
int region_waves(symamd_handle h, int k) {
  if (task_cols(k) < KG_MAX) return 1;
  return h->region_waves > 0 ? h->region_waves : 4;
}

}  // namespace

// Region cap of a solve of k right-hand sides: the handle's, else the
// plan's tuned one, else 1024 rows for single right-hand sides on the
// memory-wave kernel, 256 for tasks of eight columns run by one wave and
// 512 otherwise (measured on the IC(0) factors of the 128^3 and 64^3
// Laplacians: k = 1 with the memory wave 0.309 / 0.137 ms at 1024 vs
// 0.323 / 0.142 ms at 512 -- the 512-system batch of 2D grids prefers 512,
// which the tuner finds; k = 32 with four waves per task is fastest at 512).
int region_cap(symamd_handle h, symamd_plan p, int k) {
  if (h->region_cap > 0) return h->region_cap;
  const int tuned = p ? tuned_region_cap(p, k) : 0;
  if (tuned > 0) return tuned;
  if (k == 1 && h->region_mw) return 1024;
  return task_cols(k) == KG_MAX && region_waves(h, k) == 1 ? 256 : 512;
}

// Slices per band: the handle's choice, else 4 where the memory-wave kernel
// runs (single right-hand sides: halo values loaded band by band as the
// task reaches them) and 255 (one band for regions of up to 255 levels)
// elsewhere.
int region_band(symamd_handle h, int k) {
  const int b = h->region_band > 0 ? h->region_band : h->region_mw && k == 1 ? 4 : 255;
  return b < 255 ? b : 255;
}

// The schedule of the forward (or transposed) solve for k right-hand sides
// whose task fits a workgroup's LDS: at the default cap, else at half the
// cap, and so on down to REGION_CAP_MIN rows (a matrix whose rows read many
// distinct rows outside their region needs smaller regions, whose halos are
// smaller).  A cap the handle fixes is tried alone.  Null when none fits.
static RegionPlan *fitted_plan(symamd_handle h, symamd_plan p, int k, bool transposed) {
  if (p->n == 0 || k < 1 || k > KG_MAX * RegionPlan::MAX_GROUPS) return nullptr;
  const int vs = p->dtype == SYMAMD_F64 ? 8 : 4;
  const int kt = task_cols(k);
  for (int cap = region_cap(h, p, k);; cap /= 2) {
    RegionPlan *rp = nullptr;
    if (plan_regions(h, p, cap, k, &rp, transposed) != SYMAMD_OK) return nullptr;
    if (rp->R > 0 && region_lds_bytes(*rp, kt, vs) <= (size_t)h->lds_per_block) return rp;
    if (h->region_cap > 0 || cap / 2 < REGION_CAP_MIN) return nullptr;
  }
}

bool region_preferred(symamd_handle h, symamd_plan p, int k, int layout) {
  // column-major blocks in place would be staged through registers; they
  // are solved through the row-major workspace unless the handle asks for
  // in-place solves
  if (k > 1 && layout == SYMAMD_COL_MAJOR && h->colmajor_direct) return false;
  RegionPlan *rp = fitted_plan(h, p, k, false);
  return rp && rp->inside >= 0.6;
}

int gather_region_values(symamd_handle h, symamd_plan p) {
  for (auto *plans : {&p->rplans, &p->rplans_t})
    for (auto &kv : *plans) {
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
  return fitted_plan(h, p, k, false) != nullptr;
}

void region_schedule_info(symamd_handle h, symamd_plan p, int k, bool transposed, int *cap,
                          int *tasks) {
  RegionPlan *rp = fitted_plan(h, p, k, transposed);
  *cap = rp ? rp->cap : 0;
  *tasks = rp ? rp->R * ((k + KG_MAX - 1) / KG_MAX) : 0;
}

namespace {

RegionPlan *usable_plan(symamd_handle h, symamd_plan p, int k, bool transposed) {
  return fitted_plan(h, p, k, transposed);
}

// gL[e] = -(sum over columns c of GB[i,c]*X[j,c]) for every entry e = (i, j)
// of the caller's CSR (element (r, c) of X and GB at r*rs + c*cs), the k
// products added pairwise: in groups of G columns (G = pow2_ceil(k) when
// k <= 32, else 32; a group padded with +0.0 past column k), each group
// reduced by halving -- s[c] += s[c + h] for c < h, h = G/2, G/4, .., 1 --
// and the group sums added in ascending group order.  For k = 1 this is the
// sync-free adjoint's -(gb[i] * x[j]).
//
// Single right-hand sides: one thread per row in the caller's order.
template <typename T>
__global__ void k_grad_l(int n, int64_t rs, const int *__restrict__ orp,
                         const int *__restrict__ oci, const T *__restrict__ x,
                         const T *__restrict__ gb, T *gL) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T g = gb[(int64_t)i * rs];
  for (int e = orp[i]; e < orp[i + 1]; ++e) gL[e] = -(g * x[(int64_t)oci[e] * rs]);
}

// Blocks (k >= 2): a group of G = LPR * CPL columns (G = pow2_ceil(k) up to
// 32) is spread over LPR lanes, lane l holding columns l, l + LPR, ..,
// l + (CPL-1)*LPR; 64/LPR rows per wave, GRAD_EU entries of each row at a
// time with all their X loads in flight together.  The halving steps of
// distance h >= LPR pair two columns of one lane (c < h exactly when the
// column's slot m < h/LPR); the steps below LPR pair lanes by butterfly:
// after the step of distance h every lane holds s[l] + s[l ^ h], which for
// l < h is the halving order's s[l] + s[l + h] (addition is commutative bit
// for bit), so lane 0 of a row ends with the group's pairwise sum.
constexpr int GRAD_EU = 4;
template <typename T, int LPR, int CPL>
__global__ __launch_bounds__(256) void k_grad_l_wide(int n, int k, int64_t rs, int64_t cs,
                                                     const int *__restrict__ orp,
                                                     const int *__restrict__ oci,
                                                     const T *__restrict__ X,
                                                     const T *__restrict__ GB, T *gL) {
  constexpr int RPW = 64 / LPR, EU = GRAD_EU, G = LPR * CPL;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int r = lane / LPR, cl = lane - r * LPR;
  const int64_t i64 = ((int64_t)blockIdx.x * 4 + wid) * RPW + r;
  const bool live = i64 < n;
  const int i = live ? (int)i64 : 0;
  const int e0 = live ? orp[i] : 0, len = live ? orp[i + 1] - e0 : 0;
  int mlen = len;  // most entries among the wave's rows
#pragma unroll
  for (int m = LPR; m < 64; m <<= 1) mlen = max(mlen, __shfl_xor(mlen, m));
  const T *g = GB + (int64_t)i * rs;
  const int groups = (k + G - 1) / G;
  for (int e = 0; e < mlen; e += EU) {
    bool have[EU];
    const T *xr[EU];
#pragma unroll
    for (int u = 0; u < EU; ++u) {
      have[u] = e + u < len;
      xr[u] = X + (int64_t)(have[u] ? oci[e0 + e + u] : 0) * rs;
    }
    T acc[EU];
    for (int grp = 0; grp < groups; ++grp) {
      bool cin[CPL];
      int64_t o[CPL];
      T gv[CPL], sv[EU][CPL];
#pragma unroll
      for (int m = 0; m < CPL; ++m) {
        const int c = grp * G + m * LPR + cl;
        cin[m] = c < k;
        o[m] = (int64_t)(cin[m] ? c : 0) * cs;
        gv[m] = g[o[m]];
      }
#pragma unroll
      for (int u = 0; u < EU; ++u)
#pragma unroll
        for (int m = 0; m < CPL; ++m) sv[u][m] = xr[u][o[m]];
#pragma unroll
      for (int u = 0; u < EU; ++u)
#pragma unroll
        for (int m = 0; m < CPL; ++m) sv[u][m] = cin[m] ? gv[m] * sv[u][m] : T(0);
      // halving steps inside a lane (distance h = hm * LPR)
#pragma unroll
      for (int hm = CPL / 2; hm >= 1; hm >>= 1)
#pragma unroll
        for (int u = 0; u < EU; ++u)
#pragma unroll
          for (int m = 0; m < hm; ++m) sv[u][m] = sv[u][m] + sv[u][m + hm];
      // then across the row's lanes
#pragma unroll
      for (int h = LPR / 2; h >= 1; h >>= 1)
#pragma unroll
        for (int u = 0; u < EU; ++u) sv[u][0] = sv[u][0] + __shfl_xor(sv[u][0], h);
#pragma unroll
      for (int u = 0; u < EU; ++u) acc[u] = grp == 0 ? sv[u][0] : acc[u] + sv[u][0];
    }
    if (cl == 0) {
#pragma unroll
      for (int u = 0; u < EU; ++u)
        if (have[u]) gL[e0 + e + u] = -acc[u];
    }
  }
}

template <typename T>
int grad_dispatch(symamd_handle h, symamd_plan p, const T *X, const T *GB, T *gL, int k,
                  int64_t rs, int64_t cs) {
  const int n = p->n;
  if (k == 1) {
    const unsigned blocks = (unsigned)((n + 255) / 256);
    k_grad_l<T><<<blocks, 256, 0, h->stream>>>(n, rs, p->d_orp, p->d_oci, X, GB, gL);
  } else {
    // the group width G = pow2_ceil(k) up to 32, four columns per lane
    // (two for G = 2)
    const int G = k >= 32 ? 32 : pow2_ceil(k);
    const int lpr = G >= 4 ? G / 4 : 1;
    const int rows_per_block = 4 * (64 / lpr);
    const unsigned blocks = (unsigned)((n + rows_per_block - 1) / rows_per_block);
    switch (G) {
      case 2: k_grad_l_wide<T, 1, 2><<<blocks, 256, 0, h->stream>>>(n, k, rs, cs, p->d_orp, p->d_oci, X, GB, gL); break;
      case 4: k_grad_l_wide<T, 1, 4><<<blocks, 256, 0, h->stream>>>(n, k, rs, cs, p->d_orp, p->d_oci, X, GB, gL); break;
      case 8: k_grad_l_wide<T, 2, 4><<<blocks, 256, 0, h->stream>>>(n, k, rs, cs, p->d_orp, p->d_oci, X, GB, gL); break;
      case 16: k_grad_l_wide<T, 4, 4><<<blocks, 256, 0, h->stream>>>(n, k, rs, cs, p->d_orp, p->d_oci, X, GB, gL); break;
      default: k_grad_l_wide<T, 8, 4><<<blocks, 256, 0, h->stream>>>(n, k, rs, cs, p->d_orp, p->d_oci, X, GB, gL); break;
    }
  }
  SYMAMD_HIP(hipGetLastError());
  return SYMAMD_OK;
}

}  // namespace

int launch_regions_plan(symamd_handle h, symamd_plan p, const void *d_B, void *d_X, int k,
                        int64_t rs, int64_t cs, bool transposed, const void *d_xf,
                        void *d_gL, bool *fused) {
  bool f = false;
  if (fused) *fused = false;
  RegionPlan *rp = usable_plan(h, p, k, transposed);
  if (!rp) return fail(SYMAMD_ERR_NOT_SUPPORTED, "no region schedule for this matrix and k");
  const int kt = task_cols(k), nw = region_waves(h, k);
  int rc;
  if (p->dtype == SYMAMD_F64)
    rc = region_dispatch<double>(h, p, *rp, d_B, d_X, k, rs, cs, kt, nw, d_xf, d_gL, &f);
  else
    rc = region_dispatch<float>(h, p, *rp, d_B, d_X, k, rs, cs, kt, nw, d_xf, d_gL, &f);
  if (fused) *fused = f;
  return rc;
}

int launch_regions(symamd_handle h, symamd_plan p, void *d_B, int k, int64_t rs,
                   int64_t cs) {
  return launch_regions_plan(h, p, d_B, d_B, k, rs, cs, false);
}

bool region_adjoint_usable(symamd_handle h, symamd_plan p, int k) {
  return usable_plan(h, p, k, true) != nullptr;
}

bool region_adjoint_preferred(symamd_handle h, symamd_plan p, int k) {
  RegionPlan *rp = usable_plan(h, p, k, true);
  return rp && rp->inside >= 0.6;
}

int launch_grad_l(symamd_handle h, symamd_plan p, const void *d_X, const void *d_GB,
                  void *d_gL, int k, int64_t rs, int64_t cs) {
  if (p->n == 0 || k == 0) return SYMAMD_OK;
  SYMAMD_TRY(plan_oci(h, p));
  if (p->dtype == SYMAMD_F64)
    return grad_dispatch<double>(h, p, (const double *)d_X, (const double *)d_GB,
                                 (double *)d_gL, k, rs, cs);
  return grad_dispatch<float>(h, p, (const float *)d_X, (const float *)d_GB, (float *)d_gL,
                              k, rs, cs);
}

}  // namespace symamd
