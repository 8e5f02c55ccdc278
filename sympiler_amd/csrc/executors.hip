// executors.hip -- forward and adjoint sparse triangular solve kernels for
// gfx950.
//
// Row arithmetic (reference sparse_blas/sptrsv.cpp:17-25, sptrsv_csr):
//   x[i] -= L[i,j] * x[j] for the stored j < i in ascending order,
//   then x[i] /= L[i,i]   (diagonal last in the row).
// Transposed solve (reference sparse_blas/sptrsv_lt.cpp:11-20, ltsolve),
// on the transposed copy (rows of L^T, diagonal first):
//   gb[i] -= L[r,i] * gb[r] for the stored r > i ascending, then
//   gb[i] /= L[i,i]; rows visited in reverse level order.
// With one lane per (row, right-hand side) the lane repeats exactly that
// sequence; this file is compiled with -ffp-contract=off, so the result is
// bit-identical to the reference's serial routines.  With G > 1 lanes per
// row the off-diagonal sum is split over the G lanes and reduced with
// cross-lane shuffles, then x = (b - sum) / L[i,i].
//
// Work unit = a run of rows of one level (at most 64 rows and UNIT_EMAX
// entries, cut by the inspector).  One wave solves one unit:
//   1. stage the unit's row pointers, row ids and entries (contiguous in
//      the level-ordered copy) in the wave's LDS slab, and load the
//      right-hand sides of its rows into registers;
//   2. (sync-free only) wait until every unit its rows depend on carries
//      the launch's ready tag -- one coherent load per dependency unit;
//   3. load the dependency values (coherent loads, NR rows per lane in
//      flight), finish the rows, store them with device-scope stores;
//   4. (sync-free only) wait for those stores and publish the unit's tag.
// Only steps 3-4 lie on the dependency chain.
//
// Lane layout inside a 64-wide wavefront: lane = (row, split, column) with
// the column index fastest, so with row-major right-hand sides the KT lanes
// of one row read and write KT consecutive values of that row.
//
// Executors
//   LEVELSET  one launch per level, one wave per unit; plain loads and
//             stores (kernel boundaries order the levels).
//   SYNCFREE  one persistent launch.  Units are dealt to waves round-robin
//             from a cooperative launch (all waves co-resident), or taken
//             from a ticket counter when the launch is captured in a graph;
//             either way every unit a wave waits on is held by a running
//             wave that reaches it first, so the grid cannot deadlock.
//             Waits are bounded by a watchdog on the 100 MHz clock that
//             records an error and lets the wave proceed, so every wave
//             exits.
#include "device.h"

namespace symamd {

namespace {

constexpr int UNIT_ROWS_MAX = 64;

template <typename T>
struct Slab {
  int ci[UNIT_EMAX];
  T val[UNIT_EMAX];
  int rp[UNIT_ROWS_MAX + 1];
  int row[UNIT_ROWS_MAX];
};

// Arguments shared by the forward and the transposed solve.
template <typename T>
struct SolveArgs {
  int k;
  int64_t rs, cs;            // element (i, c) at X[i*rs + c*cs]
  const int *rp;             // level-ordered row pointers of the copy solved
  const int *ci;             // its column ids (original row numbering)
  const T *val;
  const int *perm;           // position -> row (null: rows are positions)
  const T *B;                // right-hand sides (== X for the in-place forward)
  T *X;                      // solution
  // transposed solve only: grad over L's pattern from the forward copy
  const int *frp, *fci;      // level-ordered forward copy
  const int *orp;            // caller CSR row pointers
  const T *x;                // forward solution
  T *gL;
  int *state;
  const int *ub;             // unit bounds (positions)
  const int *dptr, *dlist;   // dependency units of each unit
  int *uflags;               // unit ready tags
  // owner schedule (null: units dealt round-robin or by ticket): wave w
  // solves wlist[wptr[w] .. wptr[w+1]); dptr/dlist then hold only the
  // dependencies owned by other waves
  const int *wptr, *wlist;
  int U;
  int tiles;                 // column tiles of KT columns: work item = unit * tiles + tile
  int ticketed;              // 1: units from the ticket counter
  int reverse;               // 1: units visited from last to first
  uint64_t timeout;
  int total_waves;
  // optional per-unit trace (100 MHz clock): [3u] wave took the unit,
  // [3u+1] its dependencies were ready, [3u+2] it was published
  uint64_t *trace;
  int64_t trace_cap;
};

template <typename T, int G, int KT>
__device__ __forceinline__ T lane_sum(T v) {
#pragma unroll
  for (int off = KT; off < G * KT; off <<= 1) v += __shfl_xor(v, off);
  return v;
}

template <bool SYNC, typename T>
__device__ __forceinline__ T ld_x(const T *p) {
  if constexpr (SYNC) return ld_dev(p);
  else return *p;
}
template <bool SYNC, typename T>
__device__ __forceinline__ void st_x(T *p, T v) {
  if constexpr (SYNC) st_dev(p, v);
  else *p = v;
}

// Dependency values loaded per lane and row before any is consumed: a row
// with at most dep_chunk(G) (strided) entries needs one memory round trip.
// Eight when several lanes share a row (long rows) or a row has 32 or more
// column lanes (few rows per wave, each with every load of a chunk shared
// by the whole tile), else four: with one lane per row and few columns the
// rows are short and eight slots only add dead loads.
template <int G, int KT>
constexpr int dep_chunk() { return G >= 8 || KT >= 32 ? 8 : 4; }


// Rows a lane finished in the transposed solve, kept for the grad-L pass
// that runs after the unit is published (off the dependency chain).
template <typename T, int NR>
struct Solved {
  T x[NR];
  int i[NR];
  int q[NR];
  bool act[NR];
};

// grad over row i of L: gL[p] = -(gb[i] * x[j]) at the caller's CSR
// positions of that row (read from the level-ordered forward copy).
template <typename T, int G, int NR>
__device__ __forceinline__ void grad_rows(const SolveArgs<T> &a,
                                          const Solved<T, NR> &sv) {
  if (!a.gL) return;
  const int s = (threadIdx.x & 63) % G;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (!sv.act[r]) continue;
    const int q = sv.q[r];
    const int fb = a.frp[q], len = a.frp[q + 1] - fb;
    const int ob = a.orp[sv.i[r]];
    for (int t = s; t < len; t += G) a.gL[ob + t] = -(sv.x[r] * a.x[a.fci[fb + t]]);
  }
}

// Finish NR rows per lane for all k right-hand sides: for each column tile
// load the dependency values in chunks (all loads of a chunk in flight at
// once), accumulate them in ascending entry order, divide by the diagonal
// and store.  cis/vals index the row entries jb..je (stride G).  For the
// transposed solve the finished rows are also returned in `sv`.
// SY: SY_LEVEL (kernel boundaries order the rows), SY_FLAGS (dependency
// values are ready once their unit's tag is), SY_MARKS (marked protocol:
// a value still reading as the mark is loaded again).
constexpr int SY_LEVEL = 0, SY_FLAGS = 1, SY_MARKS = 2;

template <typename T, int G, int KT, int NR, int SY, bool ADJ>
__device__ __forceinline__ void finish_rows(const SolveArgs<T> &a,
                                            const int *cis, const T *vals,
                                            const int (&i)[NR], const int (&jb)[NR],
                                            const int (&je)[NR], const T (&diag)[NR],
                                            T (&bv)[NR], const bool (&act)[NR],
                                            int iters, const int (&qpos)[NR],
                                            Solved<T, NR> &sv, int cb, int ce) {
  constexpr bool SYNC = SY != SY_LEVEL;
  constexpr int DC = dep_chunk<G, KT>();
  const int lane = threadIdx.x & 63;
  const int s = (lane / KT) % G;
  const int c = lane % KT;
  for (int c0 = cb; c0 < ce; c0 += KT) {
    const int col = c0 + c;
    const bool colok = col < a.k;
    const int64_t coff = (int64_t)col * a.cs;
    if (c0 > cb) {
#pragma unroll
      for (int r = 0; r < NR; ++r)
        bv[r] = (act[r] && colok) ? a.B[(int64_t)i[r] * a.rs + coff] : T(0);
    }
    T acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = G == 1 ? bv[r] : T(0);
    // Branch-free chunk: dead slots load the lane's own right-hand side
    // address (always valid) and are discarded, so the compiler keeps all
    // dep_chunk<G, KT>() * NR loads in flight instead of waiting at every branch.
    const int64_t coff_safe = colok ? coff : 0;
    for (int t0 = 0; t0 < iters; t0 += DC) {
      T xd[NR][DC];
#pragma unroll
      for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int t = 0; t < DC; ++t) {
          const int j = jb[r] + (t0 + t) * G;
          const bool live = j < je[r];
          const int d = cis[live ? j : 0];
          xd[r][t] = ld_x<SYNC>(&a.X[(int64_t)(live ? d : i[r]) * a.rs + coff_safe]);
        }
      if constexpr (SY == SY_MARKS) {
        // A dependency still reading as the mark was published by its
        // unit's tag before its own store landed: load it again.
        const uint64_t t0w = wall_ticks();
        int rounds = 0;
        for (;;) {
          bool again = false;
#pragma unroll
          for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int t = 0; t < DC; ++t)
              again |= (jb[r] + (t0 + t) * G < je[r]) && colok && is_mark(xd[r][t]);
          if (!__any(again)) break;
          __builtin_amdgcn_s_sleep(1);
#pragma unroll
          for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int t = 0; t < DC; ++t) {
              const int j = jb[r] + (t0 + t) * G;
              if (j < je[r] && colok && is_mark(xd[r][t]))
                xd[r][t] = ld_dev(&a.X[(int64_t)cis[j] * a.rs + coff]);
            }
          if ((++rounds & 63) == 0 && wall_ticks() - t0w > a.timeout) {
            atomicOr(&a.state[ST_ERROR], 2);
            break;
          }
        }
      }
#pragma unroll
      for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int t = 0; t < DC; ++t) {
          const int j = jb[r] + (t0 + t) * G;
          const bool live = j < je[r];
          const T v = vals[live ? j : 0];
          const T prod = v * xd[r][t];
          if (G == 1) acc[r] = live ? acc[r] - prod : acc[r];
          else acc[r] = live ? acc[r] + prod : acc[r];
        }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      T xi;
      if (G == 1) {
        xi = acc[r] / diag[r];
      } else {
        const T sum = lane_sum<T, G, KT>(acc[r]);
        xi = (bv[r] - sum) / diag[r];
      }
      if (act[r] && colok && s == 0) st_x<SYNC>(&a.X[(int64_t)i[r] * a.rs + coff], xi);
      if constexpr (ADJ) {
        sv.x[r] = xi;
        sv.i[r] = i[r];
        sv.q[r] = qpos[r];
        sv.act[r] = act[r];
      }
    }
  }
}

// Rows of one staged unit: set up the lane's NR rows from the slab's row
// pointers and the entries in `cis`/`vals` (the LDS slab, or global memory
// for a unit too large for it), load their right-hand sides, wait for the
// dependency units, finish the rows.  Called with pointers of a single
// address space so the compiler emits LDS or global loads, never FLAT ones
// (FLAT loads count against both wait counters and would serialize the
// dependency loads).
template <typename T, int G, int KT, int NR, int SY, bool ADJ>
__device__ __forceinline__ void unit_rows(const SolveArgs<T> &a, int u, int tile, int q0,
                                          int nrows, const Slab<T> &sl,
                                          const int *cis, const T *vals,
                                          int target, Solved<T, NR> &sv) {
  constexpr int LPR = G * KT;
  constexpr int RPW = 64 / LPR;
  const int lane = threadIdx.x & 63;
  const int rr = lane / LPR;
  const int s = (lane / KT) % G;
  const int c = lane % KT;
  int i[NR], jb[NR], je[NR];
  T diag[NR], bv[NR];
  bool act[NR];
  int iters = 0;
  const int col0 = tile * KT + c;
  const bool colok0 = col0 < a.k;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int lr = rr + r * RPW;
    act[r] = lr < nrows;
    i[r] = 0;
    jb[r] = 0;
    je[r] = 0;
    diag[r] = T(1);
    bv[r] = T(0);
    if (act[r]) {
      i[r] = sl.row[lr];
      const int b0 = sl.rp[lr], b1 = sl.rp[lr + 1];
      if (ADJ) {
        jb[r] = b0 + 1 + s;
        je[r] = b1;
        diag[r] = vals[b0];
      } else {
        jb[r] = b0 + s;
        je[r] = b1 - 1;
        diag[r] = vals[b1 - 1];
      }
      const int it = (je[r] - jb[r] + G - 1) / G;
      iters = it > iters ? it : iters;
      if (colok0) bv[r] = a.B[(int64_t)i[r] * a.rs + (int64_t)col0 * a.cs];
    }
  }
  // 2. dependencies
  if constexpr (SY == SY_MARKS) {
    // mark the rows (after reading their right-hand sides) and make the
    // marks visible before this unit can publish its tag
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (act[r] && colok0)
        st_dev(&a.X[(int64_t)i[r] * a.rs + (int64_t)col0 * a.cs], mark_value<T>());
    publish_fence();
  }
  if constexpr (SY != SY_LEVEL) {
    // the same column tile of the units it depends on
    const int d0 = a.dptr[u];
    wait_list(a.dlist + d0, a.dptr[u + 1] - d0, a.uflags, target, a.state, a.timeout, a.tiles,
              tile);
    trace_mark(a.trace, a.trace_cap, u * a.tiles + tile, 1);
  }
  // 3. solve
  int qpos[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) qpos[r] = q0 + rr + r * RPW;
  finish_rows<T, G, KT, NR, SY, ADJ>(a, cis, vals, i, jb, je, diag, bv, act, iters,
                                     qpos, sv, tile * KT, tile * KT + KT);
}

// Solve unit u with one wave.  ADJ selects the transposed solve (diagonal
// first in each row of the copy); its finished rows come back in `sv` for
// the grad-L pass.
template <typename T, int G, int KT, int NR, int SY, bool ADJ>
__device__ __forceinline__ void solve_unit(const SolveArgs<T> &a, int u, int tile,
                                           Slab<T> &sl, int target,
                                           Solved<T, NR> &sv) {
  const int lane = threadIdx.x & 63;
  const int q0 = a.ub[u], q1 = a.ub[u + 1];
  const int nrows = q1 - q0;
  const int e0 = a.rp[q0];
  const int ne = a.rp[q1] - e0;
  // 1. stage the unit
  for (int t = lane; t <= nrows; t += 64) {
    sl.rp[t] = a.rp[q0 + t] - e0;
    if (t < nrows) sl.row[t] = a.perm ? a.perm[q0 + t] : q0 + t;
  }
  if (ne <= UNIT_EMAX) {
    for (int t = lane; t < ne; t += 64) {
      sl.ci[t] = a.ci[e0 + t];
      sl.val[t] = a.val[e0 + t];
    }
    wave_lds_sync();
    unit_rows<T, G, KT, NR, SY, ADJ>(a, u, tile, q0, nrows, sl, sl.ci, sl.val, target, sv);
  } else {
    wave_lds_sync();
    unit_rows<T, G, KT, NR, SY, ADJ>(a, u, tile, q0, nrows, sl, a.ci + e0, a.val + e0,
                                     target, sv);
  }
}

template <typename T, int G, int KT, int NR, bool ADJ, int SY>
__global__ __launch_bounds__(256) void k_syncfree(SolveArgs<T> a) {
  __shared__ Slab<T> slabs[4];
  Slab<T> &sl = slabs[threadIdx.x >> 6];
  const int lane = threadIdx.x & 63;
  const int target = ld_dev(&a.state[ST_EPOCH]) + 1;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  int cursor = wave;
  const int stride = a.total_waves;
  // owner schedule: this wave's own list, in order (reverse for the
  // transposed solve)
  int w0 = 0, w1 = 0;
  if (a.wlist && wave < a.total_waves) {
    w0 = a.wptr[wave];
    w1 = a.wptr[wave + 1];
  }
  cursor = a.wlist ? w0 : cursor;
  // work items: unit-major, column tile minor (a unit's tiles are solved
  // independently, each waiting for the same tile of its dependencies)
  const int items = a.U * a.tiles;
  for (;;) {
    int u, tile = 0;
    if (a.wlist) {
      if (cursor >= w1) break;
      const int t = cursor++;
      u = a.wlist[a.reverse ? w1 - 1 - (t - w0) : t];
    } else {
      const int t = next_unit(a.state, a.ticketed, cursor, stride);
      if (t >= items) break;
      const int ul = t / a.tiles;
      tile = t - ul * a.tiles;
      u = a.reverse ? a.U - 1 - ul : ul;
    }
    const int item = u * a.tiles + tile;
    trace_mark(a.trace, a.trace_cap, item, 0);
    Solved<T, NR> sv;
    solve_unit<T, G, KT, NR, SY, ADJ>(a, u, tile, sl, target, sv);
    // marked protocol: the tag is only a hint (consumers re-check each value
    // against the mark), so it need not wait for the value stores
    if constexpr (SY != SY_MARKS) publish_fence();
    if (lane == 0) st_dev(&a.uflags[item], target);
    trace_mark(a.trace, a.trace_cap, item, 2);
    if constexpr (ADJ) grad_rows<T, G, NR>(a, sv);
  }
  persistent_finish(a.state, a.total_waves);
}

// One launch per level: each group of G*KT lanes solves one row of
// positions [q0, q1), reading the row straight from the level-ordered copy
// (no staging: the kernel boundary already separates it from its
// dependencies).
template <typename T, int G, int KT, bool ADJ>
__global__ __launch_bounds__(256) void k_level(SolveArgs<T> a, int q0, int q1) {
  constexpr int LPR = G * KT;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = q0 + gid / LPR;
  if (q - (gid % 64) / LPR >= q1) return;  // whole wave past the level
  const int s = ((threadIdx.x & 63) / KT) % G;
  const int c = threadIdx.x % KT;
  int i[1] = {0}, jb[1] = {0}, je[1] = {0}, qpos[1] = {q};
  T diag[1] = {T(1)}, bv[1] = {T(0)};
  bool act[1] = {q < q1};
  int iters = 0;
  if (act[0]) {
    i[0] = a.perm ? a.perm[q] : q;
    const int b0 = a.rp[q], b1 = a.rp[q + 1];
    if (ADJ) {
      jb[0] = b0 + 1 + s;
      je[0] = b1;
      diag[0] = a.val[b0];
    } else {
      jb[0] = b0 + s;
      je[0] = b1 - 1;
      diag[0] = a.val[b1 - 1];
    }
    iters = (je[0] - jb[0] + G - 1) / G;
    if (c < a.k) bv[0] = a.B[(int64_t)i[0] * a.rs + (int64_t)c * a.cs];
  }
  Solved<T, 1> sv;
  finish_rows<T, G, KT, 1, SY_LEVEL, ADJ>(a, a.ci, a.val, i, jb, je, diag, bv, act,
                                          iters, qpos, sv, 0, a.k);
  if constexpr (ADJ) grad_rows<T, G, 1>(a, sv);
}

// One launch per level, one lane per row and column (forward solves): a
// wave solves 64/KT consecutive rows of [q0, q1) for column tile
// blockIdx.y, so the tiles of a row run side by side instead of one after
// another.  Each lane keeps DC dependency values in flight and has the
// column ids of the next chunk loading meanwhile; the products are
// subtracted in ascending entry order, the serial solve's order, so the
// result is bit-identical.  With KT == 64 a wave holds one row and every
// row-structure load is wave-uniform.  DC as for the unit executors (4, or
// 8 for tiles of 32 or more columns): deeper chunks only add dead loads on
// short rows, whose scattered addresses cost throughput.

// Row q (level-ordered position) for one column: x = (b - sum L*x) / diag
// with the products subtracted in entry order, DC loads in flight.  Rows
// with more than `max_off` off-diagonal entries are left alone (another
// workgroup solves them).
template <typename T, int DC>
__device__ __forceinline__ void serial_row(const SolveArgs<T> &a, int q, bool rowok,
                                           bool colok, int64_t coff, int max_off) {
  int i = 0, jb = 0, je = 0;
  T dg = T(1), acc = T(0);
  if (rowok) {
    jb = a.rp[q];
    je = a.rp[q + 1] - 1;  // the diagonal
    if (je - jb > max_off) rowok = false;
  }
  if (!rowok) jb = je = 0;  // no loop; nothing stored
  if (rowok) {
    i = a.perm ? a.perm[q] : q;
    dg = a.val[je];
    acc = a.B[(int64_t)i * a.rs + coff];
  }
  const int64_t own = (int64_t)i * a.rs + coff;  // a valid address for dead slots
  int d[DC];
#pragma unroll
  for (int t = 0; t < DC; ++t) d[t] = a.ci[jb + t < je ? jb + t : je];
  for (int e0 = jb; e0 < je; e0 += DC) {
    T xv[DC], vv[DC];
#pragma unroll
    for (int t = 0; t < DC; ++t) {
      const bool live = e0 + t < je;
      xv[t] = a.X[live ? (int64_t)d[t] * a.rs + coff : own];
      vv[t] = a.val[live ? e0 + t : je];
    }
    // the next chunk's column ids, loading while this chunk is summed
#pragma unroll
    for (int t = 0; t < DC; ++t) {
      const int e = e0 + DC + t;
      d[t] = a.ci[e < je ? e : je];
    }
#pragma unroll
    for (int t = 0; t < DC; ++t) {
      const T prod = vv[t] * xv[t];
      acc = e0 + t < je ? acc - prod : acc;
    }
  }
  if (rowok && colok) a.X[own] = acc / dg;
}

template <typename T, int KT>
__global__ __launch_bounds__(256) void k_level_cols(SolveArgs<T> a, int q0, int q1) {
  constexpr int RPW = 64 / KT;
  const int lane = threadIdx.x & 63;
  const int qw = q0 + (int)(blockIdx.x * 4 + (threadIdx.x >> 6)) * RPW;
  if (qw >= q1) return;  // whole wave past the level
  int q = qw + lane / KT;
  if constexpr (RPW == 1) q = __builtin_amdgcn_readfirstlane(q);
  const int col = (int)blockIdx.y * KT + lane % KT;
  const bool colok = col < a.k;
  serial_row<T, dep_chunk<1, KT>()>(a, q, q < q1, colok, (int64_t)(colok ? col : 0) * a.cs,
                                    0x7fffffff);
}

// Split level executor: one launch per level.  Blocks [0, nlong) each solve
// one row with more than max_off off-diagonal entries (records lrec[]:
// position, first entry, diagonal entry, row), its entries dealt over
// SPLIT_W wavefronts x 64/KT lane slices, LC_DEPS loads in flight per lane;
// the slice sums meet in a fixed tree (lanes, then wavefronts in order), so
// the row's sum is re-associated but the same on every run.  The remaining
// blocks solve the other rows of the level one lane per row and column in
// entry order (bit-identical), 64/KT rows per wavefront.  Column tile =
// blockIdx.y.  (Eight wavefronts per group measured slower at k = 32 and
// 128: fewer slices per long row.)
constexpr int SPLIT_W = 16;
constexpr int LC_DEPS = 16;

template <typename T, int KT>
__device__ __forceinline__ void split_rows(const SolveArgs<T> &a, int q0, int q1,
                                           const int4 *__restrict__ lrec, int nlong,
                                           int max_off) {
  constexpr int SPW = 64 / KT;
  __shared__ T part[SPLIT_W][KT];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int col = (int)blockIdx.y * KT + lane % KT;
  const bool colok = col < a.k;
  const int64_t coff = (int64_t)(colok ? col : 0) * a.cs;
  if ((int)blockIdx.x >= nlong) {
    const int qw = q0 + (((int)blockIdx.x - nlong) * SPLIT_W + wid) * SPW;
    if (qw >= q1) return;  // whole wave past the level
    int q = qw + lane / KT;
    if constexpr (SPW == 1) q = __builtin_amdgcn_readfirstlane(q);
    serial_row<T, LC_DEPS>(a, q, q < q1, colok, coff, max_off);
    return;
  }
  constexpr int S = SPLIT_W * SPW;  // entry slices of the row
  const int4 r = lrec[blockIdx.x];
  const int i = a.perm ? r.w : r.x;
  const int jb = r.y, je = r.z;
  const int64_t own = (int64_t)i * a.rs + coff;
  const T bv = a.B[own], dg = a.val[je];  // in flight with the first chunk
  T acc = T(0);
  for (int e0 = jb + wid * SPW + lane / KT; e0 < je; e0 += S * LC_DEPS) {
    T xv[LC_DEPS], vv[LC_DEPS];
#pragma unroll
    for (int t = 0; t < LC_DEPS; ++t) {
      const int e = e0 + t * S;
      const bool live = e < je;
      const int d = a.ci[live ? e : je];
      xv[t] = a.X[live ? (int64_t)d * a.rs + coff : own];
      vv[t] = a.val[live ? e : je];
    }
#pragma unroll
    for (int t = 0; t < LC_DEPS; ++t)
      if (e0 + t * S < je) acc += vv[t] * xv[t];
  }
  // the wavefront's slices of each column, then the wavefronts in order
#pragma unroll
  for (int off = KT; off < 64; off <<= 1) acc += __shfl_xor(acc, off);
  if (lane < KT) part[wid][lane] = acc;
  __syncthreads();
  if (wid == 0 && lane < KT) {
    T sum = part[0][lane];
#pragma unroll
    for (int w = 1; w < SPLIT_W; ++w) sum += part[w][lane];
    if (colok) a.X[own] = (bv - sum) / dg;
  }
}

template <typename T, int KT>
__global__ __launch_bounds__(64 * SPLIT_W) void k_level_split(SolveArgs<T> a, int q0, int q1,
                                                            const int4 *__restrict__ lrec,
                                                            int nlong, int max_off) {
  split_rows<T, KT>(a, q0, q1, lrec, nlong, max_off);
}

// ------------------------------------------------- caller schedules
// Executes a caller-given 3-array coarsened schedule with the semantics of
// sparse_blas/sptrsv.cpp:57-75 (sptrsv_csr_lbc): one launch per coarse
// level, one lane per partition, the rows of a partition solved in their
// listed order by that lane.  Bit-identical to the serial solve.
__global__ __launch_bounds__(256) void k_lbc_level(
    const int *__restrict__ Lp, const int *__restrict__ Li,
    const double *__restrict__ Lx, double *x, const int *__restrict__ par_ptr,
    const int *__restrict__ partition, int p0, int p1) {
  const int part = p0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (part >= p1) return;
  for (int k1 = par_ptr[part]; k1 < par_ptr[part + 1]; ++k1) {
    const int i = partition[k1];
    double xi = x[i];
    const int je = Lp[i + 1] - 1;
    for (int j = Lp[i]; j < je; ++j) xi -= Lx[j] * x[Li[j]];
    x[i] = xi / Lx[je];
  }
}

// The same schedule with a separate right-hand side, with the arithmetic of
// sparse_blas/sptrsv.cpp:78-106 (sptrsv_csr_lbc_redundant): the products
// are accumulated from zero and b[i] is added last,
// x[i] = ((0 - sum_j Lx[j] x[Li[j]]) + b[i]) / diag.  The reference's
// per-thread scratch row is a register here.
__global__ __launch_bounds__(256) void k_lbc_redundant_level(
    const int *__restrict__ Lp, const int *__restrict__ Li,
    const double *__restrict__ Lx, double *x, const double *__restrict__ b,
    const int *__restrict__ par_ptr, const int *__restrict__ partition, int p0, int p1) {
  const int part = p0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (part >= p1) return;
  for (int k1 = par_ptr[part]; k1 < par_ptr[part + 1]; ++k1) {
    const int i = partition[k1];
    double t = 0.0;
    const int je = Lp[i + 1] - 1;
    for (int j = Lp[i]; j < je; ++j) t -= Lx[j] * x[Li[j]];
    x[i] = (t + b[i]) / Lx[je];
  }
}

// ------------------------------------------------------------ dispatch

// Co-resident grid of 256-thread blocks for a persistent kernel.
int resident_grid(symamd_handle h, const void *kernel, int units) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  int64_t cap = (int64_t)per_cu * h->num_cus;
  if (h->max_waves > 0 && (h->max_waves + 3) / 4 < cap) cap = (h->max_waves + 3) / 4;
  int64_t want = (units + 3) / 4;
  if (want < 1) want = 1;
  return (int)(want < cap ? want : cap);
}

bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
  return st != hipStreamCaptureStatusNone;
}

// Launch a persistent kernel.  Plans with a region order use the owner
// schedule on every co-resident wave (cooperative launch); otherwise units
// are dealt round-robin (cooperative) or, inside graph capture or when
// forced, taken from the ticket counter with the full dependency lists.
template <typename T>
int launch_persistent(symamd_handle h, symamd_plan p, UnitSet *us, const void *kfn,
                      SolveArgs<T> &a) {
  a.wptr = nullptr;
  a.wlist = nullptr;
  const bool coop = !capturing(h->stream) && !h->force_tickets;
  if (coop && p->d_preg && h->owners && a.tiles == 1) {
    const int grid = resident_grid(h, kfn, 1 << 30);
    SolveArgs<T> b = a;
    b.total_waves = grid * 4;
    b.ticketed = 0;
    SYMAMD_TRY(plan_owners(h, p, us, b.total_waves));
    b.wptr = us->d_wptr;
    b.wlist = us->d_wlist;
    b.dptr = us->d_xdptr;
    b.dlist = us->d_xdlist;
    void *args[] = {&b};
    if (hipLaunchCooperativeKernel(kfn, dim3(grid), dim3(256), args, 0, h->stream) ==
        hipSuccess)
      return SYMAMD_OK;
    (void)hipGetLastError();
  }
  const int grid = resident_grid(h, kfn, a.U * a.tiles);
  a.total_waves = grid * 4;
  if (coop) {
    a.ticketed = 0;
    void *args[] = {&a};
    hipError_t e = hipLaunchCooperativeKernel(kfn, dim3(grid), dim3(256), args, 0,
                                              h->stream);
    if (e == hipSuccess) return SYMAMD_OK;
    (void)hipGetLastError();
  }
  a.ticketed = 1;
  void *args[] = {&a};
  SYMAMD_HIP(hipLaunchKernel(kfn, dim3(grid), dim3(256), args, 0, h->stream));
  return SYMAMD_OK;
}

}  // namespace

// Ready tags for `tiles` column tiles of every unit (tag unit * tiles +
// tile), grown on demand; new tags are cleared (never "ready").  Nothing is
// allocated when the capacity suffices, as inside graph capture, where the
// solve's preparation has grown it beforehand.
int unit_flags(symamd_handle h, UnitSet *us, int tiles) {
  if (tiles <= us->flag_tiles) return SYMAMD_OK;
  if (capturing(h->stream))
    return fail(SYMAMD_ERR_NOT_SUPPORTED, "unit tags for %d column tiles not prepared", tiles);
  const size_t cnt = (size_t)(us->count > 0 ? us->count : 1) * tiles;
  int *f = nullptr;
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  SYMAMD_HIP(hipMalloc(&f, sizeof(int) * cnt));
  SYMAMD_HIP(hipMemset(f, 0, sizeof(int) * cnt));
  (void)hipFree(us->d_uflags);
  us->d_uflags = f;
  us->flag_tiles = tiles;
  return SYMAMD_OK;
}

namespace {

template <typename T, int KT>
int run_split(symamd_handle h, symamd_plan p, SolveArgs<T> a) {
  const int thr = split_threshold(h, KT);
  SplitLists *sl = nullptr;
  SYMAMD_TRY(split_lists(h, p, thr, &sl));
  const int tiles = (a.k + KT - 1) / KT;
  if (tiles > 65535)
    return fail(SYMAMD_ERR_NOT_SUPPORTED, "split level executor: %d column tiles", tiles);
  constexpr int RPB = SPLIT_W * (64 / KT);  // short rows per workgroup
  for (int l = 0; l < p->levels; ++l) {
    const int q0 = p->h_lvl_ptr[l], q1 = p->h_lvl_ptr[l + 1];
    if (q1 <= q0) continue;
    const int l0 = sl->ptr[l], nl = sl->ptr[l + 1] - l0;
    const dim3 grid((unsigned)(nl + (q1 - q0 + RPB - 1) / RPB), (unsigned)tiles);
    k_level_split<T, KT><<<grid, 64 * SPLIT_W, 0, h->stream>>>(a, q0, q1, sl->d_rec + l0, nl,
                                                                thr);
  }
  SYMAMD_HIP(hipGetLastError());
  return SYMAMD_OK;
}

template <typename T, int G, int KT, int NR, bool ADJ>
int run(symamd_handle h, symamd_plan p, SolveArgs<T> a, int exec) {
  if (exec == SYMAMD_EXEC_LEVELSET_SPLIT) {
    if constexpr (G == 1 && !ADJ) return run_split<T, KT>(h, p, a);
    return fail(SYMAMD_ERR_INVALID_ARG, "the split level executor runs forward solves, one lane per row");
  }
  constexpr int R = NR * 64 / (G * KT);
  static_assert(R <= UNIT_ROWS_MAX, "unit rows exceed the LDS slab");
  UnitSet *us = nullptr;
  SYMAMD_TRY(plan_units(h, p, R, ADJ, &us));
  a.tiles = exec == SYMAMD_EXEC_SYNCFREE ? (a.k + KT - 1) / KT : 1;
  if (a.tiles < 1) a.tiles = 1;
  SYMAMD_TRY(unit_flags(h, us, a.tiles));
  a.ub = us->d_ub;
  a.U = us->count;
  a.dptr = us->d_dptr;
  a.dlist = us->d_dlist;
  a.uflags = us->d_uflags;
  a.reverse = ADJ ? 1 : 0;
  if (exec == SYMAMD_EXEC_SYNCFREE) {
    // the marked protocol is built for single right-hand sides only; with
    // 32 columns its mark writes and re-loads cost more than the store
    // acknowledgement they save
    if constexpr (KT == 1)
      if (a.k == 1 && !h->unmarked)
        return launch_persistent(
            h, p, us, (const void *)&k_syncfree<T, G, KT, NR, ADJ, SY_MARKS>, a);
    return launch_persistent(h, p, us,
                             (const void *)&k_syncfree<T, G, KT, NR, ADJ, SY_FLAGS>, a);
  }
  constexpr int LPR = G * KT;
  const int tiles = (a.k + KT - 1) / KT;
  if constexpr (G == 1 && !ADJ) {
    if (tiles <= 65535) {
      constexpr int RPB = 4 * (64 / KT);  // rows per workgroup
      for (int l = 0; l < p->levels; ++l) {
        const int q0 = p->h_lvl_ptr[l], q1 = p->h_lvl_ptr[l + 1];
        if (q1 <= q0) continue;
        const dim3 grid((unsigned)((q1 - q0 + RPB - 1) / RPB), (unsigned)tiles);
        k_level_cols<T, KT><<<grid, 256, 0, h->stream>>>(a, q0, q1);
      }
      SYMAMD_HIP(hipGetLastError());
      return SYMAMD_OK;
    }
  }
  for (int l0 = 0; l0 < p->levels; ++l0) {
    const int l = ADJ ? p->levels - 1 - l0 : l0;
    const int q0 = p->h_lvl_ptr[l], q1 = p->h_lvl_ptr[l + 1];
    const int64_t lanes = (int64_t)(q1 - q0) * LPR;
    const int grid = (int)((lanes + 255) / 256);
    k_level<T, G, KT, ADJ><<<grid, 256, 0, h->stream>>>(a, q0, q1);
  }
  SYMAMD_HIP(hipGetLastError());
  return SYMAMD_OK;
}

// Rows per lane: enough rows in flight per wave to hide memory latency.
constexpr int rows_per_lane(int lpr) {
  return lpr >= 4 ? 4 : (lpr >= 2 ? 2 : 1);
}

template <typename T, int KT, bool ADJ>
int pick_g(symamd_handle h, symamd_plan p, SolveArgs<T> a, int exec, int g) {
#define SYMAMD_G(GV)                                                           \
  case GV:                                                                     \
    if constexpr (GV * KT <= 64)                                               \
      return run<T, GV, KT, rows_per_lane(GV * KT), ADJ>(h, p, a, exec);       \
    break;
  switch (g) {
    SYMAMD_G(1)
    SYMAMD_G(2)
    SYMAMD_G(4)
    SYMAMD_G(8)
    SYMAMD_G(16)
    SYMAMD_G(32)
    SYMAMD_G(64)
    default: break;
  }
#undef SYMAMD_G
  return fail(SYMAMD_ERR_INVALID_ARG, "lanes per row %d not supported with a %d-column tile", g, KT);
}

template <typename T>
int fwd_dispatch(symamd_handle h, symamd_plan p, SolveArgs<T> a, int exec,
                 int g, int kt) {
  switch (kt) {
    case 1: return pick_g<T, 1, false>(h, p, a, exec, g);
    case 2: return pick_g<T, 2, false>(h, p, a, exec, g);
    case 4: return pick_g<T, 4, false>(h, p, a, exec, g);
    case 8: return pick_g<T, 8, false>(h, p, a, exec, g);
    case 16: return pick_g<T, 16, false>(h, p, a, exec, g);
    case 32: return pick_g<T, 32, false>(h, p, a, exec, g);
    case 64: return pick_g<T, 64, false>(h, p, a, exec, g);
    default: break;
  }
  return fail(SYMAMD_ERR_INVALID_ARG, "column tile %d not supported", kt);
}

template <typename T>
SolveArgs<T> forward_args(symamd_handle h, symamd_plan p, void *d_B, int k,
                          int64_t rs, int64_t cs) {
  SolveArgs<T> a{};
  a.k = k;
  a.rs = rs;
  a.cs = cs;
  a.rp = p->d_rp;
  a.ci = p->d_ci;
  a.val = (const T *)p->d_val;
  a.perm = p->d_perm;
  a.B = (const T *)d_B;
  a.X = (T *)d_B;
  a.state = p->d_state;
  a.timeout = h->timeout_ticks();
  a.trace = h->trace;
  a.trace_cap = h->trace_cap;
  return a;
}

template <typename T>
SolveArgs<T> adjoint_args(symamd_handle h, symamd_plan p, const void *d_x,
                          const void *d_gx, void *d_gb, void *d_gL) {
  SolveArgs<T> a{};
  a.k = 1;
  a.rs = 1;
  a.cs = 0;
  a.rp = p->d_trp;
  a.ci = p->d_tci;
  a.val = (const T *)p->d_tval;
  a.perm = p->d_perm;
  a.B = (const T *)d_gx;
  a.X = (T *)d_gb;
  a.frp = p->d_rp;
  a.fci = p->d_ci;
  a.orp = p->d_orp;
  a.x = (const T *)d_x;
  a.gL = (T *)d_gL;
  a.state = p->d_state;
  a.timeout = h->timeout_ticks();
  a.trace = h->trace;
  a.trace_cap = h->trace_cap;
  return a;
}

// Column-major right-hand sides <-> level-ordered row-major workspace,
// W[ipos[i]*k + c] = B[i + c*ld] (W[i*k + c] when ipos is null), through LDS
// tiles of 64 rows x 32 columns: the column-major side moves 64 consecutive
// rows of one column per access, the workspace side 32 consecutive columns
// of one row.
template <typename T, bool TO_WS>
__global__ __launch_bounds__(256) void k_ws_transpose(int n, int k, int64_t ld, T *B,
                                                      const int *__restrict__ ipos,
                                                      T *W) {
  __shared__ T tile[64][33];
  const int i0 = blockIdx.x * 64;
  const int c0 = blockIdx.y * 32;
  const int tr = threadIdx.x % 64, tc = threadIdx.x / 64;  // column-major side
  const int wc = threadIdx.x % 32, wr = threadIdx.x / 32;  // workspace side
  if (TO_WS) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int c = tc + 4 * m, i = i0 + tr;
      if (i < n && c0 + c < k) tile[tr][c] = B[i + (int64_t)(c0 + c) * ld];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int r = wr + 8 * m, i = i0 + r;
      if (i < n && c0 + wc < k) W[(int64_t)(ipos ? ipos[i] : i) * k + c0 + wc] = tile[r][wc];
    }
  } else {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int r = wr + 8 * m, i = i0 + r;
      if (i < n && c0 + wc < k) tile[r][wc] = W[(int64_t)(ipos ? ipos[i] : i) * k + c0 + wc];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int c = tc + 4 * m, i = i0 + tr;
      if (i < n && c0 + c < k) B[i + (int64_t)(c0 + c) * ld] = tile[tr][c];
    }
  }
}

__global__ void k_cols_to_positions(int64_t nnz, const int *__restrict__ ci,
                                    const int *__restrict__ ipos, int *cpos) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nnz) cpos[t] = ipos[ci[t]];
}

}  // namespace

// Long-row threshold of the split level executor: the handle's, else 16
// off-diagonal entries for column tiles of up to 16 columns and 32 for wider
// ones (measured on the Zipf(1.8) n = 1e6 matrix: at k = 8 the lower
// threshold took 3.4 ms against 4.8; at k = 32 and 128 it took 4.9 and 7.7 ms
// against 4.4 and 6.8).
int split_threshold(symamd_handle h, int kt) {
  if (h->split_long > 0) return h->split_long;
  return kt <= 16 ? 16 : 32;
}

// Long rows of every level for the split level executor: the positions q
// with more than `thr` off-diagonal entries in the level-ordered copy,
// grouped by level, as records {q, first entry, diagonal entry, row}.
int split_lists(symamd_handle h, symamd_plan p, int thr, SplitLists **out) {
  auto it = p->split.find(thr);
  if (it != p->split.end()) {
    *out = &it->second;
    return SYMAMD_OK;
  }
  if (capturing(h->stream))
    return fail(SYMAMD_ERR_NOT_SUPPORTED, "split level lists not prepared before capture");
  const int n = p->n;
  std::vector<int> rp(n + 1), perm(n);
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  SYMAMD_HIP(hipMemcpy(rp.data(), p->d_rp, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  if (n > 0)
    SYMAMD_HIP(hipMemcpy(perm.data(), p->d_perm, sizeof(int) * n, hipMemcpyDeviceToHost));
  std::vector<int> ptr(p->levels + 1, 0);
  std::vector<int4> pos;
  for (int l = 0; l < p->levels; ++l) {
    for (int q = p->h_lvl_ptr[l]; q < p->h_lvl_ptr[l + 1]; ++q)
      if (rp[q + 1] - 1 - rp[q] > thr) pos.push_back(make_int4(q, rp[q], rp[q + 1] - 1, perm[q]));
    ptr[l + 1] = (int)pos.size();
  }
  int4 *d = nullptr;
  if (!pos.empty()) {
    SYMAMD_HIP(hipMalloc(&d, sizeof(int4) * pos.size()));
    if (hipMemcpy(d, pos.data(), sizeof(int4) * pos.size(), hipMemcpyHostToDevice) !=
        hipSuccess) {
      (void)hipFree(d);
      return fail(SYMAMD_ERR_HIP, "split level lists: copy failed");
    }
  }
  auto &e = p->split[thr];
  e.d_rec = d;
  e.ptr = std::move(ptr);
  *out = &e;
  return SYMAMD_OK;
}

int prepare_workspace(symamd_handle h, symamd_plan p, int k) {
  if (!p->d_ipos) {
    UnitSet *us = nullptr;  // building any unit set builds the inverse map
    SYMAMD_TRY(plan_units(h, p, 64, false, &us));
  }
  if (!p->d_cpos) {
    SYMAMD_HIP(hipMalloc(&p->d_cpos, sizeof(int) * (p->nnz > 0 ? p->nnz : 1)));
    if (p->nnz > 0) {
      k_cols_to_positions<<<(unsigned)((p->nnz + 255) / 256), 256, 0, h->stream>>>(
          p->nnz, p->d_ci, p->d_ipos, p->d_cpos);
      SYMAMD_HIP(hipGetLastError());
    }
  }
  const size_t need = (size_t)p->n * k * (p->dtype == SYMAMD_F64 ? 8 : 4);
  if (need > p->ws_bytes) {
    if (p->d_ws) SYMAMD_HIP(hipFree(p->d_ws));
    p->d_ws = nullptr;
    p->ws_bytes = 0;
    SYMAMD_HIP(hipMalloc(&p->d_ws, need));
    p->ws_bytes = need;
  }
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  return SYMAMD_OK;
}

// Column-major k > 1: transpose into the level-ordered workspace, solve
// there with rows = positions (coalesced right-hand sides, dependency ids
// in position numbering), transpose back.  prepare_workspace must have run.
int launch_forward_ws(symamd_handle h, symamd_plan p, void *d_B, int k, int64_t ld,
                      int exec, int lanes, int tile) {
  if (p->n == 0 || k == 0) return SYMAMD_OK;
  const int kt = tile_cols(k, tile);
  if (lanes * kt > 64) lanes = 64 / kt;
  dim3 grid((p->n + 63) / 64, (k + 31) / 32);
  if (p->dtype == SYMAMD_F64) {
    k_ws_transpose<double, true><<<grid, 256, 0, h->stream>>>(p->n, k, ld, (double *)d_B,
                                                              p->d_ipos, (double *)p->d_ws);
    SolveArgs<double> a = forward_args<double>(h, p, p->d_ws, k, k, 1);
    a.ci = p->d_cpos;
    a.perm = nullptr;
    SYMAMD_TRY(fwd_dispatch<double>(h, p, a, exec, lanes, kt));
    k_ws_transpose<double, false><<<grid, 256, 0, h->stream>>>(p->n, k, ld, (double *)d_B,
                                                               p->d_ipos, (double *)p->d_ws);
  } else {
    k_ws_transpose<float, true><<<grid, 256, 0, h->stream>>>(p->n, k, ld, (float *)d_B,
                                                             p->d_ipos, (float *)p->d_ws);
    SolveArgs<float> a = forward_args<float>(h, p, p->d_ws, k, k, 1);
    a.ci = p->d_cpos;
    a.perm = nullptr;
    SYMAMD_TRY(fwd_dispatch<float>(h, p, a, exec, lanes, kt));
    k_ws_transpose<float, false><<<grid, 256, 0, h->stream>>>(p->n, k, ld, (float *)d_B,
                                                              p->d_ipos, (float *)p->d_ws);
  }
  SYMAMD_HIP(hipGetLastError());
  return SYMAMD_OK;
}

// Column-major k > 1 on the region executor: transpose d_in into the
// workspace in natural row order, solve it there as a row-major block (the
// region kernel's direct LDS loads), transpose back into d_out.
// prepare_workspace must have run.
int launch_regions_ws(symamd_handle h, symamd_plan p, const void *d_in, void *d_out, int k,
                      int64_t ld, bool transposed) {
  if (p->n == 0 || k == 0) return SYMAMD_OK;
  dim3 grid((p->n + 63) / 64, (k + 31) / 32);
  if (p->dtype == SYMAMD_F64) {
    k_ws_transpose<double, true><<<grid, 256, 0, h->stream>>>(
        p->n, k, ld, (double *)const_cast<void *>(d_in), nullptr, (double *)p->d_ws);
    SYMAMD_TRY(launch_regions_plan(h, p, p->d_ws, p->d_ws, k, k, 1, transposed));
    k_ws_transpose<double, false><<<grid, 256, 0, h->stream>>>(p->n, k, ld, (double *)d_out,
                                                               nullptr, (double *)p->d_ws);
  } else {
    k_ws_transpose<float, true><<<grid, 256, 0, h->stream>>>(
        p->n, k, ld, (float *)const_cast<void *>(d_in), nullptr, (float *)p->d_ws);
    SYMAMD_TRY(launch_regions_plan(h, p, p->d_ws, p->d_ws, k, k, 1, transposed));
    k_ws_transpose<float, false><<<grid, 256, 0, h->stream>>>(p->n, k, ld, (float *)d_out,
                                                              nullptr, (float *)p->d_ws);
  }
  SYMAMD_HIP(hipGetLastError());
  return SYMAMD_OK;
}

// Columns per tile: `tile` when given (a power of two, at most 64), else
// the columns of k up to 64.
int tile_cols(int k, int tile) {
  int kt = pow2_ceil(k < 64 ? (k > 0 ? k : 1) : 64);
  if (tile > 0 && tile < kt) kt = pow2_ceil(tile);
  return kt;
}

int forward_rows_per_unit(int k, int lanes, int tile) {
  const int kt = tile_cols(k, tile);
  if (lanes * kt > 64) lanes = 64 / kt;
  const int lpr = lanes * kt;
  return rows_per_lane(lpr) * 64 / lpr;
}

int launch_lbc(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
               const double *d_Lx, double *d_x, const double *d_b, int level_no,
               const int *h_level_ptr, const int *d_par_ptr,
               const int *d_partition) {
  (void)n;
  for (int l = 0; l < level_no; ++l) {
    const int p0 = h_level_ptr[l], p1 = h_level_ptr[l + 1];
    if (p1 <= p0) continue;
    if (d_b)
      k_lbc_redundant_level<<<(p1 - p0 + 255) / 256, 256, 0, h->stream>>>(
          d_Lp, d_Li, d_Lx, d_x, d_b, d_par_ptr, d_partition, p0, p1);
    else
      k_lbc_level<<<(p1 - p0 + 255) / 256, 256, 0, h->stream>>>(
          d_Lp, d_Li, d_Lx, d_x, d_par_ptr, d_partition, p0, p1);
  }
  SYMAMD_HIP(hipGetLastError());
  return SYMAMD_OK;
}

int launch_forward(symamd_handle h, symamd_plan p, void *d_B, int k, int64_t rs,
                   int64_t cs, int exec, int lanes, int tile) {
  if (p->n == 0 || k == 0) return SYMAMD_OK;
  const int kt = tile_cols(k, tile);
  if (lanes * kt > 64) lanes = 64 / kt;
  if (p->dtype == SYMAMD_F64)
    return fwd_dispatch<double>(h, p, forward_args<double>(h, p, d_B, k, rs, cs), exec,
                                lanes, kt);
  return fwd_dispatch<float>(h, p, forward_args<float>(h, p, d_B, k, rs, cs), exec,
                             lanes, kt);
}

int launch_adjoint(symamd_handle h, symamd_plan p, const void *d_x,
                   const void *d_gx, void *d_gb, void *d_gL, int exec,
                   int lanes) {
  if (p->n == 0) return SYMAMD_OK;
  SYMAMD_TRY(plan_transpose(h, p));
  if (p->dtype == SYMAMD_F64)
    return pick_g<double, 1, true>(h, p, adjoint_args<double>(h, p, d_x, d_gx, d_gb, d_gL),
                                   exec, lanes);
  return pick_g<float, 1, true>(h, p, adjoint_args<float>(h, p, d_x, d_gx, d_gb, d_gL),
                                exec, lanes);
}

}  // namespace symamd
