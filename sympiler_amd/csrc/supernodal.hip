// supernodal.hip -- forward solve on a supernodal (BCSC) lower triangle for
// gfx950, bit-identical to the reference's blocked solve.
//
// Reference: sparse_blas/sptrsv.cpp:176-199 (sptrsv_csc_block) with the dense
// kernels of sparse_blas/dense_blas/BLAS.cpp: lsolve_BLAS (:3-96) and
// matvec_BLAS (:286-349); sparse_blas/sptrsv.cpp:239-279
// (sptrsv_csc_levelset_block) is the same per-supernode routine under a level
// schedule, scattering with atomics.
//
// Layout (the reference's): supernode i owns columns s_i .. s_i + w_i - 1 and
// nrows_i block rows; its values are a column-major nrows_i x w_i block at
// Lx[Lp[i]] whose first w_i rows are its own columns and whose row r is row
// Li[Lp[i] + r] of L.  Per supernode the reference solves the dense top
// (column blocks of 8, then 4, then 2, then 1; inside a block x_j = (b_j -
// x_0 m_j0 - ... - x_{j-1} m_j,j-1) / m_jj left to right, then every later
// top row subtracts the block's products left to right), multiplies the rows
// below by the solved part (column blocks of 8, then 4, then 1; each block's
// products summed left to right and added to a zeroed temporary), and
// subtracts the temporary from x at those rows, supernode after supernode.
//
// Executor: one wavefront per task, tasks taken by ticket in order.  A
// supernode's first task solves its top and multiplies the first SN_CHUNK
// rows below; each further SN_CHUNK rows below are a task of their own that
// reads the solved top from memory, so the rows below a wide supernode are
// multiplied by several waves.  A task only waits for earlier tasks (the
// producers of its rows' temporaries, or its supernode's top), which running
// waves already hold.  The subtraction into x is pulled
// instead of pushed: each supernode writes its temporary to a buffer T and
// publishes a tag; a supernode starts by waiting for the supernodes that
// write to its own rows, then forms x at each own row as b minus their
// temporaries in ascending supernode order -- the order the reference's
// serial loop subtracts them in -- so x is bit-identical.  The top block is
// solved in the wave's LDS: lane j of a column block holds row j's entries,
// x_j is formed on lane j and broadcast; the rows below are one lane each.
#include <algorithm>
#include <vector>

#include "device.h"

namespace symamd {

namespace {

struct SnArgs {
  const int *sup, *nrows, *Lp;  // the reference's supernode arrays
  const double *Lx;
  double *x;                    // right-hand side in, solution out
  double *T;                    // temporaries: supernode i's at T[tb[i] ..]
  const int *tb;                // nrows_i - w_i values each
  const int *cptr, *cidx;       // row r subtracts T[cidx[cptr[r] .. cptr[r+1])]
  const int2 *task;             // task t: (supernode, chunk of its rows below)
  const int *dptr, *dlist;      // tasks that task t waits for
  int *flags, *state;           // one tag per (column, task)
  int ntasks, total_waves;
  // transposed solve: rows below each supernode (at its temporaries'
  // offsets tb), and the supernodes owning them that it waits for
  const int *rowb, *tdptr, *tdlist;
  const int *torder;            // supernode of each transposed ticket
  int m;
  int k;                        // right-hand sides: column c at x + c*ldx
  int64_t ldx, tstride;         // its temporaries at T + c*tstride
  uint64_t timeout;
};

// Column block at `col` of a top of width w: 8 while 8 fit with one to
// spare (matvec) or without (lsolve), else 4, else 2 (lsolve) or 1.
__device__ __forceinline__ int lsolve_block(int col, int w) {
  return col < w - 7 ? 8 : col < w - 3 ? 4 : col < w - 1 ? 2 : 1;
}
__device__ __forceinline__ int matvec_block(int col, int w) {
  return col < w - 7 ? 8 : col < w - 3 ? 4 : 1;
}

// Rows per lane processed together in the top's updates and the product
// below (their loads in flight at once).
constexpr int SN_ROWS = 2;
static_assert(SN_CHUNK == 64 * SN_ROWS, "a chunk of rows below is one pass of a wave");
// TOP_ROWS: rows per lane per pass of the top's update after each column
// block.  Each pass is a memory round trip on the chain from one block to
// the next; four rows per pass take more registers (two waves per SIMD
// instead of three), which pays for one right-hand side, where the chain of
// the widest supernodes sets the time, and not for blocks, where the (task,
// column) tickets keep every wave busy.

// A supernode's own rows: xs[r] = x[s+r] minus the temporaries listed for
// row s+r, in list (ascending supernode) order, for rows r0 <= r < w.  Lane l
// takes rows r0 + l + 64q, q < NQ, and its rows advance together, GB = 16/NQ
// list entries each per round (32 slots per lane measured slower: most of the
// extra slots are stand-ins): a round's temporaries are loaded for all of its
// rows at once (their indices were loaded in the round before), then
// subtracted row by row in list order.  So a lane pays one memory round
// trip per GB entries of its longest row, not per eight entries of each row
// in turn.  No load is behind a branch: list positions past a row's end are
// clamped into the index array (which holds at least one entry), and their
// temporaries are replaced by the row's own x (rows past w by row w-1's),
// loaded and not used; whether a slot is used depends only on the list
// bounds, so no value is waited for before its use.
template <int NQ>
__device__ __forceinline__ void gather_rows(double *xs, const double *x, const double *T,
                                            const int *cptr, const int *cidx, int s, int w,
                                            int r0, int lane) {
  static_assert(NQ >= 1 && NQ <= 8, "at most 512 rows per call");
  constexpr int GB = 16 / NQ;
  if (w <= r0) return;
  double v[NQ];
  int cp[NQ], ce[NQ], idx[NQ][GB];
  auto row = [&](int q) { return min(r0 + lane + 64 * q, w - 1); };
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const bool in = r0 + lane + 64 * q < w;
    v[q] = x[s + row(q)];
    const int c0 = cptr[s + row(q)], c1 = cptr[s + row(q) + 1];
    cp[q] = in ? c0 : 0;
    ce[q] = in ? c1 : 0;
  }
  auto load_idx = [&](int q, int c, int (&d)[GB]) {
#pragma unroll
    for (int u = 0; u < GB; ++u) d[u] = cidx[max(min(c + u, ce[q] - 1), 0)];
  };
  bool more = false;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    load_idx(q, cp[q], idx[q]);
    more = more || cp[q] < ce[q];
  }
  while (more) {
    double tv[NQ][GB];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int u = 0; u < GB; ++u)
        tv[q][u] = ld_dev(cp[q] + u < ce[q] ? &T[idx[q][u]] : &x[s + row(q)]);
    int nidx[NQ][GB];
    more = false;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      load_idx(q, cp[q] + GB, nidx[q]);
      more = more || cp[q] + GB < ce[q];
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
#pragma unroll
      for (int u = 0; u < GB; ++u)
        if (cp[q] + u < ce[q]) v[q] -= tv[q][u];
      cp[q] += GB;
#pragma unroll
      for (int u = 0; u < GB; ++u) idx[q][u] = nidx[q][u];
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q)
    if (r0 + lane + 64 * q < w) xs[r0 + lane + 64 * q] = v[q];
}

// NB full column blocks of eight from col: each block's products with the
// solved top summed left to right and added to the row's temporary, for the
// SN_ROWS rows k0 + 64u (rows past nr read row nr-1; their sums are unused)
template <int NB>
__device__ __forceinline__ void product_blocks(double (&t)[SN_ROWS], const double *xs,
                                               const double *M, int k0, int nr, int col) {
  double mb[NB][SN_ROWS][8];
#pragma unroll
  for (int h = 0; h < NB; ++h)
#pragma unroll
    for (int u = 0; u < SN_ROWS; ++u) {
      const int k = k0 + 64 * u;
#pragma unroll
      for (int c = 0; c < 8; ++c) mb[h][u][c] = M[min(k, nr - 1) + (int64_t)(col + 8 * h + c) * nr];
    }
#pragma unroll
  for (int h = 0; h < NB; ++h)
#pragma unroll
    for (int u = 0; u < SN_ROWS; ++u) {
      double sum = xs[col + 8 * h] * mb[h][u][0];
#pragma unroll
      for (int c = 1; c < 8; ++c) sum = sum + xs[col + 8 * h + c] * mb[h][u][c];
      t[u] = t[u] + sum;
    }
}

// v on lane j (j uniform), for every lane
__device__ __forceinline__ double lane_value(double v, int j) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, j);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), j);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <int TOP_ROWS>
__global__ __launch_bounds__(256, TOP_ROWS >= 4 ? 2 : 3) void k_supernodal(SnArgs a) {
  __shared__ double slab[4][SN_WMAX];
  double *xs = slab[threadIdx.x >> 6];
  const int lane = threadIdx.x & 63;
  const int target = ld_dev(&a.state[ST_EPOCH]) + 1;
  for (;;) {
    // ticket -> (task, column), task-major: the k columns of one task run
    // side by side (sharing its block in the caches) and every wait is for
    // an earlier ticket
    int tt = 0;
    if (lane == 0) tt = atomicAdd(&a.state[ST_TICKET], 1);
    tt = __shfl(tt, 0);
    if (tt >= a.ntasks * a.k) break;
    const int tk = tt / a.k, colk = tt - tk * a.k;
    double *const x = a.x + colk * a.ldx;
    double *const T = a.T + colk * a.tstride;
    int *const flags = a.flags + (int64_t)colk * a.ntasks;
    const int2 td = a.task[tk];
    const int i = __builtin_amdgcn_readfirstlane(td.x);
    const int chunk = __builtin_amdgcn_readfirstlane(td.y);
    const int s = a.sup[i], w = a.sup[i + 1] - s, nr = a.nrows[i];
    const double *M = a.Lx + a.Lp[i];
    // 1. the tasks that write to this task's inputs (for a later chunk of the
    // rows below: the supernode's own top)
    wait_units(a.dptr, a.dlist, tk, flags, target, a.state, a.timeout);
    if (chunk > 0) {
      // a later chunk of the rows below: the solved top from memory
      for (int r = lane; r < w; r += 64) xs[r] = ld_dev(&x[s + r]);
      wave_lds_sync();
    } else {
      // 2. own rows: b minus the temporaries, in ascending supernode order
      if (w <= 64)
        gather_rows<1>(xs, x, T, a.cptr, a.cidx, s, w, 0, lane);
      else if (w <= 128)
        gather_rows<2>(xs, x, T, a.cptr, a.cidx, s, w, 0, lane);
      else if (w <= 256)
        gather_rows<4>(xs, x, T, a.cptr, a.cidx, s, w, 0, lane);
      else
        for (int r0 = 0; r0 < w; r0 += 512)
          gather_rows<8>(xs, x, T, a.cptr, a.cidx, s, w, r0, lane);
      wave_lds_sync();
      // 3. the dense top, column block by column block.  Block (col, B):
      // lane j < B holds row col+j of the block's columns (and its diagonal)
      // in mr, and mk the first TOP_ROWS rows per lane of the update after
      // the block (B = 0: past the last block, nothing loaded).  The entries
      // do not depend on the solve: with PF the next block's are loaded
      // while this block is solved.
      constexpr bool PF = TOP_ROWS >= 4;
      // Addresses outside the block are clamped into the top, so the loads
      // need no branches; the values so loaded are never used (a lane's
      // triangle reads its own row up to its diagonal, the update skips
      // columns past B and rows past w).  B = 0 loads the top's last entry.
      auto load_block = [&](int col, int B, double (&mr)[8], double (&mk)[TOP_ROWS][8]) {
        const int rl = min(lane, B - 1);
#pragma unroll
        for (int c = 0; c < 8; ++c) mr[c] = M[(col + rl) + (int64_t)(col + min(c, rl)) * nr];
        const int k0 = col + B + lane;
#pragma unroll
        for (int u = 0; u < TOP_ROWS; ++u) {
          const int k = k0 + 64 * u;
#pragma unroll
          for (int c = 0; c < 8; ++c)
            mk[u][c] = M[min(k, w - 1) + (int64_t)(col + min(c, B - 1)) * nr];
        }
      };
      double mrow[8], mk[TOP_ROWS][8];
      int B = w > 0 ? lsolve_block(0, w) : 0;
      if constexpr (PF) load_block(0, B, mrow, mk);
      for (int col = 0; col < w;) {
        const int cn = col + B;
        const int Bn = cn < w ? lsolve_block(cn, w) : 0;
        double nrow[8], nk[TOP_ROWS][8];
        if constexpr (PF)
          load_block(cn, Bn, nrow, nk);
        else
          load_block(col, B, mrow, mk);
        const int k0 = col + B + lane;
        // the block's triangle: lane l < B keeps its row's running value;
        // at step j lane j divides, the quotient is broadcast, and every
        // lane subtracts its product (each row loses its products in column
        // order, as the reference computes them)
        double acc = lane < B ? xs[col + lane] : 0.0, mine = 0.0;
        double xv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          if (j < B) {
            const double q = acc / mrow[j];
            mine = lane == j ? q : mine;
            xv[j] = lane_value(q, j);
            acc = acc - xv[j] * mrow[j];
          } else {
            xv[j] = 0.0;
          }
        }
        wave_lds_sync();
        if (lane < B) xs[col + lane] = mine;
        // TOP_ROWS rows per lane at a time: all their entries in flight at once
        for (int kc = k0; kc < w; kc += 64 * TOP_ROWS) {
          if (kc != k0) {
#pragma unroll
            for (int u = 0; u < TOP_ROWS; ++u) {
              const int k = kc + 64 * u;
#pragma unroll
              for (int c = 0; c < 8; ++c)
                mk[u][c] = M[min(k, w - 1) + (int64_t)(col + min(c, B - 1)) * nr];
            }
          }
#pragma unroll
          for (int u = 0; u < TOP_ROWS; ++u) {
            const int k = kc + 64 * u;
            if (k < w) {
              double t = xs[k];
#pragma unroll
              for (int c = 0; c < 8; ++c)
                if (c < B) t = t - xv[c] * mk[u][c];
              xs[k] = t;
            }
          }
        }
        wave_lds_sync();
        if constexpr (PF) {
#pragma unroll
          for (int c = 0; c < 8; ++c) mrow[c] = nrow[c];
#pragma unroll
          for (int u = 0; u < TOP_ROWS; ++u)
#pragma unroll
            for (int c = 0; c < 8; ++c) mk[u][c] = nk[u][c];
        }
        col = cn;
        B = Bn;
      }
      for (int r = lane; r < w; r += 64) st_dev(&x[s + r], xs[r]);
    }
    // 4. this task's chunk of the rows below: products with the solved top
    // into T; SN_ROWS rows per lane, each column block's entries of those
    // rows loaded together, each row's sum kept in the reference's order.
    // Lanes past the last row load row nr-1 (no branch per load) and store
    // nothing; the tail skips the columns past its block.
    const int tb = a.tb[i];
    {
      const int k0 = w + chunk * SN_CHUNK + lane;
      double t[SN_ROWS];
#pragma unroll
      for (int u = 0; u < SN_ROWS; ++u) t[u] = 0.0;
      // full blocks of eight, WIDE (one right-hand side: 4, where the
      // registers allow; else 2) and then 2 at a time: all their entries in
      // flight together, each block's sum formed and added as the reference
      // orders it
      constexpr int WIDE = TOP_ROWS >= 4 ? 4 : 2;
      int col = 0;
      for (; w - col >= 8 * WIDE; col += 8 * WIDE) product_blocks<WIDE>(t, xs, M, k0, nr, col);
      for (; w - col >= 16; col += 16) product_blocks<2>(t, xs, M, k0, nr, col);
      // the last columns in blocks of 8, 4 and 1
      while (col < w) {
        const int B = matvec_block(col, w);
        double mk[SN_ROWS][8];
#pragma unroll
        for (int u = 0; u < SN_ROWS; ++u) {
          const int k = k0 + 64 * u;
#pragma unroll
          for (int c = 0; c < 8; ++c)
            mk[u][c] = M[min(k, nr - 1) + (int64_t)(col + min(c, B - 1)) * nr];
        }
#pragma unroll
        for (int u = 0; u < SN_ROWS; ++u) {
          double sum = xs[col] * mk[u][0];
#pragma unroll
          for (int c = 1; c < 8; ++c)
            if (c < B) sum = sum + xs[col + c] * mk[u][c];
          t[u] = t[u] + sum;
        }
        col += B;
      }
#pragma unroll
      for (int u = 0; u < SN_ROWS; ++u)
        if (k0 + 64 * u < nr) st_dev(&T[tb + k0 + 64 * u - w], t[u]);
    }
    // 5. publish
    publish_fence();
    if (lane == 0) st_dev(&flags[tk], target);
    wave_lds_sync();
  }
  persistent_finish(a.state, a.total_waves);
}


// ---------------------------------------------------- transposed solve
// L^T x = b on the same storage (the semantics of parsy's blockedLTsolve,
// sparse_blas/parsy/triangularSolve/Triangular_BCSC.cpp:125-186), supernodes
// level by level from the root (torder); one wave per (supernode, column)
// ticket, supernode-major, so every wait is for an earlier ticket.  Per supernode:
// wait for the supernodes owning its rows below (solved earlier); per top
// column c (lanes), t = sum over the rows below, ascending, of M[r, c] times
// x at that row (the rows' x staged in LDS chunk by chunk), x[s+c] -= t;
// then the top's transposed solve in LDS from the last column: x_c /= M[c, c]
// and every x_r, r < c, loses M[c, r] * x_c.  The order is the oracle's
// (orc_sptrsv_csc_block_t).
//
// NT column tiles of 64 per lane (NT * 64 >= w).  The matrix entries do
// not depend on the solve, so their loads are batched: RB rows below per
// round (RB * NT loads in flight per lane), and the top's columns in groups
// of D whose entries load while the group before is finished.  Addresses
// are clamped into the block (no branch per load); the values so loaded are
// not used.
constexpr int SN_TCHUNK = 256;  // rows below staged per pass

template <int NT>
__device__ __forceinline__ void sn_solve_t_top(double *xs, double *xb, const double *x,
                                               const double *M, const int *rows, int s,
                                               int w, int nr, int lane) {
  constexpr int RB = 32 / NT;           // rows below per round
  constexpr int D = NT >= 8 ? 32 / NT : 8;  // top columns per group
  int cl[NT];
#pragma unroll
  for (int j = 0; j < NT; ++j) cl[j] = min(lane + 64 * j, w - 1);
  // 1. the rows below: one running sum per top column, rows ascending
  double t[NT];
#pragma unroll
  for (int j = 0; j < NT; ++j) t[j] = 0.0;
  for (int r0 = w; r0 < nr; r0 += SN_TCHUNK) {
    const int cnt = min(SN_TCHUNK, nr - r0);
    for (int r = lane; r < cnt; r += 64) xb[r] = ld_dev(&x[rows[r0 - w + r]]);
    wave_lds_sync();
    for (int rb = 0; rb < cnt; rb += RB) {
      double mv[RB][NT];
#pragma unroll
      for (int u = 0; u < RB; ++u) {
        const int row = r0 + min(rb + u, cnt - 1);
#pragma unroll
        for (int j = 0; j < NT; ++j) mv[u][j] = M[row + (int64_t)cl[j] * nr];
      }
#pragma unroll
      for (int u = 0; u < RB; ++u) {
        if (rb + u < cnt) {
          const double v = xb[rb + u];
          const bool first = r0 + rb + u == w;
#pragma unroll
          for (int j = 0; j < NT; ++j) {
            const double p = mv[u][j] * v;
            t[j] = first ? p : t[j] + p;
          }
        }
      }
    }
    wave_lds_sync();
  }
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const int c = lane + 64 * j;
    if (c < w) xs[c] = nr > w ? x[s + c] - t[j] : x[s + c];
  }
  wave_lds_sync();
  // 2. the top, transposed, from the last column, D columns per group
  auto load_group = [&](int cb, double (&m)[D][NT], double (&dg)[D]) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const int c = max(cb - k, 0);
#pragma unroll
      for (int j = 0; j < NT; ++j) m[k][j] = M[c + (int64_t)cl[j] * nr];
      dg[k] = M[c + (int64_t)c * nr];
    }
  };
  double mc[D][NT], dc[D];
  load_group(w - 1, mc, dc);
  for (int cb = w - 1; cb >= 0; cb -= D) {
    double mn[D][NT], dn[D];
    load_group(cb - D, mn, dn);
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const int c = cb - k;
      if (c >= 0) {
        const double y = xs[c] / dc[k];
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          const int r = lane + 64 * j;
          if (r < c) xs[r] = xs[r] - mc[k][j] * y;
        }
        if (lane == (c & 63)) xs[c] = y;
        wave_lds_sync();
      }
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
      dc[k] = dn[k];
#pragma unroll
      for (int j = 0; j < NT; ++j) mc[k][j] = mn[k][j];
    }
  }
}

__global__ __launch_bounds__(256) void k_supernodal_t(SnArgs a) {
  __shared__ double slab[4][SN_WMAX];
  __shared__ double below[4][SN_TCHUNK];
  double *xs = slab[threadIdx.x >> 6];
  double *xb = below[threadIdx.x >> 6];
  const int lane = threadIdx.x & 63;
  const int target = ld_dev(&a.state[ST_EPOCH]) + 1;
  for (;;) {
    int tt = 0;
    if (lane == 0) tt = atomicAdd(&a.state[ST_TICKET], 1);
    tt = __shfl(tt, 0);
    if (tt >= a.m * a.k) break;
    const int ti = tt / a.k, colk = tt - ti * a.k;
    const int i = a.torder[ti];
    double *const x = a.x + colk * a.ldx;
    int *const flags = a.flags + (int64_t)colk * a.ntasks;
    wait_units(a.tdptr, a.tdlist, i, flags, target, a.state, a.timeout);
    const int s = a.sup[i], w = a.sup[i + 1] - s, nr = a.nrows[i];
    const double *M = a.Lx + a.Lp[i];
    const int *rows = a.rowb + a.tb[i];
    if (w <= 64)
      sn_solve_t_top<1>(xs, xb, x, M, rows, s, w, nr, lane);
    else if (w <= 128)
      sn_solve_t_top<2>(xs, xb, x, M, rows, s, w, nr, lane);
    else if (w <= 256)
      sn_solve_t_top<4>(xs, xb, x, M, rows, s, w, nr, lane);
    else if (w <= 512)
      sn_solve_t_top<8>(xs, xb, x, M, rows, s, w, nr, lane);
    else
      sn_solve_t_top<16>(xs, xb, x, M, rows, s, w, nr, lane);
    for (int r = lane; r < w; r += 64) st_dev(&x[s + r], xs[r]);
    publish_fence();
    if (lane == 0) st_dev(&flags[i], target);
    wave_lds_sync();
  }
  persistent_finish(a.state, a.total_waves);
}

}  // namespace

int sn_analyze(symamd_handle h, int n, int m, const int *h_sup, const int *h_nrows,
               const int *h_Lp, const int *h_Li, symamd_snplan_s **out) {
  if (n < 0 || m < 0 || (m > 0 && (!h_sup || !h_nrows || !h_Lp || !h_Li)))
    return fail(SYMAMD_ERR_INVALID_ARG, "supernodal analyze: bad arguments");
  if (m == 0 && n > 0)
    return fail(SYMAMD_ERR_STRUCTURE, "supernodal analyze: no supernodes for %d columns", n);
  if (m > 0 && (h_sup[0] != 0 || h_sup[m] != n))
    return fail(SYMAMD_ERR_STRUCTURE, "supernodes must cover columns 0 .. n-1");
  std::vector<int> tb(m + 1, 0), cnt(n + 1, 0);
  for (int i = 0; i < m; ++i) {
    const int s = h_sup[i], w = h_sup[i + 1] - s, nr = h_nrows[i];
    if (w < 1 || nr < w)
      return fail(SYMAMD_ERR_STRUCTURE, "supernode %d: width %d, %d rows", i, w, nr);
    if (w > SN_WMAX)
      return fail(SYMAMD_ERR_NOT_SUPPORTED, "supernode %d is %d columns wide (at most %d)", i,
                  w, SN_WMAX);
    if (i > 0 && h_Lp[i] < h_Lp[i - 1] + (int64_t)h_nrows[i - 1] * (h_sup[i] - h_sup[i - 1]))
      return fail(SYMAMD_ERR_STRUCTURE, "supernode %d's values overlap the previous block", i);
    // the rows below the top must lie below the supernode (the top's own
    // row ids are not read, as in the reference)
    for (int r = w; r < nr; ++r) {
      const int row = h_Li[h_Lp[i] + r];
      if (row < s + w || row >= n)
        return fail(SYMAMD_ERR_STRUCTURE, "supernode %d: block row %d is row %d", i, r, row);
      cnt[row + 1]++;
    }
    tb[i + 1] = tb[i] + (nr - w);
  }
  for (int r = 0; r < n; ++r) cnt[r + 1] += cnt[r];
  // contributions per row in ascending supernode order; dependency lists
  std::vector<int> cidx(cnt[n] > 0 ? cnt[n] : 1), fill(cnt.begin(), cnt.end() - 1);
  std::vector<int> owner(n);
  for (int i = 0; i < m; ++i)
    for (int c = h_sup[i]; c < h_sup[i + 1]; ++c) owner[c] = i;
  // tasks: supernode i's top (with the first SN_CHUNK rows below), then one
  // task per further SN_CHUNK rows below.  Tickets are dealt level by level
  // (a supernode's level is one above the highest of the supernodes writing
  // to its rows), so every leaf is taken before any supernode that waits,
  // and the chain of the last subtree does not start behind the others;
  // every task still waits only for tasks dealt before it.
  std::vector<int> lev(m, 0);
  for (int i = 0; i < m; ++i) {
    const int w = h_sup[i + 1] - h_sup[i];
    for (int r = w; r < h_nrows[i]; ++r) {
      const int o = owner[h_Li[h_Lp[i] + r]];
      lev[o] = std::max(lev[o], lev[i] + 1);
    }
  }
  std::vector<int> order(m);
  for (int i = 0; i < m; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return lev[x] < lev[y]; });
  std::vector<int> first(m + 1, 0), ntask(m);  // first[i]: ticket of i's top
  int ntasks = 0;
  for (int i : order) {
    const int below = h_nrows[i] - (h_sup[i + 1] - h_sup[i]);
    ntask[i] = std::max(1, (below + SN_CHUNK - 1) / SN_CHUNK);
    first[i] = ntasks;
    ntasks += ntask[i];
  }
  std::vector<int2> task(ntasks > 0 ? ntasks : 1);
  for (int i = 0; i < m; ++i)
    for (int c = 0; c < ntask[i]; ++c) task[first[i] + c] = make_int2(i, c);
  // a top task waits for the tasks writing its rows' temporaries, a later
  // chunk for its supernode's top
  std::vector<std::vector<int>> deps(ntasks);
  for (int i = 0; i < m; ++i) {
    const int w = h_sup[i + 1] - h_sup[i];
    for (int k = 0; k < h_nrows[i] - w; ++k) {
      const int row = h_Li[h_Lp[i] + w + k];
      cidx[fill[row]++] = tb[i] + k;
      std::vector<int> &d = deps[first[owner[row]]];
      const int producer = first[i] + k / SN_CHUNK;
      if (d.empty() || d.back() != producer) d.push_back(producer);
    }
    for (int t = first[i] + 1; t < first[i] + ntask[i]; ++t) deps[t].push_back(first[i]);
  }
  // transposed solve: the rows below of each supernode (in temporaries
  // order) and the supernodes owning them, which it waits for
  std::vector<int> rowb(tb[m] > 0 ? tb[m] : 1), tdptr(m + 1, 0), tdlist;
  for (int i = 0; i < m; ++i) {
    const int w = h_sup[i + 1] - h_sup[i];
    std::vector<int> own;
    for (int k = 0; k < h_nrows[i] - w; ++k) {
      const int row = h_Li[h_Lp[i] + w + k];
      rowb[tb[i] + k] = row;
      own.push_back(owner[row]);
    }
    std::sort(own.begin(), own.end());
    own.erase(std::unique(own.begin(), own.end()), own.end());
    tdlist.insert(tdlist.end(), own.begin(), own.end());
    tdptr[i + 1] = (int)tdlist.size();
  }
  // transposed tickets level by level from the root: a supernode's level is
  // one below the deepest of the supernodes owning its rows below (all of
  // higher index), so each waits only for tickets dealt before it, and no
  // subtree's waiting supernodes hold the waves another chain needs
  std::vector<int> tlev(m, 0), torder(m > 0 ? m : 1);
  for (int i = m - 1; i >= 0; --i)
    for (int e = tdptr[i]; e < tdptr[i + 1]; ++e) tlev[i] = std::max(tlev[i], tlev[tdlist[e]] + 1);
  for (int i = 0; i < m; ++i) torder[i] = m - 1 - i;
  std::stable_sort(torder.begin(), torder.begin() + m,
                   [&](int x, int y) { return tlev[x] < tlev[y]; });
  std::vector<int> dptr(ntasks + 1, 0), dlist;
  for (int j = 0; j < ntasks; ++j) {
    std::sort(deps[j].begin(), deps[j].end());
    deps[j].erase(std::unique(deps[j].begin(), deps[j].end()), deps[j].end());
    dlist.insert(dlist.end(), deps[j].begin(), deps[j].end());
    dptr[j + 1] = (int)dlist.size();
  }
  auto *q = new symamd_snplan_s();
  q->n = n;
  q->m = m;
  q->tcount = tb[m];
  q->ntasks = ntasks;
  int rc = SYMAMD_OK;
  auto up = [&](int **dst, const int *src, size_t count) {
    if (rc != SYMAMD_OK) return;
    if (hipMalloc(dst, sizeof(int) * (count ? count : 1)) != hipSuccess ||
        (count && hipMemcpy(*dst, src, sizeof(int) * count, hipMemcpyHostToDevice) != hipSuccess))
      rc = fail(SYMAMD_ERR_ALLOC, "supernodal plan upload failed");
  };
  up(&q->d_sup, h_sup, m + 1);
  up(&q->d_nrows, h_nrows, m);
  up(&q->d_Lp, h_Lp, m + 1);
  up(&q->d_tb, tb.data(), m + 1);
  up(&q->d_cptr, cnt.data(), n + 1);
  up(&q->d_cidx, cidx.data(), cnt[n]);
  up(&q->d_dptr, dptr.data(), ntasks + 1);
  up(&q->d_dlist, dlist.data(), dlist.size());
  up(&q->d_task, reinterpret_cast<const int *>(task.data()), 2 * (size_t)ntasks);
  up(&q->d_rowb, rowb.data(), tb[m]);
  up(&q->d_tdptr, tdptr.data(), m + 1);
  up(&q->d_tdlist, tdlist.data(), tdlist.size());
  up(&q->d_torder, torder.data(), m);
  if (rc == SYMAMD_OK &&
      (hipMalloc(&q->d_T, sizeof(double) * (q->tcount > 0 ? q->tcount : 1)) != hipSuccess ||
       hipMalloc(&q->d_flags, sizeof(int) * (ntasks > 0 ? ntasks : 1)) != hipSuccess ||
       hipMalloc(&q->d_state, sizeof(int) * ST_WORDS) != hipSuccess ||
       hipMemset(q->d_flags, 0, sizeof(int) * (ntasks > 0 ? ntasks : 1)) != hipSuccess ||
       hipMemset(q->d_state, 0, sizeof(int) * ST_WORDS) != hipSuccess))
    rc = fail(SYMAMD_ERR_ALLOC, "supernodal plan allocation failed");
  if (rc != SYMAMD_OK) {
    sn_destroy(q);
    return rc;
  }
  q->deps = (int64_t)dlist.size();
  (void)h;
  *out = q;
  return SYMAMD_OK;
}

int sn_solve(symamd_handle h, symamd_snplan_s *q, const double *d_Lx, double *d_x, int k,
             int64_t ldx, bool transposed) {
  if (q->m == 0 || k == 0) return SYMAMD_OK;
  // tickets run over (task, column) pairs in 32-bit counters
  if ((int64_t)q->ntasks * k >= (int64_t)INT_MAX)
    return fail(SYMAMD_ERR_NOT_SUPPORTED,
                "supernodal solve: %d tasks x %d columns exceed the 32-bit ticket range; "
                "solve the columns in smaller batches", q->ntasks, k);
  // temporaries and tags for k columns
  if (k > q->kcap) {
    double *T = nullptr;
    int *flags = nullptr;
    const int64_t tc = q->tcount > 0 ? q->tcount : 1;
    const int64_t nt = q->ntasks > 0 ? q->ntasks : 1;
    if (hipMalloc(&T, sizeof(double) * tc * k) != hipSuccess ||
        hipMalloc(&flags, sizeof(int) * nt * k) != hipSuccess) {
      if (T) (void)hipFree(T);
      return fail(SYMAMD_ERR_ALLOC, "supernodal temporaries for %d columns", k);
    }
    // cleared tags read as "not ready" in every epoch
    SYMAMD_HIP(hipMemsetAsync(flags, 0, sizeof(int) * nt * k, h->stream));
    SYMAMD_HIP(hipStreamSynchronize(h->stream));
    if (q->d_T) (void)hipFree(q->d_T);
    if (q->d_flags) (void)hipFree(q->d_flags);
    q->d_T = T;
    q->d_flags = flags;
    q->kcap = k;
  }
  SnArgs a;
  a.sup = q->d_sup;
  a.nrows = q->d_nrows;
  a.Lp = q->d_Lp;
  a.Lx = d_Lx;
  a.x = d_x;
  a.T = q->d_T;
  a.tb = q->d_tb;
  a.cptr = q->d_cptr;
  a.cidx = q->d_cidx;
  a.dptr = q->d_dptr;
  a.dlist = q->d_dlist;
  a.task = reinterpret_cast<const int2 *>(q->d_task);
  a.rowb = q->d_rowb;
  a.tdptr = q->d_tdptr;
  a.tdlist = q->d_tdlist;
  a.torder = q->d_torder;
  a.m = q->m;
  a.flags = q->d_flags;
  a.state = q->d_state;
  a.ntasks = q->ntasks;
  a.k = k;
  a.ldx = ldx;
  a.tstride = q->tcount > 0 ? q->tcount : 1;
  a.timeout = h->timeout_ticks();
  const void *kfn = transposed ? (const void *)&k_supernodal_t
                    : k == 1   ? (const void *)&k_supernodal<4>
                               : (const void *)&k_supernodal<2>;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, 256, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  int64_t grid = (int64_t)per_cu * h->num_cus;
  const int64_t want = ((int64_t)(transposed ? q->m : q->ntasks) * k + 3) / 4;
  if (want < grid) grid = want;
  if (grid < 1) grid = 1;
  a.total_waves = (int)grid * 4;
  // tickets are dealt in task order, so any grid size is safe
  if (transposed)
    k_supernodal_t<<<(unsigned)grid, 256, 0, h->stream>>>(a);
  else if (k == 1)
    k_supernodal<4><<<(unsigned)grid, 256, 0, h->stream>>>(a);
  else
    k_supernodal<2><<<(unsigned)grid, 256, 0, h->stream>>>(a);
  SYMAMD_HIP(hipGetLastError());
  return SYMAMD_OK;
}

int sn_check(symamd_handle h, symamd_snplan_s *q) {
  int st[ST_WORDS];
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  SYMAMD_HIP(hipMemcpy(st, q->d_state, sizeof(st), hipMemcpyDeviceToHost));
  if (st[ST_ERROR]) return fail(SYMAMD_ERR_TIMEOUT, "a supernodal wait exceeded the watchdog");
  return SYMAMD_OK;
}

void sn_destroy(symamd_snplan_s *q) {
  if (!q) return;
  void *ptrs[] = {q->d_sup,   q->d_nrows, q->d_Lp,    q->d_tb,     q->d_cptr,
                  q->d_cidx,  q->d_dptr,  q->d_dlist, q->d_T,      q->d_flags,
                  q->d_state, q->d_task,  q->d_rowb,  q->d_tdptr, q->d_tdlist,
                  q->d_torder};
  for (void *x : ptrs)
    if (x) (void)hipFree(x);
  delete q;
}

}  // namespace symamd
