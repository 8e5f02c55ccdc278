// regions.cpp -- compact row regions for the owner-scheduled sync-free solve.
//
// The sync-free executor pays one cross-wave hand-off (a tag store seen by
// another compute unit, ~1.5-2 us on MI355X) per dependent step.  When the
// rows of a compact region of the dependency graph are solved by ONE wave,
// level after level, the dependencies inside the region need no hand-off at
// all; only edges that leave the region do.  This file finds such regions.
//
// Method: multilevel heavy-edge matching on the undirected graph of L's
// off-diagonal pattern.  Each round visits the current clusters in order and
// merges every unmatched cluster with its unmatched neighbour of largest
// connecting weight (ties: the smaller id) provided the merged size stays
// within max_rows; the matched pairs are then contracted (weights summed).
// On a grid in natural order the rounds produce 2-, 4-, 8-... row blocks
// alternating between the axes (the heaviest connection after merging along
// one axis is along another), ending in b x b x b cubes.  For irregular
// graphs it is the usual heavy-edge coarsening.  Deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "common.h"

namespace symamd {

namespace {

struct WGraph {
  std::vector<int64_t> ptr;
  std::vector<int> adj;
  std::vector<int> w;
};

// Symmetric adjacency of the strictly lower pattern, weight 1 per entry.
WGraph initial_graph(int n, const int *Lp, const int *Li) {
  WGraph g;
  g.ptr.assign(n + 1, 0);
  for (int i = 0; i < n; ++i)
    for (int j = Lp[i]; j < Lp[i + 1]; ++j) {
      const int c = Li[j];
      if (c == i) continue;
      g.ptr[i + 1]++;
      g.ptr[c + 1]++;
    }
  for (int i = 0; i < n; ++i) g.ptr[i + 1] += g.ptr[i];
  g.adj.resize(g.ptr[n]);
  g.w.assign(g.ptr[n], 1);
  std::vector<int64_t> fill(g.ptr.begin(), g.ptr.end() - 1);
  for (int i = 0; i < n; ++i)
    for (int j = Lp[i]; j < Lp[i + 1]; ++j) {
      const int c = Li[j];
      if (c == i) continue;
      g.adj[fill[i]++] = c;
      g.adj[fill[c]++] = i;
    }
  return g;
}

}  // namespace

int cluster_regions(int n, const int *Lp, const int *Li, int max_rows,
                    std::vector<int> &region) {
  region.resize(n);
  std::iota(region.begin(), region.end(), 0);
  if (n == 0 || max_rows < 2) return n;
  WGraph g = initial_graph(n, Lp, Li);
  std::vector<int> size(n, 1);
  int nc = n;
  std::vector<int> match, newid, mark, slot;
  for (;;) {
    // 1. greedy heavy-edge matching in cluster order
    match.assign(nc, -1);
    int merges = 0;
    for (int a = 0; a < nc; ++a) {
      if (match[a] >= 0) continue;
      int best = -1, bw = 0;
      for (int64_t e = g.ptr[a]; e < g.ptr[a + 1]; ++e) {
        const int b = g.adj[e];
        if (match[b] >= 0 || size[a] + size[b] > max_rows) continue;
        if (g.w[e] > bw || (g.w[e] == bw && b < best)) {
          best = b;
          bw = g.w[e];
        }
      }
      if (best >= 0) {
        match[a] = best;
        match[best] = a;
        ++merges;
      } else {
        match[a] = a;
      }
    }
    if (merges == 0) break;
    // 2. contract: new ids in order of the pair's first cluster
    newid.assign(nc, -1);
    int m = 0;
    for (int a = 0; a < nc; ++a)
      if (newid[a] < 0) {
        newid[a] = m;
        newid[match[a]] = m;
        ++m;
      }
    for (int i = 0; i < n; ++i) region[i] = newid[region[i]];
    std::vector<int> nsize(m, 0);
    for (int a = 0; a < nc; ++a)
      if (match[a] == a || a < match[a]) nsize[newid[a]] = size[a] + (match[a] == a ? 0 : size[match[a]]);
    WGraph h;
    h.ptr.assign(m + 1, 0);
    h.adj.reserve(g.adj.size() / 2 + 16);
    h.w.reserve(g.adj.size() / 2 + 16);
    mark.assign(m, -1);
    slot.assign(m, 0);
    for (int a = 0; a < nc; ++a) {
      if (!(match[a] == a || a < match[a])) continue;
      const int c = newid[a];
      const int parts[2] = {a, match[a]};
      for (int s = 0; s < (match[a] == a ? 1 : 2); ++s) {
        const int x = parts[s];
        for (int64_t e = g.ptr[x]; e < g.ptr[x + 1]; ++e) {
          const int d = newid[g.adj[e]];
          if (d == c) continue;
          if (mark[d] != c) {
            mark[d] = c;
            slot[d] = (int)h.adj.size();
            h.adj.push_back(d);
            h.w.push_back(g.w[e]);
          } else {
            h.w[slot[d]] += g.w[e];
          }
        }
      }
      h.ptr[c + 1] = (int64_t)h.adj.size();
    }
    g.ptr.swap(h.ptr);
    g.adj.swap(h.adj);
    g.w.swap(h.w);
    size.swap(nsize);
    nc = m;
  }
  return nc;
}

namespace {

// f(part, lo, hi) over contiguous ranges of [0, count) on up to 16 host
// threads (one when count is small); parts are numbered in range order.
template <typename F>
int parallel_ranges(int count, F f) {
  int T = (int)std::thread::hardware_concurrency();
  T = std::max(1, std::min({T, 16, count / 64}));
  if (T == 1) {
    f(0, 0, count);
    return 1;
  }
  std::vector<std::thread> th;
  th.reserve(T);
  for (int t = 0; t < T; ++t)
    th.emplace_back(f, t, (int)((int64_t)count * t / T), (int)((int64_t)count * (t + 1) / T));
  for (auto &x : th) x.join();
  return T;
}

// Stable counting sort of `items` by key(item) in [0, nkeys).
template <typename Key>
std::vector<int> counting_sort(const std::vector<int> &items, int nkeys, Key key) {
  std::vector<int> start(nkeys + 1, 0), out(items.size());
  for (int v : items) start[key(v) + 1]++;
  for (int k = 0; k < nkeys; ++k) start[k + 1] += start[k];
  for (int v : items) out[start[key(v)]++] = v;
  return out;
}

}  // namespace

// Region order of a plan: cluster the rows, rank the regions by their first
// level (ties: region id), and order the positions of every level by
// (region rank, row).  Row results do not depend on the order of rows
// inside a level, so solves stay bit-identical.
int order_by_regions(symamd_handle h, symamd_plan p, const int *d_Lp, const int *d_Li) {
  const int n = p->n;
  if (h->regions == 0 || n < 2 * h->region_rows || p->levels < 3) return SYMAMD_OK;
  std::vector<int> Lp(n + 1), Li(p->nnz), perm(n);
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  SYMAMD_HIP(hipMemcpy(Lp.data(), d_Lp, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  SYMAMD_HIP(hipMemcpy(Li.data(), d_Li, sizeof(int) * p->nnz, hipMemcpyDeviceToHost));
  SYMAMD_HIP(hipMemcpy(perm.data(), p->d_perm, sizeof(int) * n, hipMemcpyDeviceToHost));
  std::vector<int> reg;
  const int R = cluster_regions(n, Lp.data(), Li.data(), h->region_rows, reg);
  int64_t inside = 0, off = 0;
  for (int i = 0; i < n; ++i)
    for (int j = Lp[i]; j < Lp[i + 1]; ++j)
      if (Li[j] != i) {
        ++off;
        inside += reg[Li[j]] == reg[i];
      }
  p->region_intra = off ? (double)inside / (double)off : 0.0;
  if (h->regions == 1 && p->region_intra < 0.5) return SYMAMD_OK;
  const int L = p->levels;
  std::vector<int> lev(n);
  for (int l = 0; l < L; ++l)
    for (int q = p->h_lvl_ptr[l]; q < p->h_lvl_ptr[l + 1]; ++q) lev[perm[q]] = l;
  std::vector<int> first(R, L);
  for (int i = 0; i < n; ++i) first[reg[i]] = std::min(first[reg[i]], lev[i]);
  std::vector<int> ids(R);
  std::iota(ids.begin(), ids.end(), 0);
  const std::vector<int> by_first = counting_sort(ids, L, [&](int r) { return first[r]; });
  std::vector<int> rank(R);
  for (int k = 0; k < R; ++k) rank[by_first[k]] = k;
  std::vector<int> rows(n);
  std::iota(rows.begin(), rows.end(), 0);
  const std::vector<int> by_rank = counting_sort(rows, R, [&](int i) { return rank[reg[i]]; });
  const std::vector<int> order = counting_sort(by_rank, L, [&](int i) { return lev[i]; });
  std::vector<int> preg(n);
  for (int q = 0; q < n; ++q) preg[q] = rank[reg[order[q]]];
  int *d_order = nullptr, *d_preg = nullptr;
  SYMAMD_HIP(hipMalloc(&d_order, sizeof(int) * n));
  if (hipMalloc(&d_preg, sizeof(int) * n) != hipSuccess) {
    (void)hipFree(d_order);
    return fail(SYMAMD_ERR_ALLOC, "region arrays");
  }
  SYMAMD_HIP(hipMemcpy(d_order, order.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  SYMAMD_HIP(hipMemcpy(d_preg, preg.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  p->d_lvl_set = p->d_perm;  // the canonical level set
  p->d_perm = d_order;
  p->d_preg = d_preg;
  p->nregions = R;
  return SYMAMD_OK;
}

// Owner schedule of a unit set for nw co-resident waves: unit u belongs to
// wave rank(region of u) mod nw (regions ranked by first level, so regions
// active at the same time sit on different waves while nw allows); every
// wave solves its units in increasing order, which keeps the schedule
// deadlock-free (the lowest unfinished unit always has all its
// dependencies done and its wave at it).  Dependencies on units of the same
// wave are dropped from the wait lists: program order already covers them.
int plan_owners(symamd_handle h, symamd_plan p, UnitSet *us, int nw) {
  if (us->nw == nw) return SYMAMD_OK;
  if (!p->d_preg || nw < 1) return fail(SYMAMD_ERR_INVALID_ARG, "no region order");
  void *old[] = {us->d_wptr, us->d_wlist, us->d_xdptr, us->d_xdlist};
  for (void *q : old)
    if (q) (void)hipFree(q);
  us->d_wptr = us->d_wlist = us->d_xdptr = us->d_xdlist = nullptr;
  us->nw = 0;
  const int U = us->count, n = p->n;
  std::vector<int> ub(U + 1), dptr(U + 1), dlist(us->deps > 0 ? us->deps : 1), preg(n);
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  SYMAMD_HIP(hipMemcpy(ub.data(), us->d_ub, sizeof(int) * (U + 1), hipMemcpyDeviceToHost));
  SYMAMD_HIP(hipMemcpy(dptr.data(), us->d_dptr, sizeof(int) * (U + 1), hipMemcpyDeviceToHost));
  if (us->deps > 0)
    SYMAMD_HIP(hipMemcpy(dlist.data(), us->d_dlist, sizeof(int) * us->deps,
                         hipMemcpyDeviceToHost));
  SYMAMD_HIP(hipMemcpy(preg.data(), p->d_preg, sizeof(int) * n, hipMemcpyDeviceToHost));
  std::vector<int> owner(U);
  for (int u = 0; u < U; ++u) owner[u] = preg[ub[u]] % nw;
  std::vector<int> units(U);
  std::iota(units.begin(), units.end(), 0);
  const std::vector<int> wlist = counting_sort(units, nw, [&](int u) { return owner[u]; });
  std::vector<int> wptr(nw + 1, 0);
  for (int u = 0; u < U; ++u) wptr[owner[u] + 1]++;
  for (int w = 0; w < nw; ++w) wptr[w + 1] += wptr[w];
  std::vector<int> xdptr(U + 1, 0), xdlist;
  xdlist.reserve(dlist.size());
  for (int u = 0; u < U; ++u) {
    for (int e = dptr[u]; e < dptr[u + 1]; ++e)
      if (owner[dlist[e]] != owner[u]) xdlist.push_back(dlist[e]);
    xdptr[u + 1] = (int)xdlist.size();
  }
  const size_t xd = xdlist.size() > 0 ? xdlist.size() : 1;
  SYMAMD_HIP(hipMalloc(&us->d_wptr, sizeof(int) * (nw + 1)));
  SYMAMD_HIP(hipMalloc(&us->d_wlist, sizeof(int) * (U > 0 ? U : 1)));
  SYMAMD_HIP(hipMalloc(&us->d_xdptr, sizeof(int) * (U + 1)));
  SYMAMD_HIP(hipMalloc(&us->d_xdlist, sizeof(int) * xd));
  SYMAMD_HIP(hipMemcpy(us->d_wptr, wptr.data(), sizeof(int) * (nw + 1), hipMemcpyHostToDevice));
  if (U > 0)
    SYMAMD_HIP(hipMemcpy(us->d_wlist, wlist.data(), sizeof(int) * U, hipMemcpyHostToDevice));
  SYMAMD_HIP(hipMemcpy(us->d_xdptr, xdptr.data(), sizeof(int) * (U + 1), hipMemcpyHostToDevice));
  if (!xdlist.empty())
    SYMAMD_HIP(hipMemcpy(us->d_xdlist, xdlist.data(), sizeof(int) * xdlist.size(),
                         hipMemcpyHostToDevice));
  us->xdeps = (int64_t)xdlist.size();
  us->nw = nw;
  return SYMAMD_OK;
}

void free_region_plan(RegionPlan &rp) {
  void *ptrs[] = {rp.d_rec, rp.d_rhead, rp.d_vsrc, rp.d_vdst, rp.d_rflags};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  rp = RegionPlan{};
}

namespace {

template <typename V>
int upload(V **dst, const std::vector<V> &src) {
  const size_t count = src.empty() ? 1 : src.size();
  SYMAMD_HIP(hipMalloc(dst, sizeof(V) * count));
  if (!src.empty())
    SYMAMD_HIP(hipMemcpy(*dst, src.data(), sizeof(V) * src.size(), hipMemcpyHostToDevice));
  return SYMAMD_OK;
}

}  // namespace

// Region schedule for region cap `cap`, from the plan's level-ordered copy:
//  - regions: heavy-edge clusters of at most `cap` rows;
//  - region graph: r waits for every region holding a column of one of its
//    rows; ranked by longest path (then first level, then id), so every
//    region waits only on regions of lower rank.  A cycle (possible for
//    irregular clusters) leaves the schedule empty (R = 0);
//  - inside a region the rows are ordered by (level, row) and cut into
//    slices of equal level; every off-diagonal entry is coded with the LDS
//    slot of its column: its local row, or smax + its index in the region's
//    halo (the sorted distinct rows of other regions it reads).
namespace {

// Region graph: regions region r reads from (deduplicated, in row order).
void region_graph(int n, const int *Lp, const int *Li, const std::vector<int> &reg, int R,
                  std::vector<int> &gptr, std::vector<int> &glist) {
  std::vector<int> rows(n);
  std::iota(rows.begin(), rows.end(), 0);
  const std::vector<int> by_reg = counting_sort(rows, R, [&](int i) { return reg[i]; });
  std::vector<int> start(R + 1, 0), mark(R, -1);
  for (int i = 0; i < n; ++i) start[reg[i] + 1]++;
  for (int r = 0; r < R; ++r) start[r + 1] += start[r];
  gptr.assign(R + 1, 0);
  glist.clear();
  for (int r = 0; r < R; ++r) {
    for (int t = start[r]; t < start[r + 1]; ++t) {
      const int i = by_reg[t];
      for (int j = Lp[i]; j < Lp[i + 1] - 1; ++j) {
        const int d = reg[Li[j]];
        if (d != r && mark[d] != r) {
          mark[d] = r;
          glist.push_back(d);
        }
      }
    }
    gptr[r + 1] = (int)glist.size();
  }
}

// Strongly connected components of the region graph (iterative Tarjan);
// returns the number of components, comp[r] = component of region r.
int region_sccs(int R, const std::vector<int> &gptr, const std::vector<int> &glist,
                std::vector<int> &comp) {
  std::vector<int> index(R, -1), low(R, 0), stack, cs_node, cs_edge;
  std::vector<char> on(R, 0);
  comp.assign(R, -1);
  int next = 0, ncomp = 0;
  for (int s = 0; s < R; ++s) {
    if (index[s] >= 0) continue;
    cs_node.push_back(s);
    cs_edge.push_back(gptr[s]);
    index[s] = low[s] = next++;
    stack.push_back(s);
    on[s] = 1;
    while (!cs_node.empty()) {
      const int v = cs_node.back();
      int &e = cs_edge.back();
      if (e < gptr[v + 1]) {
        const int w = glist[e++];
        if (index[w] < 0) {
          index[w] = low[w] = next++;
          stack.push_back(w);
          on[w] = 1;
          cs_node.push_back(w);
          cs_edge.push_back(gptr[w]);
        } else if (on[w]) {
          low[v] = std::min(low[v], index[w]);
        }
        continue;
      }
      if (low[v] == index[v]) {
        int w;
        do {
          w = stack.back();
          stack.pop_back();
          on[w] = 0;
          comp[w] = ncomp;
        } while (w != v);
        ++ncomp;
      }
      cs_node.pop_back();
      cs_edge.pop_back();
      if (!cs_node.empty()) {
        const int u = cs_node.back();
        low[u] = std::min(low[u], low[v]);
      }
    }
  }
  return ncomp;
}

// Makes the region graph acyclic: the regions of every strongly connected
// component are merged; a merged set larger than `cap` is cut into level
// bands (whole levels while they fit, a single level larger than `cap` in
// chunks).  Rows depend only on lower levels, so the bands of one component
// depend only on earlier bands, and edges between different components all
// point one way: the refined graph is acyclic.  Returns the region count.
int make_acyclic(int n, const int *Lp, const int *Li, const std::vector<int> &lev,
                 int cap, std::vector<int> &reg, int R) {
  std::vector<int> gptr, glist, comp;
  region_graph(n, Lp, Li, reg, R, gptr, glist);
  const int C = region_sccs(R, gptr, glist, comp);
  if (C == R) return R;
  std::vector<int> csize(C, 0);
  for (int r = 0; r < R; ++r) csize[comp[r]]++;  // regions per component
  std::vector<int> rows(n);
  std::iota(rows.begin(), rows.end(), 0);
  int maxlev = 0;
  for (int i = 0; i < n; ++i) maxlev = std::max(maxlev, lev[i]);
  // rows of each component in (level, row) order
  const std::vector<int> by_lev = counting_sort(rows, maxlev + 1, [&](int i) { return lev[i]; });
  const std::vector<int> by_comp =
      counting_sort(by_lev, C, [&](int i) { return comp[reg[i]]; });
  std::vector<int> newreg(n);
  std::vector<int> single(R, -1);  // new id of a region kept as it is
  int m = 0;
  size_t t = 0;
  while (t < by_comp.size()) {
    const int c = comp[reg[by_comp[t]]];
    size_t e = t;
    while (e < by_comp.size() && comp[reg[by_comp[e]]] == c) ++e;
    if (csize[c] == 1) {
      const int r = reg[by_comp[t]];
      if (single[r] < 0) single[r] = m++;
      for (size_t k = t; k < e; ++k) newreg[by_comp[k]] = single[r];
    } else {
      int filled = 0;
      size_t k = t;
      while (k < e) {
        size_t lend = k;  // rows of this level
        while (lend < e && lev[by_comp[lend]] == lev[by_comp[k]]) ++lend;
        const int cnt = (int)(lend - k);
        if (filled > 0 && filled + cnt > cap) {
          ++m;
          filled = 0;
        }
        for (size_t q = k; q < lend; ++q) {
          if (filled == cap) {
            ++m;
            filled = 0;
          }
          newreg[by_comp[q]] = m;
          ++filled;
        }
        k = lend;
      }
      ++m;
    }
    t = e;
  }
  reg.swap(newreg);
  return m;
}

}  // namespace

// Host part, from the level-ordered copy (perm, crp, cci: position -> row,
// row pointers and column ids of the copy; lvl_ptr: levels + 1 bounds).
// Returns false when the region graph has a cycle (not expected after
// make_acyclic; kept as a guard).
bool build_region_schedule(int n, int64_t nnz, int levels, const int *lvl_ptr,
                           const int *perm, const int *crp_, const int *cci_, int cap,
                           int band_slices, HostRegions &o) {
  if (band_slices < 1) band_slices = 1;
  const std::vector<int> crp(crp_, crp_ + n + 1), cci(cci_, cci_ + nnz);
  std::vector<int> ipos(n), lev(n), Lp(n + 1, 0), Li(nnz);
  for (int q = 0; q < n; ++q) ipos[perm[q]] = q;
  for (int l = 0; l < levels; ++l)
    for (int q = lvl_ptr[l]; q < lvl_ptr[l + 1]; ++q) lev[perm[q]] = l;
  for (int i = 0; i < n; ++i) Lp[i + 1] = Lp[i] + (crp[ipos[i] + 1] - crp[ipos[i]]);
  for (int i = 0; i < n; ++i)
    std::copy(cci.begin() + crp[ipos[i]], cci.begin() + crp[ipos[i] + 1], Li.begin() + Lp[i]);
  std::vector<int> reg;
  int R = cluster_regions(n, Lp.data(), Li.data(), cap, reg);
  R = make_acyclic(n, Lp.data(), Li.data(), lev, cap, reg, R);
  int64_t inside = 0, offd = 0;
  for (int i = 0; i < n; ++i)
    for (int j = Lp[i]; j < Lp[i + 1] - 1; ++j) {
      ++offd;
      inside += reg[Li[j]] == reg[i];
    }
  o.inside = offd ? (double)inside / (double)offd : 0.0;
  // region graph (deduplicated) and its longest-path depth
  std::vector<int> rows_nat(n);
  std::iota(rows_nat.begin(), rows_nat.end(), 0);
  const std::vector<int> by_reg = counting_sort(rows_nat, R, [&](int i) { return reg[i]; });
  std::vector<int> rstart(R + 1, 0);
  for (int i = 0; i < n; ++i) rstart[reg[i] + 1]++;
  for (int r = 0; r < R; ++r) rstart[r + 1] += rstart[r];
  std::vector<int> gptr(R + 1, 0), glist, mark(R, -1);
  for (int r = 0; r < R; ++r) {
    for (int t = rstart[r]; t < rstart[r + 1]; ++t) {
      const int i = by_reg[t];
      for (int j = Lp[i]; j < Lp[i + 1] - 1; ++j) {
        const int d = reg[Li[j]];
        if (d != r && mark[d] != r) {
          mark[d] = r;
          glist.push_back(d);
        }
      }
    }
    gptr[r + 1] = (int)glist.size();
  }
  std::vector<int> indeg(R, 0), succ_ptr(R + 1, 0), succ(glist.size());
  for (int r = 0; r < R; ++r) indeg[r] = gptr[r + 1] - gptr[r];
  for (int d : glist) succ_ptr[d + 1]++;
  for (int r = 0; r < R; ++r) succ_ptr[r + 1] += succ_ptr[r];
  {
    std::vector<int> fill(succ_ptr.begin(), succ_ptr.end() - 1);
    for (int r = 0; r < R; ++r)
      for (int e = gptr[r]; e < gptr[r + 1]; ++e) succ[fill[glist[e]]++] = r;
  }
  std::vector<int> depth(R, 0), queue;
  queue.reserve(R);
  for (int r = 0; r < R; ++r)
    if (indeg[r] == 0) queue.push_back(r);
  for (size_t qh = 0; qh < queue.size(); ++qh) {
    const int r = queue[qh];
    for (int e = succ_ptr[r]; e < succ_ptr[r + 1]; ++e) {
      const int s = succ[e];
      depth[s] = std::max(depth[s], depth[r] + 1);
      if (--indeg[s] == 0) queue.push_back(s);
    }
  }
  if ((int)queue.size() != R) return false;  // cyclic region graph
  std::vector<int> first(R, levels);
  for (int i = 0; i < n; ++i) first[reg[i]] = std::min(first[reg[i]], lev[i]);
  std::vector<int> ids(R);
  std::iota(ids.begin(), ids.end(), 0);
  int maxdepth = 0;
  for (int r = 0; r < R; ++r) maxdepth = std::max(maxdepth, depth[r]);
  {
    std::vector<int> at(maxdepth + 1, 0);
    for (int r = 0; r < R; ++r) ++at[depth[r]];
    o.width = *std::max_element(at.begin(), at.end());
  }
  const std::vector<int> o1 = counting_sort(ids, levels, [&](int r) { return first[r]; });
  const std::vector<int> order = counting_sort(o1, maxdepth + 1, [&](int r) { return depth[r]; });
  std::vector<int> rank(R);
  for (int k = 0; k < R; ++k) rank[order[k]] = k;
  // rows in (rank, level, row) order
  const std::vector<int> by_lev = counting_sort(rows_nat, levels, [&](int i) { return lev[i]; });
  const std::vector<int> rows = counting_sort(by_lev, R, [&](int i) { return rank[reg[i]]; });
  std::vector<int> rptr(R + 1, 0), lidx(n);
  for (int i = 0; i < n; ++i) rptr[rank[reg[i]] + 1]++;
  for (int r = 0; r < R; ++r) rptr[r + 1] += rptr[r];
  int smax = 0;
  for (int r = 0; r < R; ++r) smax = std::max(smax, rptr[r + 1] - rptr[r]);
  for (int r = 0; r < R; ++r)
    for (int t = rptr[r]; t < rptr[r + 1]; ++t) lidx[rows[t]] = t - rptr[r];
  // slices and bands: band b of a region covers its slices
  // [b*bs, (b+1)*bs) with bs = band_slices, or more when the region would
  // have more than REGION_BANDS_MAX bands
  std::vector<int> sptr(R + 1, 0), slices, bptr(R + 1, 0), bands, band_of(n);
  int slmax = 0;
  for (int r = 0; r < R; ++r) {
    const int a = rptr[r], b = rptr[r + 1];
    const size_t s0 = slices.size();
    for (int t = a; t < b; ++t)
      if (t == a || lev[rows[t]] != lev[rows[t - 1]]) slices.push_back(t - a);
    slices.push_back(b - a);
    sptr[r + 1] = (int)slices.size();
    const int ns = sptr[r + 1] - sptr[r] - 1;
    slmax = std::max(slmax, ns);
    const int bs = std::max(band_slices, (ns + REGION_BANDS_MAX - 1) / REGION_BANDS_MAX);
    const int nb = (ns + bs - 1) / bs;
    for (int k = 0; k <= nb; ++k) bands.push_back(std::min(k * bs, ns));
    bptr[r + 1] = (int)bands.size();
    for (int sl = 0; sl < ns; ++sl)
      for (int t = a + slices[s0 + sl]; t < a + slices[s0 + sl + 1]; ++t)
        band_of[rows[t]] = sl / bs;
  }
  // halos (ordered by the band that first reads them, then row), entries,
  // and the waits of each band: (rank of a region read, bands of it that
  // must be published), emitted where a region's requirement first rises.
  // Regions are independent: host threads take contiguous ranges of ranks,
  // each with its own output lists, joined in rank order afterwards.
  std::vector<int> hptr(R + 1, 0), halo, hband, dptr(R + 1, 0), dlist, wptr;
  // entries: the first RegionPlan::PACK of every row packed per row (pcnt,
  // pcode, psrc; padding slots read the all-zero LDS slot smax + hmax, and
  // their values are zero), the rest in the overflow list (eptr, ecode, esrc)
  constexpr int PK = RegionPlan::PACK;
  std::vector<int> eptr(n + 1, 0), ecode, esrc, dsrc(n);
  std::vector<int> pcnt(n), pcode((size_t)PK * n, 0), psrc((size_t)PK * n, 0);
  struct Part {
    std::vector<int> halo, hband, wptr, dlist, ecode, esrc;
    int hmax = 0, emax = 0;
  };
  std::vector<Part> parts(16);
  const int T = parallel_ranges(R, [&](int part, int r_lo, int r_hi) {
    Part &P = parts[part];
    std::vector<int> need_max(R, 0), need_stamp(R, -1), touched;
    std::vector<std::pair<int, int>> uses;   // (column, band of a reading row)
    std::vector<std::pair<int, int>> horder;  // (first band, column): halo order
    std::vector<std::pair<int, int>> slot;   // (column, halo slot), by column
    for (int r = r_lo; r < r_hi; ++r) {
      const int g = order[r];  // cluster id of rank r
      const int a = rptr[r], b = rptr[r + 1];
      const int nb = bptr[r + 1] - bptr[r] - 1;
      uses.clear();
      for (int t = a; t < b; ++t) {
        const int i = rows[t];
        for (int j = Lp[i]; j < Lp[i + 1] - 1; ++j)
          if (reg[Li[j]] != g) uses.emplace_back(Li[j], band_of[i]);
      }
      std::sort(uses.begin(), uses.end());
      horder.clear();
      for (size_t q = 0; q < uses.size(); ++q)
        if (q == 0 || uses[q].first != uses[q - 1].first)
          horder.emplace_back(uses[q].second, uses[q].first);  // smallest band first
      std::sort(horder.begin(), horder.end());
      slot.resize(horder.size());
      for (size_t q = 0; q < horder.size(); ++q) slot[q] = {horder[q].second, (int)q};
      std::sort(slot.begin(), slot.end());
      for (const auto &o2 : horder) P.halo.push_back(o2.second);
      hptr[r + 1] = (int)horder.size();  // per region; made cumulative below
      P.hmax = std::max(P.hmax, (int)horder.size());
      // halo bounds and waits per band
      size_t k = 0;
      const int dl0 = (int)P.dlist.size();
      for (int bb = 0; bb <= nb; ++bb) {
        P.hband.push_back((int)k);
        P.wptr.push_back((int)P.dlist.size() - dl0);
        if (bb == nb) break;
        const size_t k0 = k;
        while (k < horder.size() && horder[k].first == bb) ++k;
        // the largest requirement of this band per region read, in halo order
        touched.clear();
        for (size_t q = k0; q < k; ++q) {
          const int c = horder[q].second;
          const int u = rank[reg[c]];
          const int need = band_of[c] + 1;
          if (need_stamp[u] != r) {
            need_stamp[u] = r;
            need_max[u] = 0;
          }
          if (need > need_max[u]) {
            if (std::find(touched.begin(), touched.end(), u) == touched.end()) touched.push_back(u);
            need_max[u] = need;
          }
        }
        std::sort(touched.begin(), touched.end());
        for (int u : touched) P.dlist.push_back(u | (need_max[u] << REGION_RANK_BITS));
      }
      dptr[r + 1] = (int)P.dlist.size() - dl0;  // per region; made cumulative below
      const int e0 = (int)P.ecode.size();
      for (int t = a; t < b; ++t) {
        const int i = rows[t];
        const int q = ipos[i];
        dsrc[t] = crp[q + 1] - 1;
        pcnt[t] = Lp[i + 1] - 1 - Lp[i];
        for (int u = 0; u < PK; ++u) psrc[(size_t)PK * t + u] = dsrc[t];
        for (int j = Lp[i]; j < Lp[i + 1] - 1; ++j) {
          const int c = Li[j];
          const int u = j - Lp[i];
          int code;
          if (reg[c] == g) {
            code = lidx[c];
          } else {
            const auto it = std::lower_bound(slot.begin(), slot.end(), std::make_pair(c, INT_MIN));
            code = smax + it->second;
          }
          const int src = crp[q] + u;
          if (u < PK) {
            pcode[(size_t)PK * t + u] = code;
            psrc[(size_t)PK * t + u] = src;
          } else {
            P.ecode.push_back(code);
            P.esrc.push_back(src);
          }
        }
        eptr[t + 1] = (int)P.ecode.size();  // per part; made global below
      }
      P.emax = std::max(P.emax, (int)P.ecode.size() - e0);
    }
  });
  int hmax = 0, emax = 0;
  {
    // join the parts in rank order: list offsets become global
    std::vector<int> rbound(T + 1);
    for (int t = 0; t <= T; ++t) rbound[t] = (int)((int64_t)R * t / T);
    int64_t hbase = 0, dbase = 0, ebase = 0;
    for (int t = 0; t < T; ++t) {
      Part &P = parts[t];
      for (int r = rbound[t]; r < rbound[t + 1]; ++r) {
        hptr[r + 1] += hptr[r];
        dptr[r + 1] += dptr[r];
        for (int q = rptr[r]; q < rptr[r + 1]; ++q) eptr[q + 1] += (int)ebase;
      }
      halo.insert(halo.end(), P.halo.begin(), P.halo.end());
      hband.insert(hband.end(), P.hband.begin(), P.hband.end());
      wptr.insert(wptr.end(), P.wptr.begin(), P.wptr.end());
      dlist.insert(dlist.end(), P.dlist.begin(), P.dlist.end());
      ecode.insert(ecode.end(), P.ecode.begin(), P.ecode.end());
      esrc.insert(esrc.end(), P.esrc.begin(), P.esrc.end());
      hbase += (int64_t)P.halo.size();
      dbase += (int64_t)P.dlist.size();
      ebase += (int64_t)P.ecode.size();
      hmax = std::max(hmax, P.hmax);
      emax = std::max(emax, P.emax);
    }
    (void)hbase;
    (void)dbase;
  }
  if (R >= (1 << REGION_RANK_BITS)) return false;  // ranks do not fit the wait entries
  for (int t = 0; t < n; ++t)
    for (int u = pcnt[t]; u < PK; ++u) pcode[(size_t)PK * t + u] = smax + hmax;
  o.R = R;
  o.smax = smax;
  o.hmax = hmax;
  o.emax = emax;
  o.slmax = slmax;
  o.rptr.swap(rptr);
  o.rows = rows;
  o.sptr.swap(sptr);
  o.slices.swap(slices);
  o.hptr.swap(hptr);
  o.halo.swap(halo);
  o.dptr.swap(dptr);
  o.dlist.swap(dlist);
  o.bptr.swap(bptr);
  o.bands.swap(bands);
  o.hband.swap(hband);
  o.wptr.swap(wptr);
  o.eptr.swap(eptr);
  o.ecode.swap(ecode);
  o.esrc.swap(esrc);
  o.dsrc.swap(dsrc);
  o.pcnt.swap(pcnt);
  o.pcode.swap(pcode);
  o.psrc.swap(psrc);
  return true;
}

bool pack_region_records(const HostRegions &o, int vs, HostRecords &out,
                         const std::vector<int> *gmap) {
  const int gp = gmap ? 1 : 0;
  constexpr int PK = RegionPlan::PACK;
  const int R = o.R;
  out.rhead.assign(8 * (size_t)R, 0);
  out.recmax = 0;
  int64_t units = 0;
  // value slots of each region (packed entries, diagonals, overflow), so
  // that regions can be written in parallel at known offsets
  std::vector<int64_t> voff(R + 1, 0);
  for (int r = 0; r < R; ++r) {
    const int a = o.rptr[r], nr = o.rptr[r + 1] - a;
    const int nh = o.hptr[r + 1] - o.hptr[r];
    const int ne = o.eptr[a + nr] - o.eptr[a];
    const int ns = o.sptr[r + 1] - o.sptr[r] - 1;
    const int nd = o.dptr[r + 1] - o.dptr[r];
    const int nb = o.bptr[r + 1] - o.bptr[r] - 1;
    const RecLayout L = rec_layout(nr, nh, ne, ns, nd, nb, vs, gp);
    int *hd = &out.rhead[8 * (size_t)r];
    hd[0] = nr;
    hd[1] = nh;
    hd[2] = ne;
    hd[3] = ns;
    hd[4] = nd;
    hd[5] = (int)units;
    hd[6] = L.bytes / 16;
    hd[7] = nb;
    out.recmax = std::max(out.recmax, L.bytes);
    units += L.bytes / 16;
    int64_t vc = ne;
    for (int t = a; t < a + nr; ++t) vc += std::min(o.pcnt[t], PK) + 1;
    voff[r + 1] = voff[r] + vc;
    // value positions are counted in values of vs bytes
    if (units > INT_MAX || units * 16 / vs > INT_MAX) return false;
  }
  out.bytes.assign((size_t)units * 16, 0);
  out.vsrc.assign((size_t)voff[R], 0);
  out.vdst.assign((size_t)voff[R], 0);
  parallel_ranges(R, [&](int, int r_lo, int r_hi) {
  for (int r = r_lo; r < r_hi; ++r) {
    size_t vq = (size_t)voff[r];
    const int a = o.rptr[r], nr = o.rptr[r + 1] - a;
    const int h0 = o.hptr[r], nh = o.hptr[r + 1] - h0;
    const int e0 = o.eptr[a], ne = o.eptr[a + nr] - e0;
    const int s0 = o.sptr[r], ns = o.sptr[r + 1] - s0 - 1;
    const int d0 = o.dptr[r], nd = o.dptr[r + 1] - d0;
    const int b0 = o.bptr[r], nb = o.bptr[r + 1] - b0 - 1;
    const RecLayout L = rec_layout(nr, nh, ne, ns, nd, nb, vs, gp);
    uint8_t *rec = out.bytes.data() + (size_t)out.rhead[8 * (size_t)r + 5] * 16;
    const int64_t vbase = (int64_t)out.rhead[8 * (size_t)r + 5] * 16 / vs;
    auto put16 = [](uint8_t *p, int v) {
      const uint16_t w = (uint16_t)v;
      std::memcpy(p, &w, 2);
    };
    auto put32 = [](uint8_t *p, int v) { std::memcpy(p, &v, 4); };
    for (int l = 0; l < nr; ++l) {
      const int t = a + l;
      for (int u = 0; u < PK; ++u) {
        put16(rec + L.code + 2 * (PK * l + u), o.pcode[(size_t)PK * t + u]);
        if (u >= o.pcnt[t]) continue;  // padding: value stays +0.0
        out.vsrc[vq] = o.psrc[(size_t)PK * t + u];
        out.vdst[vq++] = (int)(vbase + (L.pval / vs) + PK * l + u);
      }
      out.vsrc[vq] = o.dsrc[t];
      out.vdst[vq++] = (int)(vbase + L.diag / vs + l);
      rec[L.cnt + l] = (uint8_t)std::min(o.pcnt[t], PK + 1);
      put32(rec + L.rows + 4 * l, o.rows[t]);
      if (ne > 0) put32(rec + L.lep + 4 * l, o.eptr[t] - e0);
    }
    if (ne > 0) put32(rec + L.lep + 4 * nr, ne);
    for (int e = 0; e < ne; ++e) {
      put16(rec + L.lcode + 2 * e, o.ecode[e0 + e]);
      out.vsrc[vq] = o.esrc[e0 + e];
      out.vdst[vq++] = (int)(vbase + L.lval / vs + e);
    }
    for (int q = 0; q < nh; ++q) put32(rec + L.halo + 4 * q, o.halo[h0 + q]);
    for (int s = 0; s <= ns; ++s) put32(rec + L.slices + 4 * s, o.slices[s0 + s]);
    for (int d = 0; d < nd; ++d) put32(rec + L.deps + 4 * d, o.dlist[d0 + d]);
    for (int k = 0; k <= nb; ++k) {
      put32(rec + L.bands + 4 * k, o.bands[b0 + k]);
      put32(rec + L.hband + 4 * k, o.hband[b0 + k]);
      put32(rec + L.wptr + 4 * k, o.wptr[b0 + k]);
    }
    if (gp) {
      const std::vector<int> &gm = *gmap;
      for (int l = 0; l < nr; ++l) {
        const int t = a + l;
        for (int u = 0; u < PK; ++u)
          put32(rec + L.gpos + 4 * ((PK + 1) * l + u),
                u < o.pcnt[t] ? gm[o.psrc[(size_t)PK * t + u]] : 0);
        put32(rec + L.gpos + 4 * ((PK + 1) * l + PK), gm[o.dsrc[t]]);
      }
      for (int e = 0; e < ne; ++e) put32(rec + L.gpos + 4 * ((PK + 1) * nr + e), gm[o.esrc[e0 + e]]);
    }
  }
  });
  return true;
}

// Inputs of the region builder for the transposed solve L^T x = g: with
// rows and positions numbered backwards (row j' = n-1-j, position
// q' = n-1-q), L^T is lower triangular.  Row j' holds column j of L: its
// off-diagonal entries in ascending row order of L (the order
// sptrsv_lt.cpp:14-16 accumulates them), then the diagonal (the builder
// takes it last).  The forward levels reversed are a valid level function
// (an entry (i, j) of L, i > j, has lev(i) > lev(j)), so position q' keeps
// the level set of q, in reverse.  vmap: entry of the reversed copy ->
// entry of the plan's transposed copy (CSC of L, diagonal first).
void reversed_transpose(int n, int levels, const std::vector<int> &lvl_ptr,
                        const std::vector<int> &perm, const std::vector<int> &trp,
                        const std::vector<int> &tci, std::vector<int> &rlvl,
                        std::vector<int> &rperm, std::vector<int> &rrp, std::vector<int> &rci,
                        std::vector<int> &vmap) {
  rlvl.assign(levels + 1, 0);
  for (int l = 0; l <= levels; ++l) rlvl[l] = n - lvl_ptr[levels - l];
  rperm.assign(n, 0);
  rrp.assign(n + 1, 0);
  rci.assign(tci.size(), 0);
  vmap.assign(tci.size(), 0);
  for (int qr = 0; qr < n; ++qr) {
    const int q = n - 1 - qr;
    rperm[qr] = n - 1 - perm[q];
    const int t0 = trp[q], len = trp[q + 1] - t0;
    const int r0 = rrp[qr];
    for (int u = 1; u < len; ++u) {
      rci[r0 + u - 1] = n - 1 - tci[t0 + u];
      vmap[r0 + u - 1] = t0 + u;
    }
    rci[r0 + len - 1] = n - 1 - tci[t0];
    vmap[r0 + len - 1] = t0;
    rrp[qr + 1] = r0 + len;
  }
}

// Device part: the schedule for cap `cap` (built once per plan and cap), of
// the forward solve or (transposed) of L^T x = g.  A transposed schedule
// names original rows in its records and gathers its values from the
// plan's transposed copy.
int plan_regions(symamd_handle h, symamd_plan p, int cap, int k, RegionPlan **out,
                 bool transposed) {
  auto &plans = transposed ? p->rplans_t : p->rplans;
  const int band = region_band(h, k);
  const int64_t key = (int64_t)cap * 256 + band;
  auto it = plans.find(key);
  if (it != plans.end()) {
    *out = &it->second;
    return SYMAMD_OK;
  }
  RegionPlan rp;
  rp.cap = cap;
  rp.band = band;
  rp.transposed = transposed;
  const int n = p->n;
  HostRegions o;
  std::vector<int> vmap;
  if (n > 0 && p->levels >= 2) {
    std::vector<int> perm(n), crp(n + 1), cci(p->nnz);
    if (transposed) SYMAMD_TRY(plan_transpose(h, p));
    SYMAMD_HIP(hipStreamSynchronize(h->stream));
    SYMAMD_HIP(hipMemcpy(perm.data(), p->d_perm, sizeof(int) * n, hipMemcpyDeviceToHost));
    SYMAMD_HIP(hipMemcpy(crp.data(), transposed ? p->d_trp : p->d_rp, sizeof(int) * (n + 1),
                         hipMemcpyDeviceToHost));
    SYMAMD_HIP(hipMemcpy(cci.data(), transposed ? p->d_tci : p->d_ci, sizeof(int) * p->nnz,
                         hipMemcpyDeviceToHost));
    bool ok;
    if (transposed) {
      std::vector<int> rlvl, rperm, rrp, rci;
      reversed_transpose(n, p->levels, p->h_lvl_ptr, perm, crp, cci, rlvl, rperm, rrp, rci,
                         vmap);
      ok = build_region_schedule(n, p->nnz, p->levels, rlvl.data(), rperm.data(), rrp.data(),
                                 rci.data(), cap, band, o);
      if (ok) {
        for (int &r : o.rows) r = n - 1 - r;
        for (int &r : o.halo) r = n - 1 - r;
      }
    } else {
      ok = build_region_schedule(n, p->nnz, p->levels, p->h_lvl_ptr.data(), perm.data(),
                                 crp.data(), cci.data(), cap, band, o);
    }
    if (!ok) o.R = 0;
  }
  rp.inside = o.inside;
  if (o.R == 0) {
    *out = &plans.emplace(key, rp).first->second;
    return SYMAMD_OK;
  }
  const int vs = p->dtype == SYMAMD_F64 ? 8 : 4;
  // single right-hand side adjoints write grad-L from the solve: their
  // records carry every entry's position in the caller's CSR
  std::vector<int> gmap;
  if (transposed && k == 1 && p->nnz > 0) {
    std::vector<int> tmap(p->nnz);
    SYMAMD_HIP(hipMemcpy(tmap.data(), p->d_tmap, sizeof(int) * p->nnz, hipMemcpyDeviceToHost));
    gmap.resize(vmap.size());
    for (size_t v = 0; v < vmap.size(); ++v) gmap[v] = tmap[vmap[v]];
  }
  HostRecords hr;
  if (o.smax + o.hmax >= REGION_SLOTS_MAX ||
      !pack_region_records(o, vs, hr, gmap.empty() ? nullptr : &gmap)) {
    *out = &plans.emplace(key, rp).first->second;  // R = 0: not usable
    return SYMAMD_OK;
  }
  if (transposed)
    for (int &v : hr.vsrc) v = vmap[v];
  rp.R = o.R;
  rp.gp = gmap.empty() ? 0 : 1;
  rp.smax = o.smax;
  rp.hmax = o.hmax;
  rp.emax = o.emax;
  rp.slmax = o.slmax;
  rp.recmax = hr.recmax;
  rp.width = o.width;
  rp.nvals = (int64_t)hr.vsrc.size();
  int rc = SYMAMD_OK;
  if ((rc = upload(reinterpret_cast<uint8_t **>(&rp.d_rec), hr.bytes)) ||
      (rc = upload(&rp.d_rhead, hr.rhead)) || (rc = upload(&rp.d_vsrc, hr.vsrc)) ||
      (rc = upload(&rp.d_vdst, hr.vdst))) {
    free_region_plan(rp);
    return rc;
  }
  if (hipMalloc(&rp.d_rflags, sizeof(int64_t) * rp.R * RegionPlan::MAX_GROUPS) !=
          hipSuccess ||
      hipMemset(rp.d_rflags, 0, sizeof(int64_t) * rp.R * RegionPlan::MAX_GROUPS) !=
          hipSuccess) {
    free_region_plan(rp);
    return fail(SYMAMD_ERR_ALLOC, "region schedule arrays");
  }
  RegionPlan *stored = &plans.emplace(key, rp).first->second;
  SYMAMD_TRY(gather_region_values(h, p));
  *out = stored;
  return SYMAMD_OK;
}

}  // namespace symamd

// ------------------------------------------------------------ diagnostics
// The region schedule of a host CSR lower triangle (diagonal last), built on
// the host by the same steps the device path takes (level sets with rows
// ascending inside a level, the level-ordered copy, and for the transposed
// solve the level-ordered CSC), packed with FP64 values.  For tests that
// emulate the region kernel on the CPU.
extern "C" int symamd_region_schedule_host(int n, const int *Lp, const int *Li,
                                           const double *Lx, int cap, int band_slices,
                                           int transposed, int *info, uint8_t *rec,
                                           int *rhead) {
  using namespace symamd;
  if (n < 1 || !Lp || !Li || !Lx || !info || cap < 2)
    return fail(SYMAMD_ERR_INVALID_ARG, "symamd_region_schedule_host: bad arguments");
  const int64_t nnz = Lp[n];
  std::vector<int> lev(n, 0);
  int levels = 0;
  for (int i = 0; i < n; ++i) {
    int l = 0;
    for (int64_t j = Lp[i]; j < Lp[i + 1] - 1; ++j) l = std::max(l, lev[Li[j]] + 1);
    lev[i] = l;
    levels = std::max(levels, l + 1);
  }
  std::vector<int> lvl_ptr(levels + 1, 0), perm(n);
  for (int i = 0; i < n; ++i) lvl_ptr[lev[i] + 1]++;
  for (int l = 0; l < levels; ++l) lvl_ptr[l + 1] += lvl_ptr[l];
  {
    std::vector<int> fill(lvl_ptr.begin(), lvl_ptr.end() - 1);
    for (int i = 0; i < n; ++i) perm[fill[lev[i]]++] = i;
  }
  // level-ordered copy (row perm[q] at position q), and its values
  std::vector<int> crp(n + 1, 0), cci(nnz);
  std::vector<double> cval(nnz);
  for (int q = 0; q < n; ++q) {
    const int i = perm[q];
    crp[q + 1] = crp[q] + (int)(Lp[i + 1] - Lp[i]);
    for (int64_t j = Lp[i]; j < Lp[i + 1]; ++j) {
      cci[crp[q] + (j - Lp[i])] = Li[j];
      cval[crp[q] + (j - Lp[i])] = Lx[j];
    }
  }
  HostRegions o;
  std::vector<int> vmap;
  std::vector<double> tval;
  bool ok;
  if (transposed) {
    // level-ordered CSC: column perm[q] at position q, rows ascending
    std::vector<int> cnt(n + 1, 0), trp(n + 1, 0), tci(nnz);
    tval.resize(nnz);
    for (int64_t j = 0; j < nnz; ++j) cnt[Li[j]]++;
    std::vector<int> ipos(n);
    for (int q = 0; q < n; ++q) ipos[perm[q]] = q;
    for (int q = 0; q < n; ++q) trp[q + 1] = trp[q] + cnt[perm[q]];
    std::vector<int> fill(trp.begin(), trp.end() - 1);
    for (int i = 0; i < n; ++i)
      for (int64_t j = Lp[i]; j < Lp[i + 1]; ++j) {
        const int t = fill[ipos[Li[j]]]++;
        tci[t] = i;
        tval[t] = Lx[j];
      }
    std::vector<int> rlvl, rperm, rrp, rci;
    reversed_transpose(n, levels, lvl_ptr, perm, trp, tci, rlvl, rperm, rrp, rci, vmap);
    ok = build_region_schedule(n, nnz, levels, rlvl.data(), rperm.data(), rrp.data(), rci.data(),
                               cap, band_slices, o);
    if (ok) {
      for (int &r : o.rows) r = n - 1 - r;
      for (int &r : o.halo) r = n - 1 - r;
    }
  } else {
    ok = build_region_schedule(n, nnz, levels, lvl_ptr.data(), perm.data(), crp.data(),
                               cci.data(), cap, band_slices, o);
  }
  HostRecords hr;
  if (!ok || o.R == 0 || !pack_region_records(o, 8, hr))
    return fail(SYMAMD_ERR_NOT_SUPPORTED, "symamd_region_schedule_host: no region schedule");
  info[0] = o.R;
  info[1] = o.smax;
  info[2] = o.hmax;
  info[3] = hr.recmax;
  info[4] = (int)hr.bytes.size();
  info[5] = o.slmax;
  info[6] = levels;
  info[7] = o.width;
  if (rec) {
    std::memcpy(rec, hr.bytes.data(), hr.bytes.size());
    double *vals = reinterpret_cast<double *>(rec);
    for (size_t t = 0; t < hr.vsrc.size(); ++t)
      vals[hr.vdst[t]] = transposed ? tval[vmap[hr.vsrc[t]]] : cval[hr.vsrc[t]];
  }
  if (rhead) std::memcpy(rhead, hr.rhead.data(), sizeof(int) * hr.rhead.size());
  return SYMAMD_OK;
}
