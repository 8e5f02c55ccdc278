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
#include <numeric>
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

}  // namespace symamd
