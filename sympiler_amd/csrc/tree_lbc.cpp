// tree_lbc.cpp -- load-balanced level coarsening of an elimination tree.
//
// Input: a forest given as a parent array (a negative parent marks a root)
// and a cost per node.  Output: the 3-array schedule that the LBC solves and
// factorizations consume -- coarse level l holds parts
// level_ptr[l] .. level_ptr[l+1]-1, part p holds the nodes
// partition[part_ptr[p] .. part_ptr[p+1]-1].  Inside a part every node comes
// after its children in it; two parts of one coarse level share no tree
// edge; every other edge goes from a coarse level to a later one.  A node's
// dependencies in a Cholesky-type factor are its descendants, so the
// schedule is valid for any solve whose dependency graph the tree covers.
//
// The decisions follow sym_lib::parsy::getCoarseLevelSet_6
// (sparse_blas/parsy/symbolic/InspectionLevel_06.cpp:17-385) so that the
// three arrays come out identical for every input that routine finishes on:
//
//  1. levels counted from the leaves (a leaf is level 0, a parent one above
//     its highest child; TreeUtils.cpp:95-146), nodes of a level by id;
//  2. the cut of the levels into clusters (heightPartitioning,
//     TreeUtils.cpp:283-352, with its default `sw = false`): from the root
//     end, skip the levels holding at most `parts` nodes, cut min_level_dist
//     levels below that when the cut is above the previous one, divide
//     `parts` by div_rate and repeat while it exceeds 1; the part counts are
//     recorded per round, also in a round that adds no cut, and cluster l
//     uses the l-th of them; the last cluster ends past the root level;
//  3. inside a cluster the edges to parents outside it are cut and each
//     remaining subtree is collected by the stack walk of postOrderSpliting
//     (PostOrderSpliting.cpp:13-120, children before parents; its stack is
//     a deque here, the reference's a vector with inserts at the front);
//  4. when a cluster has more subtrees than its part count they are sorted
//     by cost, largest first (std::sort, as in worstFitBinPack,
//     TreeUtils.cpp:190-225, so ties fall the same way), and each goes to
//     the first bin of least accumulated cost; the bins' costs carry over
//     from one cluster to the next (the reference never clears them) and
//     every bin is emitted, also an empty one.
//
// Costs, depths and levels are computed here in linear time (the reference
// walks every node to the root and inserts at the front of vectors, which
// is quadratic on deep trees); only the order of the floating-point cost
// sums, which decides the bin choices, is kept as the reference has it.
// The per-cluster integer cost and its threshold in the reference feed only
// unused parameters and assertions, so they are not formed.
#include <algorithm>
#include <cmath>
#include <deque>
#include <set>
#include <utility>
#include <vector>

#include "common.h"

using symamd::fail;

namespace {

struct Forest {
  int n;
  const int *parent;  // < 0: root
  std::vector<int> depth, level;
  int height = 0;     // longest root-to-leaf path, in edges
};

// Depth of every node without walking a path twice; fails on a cycle.
bool depths(Forest &f) {
  const int n = f.n;
  f.depth.assign(n, -1);
  std::vector<char> on_path(n, 0);
  std::vector<int> path;
  for (int i = 0; i < n; ++i) {
    if (f.depth[i] >= 0) continue;
    int v = i;
    while (v >= 0 && f.depth[v] < 0) {
      if (on_path[v]) return false;
      on_path[v] = 1;
      path.push_back(v);
      v = f.parent[v] < 0 ? -1 : f.parent[v];
    }
    int d = v < 0 ? -1 : f.depth[v];
    for (int k = (int)path.size() - 1; k >= 0; --k) {
      f.depth[path[k]] = ++d;
      on_path[path[k]] = 0;
    }
    path.clear();
  }
  for (int i = 0; i < n; ++i) f.height = std::max(f.height, f.depth[i]);
  return true;
}

// Level from the leaves: nodes by decreasing depth push 1 + their level up.
void levels(Forest &f) {
  const int n = f.n, H = f.height;
  std::vector<int> start(H + 2, 0), order(n);
  for (int i = 0; i < n; ++i) ++start[f.depth[i] + 1];
  for (int d = 0; d <= H; ++d) start[d + 1] += start[d];
  for (int i = 0; i < n; ++i) order[start[f.depth[i]]++] = i;
  f.level.assign(n, 0);
  for (int k = n - 1; k >= 0; --k) {
    const int v = order[k], p = f.parent[v];
    if (p >= 0) f.level[p] = std::max(f.level[p], f.level[v] + 1);
  }
}

// Step 2: cluster boundaries (in levels) and the part count of each round.
void cut_levels(const std::vector<int> &level_size, int H, int parts,
                int min_level_dist, int div_rate, std::vector<int> &bounds,
                std::vector<int> &parts_of) {
  const int nlev = (int)level_size.size();  // == H + 1
  bounds.assign(1, 0);
  if (nlev <= 2) {
    bounds.push_back(nlev);
    parts_of.assign(1, 1);
    return;
  }
  for (int p = parts; p > 1; p /= div_rate) {
    int cut = 0;
    // Levels near the root with no more nodes than parts are skipped.
    while (level_size[H - cut - 1] <= p && cut < nlev) {
      ++cut;
      if (H - cut - 1 < 0 || H - cut - 1 >= nlev) break;
    }
    parts_of.push_back(p);
    const int at = H - cut - min_level_dist;
    if (at > bounds.back() && at < nlev) bounds.push_back(at);
  }
  bounds.push_back(H + 1);
  parts_of.push_back(1);
}

// Step 3 on one cluster.  `members` are the cluster's nodes by id; `up[v]`
// is a member's parent inside the cluster, or -1 where the edge is cut;
// `kids`/`kid_ptr` list every node's children by id, those below the
// cluster's first level (`lo`) are skipped; `nkid` holds the members' child
// counts inside the cluster and is consumed; `done`/`queued` are false for
// the members on entry (each node belongs to one cluster only).
void collect_subtrees(const std::vector<int> &members, const std::vector<int> &up,
                      const std::vector<int> &kid_ptr, const std::vector<int> &kids,
                      const std::vector<int> &level, int lo, std::vector<int> &nkid,
                      std::vector<char> &done, std::vector<char> &queued,
                      const double *cost, std::vector<int> &flat,
                      std::vector<int> &sub_ptr, std::vector<double> &sub_cost) {
  std::deque<int> st;  // front = top
  for (int s : members) {
    if (done[s]) continue;
    double c = 0.0;
    st.push_back(s);
    while (!st.empty()) {
      const int k = st.front();
      if (nkid[k] == 0) {
        st.pop_front();
        flat.push_back(k);
        c += cost[k];
        done[k] = 1;
        const int p = up[k];
        if (p < 0) break;  // the subtree's root: this subtree is complete
        --nkid[p];
        if (!done[p]) {
          if (!queued[p]) st.push_front(p);
          queued[p] = 1;
        }
        continue;
      }
      for (int e = kid_ptr[k]; e < kid_ptr[k + 1]; ++e) {
        const int ch = kids[e];
        if (level[ch] < lo || done[ch]) continue;
        if (nkid[ch] != 0) {
          st.push_front(ch);
          continue;
        }
        if (!done[k]) {
          // k and its ancestors up to the walk's bottom entry are (or will
          // be) on the stack: do not queue them again when a child ends.
          int t = k;
          while (!st.empty() && t >= 0 && t != st.back()) {
            queued[t] = 1;
            t = up[t];
          }
          if (t >= 0) queued[t] = 1;
        }
        flat.push_back(ch);
        c += cost[ch];
        done[ch] = 1;
        --nkid[k];
      }
    }
    sub_ptr.push_back((int)flat.size());
    sub_cost.push_back(c);
  }
}

}  // namespace

extern "C" int symamd_tree_coarsen(int n, const int *parent,
                                   const double *node_cost, int inner_parts,
                                   int min_level_dist, int div_rate,
                                   int *level_ptr, int *part_ptr,
                                   int *partition, int *level_no,
                                   int *part_no) {
  if (n < 0 || !level_ptr || !part_ptr || !level_no || !part_no)
    return fail(SYMAMD_ERR_INVALID_ARG, "tree_coarsen: negative n or null output");
  if (n > 0 && (!parent || !node_cost || !partition))
    return fail(SYMAMD_ERR_INVALID_ARG, "tree_coarsen: null input or partition array");
  *level_no = 0;
  *part_no = 0;
  level_ptr[0] = 0;
  part_ptr[0] = 0;
  if (n == 0) return SYMAMD_OK;
  for (int i = 0; i < n; ++i) {
    if (parent[i] >= n || parent[i] == i)
      return fail(SYMAMD_ERR_INVALID_ARG, "tree_coarsen: parent[%d] = %d", i, parent[i]);
    if (std::isnan(node_cost[i]))
      return fail(SYMAMD_ERR_INVALID_ARG, "tree_coarsen: node_cost[%d] is NaN", i);
  }
  Forest f{n, parent};
  if (!depths(f)) return fail(SYMAMD_ERR_STRUCTURE, "tree_coarsen: parent has a cycle");
  levels(f);
  const int H = f.height, nlev = H + 1;
  // A division rate below 2 would never end the cut rounds (0 divides by
  // zero) -- the reference does not return on such input.
  if (nlev > 2 && inner_parts > 1 && (div_rate == 0 || div_rate == 1))
    return fail(SYMAMD_ERR_INVALID_ARG, "tree_coarsen: div_rate %d with inner_parts %d",
                div_rate, inner_parts);

  std::vector<int> level_size(nlev, 0);
  for (int i = 0; i < n; ++i) ++level_size[f.level[i]];
  std::vector<int> bounds, parts_of;
  cut_levels(level_size, H, inner_parts, min_level_dist, div_rate, bounds,
             parts_of);
  const int nclu = (int)bounds.size() - 1;

  // Children by id (once), the clusters' members by id (a stable bucket
  // sort over the cluster of each node's level).
  std::vector<int> kid_ptr(n + 1, 0), kids(n), fill(n);
  for (int v = 0; v < n; ++v)
    if (parent[v] >= 0) ++kid_ptr[parent[v] + 1];
  for (int v = 0; v < n; ++v) kid_ptr[v + 1] += kid_ptr[v];
  std::copy(kid_ptr.begin(), kid_ptr.end() - 1, fill.begin());
  for (int v = 0; v < n; ++v)
    if (parent[v] >= 0) kids[fill[parent[v]]++] = v;
  std::vector<int> cluster_of_level(nlev), mem_ptr(nclu + 1, 0), mem(n);
  for (int l = 0; l < nclu; ++l)
    for (int v = bounds[l]; v < bounds[l + 1]; ++v) cluster_of_level[v] = l;
  for (int v = 0; v < n; ++v) ++mem_ptr[cluster_of_level[f.level[v]] + 1];
  for (int l = 0; l < nclu; ++l) mem_ptr[l + 1] += mem_ptr[l];
  std::copy(mem_ptr.begin(), mem_ptr.end() - 1, fill.begin());
  for (int v = 0; v < n; ++v) mem[fill[cluster_of_level[f.level[v]]]++] = v;

  std::vector<int> up(n), nkid(n, 0);
  std::vector<char> done(n, 0), queued(n, 0);
  std::vector<double> bin_cost(n, 0.0);  // carried across clusters
  int parts_out = 0, nodes_out = 0;
  for (int l = 0; l < nclu; ++l) {
    const std::vector<int> members(mem.begin() + mem_ptr[l], mem.begin() + mem_ptr[l + 1]);
    for (int v : members) {
      const int p = parent[v];
      up[v] = (p >= 0 && f.level[p] < bounds[l + 1]) ? p : -1;
      if (up[v] >= 0) ++nkid[p];
    }
    std::vector<int> flat, sub_ptr(1, 0);
    std::vector<double> sub_cost;
    collect_subtrees(members, up, kid_ptr, kids, f.level, bounds[l], nkid, done, queued,
                     node_cost, flat, sub_ptr, sub_cost);
    const int nsub = (int)sub_cost.size();
    auto emit = [&](int t) {
      for (int e = sub_ptr[t]; e < sub_ptr[t + 1]; ++e) partition[nodes_out++] = flat[e];
    };

    const int want = parts_of[l];
    if (nsub > want) {
      std::vector<int> order(nsub);
      for (int i = 0; i < nsub; ++i) order[i] = i;
      std::sort(order.begin(), order.end(),
                [&](int a, int b) { return sub_cost[a] > sub_cost[b]; });
      // The first bin of least cost, as a linear scan from a bound of
      // INT_MAX finds it: the least (cost, index) while that cost is below
      // the bound, bin 0 once no bin is.
      std::set<std::pair<double, int>> by_cost;
      for (int b = 0; b < want; ++b) by_cost.insert({bin_cost[b], b});
      std::vector<int> bin_of(nsub), bin_ptr(want + 1, 0);
      for (int t : order) {
        const auto least = *by_cost.begin();
        const int best = least.first < 2147483647.0 ? least.second : 0;
        by_cost.erase({bin_cost[best], best});
        bin_cost[best] += sub_cost[t];
        by_cost.insert({bin_cost[best], best});
        bin_of[t] = best;
        ++bin_ptr[best + 1];
      }
      // each bin lists its subtrees in the order they were assigned
      for (int b = 0; b < want; ++b) bin_ptr[b + 1] += bin_ptr[b];
      std::vector<int> in_bin(nsub);
      for (int t : order) in_bin[bin_ptr[bin_of[t]]++] = t;
      int k = 0;
      for (int b = 0; b < want; ++b) {
        for (; k < bin_ptr[b]; ++k) emit(in_bin[k]);
        part_ptr[++parts_out] = nodes_out;
      }
    } else {
      for (int t = 0; t < nsub; ++t) {
        emit(t);
        part_ptr[++parts_out] = nodes_out;
      }
    }
    level_ptr[l + 1] = parts_out;
  }
  *level_no = nclu;
  *part_no = parts_out;
  return SYMAMD_OK;
}
