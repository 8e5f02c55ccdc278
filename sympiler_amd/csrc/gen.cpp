// gen.cpp -- deterministic host-side generators for the sparse triangular
// workloads: 2D/3D Laplacians, their lower triangles, IC(0) factors, the
// Zipf-irregular lower-triangular matrix and the uneven batch of systems.
//
// Random numbers: one splitmix64 stream per (seed, tag), u = (z >> 11) * 2^-53
// in [0, 1).  The stream is pure integer arithmetic, so every machine draws
// the same values; fixtures do not depend on numpy's generators.
//
// All matrices are CSR lower triangles with the diagonal LAST in each row and
// column indices ascending (the layout sym_lib::sptrsv_csr consumes,
// sparse_blas/sptrsv.cpp:17-25).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/sympiler_amd_gen.h"

namespace {

struct Stream {
  uint64_t s;
  Stream(uint64_t seed, uint64_t tag) {
    s = seed ^ (tag * 0x9E3779B97F4A7C15ull);
    next();  // decorrelate neighbouring seeds
  }
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double unit() { return (double)(next() >> 11) * 0x1.0p-53; }
};

}  // namespace

extern "C" {

void symgen_uniform(uint64_t seed, uint64_t tag, int64_t count, double lo,
                    double hi, double *out) {
  Stream st(seed, tag);
  const double w = hi - lo;
  for (int64_t i = 0; i < count; ++i) out[i] = lo + w * st.unit();
}

// ------------------------------------------------------------ Laplacians

int64_t symgen_lap2d_tril_nnz(int nx, int ny) {
  return (int64_t)nx * ny + (int64_t)(nx - 1) * ny + (int64_t)nx * (ny - 1);
}

// tril of the 5-point Laplacian on an nx-by-ny grid, row id = x + nx*y,
// diagonal `diag`, off-diagonals `off`.
void symgen_lap2d_tril(int nx, int ny, double diag, double off, int *Lp,
                       int *Li, double *Lx) {
  int64_t p = 0;
  Lp[0] = 0;
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      int i = x + nx * y;
      if (y > 0) { Li[p] = i - nx; Lx[p++] = off; }
      if (x > 0) { Li[p] = i - 1; Lx[p++] = off; }
      Li[p] = i; Lx[p++] = diag;
      Lp[i + 1] = (int)p;
    }
}

int64_t symgen_lap3d_tril_nnz(int nx, int ny, int nz) {
  int64_t n = (int64_t)nx * ny * nz;
  return n + (int64_t)(nx - 1) * ny * nz + (int64_t)nx * (ny - 1) * nz +
         (int64_t)nx * ny * (nz - 1);
}

// tril of the 7-point Laplacian on nx*ny*nz, row id = x + nx*(y + ny*z).
void symgen_lap3d_tril(int nx, int ny, int nz, double diag, double off,
                       int *Lp, int *Li, double *Lx) {
  int64_t p = 0;
  const int sxy = nx * ny;
  Lp[0] = 0;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        int i = x + nx * (y + ny * z);
        if (z > 0) { Li[p] = i - sxy; Lx[p++] = off; }
        if (y > 0) { Li[p] = i - nx; Lx[p++] = off; }
        if (x > 0) { Li[p] = i - 1; Lx[p++] = off; }
        Li[p] = i; Lx[p++] = diag;
        Lp[i + 1] = (int)p;
      }
}

// ---------------------------------------------------------------- IC(0)

// Zero-fill incomplete Cholesky on the pattern of the CSR lower triangle
// (Lp, Li), overwriting Lx (which holds tril(A) on entry).  Entry (r, c):
//   L[r,c] = (A[r,c] - sum_{s<c} L[r,s] L[c,s]) / L[c,c]       (r > c)
//   L[c,c] = sqrt(A[c,c] - sum_{s<c} L[c,s]^2)
// with s running ascending over the columns both rows store.  That is the
// subtraction order of the left-looking column algorithm in
// sparse_blas/spic0.cpp:11-46 (spic0_csc_left walks its prune set, i.e. the
// s of row c, ascending, subtracting L[r,s]*L[c,s] from every r of column s
// at or below c), so the factor is bit-identical to it.  Rows are processed
// in order, so every L[c, *] with c < r is final when row r reads it.
// Returns 0 on success or 1 + the row with a non-positive pivot.
int symgen_ic0_csr(int n, const int *Lp, const int *Li, double *Lx) {
  for (int r = 0; r < n; ++r) {
    const int rb = Lp[r], re = Lp[r + 1] - 1;  // re = diagonal position
    for (int q = rb; q < re; ++q) {
      const int c = Li[q];
      double f = Lx[q];
      // merge row r (positions rb..q-1) with row c (strict part), s < c
      int a = rb, b = Lp[c];
      const int be = Lp[c + 1] - 1;
      while (a < q && b < be) {
        const int sa = Li[a], sb = Li[b];
        if (sa == sb) { f -= Lx[a] * Lx[b]; ++a; ++b; }
        else if (sa < sb) ++a;
        else ++b;
      }
      Lx[q] = f / Lx[be];
    }
    double d = Lx[re];
    for (int q = rb; q < re; ++q) d -= Lx[q] * Lx[q];
    if (!(d > 0)) return r + 1;
    Lx[re] = std::sqrt(d);
  }
  return 0;
}

// ------------------------------------------------------------- SpMV (host)

// y = L x on a CSR lower triangle (used to form b = L * x_true).
void symgen_spmv(int n, const int *Lp, const int *Li, const double *Lx,
                 const double *x, double *y) {
  for (int i = 0; i < n; ++i) {
    double s = 0;
    for (int p = Lp[i]; p < Lp[i + 1]; ++p) s += Lx[p] * x[Li[p]];
    y[i] = s;
  }
}

// ------------------------------------------------------ Zipf lower matrix

// Irregular lower triangle: for row i the count of strictly-lower entries is
// min(Z, cap, w) with Z ~ Zipf(alpha) on {1, 2, ...} and w = min(i, window);
// the columns are distinct and uniform in [max(0, i-window), i), stored
// ascending; off-diagonal values ~ U[-1, 0); the diagonal is 1 + sum |off|.
// Z is drawn by inverting the Zipf CDF; counts at or beyond `cap` are
// clipped to `cap`.  Two passes over the same stream: count, then fill.
// Lp has n+1 entries; Li/Lx may be NULL (count only).  Returns nnz.
static double zeta_tail(double s, int from) {
  // sum_{k >= from} k^-s by Euler-Maclaurin from N = from + 2048
  const int N = from + 2048;
  double acc = 0;
  for (int k = from; k < N; ++k) acc += std::pow((double)k, -s);
  const double dN = (double)N;
  acc += std::pow(dN, 1 - s) / (s - 1) + 0.5 * std::pow(dN, -s) +
         s * std::pow(dN, -s - 1) / 12.0 -
         s * (s + 1) * (s + 2) * std::pow(dN, -s - 3) / 720.0;
  return acc;
}

int64_t symgen_zipf_lower(int n, double alpha, int cap, int window,
                          uint64_t seed, int *Lp, int *Li, double *Lx) {
  std::vector<double> cdf(cap);  // cdf[k-1] = P(Z <= k), k = 1..cap-1
  const double zeta = zeta_tail(alpha, 1);
  double run = 0;
  for (int k = 1; k < cap; ++k) {
    run += std::pow((double)k, -alpha) / zeta;
    cdf[k - 1] = run;
  }
  Stream st(seed, 1);
  std::vector<int> cols;
  cols.reserve(cap);
  int64_t p = 0;
  Lp[0] = 0;
  for (int i = 0; i < n; ++i) {
    const double u = st.unit();
    int z = (int)(std::upper_bound(cdf.begin(), cdf.begin() + (cap - 1), u) -
                  cdf.begin()) + 1;  // first k with cdf >= u ... clipped
    if (z > cap) z = cap;
    const int lo = i > window ? i - window : 0;
    const int w = i - lo;
    int cnt = std::min(z, w);
    cols.clear();
    // distinct uniform columns in [lo, i): Floyd's sampling
    for (int t = w - cnt; t < w; ++t) {
      int v = (int)(st.unit() * (t + 1));
      if (v > t) v = t;
      int c = lo + v;
      if (std::find(cols.begin(), cols.end(), c) != cols.end()) c = lo + t;
      cols.push_back(c);
    }
    std::sort(cols.begin(), cols.end());
    double rowsum = 0;
    for (int c : cols) {
      const double v = st.unit() - 1.0;  // U[-1, 0)
      if (Li) { Li[p] = c; Lx[p] = v; }
      rowsum += std::fabs(v);
      ++p;
    }
    if (Li) { Li[p] = i; Lx[p] = 1.0 + rowsum; }
    ++p;
    Lp[i + 1] = (int)p;
  }
  return p;
}

// --------------------------------------------------------- uneven batch

// Grid sizes of the uneven batch: a_s, b_s ~ U{lo..hi} for s = 0..count-1.
void symgen_batch_dims(int count, int lo, int hi, uint64_t seed, int *a,
                       int *b) {
  Stream st(seed, 2);
  const int span = hi - lo + 1;
  for (int s = 0; s < count; ++s) {
    int u = (int)(st.unit() * span);
    a[s] = lo + (u < span ? u : span - 1);
    u = (int)(st.unit() * span);
    b[s] = lo + (u < span ? u : span - 1);
  }
}

// ------------------------------------------------ random sparse lower

// A random lower triangle with about density * i strictly-lower entries in
// row i (at least `min_off` where possible), off-diagonals ~ U[-1, 0) and a
// dominant diagonal 1 + sum |off|.  Pass Li = NULL to count.  Returns nnz.
int64_t symgen_random_lower(int n, double density, int min_off, uint64_t seed,
                            int *Lp, int *Li, double *Lx) {
  Stream st(seed, 3);
  std::vector<int> cols;
  int64_t p = 0;
  Lp[0] = 0;
  for (int i = 0; i < n; ++i) {
    cols.clear();
    for (int c = 0; c < i; ++c)
      if (st.unit() < density) cols.push_back(c);
    if ((int)cols.size() < min_off && i > 0) {
      int want = std::min(min_off, i);
      while ((int)cols.size() < want) {
        int c = (int)(st.unit() * i);
        if (c >= i) c = i - 1;
        if (std::find(cols.begin(), cols.end(), c) == cols.end())
          cols.push_back(c);
      }
      std::sort(cols.begin(), cols.end());
    }
    double rowsum = 0;
    for (int c : cols) {
      const double v = st.unit() - 1.0;
      if (Li) { Li[p] = c; Lx[p] = v; }
      rowsum += std::fabs(v);
      ++p;
    }
    if (Li) { Li[p] = i; Lx[p] = 1.0 + rowsum; }
    ++p;
    Lp[i + 1] = (int)p;
  }
  return p;
}

// Supernodal (BCSC) lower triangle (structure from stream (seed, 4), values
// from (seed, 5)): supernode i owns columns s_i .. s_i+w_i-1
// (w_i ~ U{1..wmax}) and nrows_i = w_i + c_i block rows, the c_i ~
// U{0..min(rmax, rows left)} rows below it distinct and ascending, drawn from
// the next `window` rows.  Values: a column-major nrows_i x w_i block at
// Lx[Lp[i]]; on the top block's diagonal 1 + U[0, 1), below it U[-1, 1) /
// nrows_i; the top block's strictly upper part is never read by the solve
// and holds NaN.  Li[Lp[i] + j*nrows_i + r] = row of block row r (every
// column of the block repeats the row list).  With Lp == NULL only the
// counts are returned: *num_nodes and the value count.
int64_t symgen_bcsc(int n, int wmax, int rmax, int window, uint64_t seed, int *num_nodes,
                    int *supernodes, int *nrows, int *Lp, int *Li, double *Lx) {
  Stream st(seed, 4);   // structure
  Stream sv(seed, 5);   // values (drawn only when filling)
  std::vector<int> rows;
  int64_t p = 0;
  int i = 0;
  for (int s = 0; s < n; ++i) {
    int w = 1 + (int)(st.unit() * wmax);
    if (w > wmax) w = wmax;
    if (w > n - s) w = n - s;
    const int e = s + w;  // first row below the supernode
    const int span = std::min(window, n - e);
    int c = (int)(st.unit() * (std::min(rmax, span) + 1));
    if (c > span) c = span;
    rows.clear();
    while ((int)rows.size() < c) {  // c distinct rows of [e, e + span)
      int r = e + (int)(st.unit() * span);
      if (r >= e + span) r = e + span - 1;
      if (std::find(rows.begin(), rows.end(), r) == rows.end()) rows.push_back(r);
    }
    std::sort(rows.begin(), rows.end());
    const int nr = w + c;
    if (Lp) {
      supernodes[i] = s;
      nrows[i] = nr;
      Lp[i] = (int)p;
      for (int j = 0; j < w; ++j)
        for (int r = 0; r < nr; ++r) {
          const int64_t q = p + (int64_t)j * nr + r;
          Li[q] = r < w ? s + r : rows[r - w];
          double v;
          if (r < j) v = NAN;
          else if (r == j) v = 1.0 + sv.unit();
          else v = (2.0 * sv.unit() - 1.0) / nr;
          Lx[q] = v;
        }
    }
    p += (int64_t)nr * w;
    s = e;
  }
  if (Lp) {
    supernodes[i] = n;
    Lp[i] = (int)p;
  }
  *num_nodes = i;
  return p;
}

// Supernodal lower triangle shaped like a nested-dissection Cholesky factor:
// a complete binary tree of `levels` levels of supernodes numbered in
// postorder (children's columns before their parent's); a node h levels
// above the leaves is min(wmax, round(wleaf * 2^(h/2))) columns wide and
// stores below its top about rowfac x its width rows drawn from its
// ancestors' columns (all of them if fewer), ascending.  Values as
// symgen_bcsc; streams (seed, 6) for the structure and (seed, 7) for values.
int64_t symgen_bcsc_tree(int levels, int wleaf, int wmax, double rowfac, uint64_t seed,
                         int *num_nodes, int *n_out, int *supernodes, int *nrows, int *Lp,
                         int *Li, double *Lx) {
  Stream st(seed, 6), sv(seed, 7);
  const int m = (1 << levels) - 1;
  // postorder: node heights and widths, then column ranges
  std::vector<int> height, parent, order;  // order[k] = node id in postorder
  height.resize(m);
  parent.assign(m, -1);
  // heap numbering: node t has children 2t+1, 2t+2
  std::vector<int> stack{0}, visited(m, 0);
  while (!stack.empty()) {
    const int t = stack.back();
    const int l = 2 * t + 1, r = 2 * t + 2;
    if (l < m && !visited[t]) {
      visited[t] = 1;
      parent[l] = t;
      parent[r] = t;
      stack.push_back(r);
      stack.push_back(l);
    } else {
      stack.pop_back();
      order.push_back(t);
    }
  }
  for (int t = m - 1; t >= 0; --t) {
    const int l = 2 * t + 1;
    height[t] = l < m ? height[l] + 1 : 0;
  }
  std::vector<int> width(m), first(m), post(m);
  int col = 0;
  for (int k = 0; k < m; ++k) {
    const int t = order[k];
    post[t] = k;
    width[t] = std::min(wmax, std::max(1, (int)std::lround(wleaf * std::pow(2.0, 0.5 * height[t]))));
    first[t] = col;
    col += width[t];
  }
  const int n = col;
  std::vector<int> anc, rows;
  int64_t p = 0;
  for (int k = 0; k < m; ++k) {
    const int t = order[k], w = width[t], s = first[t];
    anc.clear();
    for (int a = parent[t]; a >= 0; a = parent[a])
      for (int c = first[a]; c < first[a] + width[a]; ++c) anc.push_back(c);
    std::sort(anc.begin(), anc.end());
    int want = (int)std::lround(rowfac * w);
    if (want > (int)anc.size()) want = (int)anc.size();
    // a uniform subset of the ancestors' columns: partial Fisher-Yates
    for (int j = 0; j < want; ++j) {
      const int q = j + (int)(st.unit() * (anc.size() - j));
      std::swap(anc[j], anc[q < (int)anc.size() ? q : (int)anc.size() - 1]);
    }
    rows.assign(anc.begin(), anc.begin() + want);
    std::sort(rows.begin(), rows.end());
    const int nr = w + want;
    if (Lp) {
      supernodes[k] = s;
      nrows[k] = nr;
      Lp[k] = (int)p;
      for (int j = 0; j < w; ++j)
        for (int r = 0; r < nr; ++r) {
          const int64_t q = p + (int64_t)j * nr + r;
          Li[q] = r < w ? s + r : rows[r - w];
          Lx[q] = r < j ? NAN : r == j ? 1.0 + sv.unit() : (2.0 * sv.unit() - 1.0) / nr;
        }
    }
    p += (int64_t)nr * w;
  }
  if (Lp) {
    supernodes[m] = n;
    Lp[m] = (int)p;
  }
  *num_nodes = m;
  *n_out = n;
  return p;
}

}  // extern "C"
