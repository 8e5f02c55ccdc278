// dropin.cpp -- sym_lib:: entry points with the reference's signatures
// (include/sympiler/sparse_blas_lib.h), implemented on the C ABI.
//
// Each call is blocking, as in the reference: host arrays are copied to the
// device, the HIP kernels run with one lane per row (bit-identical to the
// serial reference), and x is copied back.  The analysis of a matrix is kept
// (per thread, the DROPIN_CACHE most recently used patterns): a later call
// whose n, Lp and Li equal a kept pattern's exactly only uploads the values
// and the right-hand side, as an iterative solver calling the same factor's
// solve every step does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sym_lib_sparse_blas.h"
#include "../../include/sympiler_amd.h"

namespace {

void check(int rc, const char *what) {
  if (rc != SYMAMD_OK)
    throw std::runtime_error(std::string("sympiler_amd ") + what + ": " +
                             symamd_last_error());
}

void hcheck(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw std::runtime_error(std::string("sympiler_amd ") + what + ": " +
                             hipGetErrorString(e));
}

symamd_handle handle() {
  thread_local symamd_handle h = nullptr;
  if (!h) check(symamd_create(&h, nullptr), "create");
  return h;
}

template <typename T>
struct DevBuf {
  T *p = nullptr;
  DevBuf() = default;
  explicit DevBuf(size_t count, const T *src = nullptr) {
    hcheck(hipMalloc(&p, sizeof(T) * (count ? count : 1)), "hipMalloc");
    if (src && count) hcheck(hipMemcpy(p, src, sizeof(T) * count, hipMemcpyHostToDevice), "upload");
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  void download(T *dst, size_t count) const {
    if (count) hcheck(hipMemcpy(dst, p, sizeof(T) * count, hipMemcpyDeviceToHost), "download");
  }
};

struct Plan {
  symamd_plan p = nullptr;
  ~Plan() {
    if (p) symamd_plan_destroy(p);
  }
};

// ------------------------------------------------------- analysis cache
// A kept analysis: the pattern (host copies, for exact matching), its device
// arrays, a right-hand-side buffer and the plan.
struct Kept {
  int n = 0, nnz = 0;
  std::vector<int> Lp, Li;
  int *dLp = nullptr, *dLi = nullptr;
  double *dLx = nullptr, *dx = nullptr;
  symamd_plan plan = nullptr;
  uint64_t used = 0;
};
constexpr int DROPIN_CACHE = 4;

// Per thread; never destroyed at exit (the device memory goes with the
// process, and no HIP call runs while the runtime shuts down).
std::vector<Kept *> &kept() {
  thread_local std::vector<Kept *> *k = new std::vector<Kept *>();
  return *k;
}

void release(Kept *e) {
  if (e->plan) (void)symamd_plan_destroy(e->plan);
  void *bufs[] = {e->dLp, e->dLi, e->dLx, e->dx};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  delete e;
}

// The kept analysis of (n, Lp, Li) with the values Lx (and, when x is given,
// the right-hand side) uploaded; analysed now if no kept pattern matches,
// evicting the least recently used one beyond DROPIN_CACHE.
Kept *analysis(symamd_handle h, int n, const int *Lp, const int *Li, const double *Lx,
               const double *x) {
  thread_local uint64_t clock = 0;
  auto &ks = kept();
  const int nnz = Lp[n];
  Kept *e = nullptr;
  for (Kept *c : ks)
    if (c->n == n && c->nnz == nnz && std::memcmp(c->Lp.data(), Lp, sizeof(int) * (n + 1)) == 0 &&
        std::memcmp(c->Li.data(), Li, sizeof(int) * nnz) == 0) {
      e = c;
      break;
    }
  if (e) {
    hcheck(hipMemcpy(e->dLx, Lx, sizeof(double) * (nnz ? nnz : 1), hipMemcpyHostToDevice),
           "upload");
    check(symamd_plan_update_values(h, e->plan, e->dLx), "update values");
  } else {
    if ((int)ks.size() >= DROPIN_CACHE) {
      auto lru = std::min_element(ks.begin(), ks.end(),
                                  [](const Kept *a, const Kept *b) { return a->used < b->used; });
      release(*lru);
      ks.erase(lru);
    }
    auto *c = new Kept();
    c->n = n;
    c->nnz = nnz;
    c->Lp.assign(Lp, Lp + n + 1);
    c->Li.assign(Li, Li + nnz);
    const hipError_t a0 = hipMalloc(&c->dLp, sizeof(int) * (n + 1));
    const hipError_t a1 = hipMalloc(&c->dLi, sizeof(int) * (nnz ? nnz : 1));
    const hipError_t a2 = hipMalloc(&c->dLx, sizeof(double) * (nnz ? nnz : 1));
    const hipError_t a3 = hipMalloc(&c->dx, sizeof(double) * n);
    if (a0 != hipSuccess || a1 != hipSuccess || a2 != hipSuccess || a3 != hipSuccess) {
      release(c);
      throw std::runtime_error("sympiler_amd: device allocation for a drop-in call failed");
    }
    int rc = hipMemcpy(c->dLp, Lp, sizeof(int) * (n + 1), hipMemcpyHostToDevice) == hipSuccess &&
                     hipMemcpy(c->dLi, Li, sizeof(int) * nnz, hipMemcpyHostToDevice) ==
                         hipSuccess &&
                     hipMemcpy(c->dLx, Lx, sizeof(double) * nnz, hipMemcpyHostToDevice) ==
                         hipSuccess
                 ? SYMAMD_OK
                 : SYMAMD_ERR_HIP;
    if (rc == SYMAMD_OK) rc = symamd_analyze(h, n, c->dLp, c->dLi, c->dLx, SYMAMD_F64, &c->plan);
    if (rc == SYMAMD_OK) rc = symamd_plan_set_lanes(c->plan, 1);
    if (rc != SYMAMD_OK) {
      const std::string msg = rc == SYMAMD_ERR_HIP && !c->plan ? "upload failed"
                                                               : symamd_last_error();
      release(c);
      throw std::runtime_error("sympiler_amd analyze: " + msg);
    }
    ks.push_back(c);
    e = c;
  }
  e->used = ++clock;
  if (x) hcheck(hipMemcpy(e->dx, x, sizeof(double) * n, hipMemcpyHostToDevice), "upload");
  return e;
}

// CSC (diagonal first) -> CSR of the same lower triangle (diagonal last).
void csc_to_csr(int n, const int *Cp, const int *Ci, const double *Cx,
                std::vector<int> &Rp, std::vector<int> &Ri,
                std::vector<double> &Rx) {
  const int nnz = Cp[n];
  Rp.assign(n + 1, 0);
  Ri.resize(nnz);
  Rx.resize(nnz);
  for (int p = 0; p < nnz; ++p) Rp[Ci[p] + 1]++;
  for (int i = 0; i < n; ++i) Rp[i + 1] += Rp[i];
  std::vector<int> fill(Rp.begin(), Rp.end() - 1);
  for (int c = 0; c < n; ++c)
    for (int p = Cp[c]; p < Cp[c + 1]; ++p) {
      int q = fill[Ci[p]]++;
      Ri[q] = c;
      Rx[q] = Cx[p];
    }
}

// Solve L x = b in place on the GPU from a host CSR (diagonal last).
void solve_csr_host(int n, const int *Lp, const int *Li, const double *Lx,
                    double *x) {
  if (n <= 0) return;
  symamd_handle h = handle();
  Kept *e = analysis(h, n, Lp, Li, Lx, x);
  check(symamd_solve(h, e->plan, e->dx, 1, n, SYMAMD_COL_MAJOR, SYMAMD_EXEC_SYNCFREE), "solve");
  check(symamd_plan_check(h, e->plan), "solve");
  hcheck(hipMemcpy(x, e->dx, sizeof(double) * n, hipMemcpyDeviceToHost), "download");
}

// Checks a caller level set: every row once, dependencies in earlier levels.
void check_levels(int n, const int *Rp, const int *Ri, int levels,
                  const int *levelPtr, const int *levelSet) {
  if (levels < 0 || !levelPtr || (n > 0 && !levelSet) || levelPtr[0] != 0 ||
      levelPtr[levels] != n)
    throw std::runtime_error("sympiler_amd: level set does not cover the n rows");
  std::vector<int> lev(n, -1);
  for (int l = 0; l < levels; ++l)
    for (int k = levelPtr[l]; k < levelPtr[l + 1]; ++k) {
      int i = levelSet[k];
      if (i < 0 || i >= n || lev[i] != -1)
        throw std::runtime_error("sympiler_amd: level set lists a row twice or out of range");
      lev[i] = l;
    }
  for (int i = 0; i < n; ++i)
    for (int j = Rp[i]; j < Rp[i + 1] - 1; ++j)
      if (lev[Ri[j]] >= lev[i])
        throw std::runtime_error("sympiler_amd: level set violates a dependency");
}

}  // namespace

namespace sym_lib {

void spmv_csr(int n, const int *Ap, const int *Ai, const double *Ax,
              const double *x, double *y) {
  if (n <= 0) return;
  const int nnz = Ap[n];
  int ncol = n;
  for (int p = 0; p < nnz; ++p) ncol = Ai[p] + 1 > ncol ? Ai[p] + 1 : ncol;
  DevBuf<int> dAp(n + 1, Ap), dAi(nnz, Ai);
  DevBuf<double> dAx(nnz, Ax), dx(ncol, x), dy(n);
  symamd_handle h = handle();
  check(symamd_spmv_csr(h, n, dAp.p, dAi.p, dAx.p, dx.p, dy.p), "spmv_csr");
  hcheck(hipDeviceSynchronize(), "spmv_csr");
  dy.download(y, n);
}

void spmv_csr(int n, const int *Ap, const int *Ai, const double *Ax,
              const double *x, const double *y, double *z, double a, double b) {
  if (n <= 0) return;
  const int nnz = Ap[n];
  int ncol = n;
  for (int p = 0; p < nnz; ++p) ncol = Ai[p] + 1 > ncol ? Ai[p] + 1 : ncol;
  DevBuf<int> dAp(n + 1, Ap), dAi(nnz, Ai);
  DevBuf<double> dAx(nnz, Ax), dx(ncol, x), dy(n, y), dz(n, z);
  symamd_handle h = handle();
  check(symamd_spmv_csr_axpby(h, n, dAp.p, dAi.p, dAx.p, dx.p, dy.p, dz.p, a, b),
        "spmv_csr");
  hcheck(hipDeviceSynchronize(), "spmv_csr");
  dz.download(z, n);
}

void spmv_csr_parallel(int n, const int *Ap, const int *Ai, const double *Ax,
                       const double *x, double *y) {
  spmv_csr(n, Ap, Ai, Ax, x, y);
}

// The rows of the supernodes are one contiguous range: the 6-argument
// spmv_csr on it (row pointers stay absolute offsets into Ai/Ax).
void spmv_csr_block(int n, const int *Ap, const int *Ai, const double *Ax, int nodes,
                    const int *supernodes, const double *x, double *y) {
  if (nodes <= 0) return;
  const int r0 = supernodes[0], r1 = supernodes[nodes];
  if (r0 < 0 || r1 < r0 || r1 > n)
    throw std::invalid_argument("spmv_csr_block: supernode rows outside [0, n)");
  if (r1 == r0) return;
  const int e0 = Ap[r0], e1 = Ap[r1];
  int ncol = n;
  for (int p = e0; p < e1; ++p) ncol = Ai[p] + 1 > ncol ? Ai[p] + 1 : ncol;
  std::vector<int> rp(r1 - r0 + 1);
  for (int r = r0; r <= r1; ++r) rp[r - r0] = Ap[r] - e0;
  DevBuf<int> dAp(rp.size(), rp.data()), dAi(e1 - e0, Ai + e0);
  DevBuf<double> dAx(e1 - e0, Ax + e0), dx(ncol, x), dy(r1 - r0);
  symamd_handle h = handle();
  check(symamd_spmv_csr(h, r1 - r0, dAp.p, dAi.p, dAx.p, dx.p, dy.p), "spmv_csr_block");
  hcheck(hipDeviceSynchronize(), "spmv_csr_block");
  dy.download(y + r0, r1 - r0);
}

namespace {

// Rows of a CSC matrix: the largest row index + 1 (at least `floor`).
int csc_rows(int n, const int *Ap, const int *Ai, int floor) {
  int m = floor;
  for (int p = 0; p < Ap[n]; ++p) m = Ai[p] + 1 > m ? Ai[p] + 1 : m;
  return m;
}

// y (m entries, updated in place) += columns c0 .. c1-1 of the CSC matrix.
void csc_product(const char *who, int n, int m, const int *Ap, const int *Ai,
                 const double *Ax, const double *x, double *y, int c0, int c1) {
  const int nnz = Ap[n];
  DevBuf<int> dAp(n + 1, Ap), dAi(nnz, Ai);
  DevBuf<double> dAx(nnz, Ax), dx(n, x), dy(m, y);
  symamd_handle h = handle();
  check(symamd_spmv_csc(h, n, m, dAp.p, dAi.p, dAx.p, dx.p, dy.p, c0, c1), who);
  dy.download(y, m);
}

// z += A x, then z = a z + b y with scalars (av, bv NULL) or per-row values.
void csc_axpby(const char *who, int n, const int *Ap, const int *Ai, const double *Ax,
               const double *x, const double *y, double *z, double a, double b,
               const double *av, const double *bv) {
  if (csc_rows(n, Ap, Ai, n) > n)
    throw std::invalid_argument(std::string(who) + ": a row index lies outside [0, n)");
  const int nnz = Ap[n];
  DevBuf<int> dAp(n + 1, Ap), dAi(nnz, Ai);
  DevBuf<double> dAx(nnz, Ax), dx(n, x), dy(n, y), dz(n, z);
  DevBuf<double> da(av ? n : 0, av), db(bv ? n : 0, bv);
  symamd_handle h = handle();
  check(symamd_spmv_csc(h, n, n, dAp.p, dAi.p, dAx.p, dx.p, dz.p, 0, n), who);
  check(symamd_axpby(h, n, av ? da.p : nullptr, bv ? db.p : nullptr, a, b, dy.p, dz.p), who);
  hcheck(hipDeviceSynchronize(), who);
  dz.download(z, n);
}

void csc_lbc(const char *who, int n_parts, const int *sets, const int *ids, int n,
             const int *Ap, const int *Ai, const double *Ax, const double *x, double *y) {
  if (n_parts <= 0 || n <= 0) return;
  for (int l = 0; l < n_parts; ++l)
    if (sets[l + 1] < sets[l] || sets[l] < 0)
      throw std::invalid_argument(std::string(who) + ": part bounds must be nondecreasing");
  for (int k = sets[0]; k < sets[n_parts]; ++k)
    if (ids[k] < 0 || ids[k] >= n)
      throw std::invalid_argument(std::string(who) + ": a column id lies outside [0, n)");
  if (csc_rows(n, Ap, Ai, n) > n)
    throw std::invalid_argument(std::string(who) + ": a row index lies outside [0, n)");
  const int nnz = Ap[n], nid = sets[n_parts];
  DevBuf<int> dAp(n + 1, Ap), dAi(nnz, Ai), dsets(n_parts + 1, sets), dids(nid, ids);
  DevBuf<double> dAx(nnz, Ax), dx(n, x), dy(n, y);
  symamd_handle h = handle();
  check(symamd_spmv_csc_parts(h, n, n, dAp.p, dAi.p, dAx.p, dx.p, dy.p, n_parts, dsets.p,
                              dids.p),
        who);
  dy.download(y, n);
}

}  // namespace

void spmv_csc(int n, const int *Ap, const int *Ai, const double *Ax, const double *x,
              double *y) {
  if (n <= 0) return;
  csc_product("spmv_csc", n, csc_rows(n, Ap, Ai, 0), Ap, Ai, Ax, x, y, 0, n);
}

void spmv_csc(int n, const int *Ap, const int *Ai, const double *Ax, const double *x,
              const double *y, double *z, double a, double b) {
  if (n <= 0) return;
  csc_axpby("spmv_csc", n, Ap, Ai, Ax, x, y, z, a, b, nullptr, nullptr);
}

void spmv_csc(int n, const int *Ap, const int *Ai, const double *Ax, const double *x,
              const double *y, double *z, double *a, double *b) {
  if (n <= 0) return;
  if (!a || !b) throw std::invalid_argument("spmv_csc: null a or b");
  csc_axpby("spmv_csc", n, Ap, Ai, Ax, x, y, z, 0.0, 0.0, a, b);
}

void spmv_csc_lbc(int n_parts, const int *sets, const int *ids, int n, const int *Ap,
                  const int *Ai, const double *Ax, const double *x, double *y) {
  csc_lbc("spmv_csc_lbc", n_parts, sets, ids, n, Ap, Ai, Ax, x, y);
}

void spmv_csc_lbc(const int *set_ptr, const int *sets, const int *ids, int n, const int *Ap,
                  const int *Ai, const double *Ax, const double *x, double *y) {
  csc_lbc("spmv_csc_lbc", set_ptr[1], sets, ids, n, Ap, Ai, Ax, x, y);
}

void spmv_csc_parallel(int n, const int *Ap, const int *Ai, const double *Ax,
                       const double *x, double *y) {
  spmv_csc(n, Ap, Ai, Ax, x, y);
}

void spmv_csc_block(int n, const int *Ap, const int *Ai, const double *Ax, const double *x,
                    double *y, int nodes, const int *supernodes) {
  if (nodes <= 0 || n <= 0) return;
  const int c0 = supernodes[0], c1 = supernodes[nodes];
  if (c0 < 0 || c1 < c0 || c1 > n)
    throw std::invalid_argument("spmv_csc_block: supernode columns outside [0, n)");
  csc_product("spmv_csc_block", n, csc_rows(n, Ap, Ai, 0), Ap, Ai, Ax, x, y, c0, c1);
}

void sptrsv_csr(int n, int *Lp, int *Li, double *Lx, double *x) {
  solve_csr_host(n, Lp, Li, Lx, x);
}

void sptrsv_csr_levelset(int n, const int *Lp, const int *Li, const double *Lx,
                         double *x, int levels, const int *levelPtr,
                         const int *levelSet) {
  check_levels(n, Lp, Li, levels, levelPtr, levelSet);
  solve_csr_host(n, Lp, Li, Lx, x);
}

void sptrsv_csr_lbc(int n, int *Lp, int *Li, double *Lx, double *x,
                    int level_no, int *level_ptr, int *par_ptr,
                    int *partition) {
  if (n <= 0) return;
  symamd_handle h = handle();
  const int nnz = Lp[n];
  const int nparts = level_no > 0 ? level_ptr[level_no] : 0;
  DevBuf<int> dLp(n + 1, Lp), dLi(nnz, Li), dpp(nparts + 1, par_ptr),
      dpart(n, partition);
  DevBuf<double> dLx(nnz, Lx), dx(n, x);
  check(symamd_solve_lbc(h, n, dLp.p, dLi.p, dLx.p, dx.p, level_no, level_ptr,
                         dpp.p, dpart.p),
        "sptrsv_csr_lbc");
  hcheck(hipDeviceSynchronize(), "sptrsv_csr_lbc");
  dx.download(x, n);
}

void sptrsv_csr_lbc_redundant(int n, int *Lp, int *Li, double *Lx, double *x,
                              double *b, int level_no, int *level_ptr,
                              int *par_ptr, int *partition, double *tmp) {
  (void)tmp;
  if (n <= 0) return;
  symamd_handle h = handle();
  const int nnz = Lp[n];
  const int nparts = level_no > 0 ? level_ptr[level_no] : 0;
  DevBuf<int> dLp(n + 1, Lp), dLi(nnz, Li), dpp(nparts + 1, par_ptr),
      dpart(n, partition);
  DevBuf<double> dLx(nnz, Lx), dx(n, x), db(n, b);
  check(symamd_solve_lbc_redundant(h, n, dLp.p, dLi.p, dLx.p, dx.p, db.p, level_no,
                                   level_ptr, dpp.p, dpart.p),
        "sptrsv_csr_lbc_redundant");
  hcheck(hipDeviceSynchronize(), "sptrsv_csr_lbc_redundant");
  dx.download(x, n);
}

void sptrsv_csr_ilu(int n, const int *Lp, const int *Li, const double *Lx,
                    const int *diags, double *x) {
  if (n <= 0) return;
  symamd_handle h = handle();
  const int nnz = Lp[n];
  DevBuf<int> dAp(n + 1, Lp), dAi(nnz, Li), ddiag(n, diags), dLp(n + 1), dLi(nnz);
  DevBuf<double> dAx(nnz, Lx), dLx(nnz), dx(n, x);
  int lnnz = 0;
  check(symamd_ilu_lower(h, n, dAp.p, dAi.p, dAx.p, ddiag.p, 1, dLp.p, dLi.p, dLx.p,
                         &lnnz),
        "sptrsv_csr_ilu");
  Plan pl;
  check(symamd_analyze(h, n, dLp.p, dLi.p, dLx.p, SYMAMD_F64, &pl.p), "sptrsv_csr_ilu");
  check(symamd_plan_set_lanes(pl.p, 1), "set_lanes");
  check(symamd_solve(h, pl.p, dx.p, 1, n, SYMAMD_COL_MAJOR, SYMAMD_EXEC_SYNCFREE),
        "sptrsv_csr_ilu");
  check(symamd_plan_check(h, pl.p), "sptrsv_csr_ilu");
  dx.download(x, n);
}

void sptrsv_csr_ilu_lbc(int n, const int *Lp, const int *Li, const double *Lx,
                        const int *diags, double *x, const int level_no,
                        const int *level_ptr, const int *par_ptr,
                        const int *partition) {
  if (n <= 0) return;
  symamd_handle h = handle();
  const int nnz = Lp[n];
  const int nparts = level_no > 0 ? level_ptr[level_no] : 0;
  DevBuf<int> dAp(n + 1, Lp), dAi(nnz, Li), ddiag(n, diags), dLp(n + 1), dLi(nnz),
      dpp(nparts + 1, par_ptr), dpart(n, partition);
  DevBuf<double> dAx(nnz, Lx), dLx(nnz), dx(n, x);
  int lnnz = 0;
  check(symamd_ilu_lower(h, n, dAp.p, dAi.p, dAx.p, ddiag.p, 0, dLp.p, dLi.p, dLx.p,
                         &lnnz),
        "sptrsv_csr_ilu_lbc");
  check(symamd_solve_lbc(h, n, dLp.p, dLi.p, dLx.p, dx.p, level_no, level_ptr, dpp.p,
                         dpart.p),
        "sptrsv_csr_ilu_lbc");
  hcheck(hipDeviceSynchronize(), "sptrsv_csr_ilu_lbc");
  dx.download(x, n);
}

void sptrsv_csc(int n, int *Lp, int *Li, double *Lx, double *x) {
  if (n <= 0) return;
  std::vector<int> Rp, Ri;
  std::vector<double> Rx;
  csc_to_csr(n, Lp, Li, Lx, Rp, Ri, Rx);
  solve_csr_host(n, Rp.data(), Ri.data(), Rx.data(), x);
}

void sptrsv_csc_levelset(int n, int *Lp, int *Li, double *Lx, double *x,
                         int levels, const int *levelptr, const int *levelset) {
  if (n <= 0) return;
  std::vector<int> Rp, Ri;
  std::vector<double> Rx;
  csc_to_csr(n, Lp, Li, Lx, Rp, Ri, Rx);
  check_levels(n, Rp.data(), Ri.data(), levels, levelptr, levelset);
  solve_csr_host(n, Rp.data(), Ri.data(), Rx.data(), x);
}

void sptrsv_csc_lbc(int n, int *Lp, int *Li, double *Lx, double *x,
                    int level_no, int *level_ptr, int *par_ptr,
                    int *partition) {
  if (n <= 0) return;
  std::vector<int> Rp, Ri;
  std::vector<double> Rx;
  csc_to_csr(n, Lp, Li, Lx, Rp, Ri, Rx);
  sptrsv_csr_lbc(n, Rp.data(), Ri.data(), Rx.data(), x, level_no, level_ptr,
                 par_ptr, partition);
}

int ltsolve(int n, int *Lp, int *Li, double *Lx, double *x) {
  if (!Lp || !Li || !x) return 0;
  if (n <= 0) return 1;
  std::vector<int> Rp, Ri;
  std::vector<double> Rx;
  csc_to_csr(n, Lp, Li, Lx, Rp, Ri, Rx);
  symamd_handle h = handle();
  Kept *e = analysis(h, n, Rp.data(), Ri.data(), Rx.data(), x);
  // gb = L^-T gx with gx = b, in place; the forward solution argument is
  // not read when no grad-L is asked for
  check(symamd_adjoint(h, e->plan, e->dx, e->dx, e->dx, nullptr, SYMAMD_EXEC_SYNCFREE),
        "ltsolve");
  check(symamd_plan_check(h, e->plan), "ltsolve");
  hcheck(hipMemcpy(x, e->dx, sizeof(double) * n, hipMemcpyDeviceToHost), "download");
  return 1;
}

// ---------------------------------------------------------------- IC(0)
// The CSC lower triangle of A (diagonal first) the reference factors is the
// CSR of the same triangle transposed: symamd_ic0 factors the CSR (diagonal
// last) on the device and the values are mapped back to CSC positions.
namespace {

// IC(0) of the CSC lower triangle (Ap, Ai, Ax) with one of symamd_ic0's
// variants: the factor in CSC order; returns the smallest column whose pivot
// failed the variant's test, or -1.
int ic0_csc(int n, const int *Ap, const int *Ai, const double *Ax, symamd_ic0_variant variant,
            std::vector<double> &out) {
  const int nnz = Ap[n];
  std::vector<int> Rp, Ri;
  std::vector<double> Rx;
  csc_to_csr(n, Ap, Ai, Ax, Rp, Ri, Rx);
  // CSR position of every CSC position (csc_to_csr fills rows in column order)
  std::vector<int> fill(Rp.begin(), Rp.end() - 1), pos(nnz);
  for (int c = 0; c < n; ++c)
    for (int p = Ap[c]; p < Ap[c + 1]; ++p) pos[p] = fill[Ai[p]]++;
  symamd_handle h = handle();
  Kept *e = analysis(h, n, Rp.data(), Ri.data(), Rx.data(), nullptr);
  int bad = -1;
  // the factor overwrites the kept value buffer (every call uploads anew)
  const int rc = symamd_ic0(h, e->plan, variant, e->dLx, e->dLx, &bad);
  if (rc != SYMAMD_ERR_NOT_SPD) check(rc, "ic0");
  hcheck(hipMemcpy(Rx.data(), e->dLx, sizeof(double) * nnz, hipMemcpyDeviceToHost), "download");
  out.resize(nnz);
  for (int p = 0; p < nnz; ++p) out[p] = Rx[pos[p]];
  return bad;
}

// The prune set spic0_csc_left walks must be the row pattern of the strict
// lower part (for column c the s < c with L[c,s] != 0, ascending): the
// factor here is defined by that pattern.
void check_prune(int n, const int *Ap, const int *Ai, const int *prunePtr,
                 const int *pruneSet) {
  if (!prunePtr || (Ap[n] > n && !pruneSet) || prunePtr[0] != 0)
    throw std::invalid_argument("spic0_csc_left: bad prune set");
  std::vector<int> cnt(n + 1, 0);
  for (int c = 0; c < n; ++c)
    for (int p = Ap[c] + 1; p < Ap[c + 1]; ++p) cnt[Ai[p] + 1]++;
  for (int c = 0; c < n; ++c) {
    cnt[c + 1] += cnt[c];
    if (prunePtr[c + 1] != cnt[c + 1])
      throw std::invalid_argument(
          "spic0_csc_left: the prune set must be the row pattern of the strict lower part");
  }
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (int c = 0; c < n; ++c)
    for (int p = Ap[c] + 1; p < Ap[c + 1]; ++p)
      if (pruneSet[fill[Ai[p]]++] != c)
        throw std::invalid_argument(
            "spic0_csc_left: the prune set must be the row pattern of the strict lower "
            "part, ascending");
}

}  // namespace

bool spic0_csc_left(int n, const int *Ap, const int *Ai, const double *Ax,
                    const int *prunePtr, const int *pruneSet, double *Lx) {
  if (n <= 0) return true;
  check_prune(n, Ap, Ai, prunePtr, pruneSet);
  std::vector<double> f;
  const int bad = ic0_csc(n, Ap, Ai, Ax, SYMAMD_IC0_LEFT, f);
  // the reference returns false at the first failed pivot, having written
  // the columns before it (a prefix of the CSC arrays) and nothing after
  const int end = bad < 0 ? Ap[n] : Ap[bad];
  std::memcpy(Lx, f.data(), sizeof(double) * (size_t)end);
  return bad < 0;
}

bool spic0_csc_left_LBC(int n, const int *Ap, const int *Ai, const double *Ax,
                        const int *prunePtr, const int *pruneSet, const int nLevels,
                        const int *levelPtr, const int *parPtr, const int *partition,
                        double *Lx) {
  (void)nLevels;
  (void)levelPtr;
  (void)parPtr;
  (void)partition;
  if (n <= 0) return true;
  check_prune(n, Ap, Ai, prunePtr, pruneSet);
  std::vector<double> f;
  // the LBC routine computes every column and reports a failed pivot
  const int bad = ic0_csc(n, Ap, Ai, Ax, SYMAMD_IC0_LEFT, f);
  std::memcpy(Lx, f.data(), sizeof(double) * (size_t)Ap[n]);
  return bad < 0;
}

void spic0_csr(int n, double *val, int *rowPtr, int *rowIdx) {
  if (n <= 0) return;
  std::vector<double> f;
  // every column is computed (the reference asserts pivot > 0)
  (void)ic0_csc(n, rowPtr, rowIdx, val, SYMAMD_IC0_CSR, f);
  std::memcpy(val, f.data(), sizeof(double) * (size_t)rowPtr[n]);
}

void spic0_csr_lbc(int n, double *val, int *rowPtr, int *colIdx, int level_no,
                   int *level_ptr, int *par_ptr, int *partition) {
  (void)level_no;
  (void)level_ptr;
  (void)par_ptr;
  (void)partition;
  spic0_csr(n, val, rowPtr, colIdx);
}

// ---------------------------------------------------------------- ILU(0)
// spilu0_csr copies A into l and factors l in place; every row is computed
// (the reference asserts that each U[i,i] is non-zero).
void spilu0_csr(int n, int nnz, const int *Ap, const int *Ai, const double *Ax,
                const int *A_diag, double *l) {
  if (n <= 0) {
    if (nnz > 0) std::memcpy(l, Ax, sizeof(double) * (size_t)nnz);
    return;
  }
  symamd_handle h = handle();
  const size_t total = (size_t)(nnz > Ap[n] ? nnz : Ap[n]);
  DevBuf<int> dAp(n + 1, Ap), dAi(Ap[n], Ai), dd(n, A_diag);
  DevBuf<double> dl(total, Ax);
  int bad = -1;
  const int rc = symamd_ilu0(h, n, dAp.p, dAi.p, dl.p, dd.p, dl.p, &bad);
  if (rc != SYMAMD_ERR_ZERO_PIVOT) check(rc, "spilu0_csr");
  dl.download(l, (size_t)nnz);
}

void spilu0_csr_lbc(int n, int nnz, const int *Ap, const int *Ai, const double *Ax,
                    const int *A_diag, double *l, int level_no, const int *level_ptr,
                    const int *par_ptr, const int *partition, int *tempvecs) {
  (void)level_no;
  (void)level_ptr;
  (void)par_ptr;
  (void)partition;
  (void)tempvecs;
  spilu0_csr(n, nnz, Ap, Ai, Ax, A_diag, l);
}

// ------------------------------------------------------------ supernodal
namespace {
// One supernodal solve on the device: the structure is analysed from the
// host arrays, Lx and the k column-major right-hand sides (leading dimension
// n; k = 1: a vector) are copied in, forward (L X = B) or transposed
// (L^T X = B), and X is copied back.
void sn_device_solve(const char *who, int n, const int *Lp, const int *Li, const int *nrows,
                     const double *Lx, double *x, int num_nodes, const int *supernodes, int k,
                     bool transposed) {
  if (n <= 0 || num_nodes <= 0 || k <= 0) return;
  symamd_handle h = handle();
  int64_t vals = 0;
  for (int i = 0; i < num_nodes; ++i)
    vals = std::max<int64_t>(vals, (int64_t)Lp[i] +
                                       (int64_t)nrows[i] * (supernodes[i + 1] - supernodes[i]));
  symamd_snplan plan = nullptr;
  check(symamd_sn_analyze(h, n, num_nodes, supernodes, nrows, Lp, Li, &plan), who);
  struct Guard {
    symamd_snplan p;
    ~Guard() { symamd_sn_destroy(p); }
  } guard{plan};
  const size_t xs = (size_t)n * (size_t)k;
  DevBuf<double> dLx((size_t)vals, Lx), dx(xs, x);
  if (transposed)
    check(symamd_sn_solve_transposed(h, plan, dLx.p, dx.p, k, n), who);
  else if (k == 1)
    check(symamd_sn_solve(h, plan, dLx.p, dx.p), who);
  else
    check(symamd_sn_solve_block(h, plan, dLx.p, dx.p, k, n), who);
  check(symamd_sn_check(h, plan), who);
  dx.download(x, xs);
}
}  // namespace

void sptrsv_csc_block(int n, const int *Lp, const int *Li, const int *nrows, double *Lx,
                      double *x, int num_nodes, const int *supernodes) {
  sn_device_solve("sptrsv_csc_block", n, Lp, Li, nrows, Lx, x, num_nodes, supernodes, 1, false);
}

// The level schedule and per-thread temporaries are not needed (the
// reference scatters with atomics under it; the result here is the serial
// routine's).
void sptrsv_csc_levelset_block(int n, const int *Lp, const int *Li, const int *nrows,
                               double *Lx, double *x, const int *levelPtr, const int *levels,
                               int n_lev, const int *supernodes, const int *sup2node,
                               double **tempvecs) {
  (void)tempvecs;
  if (n_lev < 0 || (n_lev > 0 && (!levelPtr || !levels || !sup2node || !supernodes)))
    throw std::invalid_argument("sptrsv_csc_levelset_block: bad schedule arguments");
  if (n_lev == 0) return;
  // the schedule must name supernodes 0 .. m-1 once each (levels[j] being
  // supernode sup2node[j]'s first column); the serial order then gives the
  // result of any valid level schedule
  const int j0 = levelPtr[0], j1 = levelPtr[n_lev];
  if (j0 < 0 || j1 < j0)
    throw std::invalid_argument("sptrsv_csc_levelset_block: bad level pointers");
  const int m = j1 - j0;
  std::vector<char> seen(m > 0 ? m : 1, 0);
  for (int i = 0; i < n_lev; ++i)
    if (levelPtr[i + 1] < levelPtr[i])
      throw std::invalid_argument("sptrsv_csc_levelset_block: bad level pointers");
  for (int j = j0; j < j1; ++j) {
    const int s = sup2node[j];
    if (s < 0 || s >= m || seen[s] || levels[j] != supernodes[s])
      throw std::invalid_argument(
          "sptrsv_csc_levelset_block: the schedule must list every supernode once, "
          "with levels[j] == supernodes[sup2node[j]]");
    seen[s] = 1;
  }
  sptrsv_csc_block(n, Lp, Li, nrows, Lx, x, m, supernodes);
}

int build_levelSet_CSC(int n, const int *Lp, const int *Li, int *&level_ptr,
                       int *&level_set) {
  level_ptr = new int[n + 1];
  level_set = new int[n > 0 ? n : 1];
  level_ptr[0] = 0;
  if (n <= 0) return 0;
  std::vector<double> dummy(Lp[n], 1.0);
  std::vector<int> Rp, Ri;
  std::vector<double> Rx;
  csc_to_csr(n, Lp, Li, dummy.data(), Rp, Ri, Rx);
  symamd_handle h = handle();
  DevBuf<int> dLp(n + 1, Rp.data()), dLi(Rp[n], Ri.data()), dlp(n + 1), dls(n);
  int levels = 0;
  check(symamd_levelset(h, n, dLp.p, dLi.p, dlp.p, dls.p, &levels), "build_levelSet_CSC");
  dlp.download(level_ptr, levels + 1);
  dls.download(level_set, n);
  return levels;
}

// The coarsened-schedule inspector of the un-vendored aggregation submodule,
// with the arguments its call site passes (demo/sptrsv_demo_utils.h:83-88:
// the CSR arrays of L, its storage type, the three schedule outputs, then
// lp / cp / ic and the per-row cost).  The schedule is built on the device
// (symamd_coarsen); stype is accepted and not read: the arrays are the CSR
// of the lower triangle, as that call site passes them.  The three arrays
// are new[]'d for the caller to delete[].  Returns 1.
int get_coarse_levelSet_DAG_CSC_tree(size_t n, const int *lC, const int *lR, int stype,
                                     int &finaLevelNo, int *&finaLevelPtr, int &partNo,
                                     int *&finalPartPtr, int *&finalNodePtr, int innerParts,
                                     int minLevelDist, int divRate, double *nodeCost) {
  (void)stype;
  if (n > (size_t)0x7ffffffe)
    throw std::invalid_argument("get_coarse_levelSet_DAG_CSC_tree: n too large");
  const int nn = (int)n;
  symamd_handle h = handle();
  DevBuf<int> dLp(nn + 1, lC), dLi(nn > 0 ? lC[nn] : 0, lR), dlp(nn + 1), dpp(nn + 1),
      dpart(nn);
  DevBuf<double> dcost(nodeCost ? nn : 0, nodeCost);
  check(symamd_coarsen(h, nn, dLp.p, dLi.p, nodeCost ? dcost.p : nullptr, innerParts,
                       minLevelDist, divRate, dlp.p, dpp.p, dpart.p, &finaLevelNo, &partNo),
        "get_coarse_levelSet_DAG_CSC_tree");
  finaLevelPtr = new int[nn + 1];
  finalPartPtr = new int[nn + 1];
  finalNodePtr = new int[nn > 0 ? nn : 1];
  dlp.download(finaLevelPtr, finaLevelNo + 1);
  dpp.download(finalPartPtr, partNo + 1);
  dpart.download(finalNodePtr, nn);
  return 1;
}

namespace parsy {

// Host-only: the coarsening never touches the device.  The arrays are
// allocated with new[] as the reference's are (the caller deletes them);
// parLevelSet and blk2col are not used by the reference either.
int getCoarseLevelSet_6(size_t n, const int *eTree, const int *blk2col, int &finaLevelNo,
                        int *&finaLevelPtr, int *&parLevelSet, int &partNo,
                        int *&finalPartPtr, int *&finalNodePtr, int innerParts,
                        int minLevelDist, int divRate, double *nodeCost) {
  (void)blk2col;
  (void)parLevelSet;
  if (n > (size_t)0x7ffffffe) throw std::invalid_argument("getCoarseLevelSet_6: n too large");
  const int nn = (int)n;
  finaLevelPtr = new int[nn + 1];
  finalPartPtr = new int[nn + 1];
  finalNodePtr = new int[nn > 0 ? nn : 1];
  const int rc = symamd_tree_coarsen(nn, eTree, nodeCost, innerParts, minLevelDist, divRate,
                                     finaLevelPtr, finalPartPtr, finalNodePtr, &finaLevelNo,
                                     &partNo);
  if (rc == SYMAMD_ERR_STRUCTURE)
    throw std::invalid_argument("getCoarseLevelSet_6: eTree has a cycle");
  if (rc != SYMAMD_OK)
    throw std::invalid_argument(
        "getCoarseLevelSet_6: bad parent index or NaN cost, or divRate < 2 with innerParts > 1");
  return finaLevelNo;
}

// parsy's supernodal storage (Triangular_BCSC.cpp:24-56): per-column value
// and row-list offsets (size_t), a supernode's row list at Li[Li_ptr[c0]],
// supernode i owning columns sup2col[i] .. sup2col[i+1]-1.  Converted to
// sym_lib's block storage (the row lists move to the value offsets that
// format indexes them by); the values stay where they are.
namespace {
struct ParsyBlocks {
  std::vector<int> sup, nrows, off, rows_at;
};

ParsyBlocks parsy_blocks(const char *who, int n, const size_t *Lp, const int *Li,
                         const size_t *Li_ptr, const int *sup2col, int supNo) {
  ParsyBlocks b;
  b.sup.resize(supNo + 1);
  b.nrows.resize(supNo);
  b.off.resize(supNo);
  size_t span = 0;
  for (int i = 0; i < supNo; ++i) {
    const int c0 = sup2col[i], c1 = sup2col[i + 1];
    if (c0 < 0 || c1 <= c0 || c1 > n)
      throw std::invalid_argument(std::string(who) + ": bad sup2col");
    const size_t rows = Li_ptr[c1] - Li_ptr[c0];
    if (rows < (size_t)(c1 - c0) || Lp[c0] + rows > (size_t)0x7fffffff)
      throw std::invalid_argument(std::string(who) +
                                  ": bad row count or more than 2^31-1 values");
    b.sup[i] = c0;
    b.nrows[i] = (int)rows;
    b.off[i] = (int)Lp[c0];
    span = std::max(span, (size_t)Lp[c0] + rows);
  }
  b.sup[supNo] = supNo > 0 ? sup2col[supNo] : 0;
  b.rows_at.assign(span > 0 ? span : 1, 0);
  for (int i = 0; i < supNo; ++i)
    for (int r = b.sup[i + 1] - b.sup[i]; r < b.nrows[i]; ++r)
      b.rows_at[(size_t)b.off[i] + r] = Li[Li_ptr[b.sup[i]] + r];
  return b;
}

// k column-major right-hand sides of leading dimension n, forward or
// transposed.  0 for a null Lp, Li or x, as the reference returns.
int parsy_solve(const char *who, int n, const size_t *Lp, const int *Li, const double *Lx,
                const size_t *Li_ptr, const int *sup2col, int supNo, double *x, int k,
                bool transposed) {
  if (!Lp || !Li || !x) return 0;
  if (supNo < 0 || k < 0 || !Li_ptr || !sup2col || !Lx)
    throw std::invalid_argument(std::string(who) + ": bad arguments");
  ParsyBlocks b = parsy_blocks(who, n, Lp, Li, Li_ptr, sup2col, supNo);
  sn_device_solve(who, n, b.off.data(), b.rows_at.data(), b.nrows.data(), Lx, x, supNo,
                  b.sup.data(), k, transposed);
  return 1;
}

// The leveled variants take a getCoarseLevelSet_6 schedule of the supernodes
// (levelPtr / parPtr / partition); it must list every supernode once, as the
// reference needs for a complete solve.
void check_parsy_schedule(const char *who, int supNo, int levels, const int *levelPtr,
                          const int *parPtr, const int *partition) {
  if (levels < 0 || (levels > 0 && (!levelPtr || !parPtr || !partition)))
    throw std::invalid_argument(std::string(who) + ": bad schedule arguments");
  std::vector<char> seen(supNo > 0 ? supNo : 1, 0);
  long listed = 0;
  for (int l = 0; l < levels; ++l)
    for (int p = levelPtr[l]; p < levelPtr[l + 1]; ++p)
      for (int k = parPtr[p]; k < parPtr[p + 1]; ++k) {
        const int s = partition[k];
        if (s < 0 || s >= supNo || seen[s])
          throw std::invalid_argument(std::string(who) +
                                      ": the schedule must list every supernode once");
        seen[s] = 1;
        ++listed;
      }
  if (listed != supNo)
    throw std::invalid_argument(std::string(who) +
                                ": the schedule must list every supernode once");
}
}  // namespace

// Forward solve (Triangular_BCSC.cpp:23-60).  Its dense kernels
// (parsy/triangularSolve/BLAS.cpp) order every operation as sym_lib's
// lsolve_BLAS / matvec_BLAS do, so the solve is sptrsv_csc_block on the same
// blocks, bit for bit.  col2sup and NNZ are not read (as in the reference);
// the supernodes must cover the n columns.
int blockedLsolve(int n, size_t *Lp, int *Li, double *Lx, int NNZ, size_t *Li_ptr,
                  int *col2sup, int *sup2col, int supNo, double *x) {
  (void)NNZ;
  (void)col2sup;
  return parsy_solve("blockedLsolve", n, Lp, Li, Lx, Li_ptr, sup2col, supNo, x, 1, false);
}

// n_rhs column-major right-hand sides x[r + j*n] (Triangular_BCSC.cpp:62-121),
// each column bit-identical to blockedLsolve on it.  The reference's own
// multi-column code departs from that intent in both of its builds: each
// passes a transposed flag to the diagonal block's dtrsm (CblasTrans at :85,
// "T" at :88), which solves with the block's transpose in a forward solve,
// and the non-CBLAS build solves only the first column there (:88, ione).
// This entry point solves every column of L X = B.  max_col only strides the
// reference's scratch and is not read.
int blockedLsolve_mrhs(int n, size_t *Lp, int *Li, double *Lx, int NNZ, size_t *Li_ptr,
                       int *col2sup, int *sup2col, int supNo, double *x, int n_rhs,
                       int max_col) {
  (void)NNZ;
  (void)col2sup;
  (void)max_col;
  return parsy_solve("blockedLsolve_mrhs", n, Lp, Li, Lx, Li_ptr, sup2col, supNo, x, n_rhs,
                     false);
}

// Backward solve L^T x = b (Triangular_BCSC.cpp:124-184), supernodes from
// the last: each top column's value minus the products of the block below
// with the solved values of its rows, then the top's transposed triangular
// solve.  The reference hands both steps to the BLAS it links (dgemv "T",
// dtrsm "L","L","T"), whose operation order is the BLAS's; here the order is
// the library's (symamd_sn_solve_transposed), the same in every column, and
// x agrees with the reference's to rounding.
int blockedLTsolve(int n, size_t *Lp, int *Li, double *Lx, int NNZ, size_t *Li_ptr,
                   int *col2sup, int *sup2col, int supNo, double *x) {
  (void)NNZ;
  (void)col2sup;
  return parsy_solve("blockedLTsolve", n, Lp, Li, Lx, Li_ptr, sup2col, supNo, x, 1, true);
}

// m_rhs column-major right-hand sides of L^T X = B (Triangular_BCSC.cpp:
// 186-249), each column as blockedLTsolve solves it.  The reference's CBLAS
// build solves only the first column in the diagonal block's dtrsm (:233,
// ione); this entry point solves every column.  max_col is not read.
int blockedLTsolve_mrhs(int n, size_t *Lp, int *Li, double *Lx, int NNZ, size_t *Li_ptr,
                        int *col2sup, int *sup2col, int supNo, double *x, int m_rhs,
                        int max_col) {
  (void)NNZ;
  (void)col2sup;
  (void)max_col;
  return parsy_solve("blockedLTsolve_mrhs", n, Lp, Li, Lx, Li_ptr, sup2col, supNo, x, m_rhs,
                     true);
}

// The leveled variant (Triangular_BCSC.cpp:512-579): the reference solves the
// supernodes of each part of a coarse level in parallel and subtracts their
// products from x with atomics, so its subtraction order -- and the last
// bits of x -- follow the thread schedule.  Here the same supernodes are
// solved by the dependency-driven kernel in the serial order of
// blockedLsolve (x within rounding of the reference's).  levelSet and chunk
// are not read by the reference either.
int H2LeveledBlockedLsolve(int n, size_t *Lp, int *Li, double *Lx, int NNZ, size_t *Li_ptr,
                           int *col2sup, int *sup2col, int supNo, double *x, int levels,
                           int *levelPtr, int *levelSet, int parts, int *parPtr,
                           int *partition, int chunk) {
  (void)levelSet;
  (void)parts;
  (void)chunk;
  if (!Lp || !Li || !x) return 0;
  check_parsy_schedule("H2LeveledBlockedLsolve", supNo, levels, levelPtr, parPtr, partition);
  return blockedLsolve(n, Lp, Li, Lx, NNZ, Li_ptr, col2sup, sup2col, supNo, x);
}

// n_rhs columns on a leveled schedule (Triangular_BCSC.cpp:583-653): x as
// blockedLsolve_mrhs gives it.  As in blockedLsolve_mrhs, the reference's
// builds transpose the diagonal block in its dtrsm (CblasConjTrans at :616,
// "T" at :619) and the non-CBLAS one solves only the first column (:619);
// every column of L X = B is solved here.
int H2LeveledBlockedLsolve_mrhs(int n, size_t *Lp, int *Li, double *Lx, int NNZ,
                                size_t *Li_ptr, int *col2sup, int *sup2col, int supNo,
                                double *x, int n_rhs, int levels, int *levelPtr, int *levelSet,
                                int parts, int *parPtr, int *partition, int chunk,
                                int max_col) {
  (void)levelSet;
  (void)parts;
  (void)chunk;
  if (!Lp || !Li || !x) return 0;
  check_parsy_schedule("H2LeveledBlockedLsolve_mrhs", supNo, levels, levelPtr, parPtr,
                       partition);
  return blockedLsolve_mrhs(n, Lp, Li, Lx, NNZ, Li_ptr, col2sup, sup2col, supNo, x, n_rhs,
                            max_col);
}

// L^T x = b on a leveled schedule (Triangular_BCSC.cpp:741-808; levels
// walked from the last, parts' supernodes from the last): every supernode
// still waits for the supernodes its rows below belong to, so x is
// blockedLTsolve's.
int H2LeveledBlockedLTsolve(int n, size_t *Lp, int *Li, double *Lx, int NNZ, size_t *Li_ptr,
                            int *col2sup, int *sup2col, int supNo, double *x, int levels,
                            int *levelPtr, int *levelSet, int parts, int *parPtr,
                            int *partition, int chunk) {
  (void)levelSet;
  (void)parts;
  (void)chunk;
  if (!Lp || !Li || !x) return 0;
  check_parsy_schedule("H2LeveledBlockedLTsolve", supNo, levels, levelPtr, parPtr, partition);
  return blockedLTsolve(n, Lp, Li, Lx, NNZ, Li_ptr, col2sup, sup2col, supNo, x);
}

// m_rhs columns of L^T X = B on a leveled schedule (Triangular_BCSC.cpp:
// 810-880).  The reference numbers the supernode of partition entry i as
// columns sup2col[i-1] .. sup2col[i]-1 there (:836-837), unlike the other
// leveled solves (sup2col[i] .. sup2col[i+1]-1), so on a schedule of
// supernodes 0 .. supNo-1 it skips the last supernode; its CBLAS build also
// updates and solves only the first column (dgemv at :858, dtrsm with ione
// at :871).  This entry point numbers supernodes as the others do and
// solves every column: x as blockedLTsolve_mrhs gives it.
int H2LeveledBlockedLTsolve_mrhs(int n, size_t *Lp, int *Li, double *Lx, int NNZ,
                                 size_t *Li_ptr, int *col2sup, int *sup2col, int supNo,
                                 double *x, int m_rhs, int levels, int *levelPtr,
                                 int *levelSet, int parts, int *parPtr, int *partition,
                                 int chunk, int max_col) {
  (void)levelSet;
  (void)parts;
  (void)chunk;
  if (!Lp || !Li || !x) return 0;
  check_parsy_schedule("H2LeveledBlockedLTsolve_mrhs", supNo, levels, levelPtr, parPtr,
                       partition);
  return blockedLTsolve_mrhs(n, Lp, Li, Lx, NNZ, Li_ptr, col2sup, sup2col, supNo, x, m_rhs,
                             max_col);
}

}  // namespace parsy
}  // namespace sym_lib
