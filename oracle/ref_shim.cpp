// ref_shim.cpp -- C entry points onto the reference's own compiled
// functions, so tests can load oracle/_ref/libsymref.so with ctypes.
//
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile (target `ref`) together
// with the reference sources where they lie; used by tests/ to pin the C
// restatement in oracle/oracle.c and, optionally, as the cpu_baseline of
// bench.py.  Nothing in the product loads it.
//
// The prototypes below restate the reference declarations in
// include/sympiler/sparse_blas_lib.h (cited per line); that header cannot be
// included as-is because it opens a header of the un-vendored aggregation
// submodule.
#include <cstddef>
#include <cstdint>
#include <omp.h>

// dense_blas/BLAS.cpp:3 and :286 (the kernels of sptrsv_csc_block; global
// namespace in the reference)
void lsolve_BLAS(int dim, int num_col, double *M, double *b);
void matvec_BLAS(int dim, int nrow, int ncol, double *M, double *x, double *b);

namespace sym_lib {
void sptrsv_csr(int n, int *Lp, int *Li, double *Lx, double *x);  // :126-127
void sptrsv_csr_levelset(int n, const int *Lp, const int *Li,
                         const double *Lx, double *x, int levels,
                         const int *levelPtr, const int *levelSet);  // :145-147
void sptrsv_csr_lbc(int n, int *Lp, int *Li, double *Lx, double *x,
                    int level_no, int *level_ptr, int *par_ptr,
                    int *partition);  // :150-152
void sptrsv_csr_lbc_redundant(int n, int *Lp, int *Li, double *Lx, double *x,
                              double *b, int level_no, int *level_ptr,
                              int *par_ptr, int *partition,
                              double *tmp);  // :156-161
void sptrsv_csr_ilu(int n, const int *Lp, const int *Li, const double *Lx,
                    const int *diags, double *x);  // :137-138
void sptrsv_csr_ilu_lbc(int n, const int *Lp, const int *Li, const double *Lx,
                        const int *diags, double *x, const int level_no,
                        const int *level_ptr, const int *par_ptr,
                        const int *partition);  // :139-142
void sptrsv_csc(int n, int *Lp, int *Li, double *Lx, double *x);  // :169-170
void sptrsv_csc_block(int n, const int *Lp, const int *Li, const int *nrows,
                      double *Lx, double *x, int num_nodes,
                      const int *supernodes);  // :212-214
int ltsolve(int n, int *Lp, int *Li, double *Lx, double *x);      // :330
void spmv_csr(int n, const int *Ap, const int *Ai, const double *Ax,
              const double *x, double *y);  // :18-20
void spmv_csr(int n, const int *Ap, const int *Ai, const double *Ax,
              const double *x, const double *y, double *z, double a,
              double b);  // :32-33
void spmv_csr_block(int n, const int *Ap, const int *Ai, const double *Ax, int nodes,
                    const int *supernodes, const double *x, double *y);  // :39-41
void spmv_csc(int n, const int *Ap, const int *Ai, const double *Ax, const double *x,
              double *y);  // :50-52
void spmv_csc(int n, const int *Ap, const int *Ai, const double *Ax, const double *x,
              const double *y, double *z, double a, double b);  // :64-65
void spmv_csc(int n, const int *Ap, const int *Ai, const double *Ax, const double *x,
              const double *y, double *z, double *a, double *b);  // :67-69
void spmv_csc_lbc(int n_parts, const int *sets, const int *ids, int n, const int *Ap,
                  const int *Ai, const double *Ax, const double *x, double *y);  // :82-85
void spmv_csc_block(int n, const int *Ap, const int *Ai, const double *Ax, const double *x,
                    double *y, int nodes, const int *supernodes);  // :111-114
bool spic0_csc_left(int n, const int *Ap, const int *Ai, const double *Ax,
                    const int *prunePtr, const int *pruneSet,
                    double *Lx);  // :256-259
void spic0_csr(int n, double *val, int *rowPtr, int *rowIdx);  // :242
namespace parsy {
// include/sympiler/parsy/triangularSolve/BLAS.h:12-19 (the kernels of
// parsy::blockedLsolve, Triangular_BCSC.cpp:24-56)
void dlsolve_blas_nonUnit(int ldm, int ncol, double *M, double *rhs);
void dmatvec_blas(int ldm, int nrow, int ncol, double *M, double *vec, double *Mxvec);
// include/sympiler/parsy/common/TreeUtils.h (impl TreeUtils.cpp:89-141)
int getLevelSet(size_t n, const int *inTree, int *levelPtr, int *levelSet);
// include/sympiler/parsy/symbolic/InspectionLevel_06.h:13-17
int getCoarseLevelSet_6(size_t n, const int *eTree, const int *blk2col,
                        int &finaLevelNo, int *&finaLevelPtr, int *&parLevelSet,
                        int &partNo, int *&finalPartPtr, int *&finalNodePtr,
                        int innerParts, int minLevelDist, int divRate,
                        double *nodeCost);
}  // namespace parsy
}  // namespace sym_lib

extern "C" {
// Thread count of the reference's OpenMP regions (it reads the caller's
// omp_set_num_threads, sptrsv_demo.cpp:62).
void ref_set_threads(int n) { omp_set_num_threads(n); }

void ref_sptrsv_csr(int n, int *Lp, int *Li, double *Lx, double *x) {
  sym_lib::sptrsv_csr(n, Lp, Li, Lx, x);
}
void ref_sptrsv_csr_levelset(int n, const int *Lp, const int *Li,
                             const double *Lx, double *x, int levels,
                             const int *levelPtr, const int *levelSet) {
  sym_lib::sptrsv_csr_levelset(n, Lp, Li, Lx, x, levels, levelPtr, levelSet);
}
void ref_sptrsv_csr_lbc(int n, int *Lp, int *Li, double *Lx, double *x,
                        int level_no, int *level_ptr, int *par_ptr,
                        int *partition) {
  sym_lib::sptrsv_csr_lbc(n, Lp, Li, Lx, x, level_no, level_ptr, par_ptr,
                          partition);
}
// tmp: omp_get_max_threads() * n doubles, as the reference requires
void ref_sptrsv_csr_lbc_redundant(int n, int *Lp, int *Li, double *Lx,
                                  double *x, double *b, int level_no,
                                  int *level_ptr, int *par_ptr, int *partition,
                                  double *tmp) {
  sym_lib::sptrsv_csr_lbc_redundant(n, Lp, Li, Lx, x, b, level_no, level_ptr,
                                    par_ptr, partition, tmp);
}
int ref_max_threads() { return omp_get_max_threads(); }
void ref_sptrsv_csr_ilu(int n, const int *Lp, const int *Li, const double *Lx,
                        const int *diags, double *x) {
  sym_lib::sptrsv_csr_ilu(n, Lp, Li, Lx, diags, x);
}
void ref_sptrsv_csr_ilu_lbc(int n, const int *Lp, const int *Li,
                            const double *Lx, const int *diags, double *x,
                            int level_no, const int *level_ptr,
                            const int *par_ptr, const int *partition) {
  sym_lib::sptrsv_csr_ilu_lbc(n, Lp, Li, Lx, diags, x, level_no, level_ptr,
                              par_ptr, partition);
}
void ref_sptrsv_csc_block(int n, const int *Lp, const int *Li, const int *nrows,
                          double *Lx, double *x, int num_nodes, const int *supernodes) {
  sym_lib::sptrsv_csc_block(n, Lp, Li, nrows, Lx, x, num_nodes, supernodes);
}
void ref_sptrsv_csc(int n, int *Lp, int *Li, double *Lx, double *x) {
  sym_lib::sptrsv_csc(n, Lp, Li, Lx, x);
}
int ref_ltsolve(int n, int *Lp, int *Li, double *Lx, double *x) {
  return sym_lib::ltsolve(n, Lp, Li, Lx, x);
}
void ref_spmv_csr(int n, const int *Ap, const int *Ai, const double *Ax,
                  const double *x, double *y) {
  sym_lib::spmv_csr(n, Ap, Ai, Ax, x, y);
}
void ref_spmv_csr_axpby(int n, const int *Ap, const int *Ai, const double *Ax,
                        const double *x, const double *y, double *z, double a,
                        double b) {
  sym_lib::spmv_csr(n, Ap, Ai, Ax, x, y, z, a, b);
}
void ref_spmv_csr_block(int n, const int *Ap, const int *Ai, const double *Ax, int nodes,
                        const int *supernodes, const double *x, double *y) {
  sym_lib::spmv_csr_block(n, Ap, Ai, Ax, nodes, supernodes, x, y);
}
void ref_spmv_csc(int n, const int *Ap, const int *Ai, const double *Ax, const double *x,
                  double *y) {
  sym_lib::spmv_csc(n, Ap, Ai, Ax, x, y);
}
void ref_spmv_csc_axpby(int n, const int *Ap, const int *Ai, const double *Ax,
                        const double *x, const double *y, double *z, double a, double b) {
  sym_lib::spmv_csc(n, Ap, Ai, Ax, x, y, z, a, b);
}
void ref_spmv_csc_axpby_vec(int n, const int *Ap, const int *Ai, const double *Ax,
                            const double *x, const double *y, double *z, double *a,
                            double *b) {
  sym_lib::spmv_csc(n, Ap, Ai, Ax, x, y, z, a, b);
}
void ref_spmv_csc_lbc(int n_parts, const int *sets, const int *ids, int n, const int *Ap,
                      const int *Ai, const double *Ax, const double *x, double *y) {
  sym_lib::spmv_csc_lbc(n_parts, sets, ids, n, Ap, Ai, Ax, x, y);
}
void ref_spmv_csc_block(int n, const int *Ap, const int *Ai, const double *Ax,
                        const double *x, double *y, int nodes, const int *supernodes) {
  sym_lib::spmv_csc_block(n, Ap, Ai, Ax, x, y, nodes, supernodes);
}
int ref_spic0_csc_left(int n, const int *Ap, const int *Ai, const double *Ax,
                       const int *prunePtr, const int *pruneSet, double *Lx) {
  return sym_lib::spic0_csc_left(n, Ap, Ai, Ax, prunePtr, pruneSet, Lx) ? 1
                                                                        : 0;
}
void ref_spic0_csr(int n, double *val, int *rowPtr, int *rowIdx) {
  sym_lib::spic0_csr(n, val, rowPtr, rowIdx);
}
int ref_getLevelSet(int n, const int *inTree, int *levelPtr, int *levelSet) {
  return sym_lib::parsy::getLevelSet((size_t)n, inTree, levelPtr, levelSet);
}
void ref_lsolve_BLAS(int dim, int ncol, double *M, double *b) {
  ::lsolve_BLAS(dim, ncol, M, b);
}
void ref_matvec_BLAS(int dim, int nrow, int ncol, double *M, double *x, double *b) {
  ::matvec_BLAS(dim, nrow, ncol, M, x, b);
}
void ref_parsy_dlsolve(int ldm, int ncol, double *M, double *rhs) {
  sym_lib::parsy::dlsolve_blas_nonUnit(ldm, ncol, M, rhs);
}
void ref_parsy_dmatvec(int ldm, int nrow, int ncol, double *M, double *vec, double *out) {
  sym_lib::parsy::dmatvec_blas(ldm, nrow, ncol, M, vec, out);
}
// Copies the three new[]'d arrays out (sizes n+1, n+1, n) and frees them.
void ref_getCoarseLevelSet_6(int n, const int *eTree, double *nodeCost,
                             int innerParts, int minLevelDist, int divRate,
                             int *levelPtr, int *partPtr, int *partition,
                             int *levelNo, int *partNo) {
  int *lp = nullptr, *unused = nullptr, *pp = nullptr, *pt = nullptr;
  int nl = 0, np = 0;
  sym_lib::parsy::getCoarseLevelSet_6((size_t)n, eTree, nullptr, nl, lp, unused,
                                      np, pp, pt, innerParts, minLevelDist,
                                      divRate, nodeCost);
  for (int i = 0; i <= nl; ++i) levelPtr[i] = lp[i];
  for (int i = 0; i <= np; ++i) partPtr[i] = pp[i];
  for (int i = 0; i < pp[np]; ++i) partition[i] = pt[i];
  *levelNo = nl;
  *partNo = np;
  delete[] lp;
  delete[] pp;
  delete[] pt;
}
}
