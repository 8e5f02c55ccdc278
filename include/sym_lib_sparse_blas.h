/*
 * sym_lib_sparse_blas.h -- the reference's sparse triangular solve entry
 * points, re-exported by libsympiler_amd.so with the same C++ signatures
 * (namespace sym_lib, C++ linkage, same mangled names), so code written
 * against include/sympiler/sparse_blas_lib.h relinks unchanged.
 *
 * Semantics follow the reference: blocking, stateless, host arrays owned by
 * the caller, x read as the right-hand side and overwritten with the
 * solution.  Underneath, each call copies the arrays to the GPU, runs the
 * HIP kernels through the C ABI of sympiler_amd.h and copies x back.  The
 * solve and SpMV results are bit-identical to the reference's serial
 * routines (one lane per row, no FMA contraction).  Where the reference
 * returns void, a failure (bad structure, zero pivot, HIP error) throws
 * std::runtime_error with the library's message.
 */
#ifndef SYM_LIB_SPARSE_BLAS_H
#define SYM_LIB_SPARSE_BLAS_H
#ifdef __cplusplus
#include <cstddef>

namespace sym_lib {

/* sparse_blas_lib.h:18-20 */
void spmv_csr(int n, const int *Ap, const int *Ai, const double *Ax,
              const double *x, double *y);
/* sparse_blas_lib.h:32-33 */
void spmv_csr(int n, const int *Ap, const int *Ai, const double *Ax,
              const double *x, const double *y, double *z, double a, double b);
/* sparse_blas_lib.h:35-37: the same products as the 6-argument spmv_csr. */
void spmv_csr_parallel(int n, const int *Ap, const int *Ai, const double *Ax,
                       const double *x, double *y);
/* sparse_blas_lib.h:39-41: y[j] = row j of A times x for the rows of the
 * supernodes (rows supernodes[0] .. supernodes[nodes]-1), each from +0.0. */
void spmv_csr_block(int n, const int *Ap, const int *Ai, const double *Ax, int nodes,
                    const int *supernodes, const double *x, double *y);
/* sparse_blas_lib.h:50-52: y += A x for a CSC matrix (n columns); y has as
 * many entries as A has rows (the largest row index + 1). */
void spmv_csc(int n, const int *Ap, const int *Ai, const double *Ax, const double *x,
              double *y);
/* sparse_blas_lib.h:64-65: z += A x, then z = a*z + b*y (square A). */
void spmv_csc(int n, const int *Ap, const int *Ai, const double *Ax, const double *x,
              const double *y, double *z, double a, double b);
/* sparse_blas_lib.h:67-69: the same with per-row a[i], b[i]. */
void spmv_csc(int n, const int *Ap, const int *Ai, const double *Ax, const double *x,
              const double *y, double *z, double *a, double *b);
/* sparse_blas_lib.h:82-85: part l takes the columns ids[sets[l] ..
 * sets[l+1]), its products summed from +0.0, and y (n entries, square A) +=
 * the parts' sums in part order -- bit-identical to the reference. */
void spmv_csc_lbc(int n_parts, const int *sets, const int *ids, int n, const int *Ap,
                  const int *Ai, const double *Ax, const double *x, double *y);
/* sparse_blas_lib.h:86-89: the same with n_parts = set_ptr[1]. */
void spmv_csc_lbc(const int *set_ptr, const int *sets, const int *ids, int n, const int *Ap,
                  const int *Ai, const double *Ax, const double *x, double *y);
/* sparse_blas_lib.h:98-99: the reference adds with atomics in thread order;
 * here y is spmv_csc's (column order), equal to the reference's to rounding. */
void spmv_csc_parallel(int n, const int *Ap, const int *Ai, const double *Ax,
                       const double *x, double *y);
/* sparse_blas_lib.h:111-114: y += the columns supernodes[0] ..
 * supernodes[nodes]-1 of A times x, column by column. */
void spmv_csc_block(int n, const int *Ap, const int *Ai, const double *Ax, const double *x,
                    double *y, int nodes, const int *supernodes);

/* sparse_blas_lib.h:126-127 */
void sptrsv_csr(int n, int *Lp, int *Li, double *Lx, double *x);
/* sparse_blas_lib.h:145-147; the level arrays are checked for size and
 * validity; the GPU runs its own level schedule (results are identical for
 * every valid schedule). */
void sptrsv_csr_levelset(int n, const int *Lp, const int *Li, const double *Lx,
                         double *x, int levels, const int *levelPtr,
                         const int *levelSet);
/* sparse_blas_lib.h:150-152; executes the given coarsened schedule. */
void sptrsv_csr_lbc(int n, int *Lp, int *Li, double *Lx, double *x,
                    int level_no, int *level_ptr, int *par_ptr,
                    int *partition);
/* sparse_blas_lib.h:156-161; b is the right-hand side, x receives the
 * solution; tmp (threads*n doubles in the reference) is accepted and not
 * written -- its contents after the reference call depend on its thread
 * schedule. */
void sptrsv_csr_lbc_redundant(int n, int *Lp, int *Li, double *Lx, double *x,
                              double *b, int level_no, int *level_ptr,
                              int *par_ptr, int *partition, double *tmp);
/* sparse_blas_lib.h:137-138: unit-lower solve with the lower part of a
 * general CSR matrix (entries [Lp[i], diags[i]) of each row); like the
 * reference it does not divide. */
void sptrsv_csr_ilu(int n, const int *Lp, const int *Li, const double *Lx,
                    const int *diags, double *x);
/* sparse_blas_lib.h:139-142: the same rows on a caller schedule; like the
 * reference this variant divides by Lx[diags[i]]. */
void sptrsv_csr_ilu_lbc(int n, const int *Lp, const int *Li, const double *Lx,
                        const int *diags, double *x, const int level_no,
                        const int *level_ptr, const int *par_ptr,
                        const int *partition);
/* sparse_blas_lib.h:169-170 (CSC of L, diagonal first) */
void sptrsv_csc(int n, int *Lp, int *Li, double *Lx, double *x);
/* sparse_blas_lib.h:174-175 */
void sptrsv_csc_levelset(int n, int *Lp, int *Li, double *Lx, double *x,
                         int levels, const int *levelptr, const int *levelset);
/* sparse_blas_lib.h:177-178 */
void sptrsv_csc_lbc(int n, int *Lp, int *Li, double *Lx, double *x,
                    int level_no, int *level_ptr, int *par_ptr,
                    int *partition);
/* sparse_blas_lib.h:323-330: L^T x = b from the CSC of L; returns 0 on
 * null input, 1 otherwise. */
int ltsolve(int n, int *Lp, int *Li, double *Lx, double *x);

/* sparse_blas_lib.h:242-245: IC(0) of the CSC lower triangle (diagonal
 * first), in place; every column is computed (the reference asserts that
 * each pivot is positive).  Results are the serial spic0_csr's bit for bit;
 * the _lbc variant (whose reference applies its updates with atomics in a
 * thread-dependent order) returns the same factor, its schedule unused. */
void spic0_csr(int n, double *val, int *rowPtr, int *rowIdx);
void spic0_csr_lbc(int n, double *val, int *rowPtr, int *colIdx, int level_no,
                   int *level_ptr, int *par_ptr, int *partition);
/* sparse_blas_lib.h:256-265: left-looking IC(0) of the CSC lower triangle of
 * A into Lx.  The prune set must be the row pattern of the strict lower part,
 * ascending (std::invalid_argument otherwise).  spic0_csc_left returns false
 * at the first pivot <= 0 having written only the columns before it; the LBC
 * variant writes every column (its schedule unused) and returns false if any
 * pivot failed.  Bit-identical to the reference. */
bool spic0_csc_left(int n, const int *Ap, const int *Ai, const double *Ax,
                    const int *prunePtr, const int *pruneSet, double *Lx);
bool spic0_csc_left_LBC(int n, const int *Ap, const int *Ai, const double *Ax,
                        const int *prunePtr, const int *pruneSet, const int nLevels,
                        const int *levelPtr, const int *parPtr, const int *partition,
                        double *Lx);

/* sparse_blas_lib.h:276-281: ILU(0) of a general CSR matrix (A_diag[i] =
 * position of row i's diagonal) into l, bit-identical to the reference; every
 * row is computed (the reference asserts non-zero U diagonals).  The _lbc
 * variant returns the same factor; its schedule and tempvecs are unused. */
void spilu0_csr(int n, int nnz, const int *Ap, const int *Ai, const double *Ax,
                const int *A_diag, double *l);
void spilu0_csr_lbc(int n, int nnz, const int *Ap, const int *Ai, const double *Ax,
                    const int *A_diag, double *l, int level_no, const int *level_ptr,
                    const int *par_ptr, const int *partition, int *tempvecs);

/* sparse_blas_lib.h:196-214: forward solve on a supernodal (BCSC) lower
 * triangle, bit-identical to the serial sptrsv_csc_block; the levelset
 * variant (whose reference scatters with atomics) returns the same x, its
 * schedule and temporaries unused.  Supernodes up to 1024 columns wide. */
void sptrsv_csc_levelset_block(int n, const int *Lp, const int *Li, const int *nrows,
                               double *Lx, double *x, const int *levelPtr, const int *levels,
                               int n_lev, const int *supernodes, const int *sup2node,
                               double **tempvecs);
void sptrsv_csc_block(int n, const int *Lp, const int *Li, const int *nrows, double *Lx,
                      double *x, int num_nodes, const int *supernodes);

/* The level-set inspector of the un-vendored aggregation submodule, with the
 * signature its call sites use (Catch_tests/sptrsv_tests.cpp:59,
 * demo/sptrsv_demo_utils.h:44-45): input the CSC of L, outputs new[]'d for
 * the caller to delete[], returns the number of levels. */
int build_levelSet_CSC(int n, const int *Lp, const int *Li, int *&level_ptr,
                       int *&level_set);

/* The coarsened-schedule (LBC) inspector of the un-vendored aggregation
 * submodule, with the arguments of its only call site
 * (demo/sptrsv_demo_utils.h:83-88; n, lC, lR = the CSR of L, stype, then
 * the outputs and lp = threads, cp = -1, ic = 4000 from
 * demo/sptrsv_demo.cpp:37, and the row nnz as cost, :80-82).  The parameter
 * types follow parsy's getCoarseLevelSet_6 (size_t n, const int arrays), the
 * submodule's declaration being absent.  Built on the device
 * (symamd_coarsen); arrays new[]'d for the caller to delete[]; returns 1. */
int get_coarse_levelSet_DAG_CSC_tree(size_t n, const int *lC, const int *lR, int stype,
                                     int &finaLevelNo, int *&finaLevelPtr, int &partNo,
                                     int *&finalPartPtr, int *&finalNodePtr, int innerParts,
                                     int minLevelDist, int divRate, double *nodeCost);

namespace parsy {
/* parsy/symbolic/InspectionLevel_06.h:13-17: tree-level coarsening of an
 * elimination tree (host only, on symamd_tree_coarsen); the three arrays are
 * new[]'d for the caller to delete[] and equal the reference's.  Throws
 * std::invalid_argument on a cycle, a parent index >= n, a NaN cost or a
 * divRate that the reference would never return from. */
int getCoarseLevelSet_6(size_t n, const int *eTree, const int *blk2col, int &finaLevelNo,
                        int *&finaLevelPtr, int *&parLevelSet, int &partNo,
                        int *&finalPartPtr, int *&finalNodePtr, int innerParts,
                        int minLevelDist, int divRate, double *nodeCost);
/* parsy/triangularSolve/Triangular_BCSC.h:18-19: forward solve on parsy's
 * supernodal storage (per-column size_t value and row-list offsets), on the
 * device through sptrsv_csc_block; x bit-identical to the reference's
 * (its dense kernels order every operation as lsolve_BLAS/matvec_BLAS).
 * Returns 1, or 0 for a null Lp, Li or x as the reference does. */
int blockedLsolve(int n, size_t *Lp, int *Li, double *Lx, int NNZ, size_t *Li_ptr,
                  int *col2sup, int *sup2col, int supNo, double *x);
/* Triangular_BCSC.h:20-22 (impl Triangular_BCSC.cpp:62-121): n_rhs
 * column-major right-hand sides x[r + j*n], each column bit-identical to
 * blockedLsolve on it.  The reference's builds transpose the diagonal block
 * in their dtrsm and one of them solves only the first column; every column
 * of L X = B is solved here.  max_col (the reference's scratch stride) is not
 * read. */
int blockedLsolve_mrhs(int n, size_t *Lp, int *Li, double *Lx, int NNZ, size_t *Li_ptr,
                       int *col2sup, int *sup2col, int supNo, double *x, int n_rhs,
                       int max_col = -1);
/* Triangular_BCSC.h:26-27 (impl Triangular_BCSC.cpp:124-184): L^T x = b on
 * the same storage, supernodes from the last (symamd_sn_solve_transposed;
 * the operation order is the library's where the reference's is its BLAS's,
 * so x agrees with the reference's to rounding). */
int blockedLTsolve(int n, size_t *Lp, int *Li, double *Lx, int NNZ, size_t *Li_ptr,
                   int *col2sup, int *sup2col, int supNo, double *x);
/* Triangular_BCSC.h:28-30 (impl :186-249): m_rhs column-major right-hand
 * sides of L^T X = B, each column as blockedLTsolve solves it (the
 * reference's CBLAS build solves only the first column's diagonal blocks). */
int blockedLTsolve_mrhs(int n, size_t *Lp, int *Li, double *Lx, int NNZ, size_t *Li_ptr,
                        int *col2sup, int *sup2col, int supNo, double *x, int m_rhs,
                        int max_col = -1);
/* Triangular_BCSC.h (impl Triangular_BCSC.cpp:512-579): the leveled variant
 * over a getCoarseLevelSet_6 schedule of the supernodes.  The reference
 * subtracts with atomics in thread order; here x is blockedLsolve's (serial
 * order), within rounding of the reference's.  Throws std::invalid_argument
 * unless the schedule lists every supernode once. */
int H2LeveledBlockedLsolve(int n, size_t *Lp, int *Li, double *Lx, int NNZ, size_t *Li_ptr,
                           int *col2sup, int *sup2col, int supNo, double *x, int levels,
                           int *levelPtr, int *levelSet, int parts, int *parPtr,
                           int *partition, int chunk);
/* Triangular_BCSC.h:80-83 (impl :583-653): n_rhs columns on a leveled
 * schedule, x as blockedLsolve_mrhs gives it; the schedule is checked as
 * H2LeveledBlockedLsolve checks it. */
int H2LeveledBlockedLsolve_mrhs(int n, size_t *Lp, int *Li, double *Lx, int NNZ,
                                size_t *Li_ptr, int *col2sup, int *sup2col, int supNo,
                                double *x, int n_rhs, int levels, int *levelPtr, int *levelSet,
                                int parts, int *parPtr, int *partition, int chunk,
                                int max_col = -1);
/* Triangular_BCSC.h:95-98 (impl :741-808): L^T x = b on a leveled schedule
 * (walked from its last level), x as blockedLTsolve gives it; the schedule is
 * checked as H2LeveledBlockedLsolve checks it. */
int H2LeveledBlockedLTsolve(int n, size_t *Lp, int *Li, double *Lx, int NNZ, size_t *Li_ptr,
                            int *col2sup, int *sup2col, int supNo, double *x, int levels,
                            int *levelPtr, int *levelSet, int parts, int *parPtr,
                            int *partition, int chunk);
/* Triangular_BCSC.h:99-102 (impl :810-880): m_rhs columns of L^T X = B on a
 * leveled schedule, x as blockedLTsolve_mrhs gives it.  The reference numbers
 * partition entry i as the columns sup2col[i-1] .. sup2col[i]-1 here (unlike
 * its other leveled solves), which skips the last supernode; supernodes are
 * numbered here as in the other leveled solves. */
int H2LeveledBlockedLTsolve_mrhs(int n, size_t *Lp, int *Li, double *Lx, int NNZ,
                                 size_t *Li_ptr, int *col2sup, int *sup2col, int supNo,
                                 double *x, int m_rhs, int levels, int *levelPtr,
                                 int *levelSet, int parts, int *parPtr, int *partition,
                                 int chunk, int max_col = -1);
}  // namespace parsy

}  // namespace sym_lib
#endif
#endif
