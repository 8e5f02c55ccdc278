/*
 * oracle.c -- CPU restatement of Sympiler's sparse triangular solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * sympiler_amd/.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py load it, and only to compare against or to time the CPU
 * path.  The product never links, loads or calls it.
 *
 * Every routine follows the operation order of the reference function it
 * cites, so that, built with -ffp-contract=off (no FMA contraction, as the
 * reference's own x86-64 build without -mfma), it reproduces the reference
 * bit for bit.  That claim is checked by tests/test_oracle.py against
 * oracle/_ref (the reference sources compiled by oracle/Makefile) and against
 * the committed golden vectors in tests/golden/.
 *
 * Layout conventions (reference include/sympiler/sparse_blas_lib.h):
 *   CSR lower triangle: diagonal LAST in each row   (sptrsv.cpp:20,23)
 *   CSC lower triangle: diagonal FIRST in each column (sptrsv.cpp:132-133)
 *   int32 indices, FP64 values, x in/out.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- solves */

/* sparse_blas/sptrsv.cpp:17-25 (sym_lib::sptrsv_csr): serial forward
 * substitution, j ascending, multiply then subtract, divide by the
 * last-in-row diagonal. */
void orc_sptrsv_csr(int n, const int *Lp, const int *Li, const double *Lx,
                    double *x) {
  for (int i = 0; i < n; i++) {
    for (int j = Lp[i]; j < Lp[i + 1] - 1; j++) x[i] -= Lx[j] * x[Li[j]];
    x[i] /= Lx[Lp[i + 1] - 1];
  }
}

/* sparse_blas/sptrsv.cpp:40-54 (sym_lib::sptrsv_csr_levelset): the same row
 * arithmetic, rows of one level in parallel, one fork/join per level. */
void orc_sptrsv_csr_levelset(int n, const int *Lp, const int *Li,
                             const double *Lx, double *x, int levels,
                             const int *levelPtr, const int *levelSet,
                             int nthreads) {
  (void)n;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  for (int l = 0; l < levels; l++) {
#pragma omp parallel for default(shared) schedule(static)
    for (int k = levelPtr[l]; k < levelPtr[l + 1]; ++k) {
      int i = levelSet[k];
      for (int j = Lp[i]; j < Lp[i + 1] - 1; j++) x[i] -= Lx[j] * x[Li[j]];
      x[i] /= Lx[Lp[i + 1] - 1];
    }
  }
}

/* sparse_blas/sptrsv.cpp:57-75 (sym_lib::sptrsv_csr_lbc): coarse levels in
 * order, partitions of a coarse level in parallel, rows of a partition in
 * their listed order. */
void orc_sptrsv_csr_lbc(int n, const int *Lp, const int *Li, const double *Lx,
                        double *x, int level_no, const int *level_ptr,
                        const int *par_ptr, const int *partition) {
  (void)n;
  for (int i1 = 0; i1 < level_no; ++i1) {
#pragma omp parallel for schedule(static)
    for (int j1 = level_ptr[i1]; j1 < level_ptr[i1 + 1]; ++j1) {
      for (int k1 = par_ptr[j1]; k1 < par_ptr[j1 + 1]; ++k1) {
        int i = partition[k1];
        for (int j = Lp[i]; j < Lp[i + 1] - 1; j++) x[i] -= Lx[j] * x[Li[j]];
        x[i] /= Lx[Lp[i + 1] - 1];
      }
    }
  }
}

/* sparse_blas/sptrsv.cpp:78-106 (sym_lib::sptrsv_csr_lbc_redundant): the
 * LBC schedule with a separate right-hand side b; each row accumulates its
 * products into a zeroed temporary and adds b[i] last:
 * x[i] = ((0 - sum_j Lx[j]*x[Li[j]]) + b[i]) / diag.  The reference's
 * per-thread tmp array is scratch whose final contents depend on its
 * thread schedule; this restatement keeps the temporary in a register. */
void orc_sptrsv_csr_lbc_redundant(int n, const int *Lp, const int *Li,
                                  const double *Lx, double *x, const double *b,
                                  int level_no, const int *level_ptr,
                                  const int *par_ptr, const int *partition) {
  (void)n;
  for (int i1 = 0; i1 < level_no; ++i1) {
#pragma omp parallel for schedule(static)
    for (int j1 = level_ptr[i1]; j1 < level_ptr[i1 + 1]; ++j1) {
      for (int k1 = par_ptr[j1]; k1 < par_ptr[j1 + 1]; ++k1) {
        int i = partition[k1];
        double t = 0.0;
        for (int j = Lp[i]; j < Lp[i + 1] - 1; j++) t -= Lx[j] * x[Li[j]];
        t = (t + b[i]) / Lx[Lp[i + 1] - 1];
        x[i] = t;
      }
    }
  }
}

/* sparse_blas/sptrsv.cpp:28-38 (sym_lib::sptrsv_csr_ilu): unit-lower solve
 * with the L of a general CSR matrix A (the ILU(0) factor stored in A's
 * pattern): row i uses A's entries [Ap[i], diags[i]) and does NOT divide
 * (the division is commented out in the reference). */
void orc_sptrsv_csr_ilu(int n, const int *Ap, const int *Ai, const double *Ax,
                        const int *diags, double *x) {
  for (int i = 0; i < n; i++) {
    int dia = diags[i];
    for (int j = Ap[i]; j < dia; j++) x[i] -= Ax[j] * x[Ai[j]];
  }
}

/* sparse_blas/sptrsv.cpp:109-126 (sym_lib::sptrsv_csr_ilu_lbc): the same
 * rows on an LBC schedule; this variant DOES divide by Ax[diags[i]]. */
void orc_sptrsv_csr_ilu_lbc(int n, const int *Ap, const int *Ai,
                            const double *Ax, const int *diags, double *x,
                            int level_no, const int *level_ptr,
                            const int *par_ptr, const int *partition) {
  (void)n;
  for (int i1 = 0; i1 < level_no; ++i1) {
#pragma omp parallel for schedule(static)
    for (int j1 = level_ptr[i1]; j1 < level_ptr[i1 + 1]; ++j1) {
      for (int k1 = par_ptr[j1]; k1 < par_ptr[j1 + 1]; ++k1) {
        int i = partition[k1];
        int dia = diags[i];
        for (int j = Ap[i]; j < dia; j++) x[i] -= Ax[j] * x[Ai[j]];
        x[i] /= Ax[dia];
      }
    }
  }
}

/* sparse_blas/sptrsv.cpp:129-137 (sym_lib::sptrsv_csc): column-push forward
 * substitution on the CSC of L, diagonal first. */
void orc_sptrsv_csc(int n, const int *Lp, const int *Li, const double *Lx,
                    double *x) {
  for (int i = 0; i < n; i++) {
    x[i] /= Lx[Lp[i]];
    for (int j = Lp[i] + 1; j < Lp[i + 1]; j++) x[Li[j]] -= Lx[j] * x[i];
  }
}

/* sparse_blas/sptrsv_lt.cpp:11-20 (sym_lib::ltsolve): L^T x = b from the CSC
 * of L, columns j = n-1 .. 0, entries below the diagonal ascending, then the
 * divide by the first-in-column diagonal.  Returns 0 on null input. */
int orc_ltsolve(int n, const int *Lp, const int *Li, const double *Lx,
                double *x) {
  if (!Lp || !Li || !x) return 0;
  for (int j = n - 1; j >= 0; j--) {
    for (int p = Lp[j] + 1; p < Lp[j + 1]; p++) x[j] -= Lx[p] * x[Li[p]];
    x[j] /= Lx[Lp[j]];
  }
  return 1;
}

/* Multi-RHS forward solve: the reference has no CSR SpTRSM, so the oracle
 * for k right-hand sides is k independent sptrsv_csr solves, one per column
 * of X (sptrsv.cpp:17-25 applied column by column).  X element (i, c) lives
 * at X[i * rs + c * cs] (column-major: rs = 1, cs = ld; row-major: rs = ld,
 * cs = 1). */
void orc_sptrsm_csr(int n, const int *Lp, const int *Li, const double *Lx,
                    double *X, int k, int64_t rs, int64_t cs) {
#pragma omp parallel for schedule(static)
  for (int c = 0; c < k; c++) {
    double *x = X + (int64_t)c * cs;
    for (int i = 0; i < n; i++) {
      double v = x[(int64_t)i * rs];
      for (int j = Lp[i]; j < Lp[i + 1] - 1; j++)
        v -= Lx[j] * x[(int64_t)Li[j] * rs];
      v /= Lx[Lp[i + 1] - 1];
      x[(int64_t)i * rs] = v;
    }
  }
}

/* ---------------------------------------------------------------- SpMV */

/* sparse_blas/spmv.cpp:12-21 (sym_lib::spmv_csr, 6 args): y = A x with y
 * zeroed first, entries ascending. */
void orc_spmv_csr(int n, const int *Ap, const int *Ai, const double *Ax,
                  const double *x, double *y) {
  for (int i = 0; i < n; i++) {
    y[i] = 0;
    for (int j = Ap[i]; j < Ap[i + 1]; j++) y[i] += Ax[j] * x[Ai[j]];
  }
}

/* sparse_blas/spmv.cpp:23-34 (sym_lib::spmv_csr, 9 args):
 * z[i] = a * (z[i] + sum_j A[i,j] x[j]) + b * y[i].  z is accumulated into,
 * so callers pre-zero it for z = a A x + b y. */
void orc_spmv_csr_axpby(int n, const int *Ap, const int *Ai, const double *Ax,
                        const double *x, const double *y, double *z, double a,
                        double b) {
  for (int i = 0; i < n; i++) {
    for (int j = Ap[i]; j < Ap[i + 1]; j++) z[i] += Ax[j] * x[Ai[j]];
    z[i] = a * z[i] + b * y[i];
  }
}

/* sparse_blas/spmv.cpp:59-71 (sym_lib::spmv_csr_block): the rows of the
 * supernodes, each y[j] from 0 then += entries ascending. */
void orc_spmv_csr_block(int n, const int *Ap, const int *Ai, const double *Ax, int nodes,
                        const int *supernodes, const double *x, double *y) {
  (void)n;
  for (int i = 0; i < nodes; i++)
    for (int j = supernodes[i]; j < supernodes[i + 1]; j++) {
      y[j] = 0;
      for (int k = Ap[j]; k < Ap[j + 1]; k++) y[j] += Ax[k] * x[Ai[k]];
    }
}

/* sparse_blas/spmv.cpp:73-81 (sym_lib::spmv_csc, 6 args): y[Ai[j]] +=
 * Ax[j] * x[i], column by column. */
void orc_spmv_csc(int n, const int *Ap, const int *Ai, const double *Ax, const double *x,
                  double *y) {
  for (int i = 0; i < n; i++)
    for (int j = Ap[i]; j < Ap[i + 1]; j++) y[Ai[j]] += Ax[j] * x[i];
}

/* sparse_blas/spmv.cpp:85-95 and :97-109 (sym_lib::spmv_csc with scalars a, b
 * or per-row arrays av, bv): z += A x column by column, then
 * z[i] = a z[i] + b y[i]. */
void orc_spmv_csc_axpby(int n, const int *Ap, const int *Ai, const double *Ax,
                        const double *x, const double *y, double *z, double a, double b,
                        const double *av, const double *bv) {
  for (int i = 0; i < n; i++)
    for (int j = Ap[i]; j < Ap[i + 1]; j++) z[Ai[j]] += Ax[j] * x[i];
  for (int i = 0; i < n; i++)
    z[i] = (av ? av[i] : a) * z[i] + (bv ? bv[i] : b) * y[i];
}

/* sparse_blas/spmv.cpp:122-143 (sym_lib::spmv_csc_lbc): part l scatters the
 * columns ids[sets[l] .. sets[l+1]) into its own zeroed buffer, then y[i] +=
 * buffer_0[i], buffer_1[i], ... in part order (every part, every row). */
void orc_spmv_csc_lbc(int n_parts, const int *sets, const int *ids, int n, const int *Ap,
                      const int *Ai, const double *Ax, const double *x, double *y) {
  double *red = (double *)calloc((size_t)(n_parts > 0 ? n_parts : 1) * (size_t)(n > 0 ? n : 1),
                                 sizeof(double));
  for (int l = 0; l < n_parts; l++)
    for (int k = sets[l]; k < sets[l + 1]; ++k) {
      const int i = ids[k];
      for (int j = Ap[i]; j < Ap[i + 1]; j++) red[(size_t)l * n + Ai[j]] += Ax[j] * x[i];
    }
  for (int m = 0; m < n_parts; ++m)
    for (int i = 0; i < n; ++i) y[i] += red[(size_t)m * n + i];
  free(red);
}

/* sparse_blas/spmv.cpp:171-186 (sym_lib::spmv_csc_block): the columns of the
 * supernodes, column by column, scattered into y. */
void orc_spmv_csc_block(int n, const int *Ap, const int *Ai, const double *Ax,
                        const double *x, double *y, int nodes, const int *supernodes) {
  (void)n;
  for (int i = 0; i < nodes; i++)
    for (int j = supernodes[i]; j < supernodes[i + 1]; j++)
      for (int k = Ap[j]; k < Ap[j + 1]; k++) y[Ai[k]] += Ax[k] * x[j];
}

/* ---------------------------------------------------------------- IC(0) */

/* sparse_blas/spic0.cpp:11-46 (sym_lib::spic0_csc_left): left-looking IC(0)
 * on the CSC lower triangle of A (diagonal first).  prunePtr/pruneSet is the
 * row pattern of the strict lower part (for column c: the columns s < c with
 * L[c,s] != 0, ascending).  Returns 0 on a non-positive pivot. */
int orc_spic0_csc_left(int n, const int *Ap, const int *Ai, const double *Ax,
                       const int *prunePtr, const int *pruneSet, double *Lx) {
  double *f = (double *)calloc((size_t)n, sizeof(double));
  for (int col = 0; col < n; ++col) {
    for (int p = Ap[col]; p < Ap[col + 1]; ++p) f[Ai[p]] = Ax[p];
    for (int q = prunePtr[col]; q < prunePtr[col + 1]; ++q) {
      int s = pruneSet[q];
      double t = 0;
      int found = 0;
      for (int l = Ap[s]; l < Ap[s + 1]; ++l) {
        if (Ai[l] == col && !found) {
          t = Lx[l];
          found = 1;
        }
        if (found) f[Ai[l]] -= Lx[l] * t;
      }
    }
    if (f[col] <= 0) {
      free(f);
      return 0;
    }
    double d = sqrt(f[col]);
    f[col] = 0;
    Lx[Ap[col]] = d;
    for (int p = Ap[col] + 1; p < Ap[col + 1]; ++p) {
      Lx[p] = f[Ai[p]] / d;
      f[Ai[p]] = 0;
    }
  }
  free(f);
  return 1;
}

/* sparse_blas/spic0.cpp:106-127 (sym_lib::spic0_csr): right-looking IC(0)
 * in place on the CSC lower triangle (diagonal first), loop for loop,
 * including the update guard of :118 -- l advances from m only while the
 * row of the NEXT entry is <= the target row, so a match is reached only at
 * the last entry of column i (whose successor, the next column's diagonal,
 * has row i + 1).  In place of the reference's assert(pivot > 0) this
 * returns 0 when some pivot is not positive (every column is still
 * computed, as in a build without assertions), 1 otherwise. */
int orc_spic0_csr(int n, double *Lx, const int *Lp, const int *Li) {
  int ok = 1;
  for (int i = 0; i < n; i++) {
    if (!(Lx[Lp[i]] > 0)) ok = 0;
    Lx[Lp[i]] = sqrt(Lx[Lp[i]]);
    double d = Lx[Lp[i]];
    for (int m = Lp[i] + 1; m < Lp[i + 1]; m++) Lx[m] = Lx[m] / d;
    for (int m = Lp[i] + 1; m < Lp[i + 1]; m++)
      for (int k = Lp[Li[m]]; k < Lp[Li[m] + 1]; k++)
        /* l + 1 < Lp[n] holds: a column with an off-diagonal m is not the
         * last one */
        for (int l = m; l < Lp[i + 1] && Li[l + 1] <= Li[k]; l++)
          if (Li[l] == Li[k]) Lx[k] -= Lx[m] * Lx[l];
  }
  return ok;
}

/* sparse_blas/spilu0.cpp:11-51 (sym_lib::spilu0_csr): ILU(0) of a general
 * CSR matrix, row by row, in place on a copy of A (l), with the row's
 * column map iw (EMPTY = -1).  In place of the reference's assertions this
 * returns -2 when a row's loop does not stop at A_diag[i] on the diagonal,
 * else the first row whose U[i,i] is zero (every row still computed, as in
 * a build without assertions), else -1. */
int orc_spilu0_csr(int n, int nnz, const int *Ap, const int *Ai, const double *Ax,
                   const int *A_diag, double *l) {
  int *iw = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int bad = -1, jrow = 0;
  for (int j = 0; j < n; j++) iw[j] = -1;
  for (int k = 0; k < nnz; k++) l[k] = Ax[k];
  for (int i = 0; i < n; i++) {
    for (int k = Ap[i]; k < Ap[i + 1]; k++) iw[Ai[k]] = k;
    int j;
    for (j = Ap[i]; j < Ap[i + 1]; ++j) {
      jrow = Ai[j];
      if (i <= jrow) break;
      double tl = l[j] / l[A_diag[jrow]];
      l[j] = tl;
      for (int jj = A_diag[jrow] + 1; jj < Ap[jrow + 1]; jj++) {
        int jw = iw[Ai[jj]];
        if (jw != -1) l[jw] = l[jw] - tl * l[jj];
      }
    }
    if (A_diag[i] != j || jrow != i) {
      free(iw);
      return -2;
    }
    if (l[j] == 0.0 && bad < 0) bad = i;
    for (int k = Ap[i]; k < Ap[i + 1]; k++) iw[Ai[k]] = -1;
  }
  free(iw);
  return bad;
}

/* ------------------------------------------------------ supernodal (BCSC) */

/* sparse_blas/dense_blas/BLAS.cpp:3-96 (lsolve_BLAS): b <- L^-1 b for the
 * w x w lower-triangular top of a column-major block M (leading dimension
 * dim), in column blocks of 8, then 4, then one of 2, then a single column.
 * Inside a block of B columns starting at c: x_j = (b[c+j] - x_0 M[c+j,c]
 * - ... - x_{j-1} M[c+j,c+j-1]) / M[c+j,c+j], subtractions left to right;
 * then every later row k of the top: b[k] = b[k] - x_0 M[k,c] - ... -
 * x_{B-1} M[k,c+B-1]. */
static void orc_lsolve_block(int dim, int w, const double *M, double *b) {
  int col = 0;
  while (col < w) {
    int B = col < w - 7 ? 8 : col < w - 3 ? 4 : col < w - 1 ? 2 : 1;
    double xv[8];
    for (int j = 0; j < B; j++) {
      double t = b[col + j];
      for (int i = 0; i < j; i++) t = t - xv[i] * M[(col + j) + (size_t)(col + i) * dim];
      xv[j] = t / M[(col + j) + (size_t)(col + j) * dim];
    }
    for (int j = 0; j < B; j++) b[col + j] = xv[j];
    if (B > 1)
      for (int k = col + B; k < w; k++) {
        double t = b[k];
        for (int i = 0; i < B; i++) t = t - xv[i] * M[k + (size_t)(col + i) * dim];
        b[k] = t;
      }
    col += B;
  }
}

/* sparse_blas/dense_blas/BLAS.cpp:286-349 (matvec_BLAS): b[k] += sum over
 * the columns of M (nrow x ncol, leading dimension dim) times x, in column
 * blocks of 8, then 4, then single columns; a block adds its products summed
 * left to right: b[k] = b[k] + (((x_0 M[k,c] + x_1 M[k,c+1]) + ...)). */
static void orc_matvec_block(int dim, int nrow, int ncol, const double *M, const double *x,
                             double *b) {
  int col = 0;
  while (col < ncol) {
    int B = col < ncol - 7 ? 8 : col < ncol - 3 ? 4 : 1;
    for (int k = 0; k < nrow; k++) {
      double s = x[col] * M[k + (size_t)col * dim];
      for (int i = 1; i < B; i++) s = s + x[col + i] * M[k + (size_t)(col + i) * dim];
      b[k] = b[k] + s;
    }
    col += B;
  }
}

/* sparse_blas/sptrsv.cpp:176-199 (sym_lib::sptrsv_csc_block): forward solve
 * on a supernodal lower triangle.  Supernode i owns columns supernodes[i] ..
 * supernodes[i+1]-1 (width w) and nrows[i] block rows: a column-major dense
 * block at Lx[Lp[i]] (leading dimension nrows[i]) whose first w rows are
 * the supernode's own columns; row r of the block is row Li[Lp[i] + r] of L.
 * Per supernode in order: the dense triangular solve of the top, the
 * product of the rows below with the solved part into a zeroed temporary,
 * then x[row] -= temporary, row by row. */
void orc_sptrsv_csc_block(int n, const int *Lp, const int *Li, const int *nrows,
                          const double *Lx, double *x, int num_nodes, const int *supernodes) {
  double *t = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
  for (int i = 0; i < num_nodes; i++) {
    int s = supernodes[i], w = supernodes[i + 1] - s, nr = nrows[i];
    orc_lsolve_block(nr, w, Lx + Lp[i], x + s);
    orc_matvec_block(nr, nr - w, w, Lx + Lp[i] + w, x + s, t);
    for (int k = 0; k < nr - w; k++) {
      x[Li[Lp[i] + w + k]] -= t[k];
      t[k] = 0;
    }
  }
  free(t);
}

/* Transposed solve L^T x = b on the same supernodal storage: the semantics
 * of parsy's blockedLTsolve (sparse_blas/parsy/triangularSolve/
 * Triangular_BCSC.cpp:125-186), supernodes from the last to the first:
 *   1. the top's right-hand side minus the rows below times the solved
 *      values at those rows (its CBLAS path: dgemv "T", alpha -1, beta 1);
 *      here, per top column c, t = sum over the rows r below, ascending, of
 *      M[r, c] * x[row r] (the first product, then adding in order), and
 *      x[s+c] = x[s+c] - t;
 *   2. the top's transposed triangular solve (dtrsm "L","L","T","N"); here
 *      column-oriented from the last column: x[s+c] /= M[c, c], then every
 *      x[s+r], r < c, loses M[c, r] * x[s+c].
 * The reference's own result depends on the BLAS it links (absent here;
 * its SYM_BLAS fallback is marked unfinished, :163), so this order is this
 * build's definition, pinned by dense solves and exact integer cases. */
void orc_sptrsv_csc_block_t(int n, const int *Lp, const int *Li, const int *nrows,
                            const double *Lx, double *x, int num_nodes,
                            const int *supernodes) {
  (void)n;
  for (int i = num_nodes - 1; i >= 0; i--) {
    const int s = supernodes[i], w = supernodes[i + 1] - s, nr = nrows[i];
    const double *M = Lx + Lp[i];
    for (int c = 0; c < w; c++) {
      double t = 0.0;
      for (int r = w; r < nr; r++) {
        const double p = M[r + (size_t)c * nr] * x[Li[Lp[i] + r]];
        t = r == w ? p : t + p;
      }
      if (nr > w) x[s + c] = x[s + c] - t;
    }
    for (int c = w - 1; c >= 0; c--) {
      x[s + c] = x[s + c] / M[c + (size_t)c * nr];
      for (int r = 0; r < c; r++) x[s + r] = x[s + r] - M[c + (size_t)r * nr] * x[s + c];
    }
  }
}

/* --------------------------------------------------------- level sets */

/* Level-set inspector on the CSR of L (the reference's build_levelSet_CSC
 * lives in the un-vendored aggregation submodule; its call sites are
 * Catch_tests/sptrsv_tests.cpp:59 and demo/sptrsv_demo_utils.h:44).  The
 * definition restated here is the one the in-tree analog
 * parsy/common/TreeUtils.cpp:89-141 (getLevelSet) implements: level(i) = 0
 * when row i has no off-diagonal entry, else 1 + max level(j) over
 * L[i,j] != 0; levelSet lists rows by level, ascending row id inside a
 * level (getLevelSet scans i ascending each round).  Returns the number of
 * levels.  levelPtr has room for n+1 entries. */
int orc_levelset_csr(int n, const int *Lp, const int *Li, int *levelPtr,
                     int *levelSet, int *level_of) {
  int *lev = level_of ? level_of : (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int nlev = 0;
  for (int i = 0; i < n; i++) {
    int m = -1;
    for (int j = Lp[i]; j < Lp[i + 1] - 1; j++)
      if (lev[Li[j]] > m) m = lev[Li[j]];
    lev[i] = m + 1;
    if (lev[i] + 1 > nlev) nlev = lev[i] + 1;
  }
  memset(levelPtr, 0, sizeof(int) * (size_t)(nlev + 1));
  for (int i = 0; i < n; i++) levelPtr[lev[i] + 1]++;
  for (int l = 0; l < nlev; l++) levelPtr[l + 1] += levelPtr[l];
  int *fill = (int *)malloc(sizeof(int) * (size_t)(nlev + 1));
  memcpy(fill, levelPtr, sizeof(int) * (size_t)(nlev + 1));
  for (int i = 0; i < n; i++) levelSet[fill[lev[i]]++] = i;
  free(fill);
  if (!level_of) free(lev);
  return nlev;
}

/* The same levels computed from the CSC of L (diagonal first), which is the
 * input the reference inspector takes (sptrsv_demo_utils.h:44 passes the CSC
 * of L).  Push form: level(r) >= level(c) + 1 for every L[r,c] != 0. */
int orc_levelset_csc(int n, const int *Lp, const int *Li, int *levelPtr,
                     int *levelSet) {
  int *lev = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
  int nlev = n > 0 ? 1 : 0;
  for (int c = 0; c < n; c++) {
    for (int p = Lp[c] + 1; p < Lp[c + 1]; p++) {
      int r = Li[p];
      if (lev[c] + 1 > lev[r]) lev[r] = lev[c] + 1;
    }
    if (lev[c] + 1 > nlev) nlev = lev[c] + 1;
  }
  memset(levelPtr, 0, sizeof(int) * (size_t)(nlev + 1));
  for (int i = 0; i < n; i++) levelPtr[lev[i] + 1]++;
  for (int l = 0; l < nlev; l++) levelPtr[l + 1] += levelPtr[l];
  int *fill = (int *)malloc(sizeof(int) * (size_t)(nlev + 1));
  memcpy(fill, levelPtr, sizeof(int) * (size_t)(nlev + 1));
  for (int i = 0; i < n; i++) levelSet[fill[lev[i]]++] = i;
  free(fill);
  free(lev);
  return nlev;
}

/* Validity of a 3-array coarsened schedule (level_ptr / par_ptr / partition,
 * the format sptrsv.cpp:57-75 consumes).  Returns 0 when valid, else a
 * positive code:
 *   1 shape (pointers not monotone or out of range)
 *   2 some row missing or listed twice
 *   3 a dependency j -> i with coarse(j) > coarse(i)
 *   4 a dependency inside one coarse level across two partitions
 *   5 a dependency inside one partition listed after its consumer */
int orc_check_schedule(int n, const int *Lp, const int *Li, int level_no,
                       const int *level_ptr, const int *par_ptr,
                       const int *partition) {
  if (level_no < 0 || level_ptr[0] != 0) return 1;
  for (int l = 0; l < level_no; l++)
    if (level_ptr[l + 1] < level_ptr[l]) return 1;
  int nparts = level_ptr[level_no];
  if (par_ptr[0] != 0) return 1;
  for (int p = 0; p < nparts; p++)
    if (par_ptr[p + 1] < par_ptr[p]) return 1;
  if (par_ptr[nparts] != n) return 1;
  int *coarse = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int *part = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int *pos = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; i++) coarse[i] = -1;
  int rc = 0;
  for (int l = 0; l < level_no && !rc; l++)
    for (int p = level_ptr[l]; p < level_ptr[l + 1] && !rc; p++)
      for (int k = par_ptr[p]; k < par_ptr[p + 1]; k++) {
        int i = partition[k];
        if (i < 0 || i >= n || coarse[i] != -1) {
          rc = 2;
          break;
        }
        coarse[i] = l;
        part[i] = p;
        pos[i] = k;
      }
  for (int i = 0; i < n && !rc; i++)
    for (int j = Lp[i]; j < Lp[i + 1] - 1 && !rc; j++) {
      int d = Li[j];
      if (coarse[d] > coarse[i]) rc = 3;
      else if (coarse[d] == coarse[i] && part[d] != part[i]) rc = 4;
      else if (coarse[d] == coarse[i] && pos[d] > pos[i]) rc = 5;
    }
  free(coarse);
  free(part);
  free(pos);
  return rc;
}

/* ------------------------------------------------------------- adjoint */

/* Adjoint of x = L^{-1} b for a loss with gradient gx = dloss/dx.  The
 * reference has no adjoint; this restates the closed form
 *   gb = L^{-T} gx           (ltsolve, sptrsv_lt.cpp:11-20, on the CSC of L)
 *   gL[p] = -(gb[i] * x[j])  for every stored entry p = (i, j) of the CSR of L
 * (from d x = -L^{-1} dL x).  Lcsc_* is the CSC of the same L, diagonal
 * first, rows ascending in each column.  gb may alias gx. */
void orc_adjoint(int n, const int *Lp, const int *Li, const int *Cp,
                 const int *Ci, const double *Cx, const double *x,
                 const double *gx, double *gb, double *gL) {
  if (gb != gx) memcpy(gb, gx, sizeof(double) * (size_t)n);
  orc_ltsolve(n, Cp, Ci, Cx, gb);
  for (int i = 0; i < n; i++)
    for (int p = Lp[i]; p < Lp[i + 1]; p++) gL[p] = -(gb[i] * x[Li[p]]);
}

/* The pairwise sum of the k products g[c]*x[c]: columns in groups of G
 * (G = the smallest power of two >= k when k <= 32, else 32), a group's
 * values s[0..G) (+0.0 past column k) reduced by halving -- for h = G/2,
 * G/4, .., 1: s[c] = s[c] + s[c+h] for c < h -- and the group sums added in
 * ascending group order.  The reference has no adjoint, so this order is
 * the library's own choice (a tree fits a wavefront's lanes). */
static double pairwise_dot(const double *g, const double *x, int k) {
  int G = 1;
  if (k > 32)
    G = 32;
  else
    while (G < k) G <<= 1;
  double acc = 0.0, s[32];
  for (int g0 = 0; g0 < k; g0 += G) {
    for (int c = 0; c < G; c++) s[c] = g0 + c < k ? g[g0 + c] * x[g0 + c] : 0.0;
    for (int h = G / 2; h >= 1; h /= 2)
      for (int c = 0; c < h; c++) s[c] = s[c] + s[c + h];
    acc = g0 == 0 ? s[0] : acc + s[0];
  }
  return acc;
}

/* Adjoint of X = L^{-1} B for a row-major block of k right-hand sides
 * (element (i, c) at [i*k + c]): GB = L^{-T} GX column by column (ltsolve,
 * sptrsv_lt.cpp:11-20, on the CSC of L) and, for every stored entry
 * p = (i, j) of the CSR of L,
 *   gL[p] = -(pairwise sum over c of GB[i,c]*X[j,c])
 * (the gradient of a loss over the whole block; pairwise_dot's order).  For
 * k = 1 this is orc_adjoint. */
void orc_adjoint_block(int n, int k, const int *Lp, const int *Li, const int *Cp,
                       const int *Ci, const double *Cx, const double *X,
                       const double *GX, double *GB, double *gL) {
  double *col = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int c = 0; c < k; c++) {
    for (int i = 0; i < n; i++) col[i] = GX[(size_t)i * k + c];
    orc_ltsolve(n, Cp, Ci, Cx, col);
    for (int i = 0; i < n; i++) GB[(size_t)i * k + c] = col[i];
  }
  free(col);
  for (int i = 0; i < n; i++)
    for (int p = Lp[i]; p < Lp[i + 1]; p++)
      gL[p] = -pairwise_dot(GB + (size_t)i * k, X + (size_t)Li[p] * k, k);
}

/* k values g[c]*x[c] added in pairwise_dot's order (exposed for tests). */
double orc_pairwise_dot(const double *g, const double *x, int k) {
  return pairwise_dot(g, x, k);
}

/* CSR -> CSC transpose (stable: entries of a column in ascending row order,
 * which puts the diagonal of a lower triangle first).  map[p] gives the CSC
 * position of CSR entry p (may be NULL). */
void orc_csr_to_csc(int n, const int *Ap, const int *Ai, const double *Ax,
                    int *Cp, int *Ci, double *Cx, int *map) {
  memset(Cp, 0, sizeof(int) * (size_t)(n + 1));
  for (int i = 0; i < n; i++)
    for (int p = Ap[i]; p < Ap[i + 1]; p++) Cp[Ai[p] + 1]++;
  for (int c = 0; c < n; c++) Cp[c + 1] += Cp[c];
  int *fill = (int *)malloc(sizeof(int) * (size_t)(n + 1));
  memcpy(fill, Cp, sizeof(int) * (size_t)(n + 1));
  for (int i = 0; i < n; i++)
    for (int p = Ap[i]; p < Ap[i + 1]; p++) {
      int q = fill[Ai[p]]++;
      Ci[q] = i;
      if (Cx) Cx[q] = Ax[p];
      if (map) map[p] = q;
    }
  free(fill);
}
