/*
 * sympiler_amd_gen.h -- deterministic host generators for the sparse
 * lower-triangular workloads (libsymamd_gen.so).
 *
 * These build the inputs the solve path is exercised on; they are host C++
 * and carry no GPU code.  Matrices are CSR lower triangles, diagonal LAST in
 * each row, columns ascending, int32 indices, FP64 values -- the layout of
 * the reference's sym_lib::sptrsv_csr (sparse_blas/sptrsv.cpp:17-25).
 *
 * Random numbers come from a splitmix64 stream keyed by (seed, tag):
 * u = (next >> 11) * 2^-53 in [0, 1).
 */
#ifndef SYMPILER_AMD_GEN_H
#define SYMPILER_AMD_GEN_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* out[i] = lo + (hi - lo) * u_i for i < count, from stream (seed, tag). */
void symgen_uniform(uint64_t seed, uint64_t tag, int64_t count, double lo,
                    double hi, double *out);

/* tril of the 2D 5-point Laplacian on nx*ny (row id x + nx*y). */
int64_t symgen_lap2d_tril_nnz(int nx, int ny);
void symgen_lap2d_tril(int nx, int ny, double diag, double off, int *Lp,
                       int *Li, double *Lx);

/* tril of the 3D 7-point Laplacian on nx*ny*nz (row id x + nx*(y + ny*z)). */
int64_t symgen_lap3d_tril_nnz(int nx, int ny, int nz);
void symgen_lap3d_tril(int nx, int ny, int nz, double diag, double off,
                       int *Lp, int *Li, double *Lx);

/* Zero-fill incomplete Cholesky in place on a CSR lower triangle holding
 * tril(A); bit-identical to the reference's spic0_csc_left
 * (sparse_blas/spic0.cpp:11-46).  Returns 0, or 1 + the first row whose
 * pivot is not positive. */
int symgen_ic0_csr(int n, const int *Lp, const int *Li, double *Lx);

/* y = L x on a CSR matrix (host). */
void symgen_spmv(int n, const int *Lp, const int *Li, const double *Lx,
                 const double *x, double *y);

/* Irregular lower triangle: strictly-lower count per row min(Zipf(alpha),
 * cap, min(i, window)), distinct columns uniform in [max(0, i-window), i),
 * off-diagonals U[-1, 0), diagonal 1 + sum|off|.  Li = NULL counts only.
 * Returns nnz (diagonal included). */
int64_t symgen_zipf_lower(int n, double alpha, int cap, int window,
                          uint64_t seed, int *Lp, int *Li, double *Lx);

/* Grid sizes a[s], b[s] ~ U{lo..hi} of an uneven batch of systems. */
void symgen_batch_dims(int count, int lo, int hi, uint64_t seed, int *a,
                       int *b);

/* Random lower triangle with Bernoulli(density) strictly-lower entries per
 * (i, j < i), at least min(min_off, i) of them, off-diagonals U[-1, 0),
 * diagonal 1 + sum|off|.  Li = NULL counts only.  Returns nnz. */
int64_t symgen_random_lower(int n, double density, int min_off, uint64_t seed,
                            int *Lp, int *Li, double *Lx);

/* Supernodal (BCSC) lower triangle for sym_lib::sptrsv_csc_block
 * (sparse_blas/sptrsv.cpp:176-199): supernode i owns columns
 * supernodes[i] .. supernodes[i+1]-1 (widths ~ U{1..wmax}) and nrows[i] block
 * rows (its own columns, then up to rmax distinct ascending rows drawn from
 * the next `window` rows); its column-major block of values is at
 * Lx[Lp[i]] with row ids Li[Lp[i] + j*nrows[i] + r].  Top-block diagonal
 * 1 + U[0,1), entries below it U[-1,1)/nrows[i], strictly upper part NaN
 * (never read by the solve).  With Lp == NULL only *num_nodes and the
 * returned value count are produced; then supernodes/nrows need
 * num_nodes+1 / num_nodes entries, Lp num_nodes+1, Li/Lx the count. */
int64_t symgen_bcsc(int n, int wmax, int rmax, int window, uint64_t seed,
                    int *num_nodes, int *supernodes, int *nrows, int *Lp,
                    int *Li, double *Lx);

/* Supernodal lower triangle shaped like a nested-dissection Cholesky factor:
 * a complete binary tree of `levels` levels of supernodes in postorder; a
 * node h levels above the leaves is min(wmax, round(wleaf 2^(h/2))) columns
 * wide and has about rowfac x its width rows below its top, drawn from its
 * ancestors' columns.  Layout and values as symgen_bcsc.  With Lp == NULL
 * only *num_nodes, *n_out and the returned value count are produced. */
int64_t symgen_bcsc_tree(int levels, int wleaf, int wmax, double rowfac,
                         uint64_t seed, int *num_nodes, int *n_out,
                         int *supernodes, int *nrows, int *Lp, int *Li,
                         double *Lx);

#ifdef __cplusplus
}
#endif
#endif
