/*
 * sympiler_amd.h -- C ABI of the MI355X (gfx950) sparse lower-triangular
 * solve library libsympiler_amd.so.
 *
 * Everything here is `extern "C"`, plain pointers and sizes.  Matrix and
 * vector arguments named d_* are DEVICE pointers (hipMalloc'd or torch CUDA
 * tensors); h_* are host pointers.  Streams are hipStream_t passed as void*.
 *
 * Matrix layout is the reference's (include/sympiler/sparse_blas_lib.h):
 *   CSR lower triangle L: Lp[n+1], Li[nnz], Lx[nnz]; column indices
 *   ascending, diagonal LAST in each row (sparse_blas/sptrsv.cpp:20,23);
 *   int32 indices, FP64 (or FP32) values.
 *
 * The reference's own blocking, stateless C++ entry points
 * (namespace sym_lib, C++ linkage) are re-exported by the same library on
 * top of this ABI; their declarations are in sym_lib_sparse_blas.h.
 *
 * Errors: every call returns a symamd_status; symamd_last_error() returns a
 * thread-local message for the last failing call.  A handle is bound to one
 * stream and is not thread-safe; separate handles are independent.
 */
#ifndef SYMPILER_AMD_H
#define SYMPILER_AMD_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  SYMAMD_OK = 0,
  SYMAMD_ERR_INVALID_ARG = 1,   /* null pointer, negative size, bad enum   */
  SYMAMD_ERR_STRUCTURE = 2,     /* not lower / diagonal not last / unsorted */
  SYMAMD_ERR_ZERO_PIVOT = 3,    /* a zero diagonal entry (or ILU(0) pivot)  */
  SYMAMD_ERR_HIP = 4,           /* a HIP runtime call failed                */
  SYMAMD_ERR_ALLOC = 5,         /* device or host allocation failed         */
  SYMAMD_ERR_TIMEOUT = 6,       /* a sync-free wait exceeded its watchdog   */
  SYMAMD_ERR_NOT_SUPPORTED = 7,
  SYMAMD_ERR_NOT_SPD = 8        /* IC(0): a pivot that is not positive      */
} symamd_status;

typedef enum { SYMAMD_F64 = 0, SYMAMD_F32 = 1 } symamd_dtype;

/* Dense right-hand-side layouts: element (i, c) of an n-by-k block lives at
 *   COL_MAJOR: B[i + c*ld]   (ld >= n; the reference's multi-RHS convention,
 *                             parsy Triangular_BCSC.cpp:114)
 *   ROW_MAJOR: B[i*ld + c]   (ld >= k; right-hand sides interleaved)     */
typedef enum { SYMAMD_COL_MAJOR = 0, SYMAMD_ROW_MAJOR = 1 } symamd_layout;

/* Executors for the forward (and adjoint) solve. */
typedef enum {
  SYMAMD_EXEC_AUTO = 0,      /* tuned choice if symamd_tune ran, else heuristic */
  SYMAMD_EXEC_LEVELSET = 1,  /* one kernel launch per level (plain level set)   */
  SYMAMD_EXEC_SYNCFREE = 2,  /* one persistent launch, per-unit ready flags     */
  SYMAMD_EXEC_LEVELSET_GRAPH = 3, /* the per-level launches replayed as a hipGraph */
  SYMAMD_EXEC_REGIONS = 4,   /* one persistent launch over compact regions of rows,
                                each solved by one workgroup in LDS (forward
                                and adjoint solves, 1 <= k <= 64, at the largest
                                region cap whose task fits the LDS; other cases,
                                and matrices with no such schedule, run
                                SYNCFREE) */
  SYMAMD_EXEC_WIDE = 5,      /* disabled: returns SYMAMD_ERR_NOT_SUPPORTED (a
                                persistent executor with one workgroup per row;
                                its first run exceeded the watchdog and the
                                cause is not established) */
  SYMAMD_EXEC_LEVELSET_SPLIT = 6 /* one launch per level, replayed as a hipGraph:
                                rows with at most T off-diagonal entries one
                                lane per row and column in entry order
                                (bit-identical), each longer row by a
                                1024-thread workgroup over its entries (sum
                                re-associated, within 1e-12); T from
                                SYMAMD_OPT_SPLIT_LONG; forward solves
                                (adjoints run LEVELSET) */
} symamd_executor;

typedef struct symamd_handle_s *symamd_handle;
typedef struct symamd_plan_s *symamd_plan;
typedef struct symamd_graph_s *symamd_graph;

/* ---------------------------------------------------------------- handle */
int symamd_create(symamd_handle *h, void *stream);
int symamd_destroy(symamd_handle h);
int symamd_set_stream(symamd_handle h, void *stream);
/* Sync-free watchdog: a wait longer than `ms` milliseconds aborts the solve
 * with SYMAMD_ERR_TIMEOUT (default 20000). */
int symamd_set_timeout_ms(symamd_handle h, double ms);
/* Handle options.
 *   SYMAMD_OPT_TICKETS: 1 = persistent kernels take work units from a device
 *     ticket counter instead of a co-resident round-robin deal (the mode
 *     used automatically inside graph capture); 0 = automatic (default).
 *   SYMAMD_OPT_TRACE_PTR / SYMAMD_OPT_TRACE_CAP: device pointer (as an
 *     integer; 0 disables) and capacity in uint64 entries of a trace
 *     buffer.  Sync-free launches then record, per work unit u, three
 *     100 MHz clock stamps: [3u] unit taken, [3u+1] dependencies ready,
 *     [3u+2] unit published (units beyond the capacity are not traced). */
/*   SYMAMD_OPT_MAX_WAVES: cap on the wavefronts of a persistent launch
 *     (0 = as many as fit the GPU, the default).
 *   SYMAMD_OPT_COLMAJOR_DIRECT: 1 = solve column-major blocks of k > 1
 *     right-hand sides in place (one cache line per row and column); 0 =
 *     (default) transpose them into a row-major workspace owned by the plan
 *     (level-ordered for the unit executors, natural order for the region
 *     executor), solve there and transpose back. */
/*   SYMAMD_OPT_MARKS: 1 (default) = sync-free solves of a single
 *     right-hand side mark each row's right-hand side with a signalling-NaN
 *     pattern after reading it and publish a unit's ready tag without waiting
 *     for its value stores; readers re-load any value still marked.  0 = the
 *     tag is published after the stores are acknowledged.  The solve
 *     overwrites B in place either way; intermediate contents of B are
 *     unspecified while a solve runs. */
/*   SYMAMD_OPT_REGIONS (read by symamd_analyze*): 1 = cluster the rows into
 *     compact regions of the dependency graph (heavy-edge matching, at most
 *     SYMAMD_OPT_REGION_ROWS rows each) when at least half of L's
 *     off-diagonal entries fall inside a region; the units of each region
 *     are then solved level after level by one wavefront of the sync-free
 *     executor (the owner schedule), which waits only on dependencies owned
 *     by other wavefronts.  0 = never (default: on MI355X the owner
 *     schedule serializes each unit's staging and loads behind the previous
 *     one and measured slower than the round-robin deal); 2 = always (no
 *     threshold).  Results are identical either way.
 *   SYMAMD_OPT_REGION_ROWS: region size cap (default 512).
 *   SYMAMD_OPT_OWNERS (read at solve time): 1 (default) = plans with
 *     regions run the owner schedule; 0 = they deal units like other plans
 *     (for comparison; results are identical).
 *   SYMAMD_OPT_REGION_CAP (read at solve time): region size cap of the
 *     SYMAMD_EXEC_REGIONS schedule (0, the default: the cap symamd_tune
 *     chose for the plan, else 1024 rows for single right-hand sides on the
 *     memory-wavefront kernel, 256 for tasks of eight columns run by one
 *     wavefront and 512 otherwise; halved -- down to 32 -- until a task's
 *     rows, halo and record fit a workgroup's LDS; a cap set here is used
 *     alone); each cap's schedule is built once per plan.
 *   SYMAMD_OPT_REGION_WAVES (read at solve time): wavefronts per task of
 *     the SYMAMD_EXEC_REGIONS schedule, 1, 2 or 4 for tasks of eight
 *     columns (k >= 8; narrower tasks use one); 0, the default, picks it.
 *     Results are identical either way.
 *   SYMAMD_OPT_REGION_BAND (read at solve time): a region task publishes
 *     its progress every this many internal levels (slices), so a region
 *     that reads it starts on its first levels while the rest is solved
 *     (0, the default, = 255: one band; a region has at most 31 bands, wider
 *     ones when it has more levels); each value's schedule is built once per
 *     plan.
 *     Results are identical either way.
 *   SYMAMD_OPT_REGION_MW (read at solve time): 1 (default) = region tasks
 *     whose right-hand sides are loaded directly (row-major blocks, single
 *     vectors) run with one extra wavefront per task that does all of its
 *     waiting, loading, storing and publishing, so the solving wavefronts
 *     never wait for memory; the default band is then 4 slices for single
 *     right-hand sides.  0 = every wavefront of a task does both (default
 *     band: one per region).
 *     Results are identical either way.
 *   SYMAMD_OPT_SPLIT_LONG (read at solve time): SYMAMD_EXEC_LEVELSET_SPLIT
 *     solves rows with more than this many off-diagonal entries with a
 *     whole workgroup each (0, the default: 16 for column tiles of up to 16
 *     columns, 32 for wider ones); each value's lists are built once per
 *     plan.  Rows at or under it are bit-identical to the serial solve. */
enum { SYMAMD_OPT_TICKETS = 1, SYMAMD_OPT_TRACE_PTR = 2, SYMAMD_OPT_TRACE_CAP = 3,
       SYMAMD_OPT_MAX_WAVES = 4, SYMAMD_OPT_COLMAJOR_DIRECT = 5, SYMAMD_OPT_MARKS = 6,
       SYMAMD_OPT_REGIONS = 7, SYMAMD_OPT_REGION_ROWS = 8, SYMAMD_OPT_OWNERS = 9,
       SYMAMD_OPT_REGION_CAP = 10, SYMAMD_OPT_REGION_WAVES = 11,
       SYMAMD_OPT_REGION_BAND = 12, SYMAMD_OPT_REGION_MW = 13,
       SYMAMD_OPT_SPLIT_LONG = 14 };
int symamd_set_option(symamd_handle h, int option, int64_t value);
const char *symamd_last_error(void);
/* Library version string. */
const char *symamd_version(void);

/* ------------------------------------------------------------- inspector */
/* Level sets of the dependency DAG of a device CSR lower triangle, computed
 * on the GPU.  level(i) = 0 if row i has no off-diagonal, else
 * 1 + max level(j) over L[i,j] != 0; d_level_set lists rows level by level,
 * ascending row id inside a level.  d_level_ptr needs n+1 entries.
 * Replaces the un-vendored build_levelSet_CSC (call sites
 * Catch_tests/sptrsv_tests.cpp:59, demo/sptrsv_demo_utils.h:44-45), whose
 * input is the CSC of L; this entry point takes the CSR of the same L. */
int symamd_levelset(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
                    int *d_level_ptr, int *d_level_set, int *levels);

/* symamd_levelset without any allocation.  No counterpart in the reference;
 * for callers that cannot allocate device memory (code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * results, errors and blocking as symamd_levelset, for n <= 2,097,152 rows
 * (SYMAMD_ERR_NOT_SUPPORTED above that, before any device work).  Its
 * working memory (counters, per-row levels, sort keys and sort storage) is
 * one fixed 48 MiB device array per GPU, held until the process ends and
 * shared by every call on that GPU: calls that overlap on one GPU (other
 * threads or streams) race and must be serialised by the caller. */
int symamd_levelset_noalloc(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
                            int *d_level_ptr, int *d_level_set, int *levels);

/* Coarsened 3-array schedule (level_ptr / par_ptr / partition), the format
 * sym_lib::sptrsv_csr_lbc consumes (sparse_blas/sptrsv.cpp:57-75), built on
 * the GPU from a device CSR lower triangle (diagonal last).  Replaces the
 * un-vendored get_coarse_levelSet_DAG_CSC_tree (call site
 * demo/sptrsv_demo_utils.h:83-88; the demo passes inner_parts = threads,
 * min_level_dist = -1, div_rate = 4000, demo/sptrsv_demo.cpp:37, and the row
 * nnz as cost, :80-82), with the same parameters:
 *   d_cost          per-row cost (device, n doubles); NULL = row nnz
 *   inner_parts     parts (threads) per coarse level, 1..4096
 *   min_level_dist  > 0: that many consecutive dependency levels form each
 *                   coarse level; <= 0: chosen from the DAG -- the largest
 *                   block of 2^t levels starting at a coarse level's first
 *                   level whose weakly connected components number at least
 *                   inner_parts, none costlier than cost/inner_parts
 *   div_rate        the largest block tried is the largest 2^t <= div_rate
 * Inside a coarse level the components (dependency edges that stay inside
 * it) are packed worst-fit (largest first, into the first least-loaded
 * part); a part lists its rows in (level, row id) order; empty parts are
 * dropped.  Every dependency then points to an earlier coarse level or to
 * an earlier row of the same part.  Outputs are device arrays of n+1
 * (level_ptr, par_ptr) and n (partition) ints; *level_no coarse levels,
 * *part_no parts.  Deterministic (costs summed in 1/1024 fixed point). */
int symamd_coarsen(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
                   const double *d_cost, int inner_parts, int min_level_dist,
                   int div_rate, int *d_level_ptr, int *d_par_ptr, int *d_partition,
                   int *level_no, int *part_no);

/* Load-balanced level coarsening of an elimination tree (host arrays; no
 * device work): `parent[i]` is node i's parent, negative for a root;
 * `node_cost[i]` its cost.  Writes the same 3-array schedule as
 * symamd_coarsen -- level_ptr (n+1 ints), part_ptr (n+1) and partition (n)
 * -- with inside each part every node after its children, and the counts of
 * coarse levels and parts.  The arrays equal those of
 * sym_lib::parsy::getCoarseLevelSet_6(n, parent, -, ..., inner_parts,
 * min_level_dist, div_rate, node_cost)
 * (sparse_blas/parsy/symbolic/InspectionLevel_06.cpp:17; callers
 * symbolic_phase.cpp:692 on the supernodal tree and :787 on the column
 * tree).  SYMAMD_ERR_STRUCTURE if parent has a cycle; SYMAMD_ERR_INVALID_ARG
 * for a parent index >= n or equal to its node, a NaN cost, or div_rate 0
 * or 1 with inner_parts > 1 on a tree of more than two levels (input the
 * reference never returns from).  n = 0 gives no levels. */
int symamd_tree_coarsen(int n, const int *parent, const double *node_cost,
                        int inner_parts, int min_level_dist, int div_rate,
                        int *level_ptr, int *part_ptr, int *partition,
                        int *level_no, int *part_no);

/* Execute a caller-given coarsened schedule on a device CSR of L (FP64), in
 * place on d_x, with the semantics of sym_lib::sptrsv_csr_lbc
 * (sparse_blas/sptrsv.cpp:57-75): coarse levels in order, partitions of a
 * coarse level in parallel, rows of a partition in listed order.
 * h_level_ptr is a HOST array (level_no+1); d_par_ptr and d_partition are
 * device arrays. */
int symamd_solve_lbc(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
                     const double *d_Lx, double *d_x, int level_no,
                     const int *h_level_ptr, const int *d_par_ptr,
                     const int *d_partition);

/* The same schedule with a separate right-hand side d_b, with the
 * arithmetic of sym_lib::sptrsv_csr_lbc_redundant (decl
 * sparse_blas_lib.h:156-161, impl sparse_blas/sptrsv.cpp:78-106): each row
 * accumulates its products from zero and adds b[i] last,
 * x[i] = ((0 - sum_j Lx[j] x[Li[j]]) + b[i]) / diag.  The reference's
 * threads*n scratch array is not needed. */
int symamd_solve_lbc_redundant(symamd_handle h, int n, const int *d_Lp,
                               const int *d_Li, const double *d_Lx, double *d_x,
                               const double *d_b, int level_no,
                               const int *h_level_ptr, const int *d_par_ptr,
                               const int *d_partition);

/* Extract the lower factor stored inside a general CSR matrix A (the L of
 * an ILU(0) factor kept in A's pattern) as a CSR lower triangle with the
 * diagonal last: row i = A's entries [Ap[i], diags[i]) followed by column i
 * with value Ax[diags[i]] (unit_diag == 0) or 1.0 (unit_diag != 0).  Solving
 * with it reproduces, bit for bit, sym_lib::sptrsv_csr_ilu (unit_diag = 1:
 * that routine does not divide, sparse_blas/sptrsv.cpp:28-38) and
 * sym_lib::sptrsv_csr_ilu_lbc (unit_diag = 0: it divides by Ax[diags[i]],
 * sparse_blas/sptrsv.cpp:109-126); decls sparse_blas_lib.h:137-142.
 * d_Lp needs n+1 entries, d_Li/d_Lx at least Ap[n]; *nnz gets Lp[n].
 * SYMAMD_ERR_INVALID_ARG when some diags[i] lies outside row i. */
int symamd_ilu_lower(symamd_handle h, int n, const int *d_Ap, const int *d_Ai,
                     const double *d_Ax, const int *d_diags, int unit_diag,
                     int *d_Lp, int *d_Li, double *d_Lx, int *nnz);

/* ------------------------------------------------------------------ plan */
/* Analyse L (device CSR, diagonal last) once: level sets, a level-ordered
 * copy of L, work units and ready flags.  The plan owns its device memory;
 * d_Lp/d_Li/d_Lx may be freed after the call. */
int symamd_analyze(symamd_handle h, int n, const int *d_Lp, const int *d_Li,
                   const void *d_Lx, symamd_dtype dtype, symamd_plan *plan);
/* Analyse an uneven batch of independent systems given as one concatenated
 * CSR: system s owns rows h_row_off[s] .. h_row_off[s+1]-1; d_Li holds
 * column ids LOCAL to each system.  The plan solves all of them in one
 * launch (block-diagonal union). */
int symamd_analyze_batched(symamd_handle h, int nsys, const int *h_row_off,
                           const int *d_Lp, const int *d_Li, const void *d_Lx,
                           symamd_dtype dtype, symamd_plan *plan);
/* Refresh the plan's copy of the values (same pattern). */
int symamd_plan_update_values(symamd_handle h, symamd_plan plan,
                              const void *d_Lx);
/* n, nnz, number of levels, max rows in one level. */
int symamd_plan_info(symamd_plan plan, int *n, int64_t *nnz, int *levels,
                     int *max_level_rows);
/* Copy the plan's level sets to host arrays (level_ptr: levels+1, level_set: n). */
int symamd_plan_levels(symamd_plan plan, int *h_level_ptr, int *h_level_set);
/* Regions of the plan (0 when it has no region order) and the fraction of
 * L's off-diagonal entries whose two rows share a region (computed whenever
 * the clustering ran). */
int symamd_plan_regions(symamd_plan plan, int *count, double *inside);
/* The region schedule SYMAMD_EXEC_REGIONS runs for a forward (transposed =
 * 0) or adjoint (1) solve of k right-hand sides: its region cap and its task
 * count (regions times column groups of up to eight columns), or 0 and 0
 * when such a solve runs SYMAMD_EXEC_SYNCFREE.  May build the schedule. */
int symamd_region_schedule(symamd_handle h, symamd_plan plan, int k, int transposed,
                           int *cap, int *tasks);
/* The executor (and lanes per row) SYMAMD_EXEC_AUTO takes for a forward
 * solve of k right-hand sides in `layout`: the tuned choice if symamd_tune
 * ran, else SYMAMD_EXEC_REGIONS when the plan has a usable region schedule
 * with at least 60 % of L's off-diagonal entries inside regions (column-major
 * blocks only through the workspace, see SYMAMD_OPT_COLMAJOR_DIRECT), else
 * SYMAMD_EXEC_SYNCFREE.  May build the region schedule. */
int symamd_choice(symamd_handle h, symamd_plan plan, int k, symamd_layout layout,
                  int *exec, int *lanes);
int symamd_plan_destroy(symamd_plan plan);

/* ----------------------------------------------------------------- solve */
/* Solve L X = B in place for k right-hand sides (X overwrites B), on the
 * handle's stream, asynchronously.  dtype of B is the plan's dtype. */
int symamd_solve(symamd_handle h, symamd_plan plan, void *d_B, int k,
                 int64_t ld, symamd_layout layout, symamd_executor exec);
/* Synchronise the handle's stream and report a sync-free watchdog expiry
 * recorded by any earlier solve on this plan (SYMAMD_ERR_TIMEOUT). */
int symamd_plan_check(symamd_handle h, symamd_plan plan);

/* Capture one solve (fixed B pointer, k, layout, executor) as a hipGraph
 * and replay it with symamd_graph_launch on the handle's stream. */
int symamd_graph_create(symamd_handle h, symamd_plan plan, void *d_B, int k,
                        int64_t ld, symamd_layout layout,
                        symamd_executor exec, symamd_graph *g);
int symamd_graph_launch(symamd_handle h, symamd_graph g);
int symamd_graph_destroy(symamd_graph g);

/* --------------------------------------------------------------- adjoint */
/* For x = L^{-1} b and an upstream gradient gx = dloss/dx:
 *   gb = L^{-T} gx                      (transposed solve, reverse levels;
 *                                        the reference's ltsolve,
 *                                        sparse_blas/sptrsv_lt.cpp:11-20)
 *   gLx[p] = -(gb[i] * x[j])            for every stored entry p = (i, j) of
 *                                        the caller's CSR of L (same order
 *                                        as the Lx given to analyze).
 * d_gb may alias d_gx.  d_gLx may be NULL (gb only). */
int symamd_adjoint(symamd_handle h, symamd_plan plan, const void *d_x,
                   const void *d_gx, void *d_gb, void *d_gLx,
                   symamd_executor exec);
/* The same for a block of k right-hand sides, X = L^{-1} B, with the
 * upstream gradient GX = dloss/dX (no counterpart in the reference, which
 * has no adjoint; the batched form of symamd_adjoint):
 *   GB = L^{-T} GX                      column by column
 *   gLx[p] = -(GB[i,0]*X[j,0] + ... + GB[i,k-1]*X[j,k-1])
 *                                       for every stored entry p = (i, j),
 * the k products added pairwise: in groups of G columns (G = the smallest
 * power of two >= k when k <= 32, else 32; a group padded with +0.0 past
 * column k), each reduced by halving (s[c] += s[c + h] for c < h, h = G/2,
 * .., 1), group sums added in ascending order (k = 1: -(gb[i] * x[j])).
 * X, GX and GB share k, ld and layout; d_GB may alias d_GX; d_gLx may be
 * NULL.  The region executor (SYMAMD_EXEC_AUTO when the matrix has a
 * usable region schedule, or SYMAMD_EXEC_REGIONS) solves the whole block in
 * one launch; otherwise the columns are solved one after another. */
int symamd_adjoint_block(symamd_handle h, symamd_plan plan, const void *d_X,
                         const void *d_GX, void *d_GB, void *d_gLx, int k, int64_t ld,
                         symamd_layout layout, symamd_executor exec);

/* ------------------------------------------------------------------ SpMV */
/* y = A x (sparse_blas/spmv.cpp:12-21). FP64. */
int symamd_spmv_csr(symamd_handle h, int n, const int *d_Ap, const int *d_Ai,
                    const double *d_Ax, const double *d_x, double *d_y);
/* z = a * (z + A x) + b * y (sparse_blas/spmv.cpp:23-34). FP64. */
int symamd_spmv_csr_axpby(symamd_handle h, int n, const int *d_Ap,
                          const int *d_Ai, const double *d_Ax,
                          const double *d_x, const double *d_y, double *d_z,
                          double a, double b);
/* y += A x over the columns c0 .. c1-1 of a CSC matrix A (n columns, m rows:
 * column c's rows Ai[Ap[c] .. Ap[c+1]), values Ax), each y[r] receiving its
 * products in ascending column order -- the order of the reference's
 * column-by-column scatter, so y is bit-identical to sym_lib::spmv_csc
 * (sparse_blas/spmv.cpp:73-81; c0 = 0, c1 = n) and spmv_csc_block (:171-186;
 * the supernodes' column range).  FP64, blocking.  SYMAMD_ERR_STRUCTURE for a
 * row index outside [0, m). */
int symamd_spmv_csc(symamd_handle h, int n, int m, const int *d_Ap, const int *d_Ai,
                    const double *d_Ax, const double *d_x, double *d_y, int c0, int c1);
/* The LBC form of the same product (sym_lib::spmv_csc_lbc, spmv.cpp:122-167):
 * part l takes the columns ids[sets[l] .. sets[l+1]) (d_sets: nparts + 1
 * nondecreasing bounds, d_ids: column ids, device arrays); each part's
 * products are summed from +0.0 in that order, then y[r] += the parts' sums in
 * part order (every part, as the reference's reduction does).  Bit-identical
 * to the reference's result.  FP64, blocking. */
int symamd_spmv_csc_parts(symamd_handle h, int n, int m, const int *d_Ap, const int *d_Ai,
                          const double *d_Ax, const double *d_x, double *d_y, int nparts,
                          const int *d_sets, const int *d_ids);
/* z[i] = a_i * z[i] + b_i * y[i] for i < n, a_i = d_a[i] (or a when d_a is
 * NULL), b_i likewise: the finishing step of the reference's axpby forms of
 * spmv_csc (spmv.cpp:85-109).  FP64, on the stream. */
int symamd_axpby(symamd_handle h, int n, const double *d_a, const double *d_b, double a,
                 double b, const double *d_y, double *d_z);
/* Relative residual ||L X - B||_2 / ||B||_2 of every column c < k, written to
 * h_rel[c], using the plan's copy of L (blocking). */
int symamd_residual(symamd_handle h, symamd_plan plan, const void *d_X,
                    const void *d_B, int k, int64_t ld, symamd_layout layout,
                    double *h_rel);

/* ----------------------------------------------------------------- IC(0) */
/* Zero-fill incomplete Cholesky on the plan's pattern (FP64 plans).  d_Ax
 * holds the lower triangle of the SPD matrix A on the plan's CSR pattern
 * (diagonal last, the caller's entry order); d_Lx (may equal d_Ax) receives
 * L on the same pattern, ready for symamd_plan_update_values.  Variants:
 *   SYMAMD_IC0_LEFT  sym_lib::spic0_csc_left and spic0_csc_left_LBC
 *                    (sparse_blas/spic0.cpp:11-46, :48-99; textbook IC(0):
 *                    L[r,c] = (A[r,c] - sum_s L[r,s] L[c,s]) / L[c,c] over
 *                    every s < c both rows store, ascending)
 *   SYMAMD_IC0_CSR   sym_lib::spic0_csr and spic0_csr_lbc
 *                    (sparse_blas/spic0.cpp:106-127, :130-160; the same sum
 *                    restricted to the s whose last stored row is r)
 * Every entry is bit-identical to the serial reference routine.  Every row
 * is computed; a pivot that fails the variant's test (LEFT: <= 0, as
 * spic0_csc_left returns false; CSR: not > 0, spic0_csr's assertion) sets
 * *bad_row to the smallest such row and the call returns SYMAMD_ERR_NOT_SPD
 * (rows that read it hold NaN or inf); otherwise *bad_row = -1.  Blocks until
 * the factor is complete.  The update lists are built on the first call of a
 * variant and kept by the plan. */
typedef enum { SYMAMD_IC0_LEFT = 0, SYMAMD_IC0_CSR = 1 } symamd_ic0_variant;
int symamd_ic0(symamd_handle h, symamd_plan plan, symamd_ic0_variant variant,
               const double *d_Ax, double *d_Lx, int *bad_row);

/* ILU(0) of a general CSR matrix A (n rows; columns ascending in each row;
 * d_Adiag[i] = position of row i's diagonal), bit-identical to
 * sym_lib::spilu0_csr and spilu0_csr_lbc (sparse_blas/spilu0.cpp:11-51,
 * :54-110; decls sparse_blas_lib.h:276-281): d_l (may equal d_Ax) receives,
 * on A's pattern, L strictly below the diagonal (unit diagonal implied) and U
 * on and above it:
 *   L[i,c] = (A[i,c] - sum_k L[i,k] U[k,c]) / U[c,c]   (c < i)
 *   U[i,c] =  A[i,c] - sum_k L[i,k] U[k,c]             (c >= i)
 * over the k < min(i, c) where A stores (i, k) and (k, c), ascending.  Every
 * row is computed; a zero U[i,i] (the reference's assertion) sets *bad_row to
 * the smallest such row and returns SYMAMD_ERR_ZERO_PIVOT, else *bad_row =
 * -1.  SYMAMD_ERR_STRUCTURE when a row is not ascending or d_Adiag[i] is not
 * its diagonal.  Builds its schedule from A's lower part on each call and
 * blocks until the factor is complete. */
int symamd_ilu0(symamd_handle h, int n, const int *d_Ap, const int *d_Ai,
                const double *d_Ax, const int *d_Adiag, double *d_l, int *bad_row);
/* The same in two steps for repeated factorizations of one pattern:
 * symamd_ilu0_analyze checks the structure and builds the schedule and the
 * update lists (the plan copies the pattern; the arrays may be freed after
 * the call); symamd_ilu0_factor computes the factor (same results, errors
 * and blocking as symamd_ilu0). */
typedef struct symamd_ilu0_s *symamd_ilu0_plan;
int symamd_ilu0_analyze(symamd_handle h, int n, const int *d_Ap, const int *d_Ai,
                        const int *d_Adiag, symamd_ilu0_plan *plan);
int symamd_ilu0_factor(symamd_handle h, symamd_ilu0_plan plan, const double *d_Ax,
                       double *d_l, int *bad_row);
int symamd_ilu0_destroy(symamd_ilu0_plan plan);

/* ------------------------------------------------------ supernodal solve */
/* Forward solve on a supernodal (BCSC) lower triangle, bit-identical to
 * sym_lib::sptrsv_csc_block (sparse_blas/sptrsv.cpp:176-199, dense kernels
 * dense_blas/BLAS.cpp lsolve_BLAS :3-96 and matvec_BLAS :286-349; decl
 * sparse_blas_lib.h:212-214).  Layout (the reference's): supernode i owns
 * columns h_sup[i] .. h_sup[i+1]-1 (width w, at most 1024) and
 * h_nrows[i] >= w block rows; its values are a column-major h_nrows[i] x w
 * block at Lx[h_Lp[i]] whose first w rows are its own columns, and row r >= w
 * of the block is row h_Li[h_Lp[i] + r] of L, which must lie below the
 * supernode (SYMAMD_ERR_STRUCTURE otherwise).  symamd_sn_analyze reads the
 * HOST structure arrays (m = number of supernodes, h_sup[m] = n) and builds
 * the schedule on the device; symamd_sn_solve overwrites d_x (b in, x out,
 * FP64) asynchronously on the handle's stream; symamd_sn_check synchronises
 * and reports a watchdog expiry. */
typedef struct symamd_snplan_s *symamd_snplan;
int symamd_sn_analyze(symamd_handle h, int n, int m, const int *h_sup, const int *h_nrows,
                      const int *h_Lp, const int *h_Li, symamd_snplan *plan);
int symamd_sn_solve(symamd_handle h, symamd_snplan plan, const double *d_Lx, double *d_x);
/* k right-hand sides at once, column-major (column c at d_X + c*ld, ld >= n;
 * the layout the reference's multi-RHS supernodal solves use,
 * parsy/triangularSolve/Triangular_BCSC.cpp:62-121): each column's result is
 * bit-identical to sptrsv_csc_block on that column. */
int symamd_sn_solve_block(symamd_handle h, symamd_snplan plan, const double *d_Lx,
                          double *d_X, int k, int64_t ld);
/* The transposed solve L^T X = B on the same plan and storage, k
 * column-major right-hand sides (k = 1: a vector): the semantics of parsy's
 * blockedLTsolve (parsy/triangularSolve/Triangular_BCSC.cpp:125-186) and of
 * its multi-RHS and level-scheduled variants (:187-249, :742-808,
 * :811-880), supernodes from the last to the first: per supernode, each top
 * column's right-hand side minus the sum over the rows below, ascending, of
 * the block entry times the solved value at that row; then the top's
 * transposed triangular solve from its last column (x_c /= M[c,c]; each
 * x_r, r < c, loses M[c,r] x_c).  The reference's own order depends on the
 * BLAS it links (none is present here), so this order is the library's,
 * the same in every column. */
int symamd_sn_solve_transposed(symamd_handle h, symamd_snplan plan, const double *d_Lx,
                               double *d_X, int k, int64_t ld);
int symamd_sn_check(symamd_handle h, symamd_snplan plan);
int symamd_sn_destroy(symamd_snplan plan);

/* ----------------------------------------------------------------- tuner */
/* Time the candidate schedules for k right-hand sides in `layout` on
 * scratch data and remember the fastest in the plan (and in a process-wide
 * cache keyed by the matrix fingerprint); SYMAMD_EXEC_AUTO then uses it.
 * Candidates: the sync-free unit executor (lanes per row, and for k > 8
 * column tiles of 8 to 32 columns solved as separate work items), the
 * level-set executor replayed from a graph (lanes per row) and the region
 * executor (for one right-hand side at region caps 512 and 1024; the chosen
 * cap also serves the plan's adjoint).  Writes a human-readable summary
 * (may be NULL). */
int symamd_tune(symamd_handle h, symamd_plan plan, int k,
                symamd_layout layout, char *summary, int summary_len);

/* Low-level schedule override used by the tuner and by tests: lanes per row
 * (1..64, power of two) for the given executor; 0 restores the heuristic. */
int symamd_plan_set_lanes(symamd_plan plan, int lanes_per_row);

/* ----------------------------------------------------------- diagnostics */
/* The region executor's schedule for a HOST CSR lower triangle (diagonal
 * last, FP64 values), built on the host by the steps the device path takes:
 * region cap `cap`, `band_slices` slices per published band, the forward
 * solve or (transposed != 0) L^T x = g.  info[8] receives {regions, max
 * rows per region, max halo rows, largest record (bytes), total record
 * bytes, max slices per region, levels, widest region depth}; when rec is
 * not NULL it receives the records (info[4] bytes, values filled) and rhead
 * the 8-int header of every region.  No device work; for tests that emulate
 * the region kernel on the CPU.  SYMAMD_ERR_NOT_SUPPORTED when the matrix
 * has no usable region schedule. */
int symamd_region_schedule_host(int n, const int *Lp, const int *Li, const double *Lx,
                                int cap, int band_slices, int transposed, int *info,
                                uint8_t *rec, int *rhead);

#ifdef __cplusplus
}
#endif
#endif
