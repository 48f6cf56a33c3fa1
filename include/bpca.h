/*
 * bpca.h -- C ABI of the MI355X sparse-PCA (IRLBA) library, libbrisc_pca.so.
 *
 * The library computes the top-k principal components of the implicitly
 * centred and scaled cells x genes matrix Z = (X - 1 mu^T) diag(1/sd), where X
 * is a CSR float32 matrix of log-normalised counts, by the implicitly
 * restarted Lanczos bidiagonalisation the reference implements in
 * brisc/pca.pyx:661-1257.  X.v and X^T.u stream FP32 values with FP32
 * accumulation; reorthogonalisation, restart and the small SVD run in FP64.
 * Every reduction has a fixed order keyed to data blocks, so results are
 * bit-identical from run to run and across launch shapes.
 *
 * Conventions: every entry point returns a status code (BPCA_OK on success)
 * and never throws across the ABI; bpca_last_error() returns a message for
 * the calling thread's last failure.  Pointers named d_* are device pointers
 * on the context's device; all other pointers are host pointers.  Calls are
 * synchronous: when they return, outputs are complete.
 */
#ifndef BPCA_H
#define BPCA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum bpca_status {
    BPCA_OK = 0,
    BPCA_ERR_INVALID_ARGUMENT = 1, /* ValueError/TypeError in pca() */
    BPCA_ERR_OUT_OF_MEMORY = 2,    /* MemoryError (bad_alloc via except +) */
    BPCA_ERR_DEVICE = 3,           /* HIP runtime failure */
    BPCA_ERR_NONCANONICAL = 4,     /* CSR row indices not strictly increasing
                                      or out of range */
    BPCA_ERR_NUMERICAL = 5,        /* small SVD did not converge */
    BPCA_ERR_CANCELLED = 6,        /* KeyboardInterrupt (PyErr_CheckSignals) */
    BPCA_ERR_BUDGET = 7,           /* request larger than the HBM budget */
    BPCA_ERR_NOT_FOUND = 8         /* unknown request id */
};

/* Message describing the calling thread's most recent failure. */
const char *bpca_last_error(void);

/* Library version, major*10000 + minor*100 + patch. */
int bpca_version(void);

/* ---------------------------------------------------------------- context */

/* Device, stream, HBM budget and reusable workspace.  No counterpart in the
 * reference, which has no device; it carries the GPU-specific knobs that are
 * kept out of pca()'s signature. */
typedef struct bpca_context bpca_context;

int bpca_context_create(int device, uint64_t hbm_budget_bytes,
                        bpca_context **out);
int bpca_context_destroy(bpca_context *ctx);
/* Queue all work of this context on `hip_stream` (a hipStream_t; NULL = the
 * context's own stream). */
int bpca_context_set_stream(bpca_context *ctx, void *hip_stream);
/* Request that the solves running on this context stop: every call that is
 * advancing solves when the request arrives (bpca_irlba_f32, _device,
 * _batched, bpca_engine_step) returns BPCA_ERR_CANCELLED within its next
 * few Lanczos steps (the role of the reference's PyErr_CheckSignals per
 * step, pca.pyx:1173-1174).  A call that starts later is not affected. */
int bpca_context_cancel(bpca_context *ctx);
/* Charge `bytes` of device memory held outside the library (for example a
 * request's input CSR waiting to be admitted) to the context's HBM budget,
 * so that admission counts it: BPCA_ERR_BUDGET if it does not fit beside
 * what is charged now (force = 1 charges it regardless).  No counterpart
 * in the reference (a serving-loop knob). */
int bpca_context_reserve(bpca_context *ctx, uint64_t bytes, int force);
/* Return bytes charged by bpca_context_reserve. */
int bpca_context_release(bpca_context *ctx, uint64_t bytes);
/* Budget (0 = none), bytes charged now (workspaces in use plus reserved
 * bytes) and workspace bytes cached for reuse. */
int bpca_context_budget(bpca_context *ctx, uint64_t *budget,
                        uint64_t *charged, uint64_t *cached);

/* Kernel timing with HIP events recorded on the context's stream around each
 * class of launches.  Classes: */
enum bpca_kernel_class {
    BPCA_K_SPMV = 0,     /* u = X v'                 */
    BPCA_K_SPMV_T = 1,   /* partials of X^T u        */
    BPCA_K_REORTH = 2,   /* U-side CGS, mean, alpha  */
    BPCA_K_VSIDE = 3,    /* group sums, V-side CGS   */
    BPCA_K_RESTART = 4,  /* U, V restart products    */
    BPCA_K_FINAL = 5,    /* PCs = U (Ub S)           */
    BPCA_K_SETUP = 6,    /* validation, column stats */
    BPCA_NUM_KERNEL_CLASSES = 7
};
int bpca_context_set_profiling(bpca_context *ctx, int enable);
/* The classes timed while profiling is on: bit c for class c (default all).
 * Each timed class adds two event records per macro step, and each record
 * waits for the launches before it to finish, so timing fewer classes
 * perturbs the solve less. */
int bpca_context_set_profiled_classes(bpca_context *ctx, uint32_t mask);
/* Accumulated milliseconds and launch counts per class since the last reset
 * (arrays of BPCA_NUM_KERNEL_CLASSES entries). */
int bpca_context_kernel_times(bpca_context *ctx, double *total_ms,
                              int64_t *launches);
int bpca_context_reset_kernel_times(bpca_context *ctx);

/* Device facts the kernels adapt to: dynamic LDS per workgroup, compute
 * units, and whether macro steps are replayed as a captured graph. */
int bpca_context_info(bpca_context *ctx, int64_t *max_lds_per_workgroup,
                      int32_t *num_cus, int32_t *uses_graph);

/* Load-group sizes of the X v and X^T u kernels (rows whose loads a wave
 * keeps in flight: 2, 4 or 8).  They change speed only: every order of
 * summation, hence every result bit, is the same for all sizes. */
int bpca_context_set_kernel_config(bpca_context *ctx, int32_t spmv_group,
                                   int32_t spmvt_group);
int bpca_context_get_kernel_config(bpca_context *ctx, int32_t *spmv_group,
                                   int32_t *spmvt_group);
/* Caps on the workgroups per compute unit of the X v and X^T u launches
 * (0 = no cap, else 1 to 16), applied by raising the launch's LDS request
 * so that one more workgroup no longer fits.  Speed only, like the load
 * groups: no order of summation depends on them.  No counterpart in the
 * reference (a device-occupancy knob); bpca_autotune sets them. */
int bpca_context_set_occupancy(bpca_context *ctx, int32_t spmv_wg,
                               int32_t spmvt_wg);
int bpca_context_get_occupancy(bpca_context *ctx, int32_t *spmv_wg,
                               int32_t *spmvt_wg);
/* Whether problems admitted from now on with num_genes <= 4096 compute
 * X^T u from a gene-sorted copy of X made at admission (1, the default) or
 * from the packed CSR (0).  The two sum each gene's block partial in
 * different fixed orders, so results agree to rounding, not bit for bit; a
 * problem keeps the choice it was admitted with.  The copy adds about 5
 * bytes per nonzero to the workspace.  No counterpart in the reference (a
 * data-layout knob of this library). */
int bpca_context_set_sorted_xtu(bpca_context *ctx, int32_t enable);
int bpca_context_get_sorted_xtu(bpca_context *ctx, int32_t *enable);
/* How problems admitted from now on reorthogonalise each new left Lanczos
 * vector (the reference's classical Gram-Schmidt, pca.pyx:997-1090):
 * 2 = two passes over U per step (dots, then the update), as the reference;
 * 1 = one pass per step wherever num_cells >= 8 num_genes (the update of a
 * step is made in the next step's pass over U, and X^T U is carried in gene
 * space); 0 (the default) = one pass where num_cells >= max(8 num_genes,
 * 65536), else two.  The two agree to rounding, not bit for bit; a problem
 * keeps the choice it was admitted with.  No counterpart in the reference
 * (a memory-traffic knob of this library). */
int bpca_context_set_reorth_passes(bpca_context *ctx, int32_t passes);
int bpca_context_get_reorth_passes(bpca_context *ctx, int32_t *passes);
/* Whether problems admitted from now on that reorthogonalise in one pass,
 * with num_genes <= 4096, form X v and X^T u of a Lanczos step from one read
 * of X (1, the default; no gene-sorted copy is made for them) or in two
 * passes (0).  The one-pass scheme applies X^T to the raw step vector less
 * its mean, which the pass knows beforehand from the column sums of X, so
 * the two products need not wait for the reorthogonalisation in between.
 * Results agree with the two-pass products to rounding, not bit for bit; a
 * problem keeps the choice it was admitted with.  No counterpart in the
 * reference (a memory-traffic knob of this library). */
int bpca_context_set_pair_products(bpca_context *ctx, int32_t enable);
int bpca_context_get_pair_products(bpca_context *ctx, int32_t *enable);
/* Bytes of device workspace a solve of this shape allocates. */
uint64_t bpca_workspace_bytes(uint64_t num_cells, uint32_t num_genes,
                              uint32_t k, uint32_t subspace_size);

/* Per-solve statistics. */
typedef struct bpca_solve_stats {
    int32_t converged;      /* all k Ritz values met the tolerance */
    int32_t restarts;       /* IRLBA iterations run (1 = no restart) */
    int64_t lanczos_steps;  /* Lanczos steps over all iterations */
    int64_t matvecs;        /* X.v plus X^T.u products (2 per step) */
    double seconds;         /* wall time from admission to collection */
    int32_t svd_sweeps;     /* Jacobi sweeps of the last small SVD */
    int32_t slot;           /* engine slot the problem ran in */
    double smax;            /* largest Ritz value seen (convergence scale) */
    int32_t reorth_passes;  /* passes over U per Lanczos step (1 or 2) */
    int32_t pair_products;  /* 1: X v and X^T u formed in one read of X */
} bpca_solve_stats;

/* ------------------------------------------------------- single solve */

/* Drop-in for brisc.pca.irlba (brisc/pca.pyx:661-678).
 *
 * Same arguments in the same order and with the same meaning; the index
 * arrays are all int32 (index_width 4) or all int64 (index_width 8), as the
 * reference's fused type `signed_integer` requires (cyutils.pxd:25-27).
 * The matvec operand is the CSR of X, except on the reference's serial CSC
 * path (is_csr = 0, num_threads = 1, match_parallel = 0;
 * single_cell.py:12054-12069), where it is the CSC of X (num_genes + 1
 * offsets, cell indices), which is transposed to the CSR on the device.  The
 * rmatvec operand may be NULL because X^T.u is computed from the CSR.  match_parallel,
 * is_csr, num_threads and chunk_size are validated and otherwise do not
 * change the result.  PCs (num_cells x k, float32, C order, caller-owned)
 * receives U_k S_k; *converged receives the return value of the reference
 * function.  Runs on the process-wide default context (current device). */
int bpca_irlba_f32(const float *data_matvec, const void *indices_matvec,
                   const void *indptr_matvec, const float *data_rmatvec,
                   const void *indices_rmatvec, const void *indptr_rmatvec,
                   int index_width, uint64_t num_cells, uint32_t num_genes,
                   uint32_t k, uint32_t subspace_size, float tolerance,
                   uint32_t max_iterations, uint64_t seed, int match_parallel,
                   int is_csr, uint32_t num_threads, uint64_t chunk_size,
                   float *PCs, int *converged);

/* Device-resident variant of bpca_irlba_f32 for data already in HBM.
 * d_indices are int32 column indices, d_indptr int32 or int64 (indptr_width
 * 4 or 8) row offsets with d_indptr[0] == 0.  d_PCs: num_cells x k float32,
 * C order.  singular_values (host, k doubles) and stats may be NULL.
 * Admission always checks on the device that indptr starts at 0 and never
 * decreases and that each row's column indices are strictly increasing and
 * below num_genes (BPCA_ERR_NONCANONICAL if not), in the same pass that packs
 * the indices to 16 bits for the solve; `validate` is accepted for
 * compatibility and has no further effect. */
int bpca_irlba_device(bpca_context *ctx, const float *d_data,
                      const int32_t *d_indices, const void *d_indptr,
                      int indptr_width, uint64_t num_cells, uint32_t num_genes,
                      uint32_t k, uint32_t subspace_size, float tolerance,
                      uint32_t max_iterations, uint64_t seed, int validate,
                      float *d_PCs, double *singular_values,
                      bpca_solve_stats *stats);

/* Allocation-free variant of bpca_irlba_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as bpca_irlba_device (validate is always on,
 * PCs bit-identical to bpca_irlba_device's for the same inputs); the solver
 * state lives in a library-held device array of 256 MiB per device instead
 * of allocated workspace, so it calls no hipMalloc / hipFree and no
 * allocating library routine.  A problem whose descriptor, work lists and
 * workspace (two FP64 num_cells x subspace_size Lanczos bases, a packed copy
 * of the CSR and small vectors) exceed 256 MiB is refused with
 * BPCA_ERR_INVALID_ARGUMENT (about 100k cells at subspace 100, 2000 genes
 * and 7% nonzeros).  All calls on one device share that array, which keeps
 * the last call's contents; calls are serialised per device inside the
 * process, so concurrent calls wait for each other instead of racing. */
int bpca_irlba_device_noalloc(bpca_context *ctx, const float *d_data,
                              const int32_t *d_indices, const void *d_indptr,
                              int indptr_width, uint64_t num_cells,
                              uint32_t num_genes, uint32_t k,
                              uint32_t subspace_size, float tolerance,
                              uint32_t max_iterations, uint64_t seed,
                              float *d_PCs, double *singular_values,
                              bpca_solve_stats *stats);

/* ------------------------------------------------- batched / engine */

/* One IRLBA problem on device-resident CSR (the arguments of
 * bpca_irlba_device gathered in a struct).  No counterpart in the reference,
 * which solves one matrix per call; SingleCell.pca() per cluster is the use
 * it batches. */
typedef struct bpca_problem {
    const float *data;        /* device, nnz float32 */
    const int32_t *indices;   /* device, nnz int32 column indices */
    const void *indptr;       /* device, num_cells+1 int32 or int64 */
    int32_t indptr_width;     /* 4 or 8 */
    int32_t validate;         /* accepted; the CSR is always checked */
    uint64_t num_cells;
    uint32_t num_genes;
    uint32_t k;
    uint32_t subspace_size;   /* <= min(num_cells, num_genes) */
    uint32_t max_iterations;
    float tolerance;
    uint32_t pad;
    uint64_t seed;
    float *PCs;               /* device, num_cells x k float32, C order */
} bpca_problem;

/* Solve many independent problems together: every launch advances all of
 * them (one Lanczos step each; restarts and final products inside the same
 * launch sequence), replayed as a captured HIP graph.  Each problem's result
 * is bit-identical to solving it alone.  stats (num_problems entries) and
 * singular_values (num_problems rows of sv_stride doubles) may be NULL. */
int bpca_irlba_batched(bpca_context *ctx, int32_t num_problems,
                       const bpca_problem *problems, bpca_solve_stats *stats,
                       double *singular_values, int32_t sv_stride);

/* Autotuner: time each load-group size of both kernels on `problem` (a
 * throwaway solve state; the problem's arrays are only read and its PCs
 * buffer is not written), each with no cap and with caps of 4, 3 and 2
 * workgroups per compute unit, and make the fastest pair the context's
 * configuration (bpca_context_get_kernel_config / _get_occupancy).
 * timings_ms (may be NULL) receives 6 averages: X v at 2, 4, 8 then X^T u
 * at 2, 4, 8, all without a cap. */
int bpca_autotune(bpca_context *ctx, const bpca_problem *problem, int32_t reps,
                  double *timings_ms);

/* Continuous-batching engine (the serving loop drives it): problems are
 * admitted while others run, subject to the context's HBM budget
 * (BPCA_ERR_BUDGET when a problem's workspace does not fit now). */
typedef struct bpca_engine bpca_engine;
int bpca_engine_create(bpca_context *ctx, bpca_engine **out);
int bpca_engine_destroy(bpca_engine *engine);
/* Admit a problem (set-up kernels run at the next step); *slot names it. */
int bpca_engine_add(bpca_engine *engine, const bpca_problem *problem,
                    int32_t *slot);
/* Advance every admitted problem by up to max_steps Lanczos steps; returns
 * early once a problem has finished.  *num_done = finished, uncollected. */
int bpca_engine_step(bpca_engine *engine, int32_t max_steps,
                     int32_t *num_done);
/* Retire up to max_out finished problems: their slots, statistics and
 * singular values (rows of sv_stride doubles); frees their workspace. */
int bpca_engine_collect(bpca_engine *engine, int32_t max_out, int32_t *slots,
                        bpca_solve_stats *stats, double *singular_values,
                        int32_t sv_stride, int32_t *num_out);
/* Number of admitted, uncollected problems. */
int bpca_engine_active(bpca_engine *engine, int32_t *active);

/* The device's small SVD on one m x (m+1) matrix B (host, column-major):
 * B = Ub diag(Sb) Vb^T with Ub, Vb m x m column-major.  Test entry point of
 * the kernel the restart uses (role of sgesdd, brisc/pca.pyx:514-545). */
int bpca_test_small_svd(bpca_context *ctx, int32_t m, const double *B,
                        double *Ub, double *Sb, double *Vb, int32_t *sweeps);

/* out (n x kout) = A (n x m) C (m x kout), all column-major float64 host
 * arrays, through the MFMA tile code of the restart product (test entry
 * point; m <= 128, kout <= 128). */
int bpca_test_gemm(bpca_context *ctx, int64_t n, int32_t m, int32_t kout,
                   const double *A, const double *C, double *out);
/* The same product on device arrays (column-major d_A n x m, d_C m x kout,
 * d_out n x kout); synchronous. */
int bpca_test_gemm_device(bpca_context *ctx, int64_t n, int32_t m,
                          int32_t kout, const double *d_A, const double *d_C,
                          double *d_out);
/* Calibration: the FP64 rate this device sustains on every compute unit with
 * register operands only (which = 0: v_mfma_f64_16x16x4f64, the restart and
 * final products' instruction; 1: v_fma_f64 on the vector ALUs), in
 * TFLOP/s, timed with HIP events over `reps` launches.  No counterpart in
 * the reference; bench.py prices the FP64 products against it beside the
 * published peak. */
int bpca_test_fp64_rate(bpca_context *ctx, int32_t which, int32_t reps,
                        double *tflops);

/* ------------------------------------------------ synthetic workloads */

/* Device generator of the synthetic NB count workloads (no counterpart in the
 * reference; benchmark input).  Rows [row0, row0+num_rows) of the workload
 * with per-type gene means type_gene_mean (num_types x num_genes, float64),
 * per-row cell_type (int32) and cell_factor (float64), all device arrays
 * indexed from row0.  synth_row_nnz writes the nonzero count of each row;
 * synth_fill writes int32 column indices and int32 counts at the offsets
 * d_indptr (int64, num_rows+1). */
int bpca_synth_row_nnz(bpca_context *ctx, uint64_t seed, uint64_t row0,
                       uint64_t num_rows, uint32_t num_genes,
                       const double *d_type_gene_mean,
                       const int32_t *d_cell_type, const double *d_cell_factor,
                       int64_t *d_row_nnz);
int bpca_synth_fill(bpca_context *ctx, uint64_t seed, uint64_t row0,
                    uint64_t num_rows, uint32_t num_genes,
                    const double *d_type_gene_mean, const int32_t *d_cell_type,
                    const double *d_cell_factor, const int64_t *d_indptr,
                    int32_t *d_indices, int32_t *d_counts);

/* SingleCell.normalize() on the device: brisc/normalize.pyx:9-268
 * (normalize_csr, is_csr = 1) and 271-460 (normalize_csc, is_csr = 0).
 *   d_data      counts, data_type 0 int32, 1 uint32, 2 int64, 3 uint64,
 *               4 float32, 5 float64 (the reference's `numeric` types)
 *   d_indices   CSR: gene of each entry; CSC: cell of each entry
 *   d_indptr    CSR: num_cells + 1 offsets; CSC: num_genes + 1 offsets
 *               (index_width / indptr_width 4 or 8)
 *   d_qc        uint8 per cell, nonzero = passed QC (NULL: all cells); the
 *               QC cells alone define the mean library size
 *   method      0 logCP10k, 1 log1pPF, 2 PFlog1pPF
 *   d_normalized  float32 output per entry (may be d_data for float32
 *               counts: the reference's inplace=True)
 *   d_row_sums  uint64 per cell: its library size, as the reference returns
 * Each value is the reference's, bit for bit: the same single-precision
 * operations in the same order (PFlog1pPF's per-cell sums and their total
 * included) and log1p bit-identical to the host C library's log1pf.
 * Synchronous.  When d_qc selects no cell the reference's pca() raises
 * before calling (no cells pass QC); here the factor is then NaN. */
int bpca_normalize(bpca_context *ctx, const void *d_data, int data_type,
                   const void *d_indices, int index_width, const void *d_indptr,
                   int indptr_width, int is_csr, int64_t num_cells,
                   int64_t num_genes, const uint8_t *d_qc, int method,
                   float *d_normalized, uint64_t *d_row_sums);

/* y = log1pf(x) elementwise on the device with the restated host-library
 * log1pf that bpca_normalize uses (test entry point). */
int bpca_test_log1pf(bpca_context *ctx, const float *d_x, int64_t n,
                     float *d_y);

/* log1pPF normalisation of integer counts on the device
 * (brisc/normalize.pyx:9-128, method 'log1pPF', no QC column):
 * x' = log1p(x * (mean_row_sum / row_sum)) in float32. */
int bpca_normalize_log1ppf(bpca_context *ctx, const int32_t *d_counts,
                           const int64_t *d_indptr, uint64_t num_cells,
                           float *d_out);

/* QC x HVG subset of a device-resident CSR matrix, the operand preparation
 * of SingleCell.pca() (brisc/single_cell.py:11853-11908, which takes
 * X[np.ix_(QC rows, HVG columns)] through brisc/cysparse.pyx:926-1029), done
 * on the device in two calls.  All pointers are device pointers.
 *   rows      int64[num_out_rows]: output row r is input row rows[r]
 *             (NULL: all num_rows rows, num_out_rows = num_rows)
 *   col_map   int32[num_cols]: input column c becomes column col_map[c] of
 *             the output, or is dropped when col_map[c] < 0 (NULL: identity)
 * bpca_csr_subset_count writes the kept entries of every output row to
 * row_nnz (int64[num_out_rows]) after checking the rows, row offsets and
 * column indices it reads (BPCA_ERR_NONCANONICAL if any is out of range).
 * The caller forms out_indptr (int64[num_out_rows+1], exclusive prefix sum
 * of row_nnz) and allocates the outputs; bpca_csr_subset_fill writes the
 * values and int32 column indices of each output row in ascending new
 * column order (the input rows' column indices must be strictly
 * increasing).  col_map_increasing = 1 promises that col_map increases over
 * the kept columns (one dataset's HVG positions); otherwise each row is
 * sorted on the device and max_row_nnz (the largest row_nnz) must not
 * exceed 4096.  Both calls are synchronous on the context's stream. */
int bpca_csr_subset_count(bpca_context *ctx, const void *d_indptr,
                          int indptr_width, const void *d_indices,
                          int index_width, int64_t num_rows, int64_t nnz,
                          int64_t num_cols, const int64_t *d_rows,
                          int64_t num_out_rows, const int32_t *d_col_map,
                          int64_t *d_row_nnz);
int bpca_csr_subset_fill(bpca_context *ctx, const float *d_data,
                         const void *d_indptr, int indptr_width,
                         const void *d_indices, int index_width,
                         const int64_t *d_rows, int64_t num_out_rows,
                         const int32_t *d_col_map, int col_map_increasing,
                         const int64_t *d_out_indptr, int64_t max_row_nnz,
                         float *d_out_data, int32_t *d_out_indices);

/* Transpose of a sparse matrix stored by rows: CSR -> CSC or CSC -> CSR, the
 * role of brisc's csr_tocsc / csc_tocsr (brisc/cysparse.pyx:1179-1295,
 * sparse.py:928-947), on the device.  Input: num_rows + 1 offsets and
 * indices in [0, num_cols), int32 or int64 (widths 4 or 8).  Output:
 * num_cols + 1 int64 offsets, int32 indices and the values, each output row
 * listing its entries in ascending input-row order (as the reference's
 * stable scatter and scipy's tocsc/tocsr).  Deterministic (stable radix sort,
 * no atomics).  BPCA_ERR_NONCANONICAL for offsets or indices out of range.
 * Synchronous on the context's stream; allocates its work space. */
int bpca_csr_transpose(bpca_context *ctx, const float *d_data,
                       const void *d_indices, int index_width,
                       const void *d_indptr, int indptr_width,
                       int64_t num_rows, int64_t num_cols, float *d_out_data,
                       int32_t *d_out_indices, int64_t *d_out_indptr);

/* The exact k nearest neighbours of every cell among all cells, from PCs
 * held on the device: the exhaustive form of SingleCell.neighbors()
 * (brisc/single_cell.py:12139-12665, whose knn_self in brisc/knn.pyx:13-516
 * returns the exact neighbours when every cluster is searched).
 *   d_X          float32 num_cells x num_dims, row-major (num_dims <= 256)
 *   d_neighbors  uint32 num_cells x k: neighbour indices (rows of X)
 *   d_distances  float32 num_cells x k: squared Euclidean distances formed
 *                as ||q||^2 + (||c||^2 - 2 q.c) in FP32, as the reference
 * A cell is not its own neighbour; each row is in ascending distance, ties
 * in ascending index.  1 <= k <= num_cells - 1; more than 64 neighbours
 * run in passes of 64, each pass keeping only
 * candidates after the previous pass's last (distance, index).  Synchronous. */
int bpca_knn_exact(bpca_context *ctx, const float *d_X, int64_t num_cells,
                   int32_t num_dims, int32_t k, uint32_t *d_neighbors,
                   float *d_distances);

/* SingleCell.neighbors() with its default inverted-file search
 * (brisc/single_cell.py:12139-12665: kmeans() of brisc/kmeans.pyx:889-1100
 * with random initialisation, then knn_self of brisc/knn.pyx:13-516).
 *   d_X            float32 num_cells x num_dims PCs (num_dims <= 256)
 *   num_clusters   k-means clusters, 1 <= num_clusters <= min(num_cells, 2^26)
 *   num_clusters_searched  nearest clusters searched per cell (raised, as
 *                  the reference does, until k + 1 candidates are certain)
 *   num_kmeans_iterations, kmeans_tolerance, chunk_size_kmeans, seed: as
 *                  the reference's arguments of the same names
 *   d_neighbors / d_distances: as bpca_knn_exact (indices are rows of X)
 *   info           int64[5] out: non-empty clusters, clusters searched,
 *                  k-means iterations run, 1 if every cluster was searched
 *                  (the exhaustive search ran), (workgroup, cluster) pairs
 * Synchronous. */
int bpca_knn_ivf(bpca_context *ctx, const float *d_X, int64_t num_cells,
                 int32_t num_dims, int32_t k, int32_t num_clusters,
                 int32_t num_clusters_searched, int32_t num_kmeans_iterations,
                 float kmeans_tolerance, int64_t chunk_size_kmeans,
                 uint64_t seed, uint32_t *d_neighbors, float *d_distances,
                 int64_t *info);

/* bpca_knn_ivf with the initialisation choice of SingleCell.neighbors()
 * (brisc/single_cell.py:12151-12155, 12532-12542): kmeans_barbar = 1 starts
 * k-means from k-means|| (brisc/kmeans.pyx:41-720; see
 * bpca_kmeans_barbar_init) with num_init_iterations (>= 1) and
 * oversampling_factor (> 0); 0 draws random cells (the two other arguments
 * unused).  info is int64[6]: bpca_knn_ivf's five, then the k-means||
 * candidates selected (0 for random cells).  Synchronous. */
int bpca_knn_ivf_ex(bpca_context *ctx, const float *d_X, int64_t num_cells,
                    int32_t num_dims, int32_t k, int32_t num_clusters,
                    int32_t num_clusters_searched, int32_t num_kmeans_iterations,
                    float kmeans_tolerance, int64_t chunk_size_kmeans,
                    uint64_t seed, int32_t kmeans_barbar,
                    int32_t num_init_iterations, float oversampling_factor,
                    uint32_t *d_neighbors, float *d_distances, int64_t *info);

/* kmeans() with random initialisation (brisc/kmeans.pyx:889-1432), the
 * first stage of bpca_knn_ivf, on its own: d_labels (uint32 num_cells,
 * from the final E-step against the final centroids), d_centroids (float32
 * num_clusters x num_dims) and d_counts (uint32 num_clusters, of those
 * labels) as the reference's kmeans leaves them; *iterations_run =
 * iterations done.
 * Arguments as bpca_knn_ivf.  Synchronous. */
int bpca_kmeans(bpca_context *ctx, const float *d_X, int64_t num_cells,
                int32_t num_dims, int32_t num_clusters,
                int32_t num_kmeans_iterations, float kmeans_tolerance,
                int64_t chunk_size_kmeans, uint64_t seed, uint32_t *d_labels,
                float *d_centroids, uint32_t *d_counts,
                int32_t *iterations_run);

/* bpca_kmeans with kmeans()'s kmeans_barbar, num_init_iterations and
 * oversampling_factor arguments (brisc/kmeans.pyx:889-903, 938-948),
 * meaning as in bpca_knn_ivf_ex.  Synchronous. */
int bpca_kmeans_ex(bpca_context *ctx, const float *d_X, int64_t num_cells,
                   int32_t num_dims, int32_t num_clusters,
                   int32_t num_kmeans_iterations, float kmeans_tolerance,
                   int64_t chunk_size_kmeans, uint64_t seed, int32_t kmeans_barbar,
                   int32_t num_init_iterations, float oversampling_factor,
                   uint32_t *d_labels, float *d_centroids, uint32_t *d_counts,
                   int32_t *iterations_run);

/* The cells k-means|| initialisation takes as centroids
 * (kmeans_barbar_init, brisc/kmeans.pyx:41-360; kmeans_barbar_init_parallel,
 * 362-720, takes the same), from d_X (float32 num_cells x num_dims,
 * num_dims <= 256): rounds of draws from srand(seed + cell * round), each
 * cell taken with probability oversampling_factor * num_clusters *
 * distance / cost, until num_init_iterations rounds and at least
 * num_clusters candidates (or every cell); then weighted k-means++ over the
 * candidates.  Distances by explicit differences, fused, dimensions in
 * order; the rounds' candidate updates by partial distances on the matrix
 * cores; costs as chunk_size_kmeans chunk sums in order (k-means++: sums of
 * 256-candidate halving trees in order); the k-means++ draw the reference's
 * fused sequential chain.  cells: host uint32[num_clusters] out;
 * *num_candidates: candidates selected.  The reference's errors (every
 * cell at one point; fewer distinct cells than clusters) are
 * BPCA_ERR_INVALID_ARGUMENT with its message.  Synchronous. */
int bpca_kmeans_barbar_init(bpca_context *ctx, const float *d_X, int64_t num_cells,
                            int32_t num_dims, int32_t num_clusters,
                            int32_t num_init_iterations, float oversampling_factor,
                            int64_t chunk_size_kmeans, uint64_t seed, uint32_t *cells,
                            int64_t *num_candidates);

/* The IRLBA start vector for `seed`: the num_genes FP32 normal draws that
 * brisc/pca.pyx:801-806 puts in V[:, 0] (srand and random_normal of
 * brisc/cyutils.pyx:1153-1191, 1228-1232: PCG-XSH-RR, FP32 Box-Muller) go to
 * `draws` and the same vector normalised in FP64 (the reference normalises in
 * FP32) to `v0`; either may be NULL.  Host only, no device. */
int bpca_start_vector(uint64_t seed, uint32_t num_genes, float *draws,
                      double *v0);

/* The cells kmeans() takes as initial centroids with random initialisation
 * (kmeans_random_init, brisc/kmeans.pyx:13-39): num_clusters distinct cells
 * drawn with the reference's PCG32 generator and Lemire's bounded draw
 * (brisc/cyutils.pyx:1153-1232) from srand(seed).  Host only, no device. */
int bpca_kmeans_init_cells(int64_t num_cells, int32_t num_clusters,
                           uint64_t seed, uint32_t *cells);

#ifdef __cplusplus
}
#endif

#endif /* BPCA_H */
