// Launchers of the device kernels of the IRLBA path (kernels.hip).
//
// Data layout in HBM (n cells, g genes, m subspace size, k PCs):
//   X       CSR: data float32[nnz], indices int32[nnz], indptr int32|int64[n+1]
//   U, U2   float64 column-major n x m (ld n): left Lanczos vectors
//   V, V2   float64 column-major g x (m+1) (ld g): right Lanczos vectors
//   u32     float32[n]: raw X.v' of the current step
//   vs      float32[g]: v_j / sd, the SpMV input
//   B       float64 column-major m x (m+1): projected matrix
//   Every reduction over cells is keyed to fixed data blocks of kRowBlock
//   rows, never to the launch geometry.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace bpca {

constexpr int kWave = 64;
constexpr int kRowBlock = 1024;   // rows per data block
constexpr int kSubRows = 256;     // rows per wave sub-block of X^T u
constexpr int kGroup = 64;        // data blocks per first-level group
constexpr int kMaxGenesTranspose = 8192;  // LDS bound of the X^T u kernel

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

struct Csr {
    const float *data;
    const int32_t *indices;
    const void *indptr;
    int indptr_width;  // 4 or 8
    int64_t n;
    int g;
};

// CSR sanity: indptr monotone from 0, each row's indices strictly increasing
// and in [0, g).  Sets *d_flag nonzero on violation.
hipError_t launch_validate(const Csr &X, int *d_flag, hipStream_t s);

// Narrow int64 column indices to int32, flagging values outside [0, g).
hipError_t launch_narrow_indices(const int64_t *in, int64_t nnz, int g,
                                 int32_t *out, int *d_flag, hipStream_t s);

// u32 = X vs  (FP32 gather of vs from LDS, FP32 accumulation).
hipError_t launch_spmv(const Csr &X, const float *vs, float *u32, int grid,
                       hipStream_t s);

// Per-block partials of the reorthogonalisation: P[c*nb + b] with c = 0 the
// block sum of u32 and c = 1..j the dot of U[:, c-1] with u32.
hipError_t launch_reorth_partials(const float *u32, const double *U,
                                  int64_t n, int j, double *P, hipStream_t s);
// red[c] = sum over blocks of P[c*nb + b], c = 0..ncols-1.
hipError_t launch_reduce_columns(const double *P, int64_t nb, int ncols,
                                 double *red, hipStream_t s);
// U[:, j] = u32 - (red[0]/n + U[:, :j] red[1..j]); sq[b] = block sum of squares.
hipError_t launch_reorth_apply(const float *u32, double *U, int64_t n, int j,
                               const double *red, double *sq, hipStream_t s);
// alpha = max(sqrt(sum sq), 1e-8): scal[0] = alpha, scal[1] = 1/alpha,
// B[j + j*m] = alpha.
hipError_t launch_alpha(const double *sq, int64_t nb, double *scal, double *B,
                        int m, int j, hipStream_t s);

// X^T u with u = U[:, j] * scal[1] (written back normalised):
// partial[b*g + gene] over data blocks b.
hipError_t launch_spmv_t(const Csr &X, double *Ucol, const double *scal,
                         float *partial, hipStream_t s);
// Column statistics: partial[b*2g + gene] = block sum of x,
// partial[b*2g + g + gene] = block sum of x^2.
hipError_t launch_colstats(const Csr &X, float *partial, hipStream_t s);
// Largest dynamic LDS request of any kernel for g genes and subspace m.
size_t lds_bytes_needed(int g, int m);
// First-level group sums: grp[G*width + i] = sum of partial rows in group G.
hipError_t launch_reduce_groups(const float *partial, int64_t nblocks,
                                int width, double *grp, hipStream_t s);
// sd[gene] from the group sums of colstats.
hipError_t launch_finish_stddev(const double *grp, int64_t ngroups, int g,
                                int64_t n, double *sd, hipStream_t s);

// V-side step j: w = (X^T u)/sd - alpha V[:, j], CGS against V[:, :j+1]
// (all m columns on the last step), beta, and either V[:, j+1] + vs + B, or
// the normalised remainder + its norm (rscal[0]) on the last step.
hipError_t launch_vside(const double *grp, int64_t ngroups, int g,
                        const double *sd, double *V, int j, int m, double *B,
                        const double *scal, double *rem, double *rscal,
                        float *vs, hipStream_t s);
// vs = V[:, 0] / sd.
hipError_t launch_init_vs(const double *V, const double *sd, int g, float *vs,
                          hipStream_t s);

// out[:, c] = A coef[:, c] for a tall column-major A (n x kin, ld lda) and a
// small column-major coef (kin x kout): column-major float64 output (ld ldo).
hipError_t launch_tall_gemm_f64(const double *A, int64_t lda, int64_t n,
                                int kin, const double *coef, int kout,
                                double *out, int64_t ldo, hipStream_t s);
// Same product written as a row-major float32 n x kout matrix (the PCs).
hipError_t launch_tall_gemm_rowmajor_f32(const double *A, int64_t lda,
                                         int64_t n, int kin,
                                         const double *coef, int kout,
                                         float *out, hipStream_t s);
// Small V restart: V2[:, c] = V[:, :m] Vbt[c, :m]^T for c < kc (Vbt is
// column-major m x (m+1)), then V2[:, kc] = rem.
hipError_t launch_v_restart(const double *V, int g, int m, const double *Vbt,
                            int kc, const double *rem, double *V2,
                            hipStream_t s);

// Synthetic workload generator and log1pPF (synth.hip).
hipError_t launch_synth_row_nnz(uint64_t seed, uint64_t row0, int64_t rows,
                                int g, const double *tgm, const int32_t *type,
                                const double *factor, int64_t *row_nnz,
                                hipStream_t s);
hipError_t launch_synth_fill(uint64_t seed, uint64_t row0, int64_t rows, int g,
                             const double *tgm, const int32_t *type,
                             const double *factor, const int64_t *indptr,
                             int32_t *indices, int32_t *counts, hipStream_t s);
hipError_t launch_row_sums(const int32_t *counts, const int64_t *indptr,
                           int64_t n, uint64_t *row_sums, hipStream_t s);
hipError_t launch_sum_u64(const uint64_t *x, int64_t n, uint64_t *out,
                          hipStream_t s);
hipError_t launch_log1ppf(const int32_t *counts, const int64_t *indptr,
                          int64_t n, const uint64_t *row_sums,
                          const uint64_t *total, float *out, hipStream_t s);

}  // namespace bpca
