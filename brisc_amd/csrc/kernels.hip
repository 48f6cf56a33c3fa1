// Device kernels of the IRLBA path for gfx950 (CDNA4, wave64).
//
// Reference loops they replace (brisc/pca.pyx):
//   spmv_kernel            matvec_csr_parallel_raw       440-475
//   reorth_* / alpha       fused passes 1 and 2, alpha   1006-1090
//   spmv_t_kernel          rmatvec_csc_parallel_raw      478-511 (from CSR)
//   colstats_kernel        clipped_stddev_csc            51-103
//   vside_kernel           V update, CGS, beta           1103-1171
//   tall_gemm_*            matmul_nn_f_out / _c_out      578-658, 1211-1254
//
// Determinism: no floating-point atomics anywhere.  Every sum has a fixed
// order defined by the data (row position, fixed row blocks of kRowBlock
// rows, fixed sub-blocks of kSubRows rows, fixed groups of kGroup blocks),
// never by grid size or scheduling, so results are bit-identical from run to
// run, for any grid size, and when problems share the device.
#include "kernels.h"

namespace bpca {

namespace {

__device__ inline float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Sum over a block of NW waves: xor-tree inside each wave, then a pairwise
// tree over the wave sums in wave order.  `red` holds NW doubles.
template <int NW>
__device__ inline double block_sum(double v, double *red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    if (lane == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int s = 1; s < NW; s <<= 1)
            for (int i = 0; i + s < NW; i += 2 * s) red[i] = red[i] + red[i + s];
    }
    __syncthreads();
    double r = red[0];
    __syncthreads();
    return r;
}

__device__ inline int wave_id() {
    return __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
}

}  // namespace

// ------------------------------------------------------------ validation

template <typename IP>
__global__ __launch_bounds__(256) void validate_kernel(
    const int32_t *__restrict__ indices, const IP *__restrict__ indptr,
    int64_t n, int g, int *flag) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t row = (int64_t)blockIdx.x * 4 + wave_id();
    if (row >= n) return;
    const int64_t b = indptr[row], e = indptr[row + 1];
    int bad = 0;
    if (row == 0 && b != 0) bad = 1;
    if (e < b) bad = 1;
    for (int64_t p = b + lane; p < e && !bad; p += kWave) {
        int32_t c = indices[p];
        int32_t prev = p > b ? indices[p - 1] : -1;
        if (c < 0 || c >= g || c <= prev) bad = 1;
    }
    if (bad) atomicOr(flag, 1);
}

hipError_t launch_validate(const Csr &X, int *d_flag, hipStream_t s) {
    dim3 grid((unsigned)ceil_div(X.n, 4));
    if (X.indptr_width == 8)
        validate_kernel<int64_t><<<grid, 256, 0, s>>>(
            X.indices, (const int64_t *)X.indptr, X.n, X.g, d_flag);
    else
        validate_kernel<int32_t><<<grid, 256, 0, s>>>(
            X.indices, (const int32_t *)X.indptr, X.n, X.g, d_flag);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void narrow_indices_kernel(
    const int64_t *__restrict__ in, int64_t nnz, int g,
    int32_t *__restrict__ out, int *flag) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    int bad = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nnz; i += stride) {
        const int64_t v = in[i];
        if (v < 0 || v >= g) bad = 1;
        out[i] = (int32_t)v;
    }
    if (bad) atomicOr(flag, 1);
}

hipError_t launch_narrow_indices(const int64_t *in, int64_t nnz, int g,
                                 int32_t *out, int *d_flag, hipStream_t s) {
    if (nnz <= 0) return hipSuccess;
    const int grid = (int)std::min<int64_t>(ceil_div(nnz, 256), 65536);
    narrow_indices_kernel<<<grid, 256, 0, s>>>(in, nnz, g, out, d_flag);
    return hipGetLastError();
}

// ------------------------------------------------------------ u = X vs
//
// One wave per row (grid-stride over rows).  Canonical order of a row's sum:
// lane l accumulates the row's elements at positions l, l+64, l+128, ...
// (relative to the row start) with FP32 fma in that order, then the 64 lane
// sums are combined by the xor butterfly 32,16,8,4,2,1.

template <typename IP>
__global__ __launch_bounds__(256) void spmv_kernel(
    const float *__restrict__ data, const int32_t *__restrict__ indices,
    const IP *__restrict__ indptr, const float *__restrict__ vs, int g,
    int64_t n, float *__restrict__ u) {
    extern __shared__ float lds_v[];
    for (int i = threadIdx.x; i < g; i += blockDim.x) lds_v[i] = vs[i];
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t stride = (int64_t)gridDim.x * 4;
    for (int64_t row = (int64_t)blockIdx.x * 4 + wave_id(); row < n;
         row += stride) {
        const int64_t b = indptr[row], e = indptr[row + 1];
        float acc = 0.0f;
        for (int64_t base = b; base < e; base += 4 * kWave) {
            float d[4];
            int ix[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int64_t p = base + lane + t * kWave;
                const bool ok = p < e;
                d[t] = ok ? data[p] : 0.0f;
                ix[t] = ok ? indices[p] : 0;
            }
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const bool ok = base + lane + t * kWave < e;
                if (ok) acc = fmaf(d[t], lds_v[ix[t]], acc);
            }
        }
        acc = wave_sum(acc);
        if (lane == 0) u[row] = acc;
    }
}

hipError_t launch_spmv(const Csr &X, const float *vs, float *u32, int grid,
                       hipStream_t s) {
    if (grid <= 0) grid = (int)std::min<int64_t>(ceil_div(X.n, 4), 8192);
    size_t lds = sizeof(float) * (size_t)X.g;
    if (X.indptr_width == 8)
        spmv_kernel<int64_t><<<grid, 256, lds, s>>>(
            X.data, X.indices, (const int64_t *)X.indptr, vs, X.g, X.n, u32);
    else
        spmv_kernel<int32_t><<<grid, 256, lds, s>>>(
            X.data, X.indices, (const int32_t *)X.indptr, vs, X.g, X.n, u32);
    return hipGetLastError();
}

// ------------------------------------------- reorthogonalisation (FP64)
//
// Block b covers rows [b*1024, b*1024+1024).  Within a block, lane l of a
// wave handles rows l, l+64, ..., l+960 in that order, then xor butterfly.

__global__ __launch_bounds__(256) void reorth_partials_kernel(
    const float *__restrict__ u32, const double *__restrict__ U, int64_t n,
    int j, double *__restrict__ P, int64_t nb) {
    constexpr int R = kRowBlock / kWave;  // 16 rows per lane
    const int lane = threadIdx.x & (kWave - 1);
    const int w = wave_id();
    const int64_t b = blockIdx.x;
    const int64_t r0 = b * kRowBlock;
    const int64_t nrows = std::min<int64_t>(kRowBlock, n - r0);
    double uv[R];
#pragma unroll
    for (int t = 0; t < R; t++) {
        const int r = lane + t * kWave;
        uv[t] = r < nrows ? (double)u32[r0 + r] : 0.0;
    }
    for (int c = w; c <= j; c += 4) {
        double acc = 0.0;
        if (c == 0) {
#pragma unroll
            for (int t = 0; t < R; t++) acc += uv[t];
        } else {
            const double *Uc = U + (int64_t)(c - 1) * n + r0;
#pragma unroll
            for (int t = 0; t < R; t++) {
                const int r = lane + t * kWave;
                const double x = r < nrows ? Uc[r] : 0.0;
                acc = fma(x, uv[t], acc);
            }
        }
        acc = wave_sum(acc);
        if (lane == 0) P[(int64_t)c * nb + b] = acc;
    }
}

hipError_t launch_reorth_partials(const float *u32, const double *U,
                                  int64_t n, int j, double *P, hipStream_t s) {
    const int64_t nb = ceil_div(n, kRowBlock);
    reorth_partials_kernel<<<(unsigned)nb, 256, 0, s>>>(u32, U, n, j, P, nb);
    return hipGetLastError();
}

// Thread t sums blocks t, t+256, ... in order; then block_sum<4>.
__global__ __launch_bounds__(256) void reduce_columns_kernel(
    const double *__restrict__ P, int64_t nb, double *__restrict__ red) {
    __shared__ double tmp[4];
    const double *p = P + (int64_t)blockIdx.x * nb;
    double acc = 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += 256) acc += p[b];
    acc = block_sum<4>(acc, tmp);
    if (threadIdx.x == 0) red[blockIdx.x] = acc;
}

hipError_t launch_reduce_columns(const double *P, int64_t nb, int ncols,
                                 double *red, hipStream_t s) {
    reduce_columns_kernel<<<ncols, 256, 0, s>>>(P, nb, red);
    return hipGetLastError();
}

// Thread t handles rows r0+t, r0+t+256, r0+t+512, r0+t+768 in order; the
// correction is mean + sum_c U[r,c] coef[c] accumulated over c ascending.
__global__ __launch_bounds__(256) void reorth_apply_kernel(
    const float *__restrict__ u32, double *__restrict__ U, int64_t n, int j,
    const double *__restrict__ red, double *__restrict__ sq) {
    extern __shared__ double coef[];  // j+1 values
    __shared__ double tmp[4];
    for (int i = threadIdx.x; i <= j; i += 256) coef[i] = red[i];
    __syncthreads();
    const double mean = coef[0] / (double)n;
    const int64_t r0 = (int64_t)blockIdx.x * kRowBlock;
    double acc = 0.0;
    double *Uj = U + (int64_t)j * n;
#pragma unroll
    for (int t = 0; t < kRowBlock / 256; t++) {
        const int64_t r = r0 + threadIdx.x + t * 256;
        if (r < n) {
            double d = mean;
            for (int c = 0; c < j; c++) d = fma(U[(int64_t)c * n + r], coef[c + 1], d);
            const double x = (double)u32[r] - d;
            Uj[r] = x;
            acc = fma(x, x, acc);
        }
    }
    acc = block_sum<4>(acc, tmp);
    if (threadIdx.x == 0) sq[blockIdx.x] = acc;
}

hipError_t launch_reorth_apply(const float *u32, double *U, int64_t n, int j,
                               const double *red, double *sq, hipStream_t s) {
    const int64_t nb = ceil_div(n, kRowBlock);
    reorth_apply_kernel<<<(unsigned)nb, 256, sizeof(double) * (j + 1), s>>>(
        u32, U, n, j, red, sq);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void alpha_kernel(
    const double *__restrict__ sq, int64_t nb, double *__restrict__ scal,
    double *__restrict__ B, int m, int j) {
    __shared__ double tmp[4];
    double acc = 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += 256) acc += sq[b];
    acc = block_sum<4>(acc, tmp);
    if (threadIdx.x == 0) {
        double alpha = sqrt(acc);
        if (alpha < 1e-8) alpha = 1e-8;
        scal[0] = alpha;
        scal[1] = 1.0 / alpha;
        B[(int64_t)j + (int64_t)j * m] = alpha;
    }
}

hipError_t launch_alpha(const double *sq, int64_t nb, double *scal, double *B,
                        int m, int j, hipStream_t s) {
    alpha_kernel<<<1, 256, 0, s>>>(sq, nb, scal, B, m, j);
    return hipGetLastError();
}

// ------------------------------------------------------------ X^T u
//
// Block b (kRowBlock rows) is processed by one workgroup of 4 waves; wave w
// owns sub-block rows [b*1024 + 256w, +256) and a private LDS accumulator of
// g floats.  Each row's column indices are unique (validated), so the 64
// lanes of a wave update distinct accumulator entries and the per-gene sum of
// a sub-block is the FP32 fma chain over its rows in ascending order.  The
// four sub-block accumulators are combined as (w0 + w1) + (w2 + w3).
// Loads of the next row are issued before the current row's LDS updates.

template <typename IP>
__global__ __launch_bounds__(256) void spmv_t_kernel(
    const float *__restrict__ data, const int32_t *__restrict__ indices,
    const IP *__restrict__ indptr, double *__restrict__ Ucol,
    const double *__restrict__ scal, int64_t n, int g,
    float *__restrict__ partial) {
    extern __shared__ float lds[];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = wave_id();
    float *acc = lds + (size_t)w * g;
    float *uval = lds + (size_t)4 * g + (size_t)w * kSubRows;
    for (int i = lane; i < g; i += kWave) acc[i] = 0.0f;
    const int64_t rs = (int64_t)blockIdx.x * kRowBlock + (int64_t)w * kSubRows;
    const int64_t nrows = std::max<int64_t>(0, std::min<int64_t>(kSubRows, n - rs));
    const double inv_alpha = scal[1];
    for (int t = lane; t < nrows; t += kWave) {
        const double x = Ucol[rs + t] * inv_alpha;
        Ucol[rs + t] = x;
        uval[t] = (float)x;
    }
    __syncthreads();
    for (int64_t rr = 0; rr < nrows; rr++) {
        const int64_t r = rs + rr;
        const int64_t b = indptr[r], e = indptr[r + 1];
        const float ur = uval[rr];
        for (int64_t base = b; base < e; base += 4 * kWave) {
            float d[4];
            int ix[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int64_t p = base + lane + t * kWave;
                const bool ok = p < e;
                d[t] = ok ? data[p] : 0.0f;
                ix[t] = ok ? indices[p] : 0;
            }
#pragma unroll
            for (int t = 0; t < 4; t++) {
                if (base + lane + t * kWave < e) acc[ix[t]] = fmaf(d[t], ur, acc[ix[t]]);
            }
        }
    }
    __syncthreads();
    float *out = partial + (int64_t)blockIdx.x * g;
    for (int i = threadIdx.x; i < g; i += 256)
        out[i] = (lds[i] + lds[g + i]) + (lds[2 * g + i] + lds[3 * g + i]);
}

hipError_t launch_spmv_t(const Csr &X, double *Ucol, const double *scal,
                         float *partial, hipStream_t s) {
    const int64_t nb = ceil_div(X.n, kRowBlock);
    size_t lds = sizeof(float) * ((size_t)4 * X.g + 4 * kSubRows);
    if (X.indptr_width == 8)
        spmv_t_kernel<int64_t><<<(unsigned)nb, 256, lds, s>>>(
            X.data, X.indices, (const int64_t *)X.indptr, Ucol, scal, X.n, X.g,
            partial);
    else
        spmv_t_kernel<int32_t><<<(unsigned)nb, 256, lds, s>>>(
            X.data, X.indices, (const int32_t *)X.indptr, Ucol, scal, X.n, X.g,
            partial);
    return hipGetLastError();
}

// Column sums of x (kSquare = false) or x^2 (true) with the same block /
// sub-block structure; written to partial[b*2g + (kSquare ? g : 0) + gene].
template <typename IP, bool kSquare>
__global__ __launch_bounds__(256) void colstats_kernel(
    const float *__restrict__ data, const int32_t *__restrict__ indices,
    const IP *__restrict__ indptr, int64_t n, int g,
    float *__restrict__ partial) {
    extern __shared__ float lds[];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = wave_id();
    float *acc = lds + (size_t)w * g;
    for (int i = lane; i < g; i += kWave) acc[i] = 0.0f;
    __syncthreads();
    const int64_t rs = (int64_t)blockIdx.x * kRowBlock + (int64_t)w * kSubRows;
    const int64_t nrows = std::max<int64_t>(0, std::min<int64_t>(kSubRows, n - rs));
    for (int64_t rr = 0; rr < nrows; rr++) {
        const int64_t r = rs + rr;
        const int64_t b = indptr[r], e = indptr[r + 1];
        for (int64_t p = b + lane; p < e; p += kWave) {
            const int c = indices[p];
            const float x = data[p];
            acc[c] = kSquare ? fmaf(x, x, acc[c]) : acc[c] + x;
        }
    }
    __syncthreads();
    float *out = partial + (int64_t)blockIdx.x * 2 * g + (kSquare ? g : 0);
    for (int i = threadIdx.x; i < g; i += 256)
        out[i] = (lds[i] + lds[g + i]) + (lds[2 * g + i] + lds[3 * g + i]);
}

template <bool kSquare>
static hipError_t colstats_pass(const Csr &X, float *partial, hipStream_t s) {
    const int64_t nb = ceil_div(X.n, kRowBlock);
    size_t lds = sizeof(float) * (size_t)4 * X.g;
    if (X.indptr_width == 8)
        colstats_kernel<int64_t, kSquare><<<(unsigned)nb, 256, lds, s>>>(
            X.data, X.indices, (const int64_t *)X.indptr, X.n, X.g, partial);
    else
        colstats_kernel<int32_t, kSquare><<<(unsigned)nb, 256, lds, s>>>(
            X.data, X.indices, (const int32_t *)X.indptr, X.n, X.g, partial);
    return hipGetLastError();
}

hipError_t launch_colstats(const Csr &X, float *partial, hipStream_t s) {
    hipError_t e = colstats_pass<false>(X, partial, s);
    if (e != hipSuccess) return e;
    return colstats_pass<true>(X, partial, s);
}

size_t lds_bytes_needed(int g, int m) {
    size_t need = sizeof(float) * ((size_t)4 * g + 4 * kSubRows);    // spmv_t
    need = std::max(need, sizeof(double) * ((size_t)g + m));          // vside
    need = std::max(need, sizeof(double) * (size_t)m * 32);           // gemm
    return need;
}

// grp[G*width + i] = sum over blocks G*64 .. G*64+63 (ascending) of
// partial[b*width + i], accumulated in FP64.
__global__ __launch_bounds__(256) void reduce_groups_kernel(
    const float *__restrict__ partial, int64_t nblocks, int width,
    double *__restrict__ grp) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= width) return;
    const int64_t G = blockIdx.y;
    const int64_t b0 = G * kGroup, b1 = std::min<int64_t>(b0 + kGroup, nblocks);
    double acc = 0.0;
    for (int64_t b = b0; b < b1; b++) acc += (double)partial[b * width + i];
    grp[G * width + i] = acc;
}

hipError_t launch_reduce_groups(const float *partial, int64_t nblocks,
                                int width, double *grp, hipStream_t s) {
    dim3 grid((unsigned)ceil_div(width, 256), (unsigned)ceil_div(nblocks, kGroup));
    reduce_groups_kernel<<<grid, 256, 0, s>>>(partial, nblocks, width, grp);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void finish_stddev_kernel(
    const double *__restrict__ grp, int64_t ngroups, int g, int64_t n,
    double *__restrict__ sd) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= g) return;
    double s = 0.0, ss = 0.0;
    for (int64_t G = 0; G < ngroups; G++) {
        s += grp[G * 2 * g + i];
        ss += grp[G * 2 * g + g + i];
    }
    double var = (ss - s * s / (double)n) / (double)(n - 1);
    double v = sqrt(var > 0.0 ? var : 0.0);
    sd[i] = v < 1e-8 ? 1e-8 : v;
}

hipError_t launch_finish_stddev(const double *grp, int64_t ngroups, int g,
                                int64_t n, double *sd, hipStream_t s) {
    finish_stddev_kernel<<<(unsigned)ceil_div(g, 256), 256, 0, s>>>(grp, ngroups,
                                                                    g, n, sd);
    return hipGetLastError();
}

// ------------------------------------------------------------ V side
//
// One workgroup of 16 waves.  w[gene] sums the group partials in group order
// (FP64); dots with V columns: lane-strided over genes then xor butterfly,
// wave c%16 owning column c; corrections accumulate over columns ascending.

__global__ __launch_bounds__(1024) void vside_kernel(
    const double *__restrict__ grp, int64_t ngroups, int g,
    const double *__restrict__ sd, double *__restrict__ V, int j, int m,
    double *__restrict__ B, const double *__restrict__ scal,
    double *__restrict__ rem, double *__restrict__ rscal,
    float *__restrict__ vs) {
    extern __shared__ double sh[];
    double *w = sh;            // g
    double *work = sh + g;     // m
    __shared__ double tmp[16];
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = wave_id();
    const double alpha = scal[0];
    const double *Vj = V + (int64_t)j * g;
    for (int i = threadIdx.x; i < g; i += 1024) {
        double acc = 0.0;
        for (int64_t G = 0; G < ngroups; G++) acc += grp[G * g + i];
        w[i] = acc / sd[i] - alpha * Vj[i];
    }
    __syncthreads();
    const bool last = (j == m - 1);
    const int ncols = last ? m : j + 1;
    for (int c = wv; c < ncols; c += 16) {
        const double *Vc = V + (int64_t)c * g;
        double acc = 0.0;
        for (int i = lane; i < g; i += kWave) acc = fma(Vc[i], w[i], acc);
        acc = wave_sum(acc);
        if (lane == 0) work[c] = acc;
    }
    __syncthreads();
    double nrm = 0.0;
    for (int i = threadIdx.x; i < g; i += 1024) {
        double d = 0.0;
        for (int c = 0; c < ncols; c++) d = fma(V[(int64_t)c * g + i], work[c], d);
        const double x = w[i] - d;
        w[i] = x;
        nrm = fma(x, x, nrm);
    }
    nrm = block_sum<16>(nrm, tmp);
    const double beta = sqrt(nrm);
    if (!last) {
        const double bclip = beta < 1e-8 ? 1e-8 : beta;
        const double inv = 1.0 / bclip;
        double *Vn = V + (int64_t)(j + 1) * g;
        for (int i = threadIdx.x; i < g; i += 1024) {
            const double x = w[i] * inv;
            Vn[i] = x;
            vs[i] = (float)(x / sd[i]);
        }
        if (threadIdx.x == 0) B[(int64_t)j + (int64_t)(j + 1) * m] = bclip;
    } else {
        const double inv = beta > 0.0 ? 1.0 / beta : 0.0;
        for (int i = threadIdx.x; i < g; i += 1024) rem[i] = w[i] * inv;
        if (threadIdx.x == 0) rscal[0] = beta;
    }
}

hipError_t launch_vside(const double *grp, int64_t ngroups, int g,
                        const double *sd, double *V, int j, int m, double *B,
                        const double *scal, double *rem, double *rscal,
                        float *vs, hipStream_t s) {
    size_t lds = sizeof(double) * ((size_t)g + m);
    vside_kernel<<<1, 1024, lds, s>>>(grp, ngroups, g, sd, V, j, m, B, scal,
                                     rem, rscal, vs);
    return hipGetLastError();
}

__global__ void init_vs_kernel(const double *__restrict__ V,
                               const double *__restrict__ sd, int g,
                               float *__restrict__ vs) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < g) vs[i] = (float)(V[i] / sd[i]);
}

hipError_t launch_init_vs(const double *V, const double *sd, int g, float *vs,
                          hipStream_t s) {
    init_vs_kernel<<<(unsigned)ceil_div(g, 256), 256, 0, s>>>(V, sd, g, vs);
    return hipGetLastError();
}

// ------------------------------------------------------------ restart GEMMs
//
// out[r, c] = sum_l A[r, l] coef[l, c], thread per row, 32 output columns per
// pass with the coefficient slab in LDS; the sum runs over l ascending.

template <bool kRowMajorF32>
__global__ __launch_bounds__(256) void tall_gemm_kernel(
    const double *__restrict__ A, int64_t lda, int64_t n, int kin,
    const double *__restrict__ coef, int kout, void *__restrict__ out,
    int64_t ldo) {
    constexpr int CW = 32;
    extern __shared__ double cs[];  // kin * CW
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    for (int c0 = 0; c0 < kout; c0 += CW) {
        const int nc = std::min(CW, kout - c0);
        __syncthreads();
        for (int i = threadIdx.x; i < kin * CW; i += 256) {
            const int l = i / CW, cc = i % CW;
            cs[i] = cc < nc ? coef[l + (int64_t)(c0 + cc) * kin] : 0.0;
        }
        __syncthreads();
        if (r < n) {
            double acc[CW];
#pragma unroll
            for (int cc = 0; cc < CW; cc++) acc[cc] = 0.0;
            for (int l = 0; l < kin; l++) {
                const double x = A[(int64_t)l * lda + r];
#pragma unroll
                for (int cc = 0; cc < CW; cc++) acc[cc] = fma(x, cs[l * CW + cc], acc[cc]);
            }
            if (kRowMajorF32) {
                float *o = (float *)out + r * kout + c0;
#pragma unroll
                for (int cc = 0; cc < CW; cc++)
                    if (cc < nc) o[cc] = (float)acc[cc];
            } else {
                double *o = (double *)out;
#pragma unroll
                for (int cc = 0; cc < CW; cc++)
                    if (cc < nc) o[(int64_t)(c0 + cc) * ldo + r] = acc[cc];
            }
        }
    }
}

hipError_t launch_tall_gemm_f64(const double *A, int64_t lda, int64_t n,
                                int kin, const double *coef, int kout,
                                double *out, int64_t ldo, hipStream_t s) {
    if (kout <= 0) return hipSuccess;
    tall_gemm_kernel<false><<<(unsigned)ceil_div(n, 256), 256,
                              sizeof(double) * kin * 32, s>>>(
        A, lda, n, kin, coef, kout, out, ldo);
    return hipGetLastError();
}

hipError_t launch_tall_gemm_rowmajor_f32(const double *A, int64_t lda,
                                         int64_t n, int kin,
                                         const double *coef, int kout,
                                         float *out, hipStream_t s) {
    tall_gemm_kernel<true><<<(unsigned)ceil_div(n, 256), 256,
                             sizeof(double) * kin * 32, s>>>(
        A, lda, n, kin, coef, kout, out, 0);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void v_restart_kernel(
    const double *__restrict__ V, int g, int m, const double *__restrict__ Vbt,
    int kc, const double *__restrict__ rem, double *__restrict__ V2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= g) return;
    if (c == kc) {
        V2[(int64_t)kc * g + i] = rem[i];
        return;
    }
    double acc = 0.0;
    for (int l = 0; l < m; l++)
        acc = fma(V[(int64_t)l * g + i], Vbt[c + (int64_t)l * m], acc);
    V2[(int64_t)c * g + i] = acc;
}

hipError_t launch_v_restart(const double *V, int g, int m, const double *Vbt,
                            int kc, const double *rem, double *V2,
                            hipStream_t s) {
    dim3 grid((unsigned)ceil_div(g, 256), (unsigned)(kc + 1));
    v_restart_kernel<<<grid, 256, 0, s>>>(V, g, m, Vbt, kc, rem, V2);
    return hipGetLastError();
}

}  // namespace bpca
