// SingleCell.normalize() on the device: brisc/normalize.pyx:9-268
// (normalize_csr) and 271-460 (normalize_csc), methods logCP10k (0),
// log1pPF (1) and PFlog1pPF (2), with an optional QC column and counts of
// any of the reference's `numeric` types (cyutils.pxd:16-22).
//
// Per cell i: row_sum = sum of (uint64) count over the row (integer, exact);
// normalisation factor nf = 10000 (logCP10k) or (float) sum of the row sums
// of the QC cells / (float) their number; per entry
// x' = log1p(count * (nf / (float) row_sum)) in float (double for float64
// counts, rounded to float), with the host library's log1pf (log1pf.h).
// PFlog1pPF then forms each cell's FP32 sum of its x' in column order, their
// FP32 sum over the QC cells in cell order, nnf = that sum / (float) the
// number of QC cells, and x'' = x' * (nnf / row sum).  Every step is the
// reference's single-precision operation in the reference's order, so the
// values are the reference's bit for bit.
//
// This file is compiled with -ffp-contract=off so that no multiply and add
// fuse.
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>

#include "kernels.h"
#include "log1pf.h"

namespace bpca {

namespace {

// count as the reference's row-sum cast reads it: (unsigned long long) x.
template <typename T>
__device__ __forceinline__ uint64_t count_u64(T x) { return (uint64_t)x; }

// count * isf in the reference's arithmetic: float for every count type but
// double, whose product and log1p are double (rounded to float on store).
template <typename T>
__device__ __forceinline__ float scaled_log1p(T x, float isf) {
    return log1pf_ref((float)x * isf);
}
template <>
__device__ __forceinline__ float scaled_log1p<double>(double x, float isf) {
    return (float)log1p(x * (double)isf);
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// CSR: one wave per row; row_sums[i]; QC cells' totals into acc[0] (sum)
// and acc[1] (count) by integer atomics (exact, order-free).
template <typename T, typename IP>
__global__ __launch_bounds__(256) void k_rowsum_csr(const T *__restrict__ data,
                                                    const IP *__restrict__ indptr,
                                                    int64_t n, const uint8_t *__restrict__ qc,
                                                    uint64_t *__restrict__ row_sums,
                                                    unsigned long long *acc) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t i = (int64_t)blockIdx.x * 4 + threadIdx.x / kWave;
    if (i >= n) return;
    uint64_t s = 0;
    for (int64_t p = (int64_t)indptr[i] + lane; p < (int64_t)indptr[i + 1]; p += kWave) s += count_u64(data[p]);
    s = wave_sum_u64(s);
    if (lane == 0) {
        row_sums[i] = s;
        if (!qc || qc[i]) {
            atomicAdd(&acc[0], (unsigned long long)s);
            atomicAdd(&acc[1], 1ull);
        }
    }
}

// CSC: row sums by integer atomics over the entries (exact), then the QC
// totals over the rows.
template <typename T, typename IC>
__global__ __launch_bounds__(256) void k_rowsum_csc(const T *__restrict__ data,
                                                    const IC *__restrict__ indices,
                                                    int64_t nnz, int64_t n,
                                                    unsigned long long *row_sums, int *flag) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < nnz; j += stride) {
        const int64_t r = (int64_t)indices[j];
        if (r < 0 || r >= n) {
            atomicOr(flag, 1);
            continue;
        }
        atomicAdd(&row_sums[r], (unsigned long long)count_u64(data[j]));
    }
}

__global__ __launch_bounds__(256) void k_qc_totals(const uint64_t *__restrict__ row_sums,
                                                   int64_t n, const uint8_t *__restrict__ qc,
                                                   unsigned long long *acc) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n || (qc && !qc[i])) return;
    atomicAdd(&acc[0], (unsigned long long)row_sums[i]);
    atomicAdd(&acc[1], 1ull);
}

// nf = 10000 (logCP10k) or (float) QC total / (float) QC count.
__global__ void k_norm_factor(const unsigned long long *acc, int method, float *nf) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    *nf = method == 0 ? 10000.0f : (float)acc[0] / (float)acc[1];
}

template <typename T, typename IP>
__global__ __launch_bounds__(256) void k_log1p_csr(const T *__restrict__ data,
                                                   const IP *__restrict__ indptr, int64_t n,
                                                   const uint64_t *__restrict__ row_sums,
                                                   const float *__restrict__ nf,
                                                   float *__restrict__ out) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t i = (int64_t)blockIdx.x * 4 + threadIdx.x / kWave;
    if (i >= n) return;
    const float isf = *nf / (float)row_sums[i];
    for (int64_t p = (int64_t)indptr[i] + lane; p < (int64_t)indptr[i + 1]; p += kWave)
        out[p] = scaled_log1p(data[p], isf);
}

template <typename T, typename IC>
__global__ __launch_bounds__(256) void k_log1p_csc(const T *__restrict__ data,
                                                   const IC *__restrict__ indices, int64_t nnz,
                                                   const uint64_t *__restrict__ row_sums,
                                                   const float *__restrict__ nf,
                                                   float *__restrict__ out) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < nnz; j += stride) {
        const float isf = *nf / (float)row_sums[indices[j]];
        out[j] = scaled_log1p(data[j], isf);
    }
}

// Per row, the FP32 sum of its entries in position order (one thread per
// row, sequential, as the reference's loop).
template <typename IP>
__global__ __launch_bounds__(256) void k_row_seq_sum(const float *__restrict__ x,
                                                     const IP *__restrict__ indptr, int64_t n,
                                                     float *__restrict__ sums) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int64_t p = (int64_t)indptr[i]; p < (int64_t)indptr[i + 1]; p++) s = s + x[p];
    sums[i] = s;
}

// nnf = (FP32 sum of the QC cells' row sums in cell order) / (float) their
// number: the workgroup stages chunks of the sums in LDS, one thread adds
// them in order.
__global__ __launch_bounds__(256) void k_pf_factor(const float *__restrict__ sums, int64_t n,
                                                   const uint8_t *__restrict__ qc,
                                                   const unsigned long long *acc,
                                                   float *__restrict__ nnf) {
    constexpr int kChunk = 4096;
    __shared__ float chunk[kChunk];
    float total = 0.0f;
    for (int64_t c0 = 0; c0 < n; c0 += kChunk) {
        const int len = (int)(n - c0 < kChunk ? n - c0 : kChunk);
        for (int t = threadIdx.x; t < len; t += 256) {
            const int64_t i = c0 + t;
            // a cell failing QC contributes nothing (an exact +0.0f is not
            // the same as skipping it only for a -0.0f total, which a sum of
            // non-negative values never is)
            chunk[t] = (!qc || qc[i]) ? sums[i] : 0.0f;
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int t = 0; t < len; t++) total = total + chunk[t];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        volatile float t = total;
        *nnf = t / (float)acc[1];
    }
}

template <typename IP>
__global__ __launch_bounds__(256) void k_pf_scale_csr(const IP *__restrict__ indptr, int64_t n,
                                                      const float *__restrict__ sums,
                                                      const float *__restrict__ nnf,
                                                      float *__restrict__ out) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t i = (int64_t)blockIdx.x * 4 + threadIdx.x / kWave;
    if (i >= n) return;
    const float f = *nnf / sums[i];
    for (int64_t p = (int64_t)indptr[i] + lane; p < (int64_t)indptr[i + 1]; p += kWave) out[p] = out[p] * f;
}

template <typename IC>
__global__ __launch_bounds__(256) void k_pf_scale_csc(const IC *__restrict__ indices, int64_t nnz,
                                                      const float *__restrict__ sums,
                                                      const float *__restrict__ nnf,
                                                      float *__restrict__ out) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < nnz; j += stride)
        out[j] = out[j] * (*nnf / sums[indices[j]]);
}

template <typename T>
struct Dev {
    T *p = nullptr;
    ~Dev() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t count) { return hipMalloc((void **)&p, sizeof(T) * std::max<size_t>(count, 1)); }
};

#define NZ_TRY(expr)            \
    do {                        \
        hipError_t _e = (expr); \
        if (_e != hipSuccess) { \
            *herr = _e;         \
            return 2;           \
        }                       \
    } while (0)

unsigned grid_rows(int64_t n) { return (unsigned)std::max<int64_t>(1, ceil_div(n, 4)); }
unsigned grid_elems(int64_t m) { return (unsigned)std::min<int64_t>(std::max<int64_t>(1, ceil_div(m, 256)), 65536); }

template <typename T, typename IX, typename IP>
int normalize_typed(const NormalizeArgs &A, hipError_t *herr, hipStream_t s) {
    const T *data = (const T *)A.data;
    const IX *indices = (const IX *)A.indices;
    const IP *indptr = (const IP *)A.indptr;
    const int64_t n = A.num_cells;
    Dev<unsigned long long> acc;
    Dev<float> nf;
    Dev<int> flag;
    NZ_TRY(acc.alloc(2));
    NZ_TRY(nf.alloc(2));
    NZ_TRY(flag.alloc(1));
    NZ_TRY(launch_zero(acc.p, 2 * sizeof(unsigned long long), s));
    NZ_TRY(launch_zero(flag.p, sizeof(int), s));
    if (A.is_csr) {
        if (n > 0) k_rowsum_csr<T, IP><<<grid_rows(n), 256, 0, s>>>(data, indptr, n, A.qc, A.row_sums, acc.p);
    } else {
        NZ_TRY(launch_zero(A.row_sums, sizeof(uint64_t) * n, s));
        if (A.nnz > 0)
            k_rowsum_csc<T, IX><<<grid_elems(A.nnz), 256, 0, s>>>(data, indices, A.nnz, n,
                                                                  (unsigned long long *)A.row_sums, flag.p);
        if (n > 0) k_qc_totals<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(A.row_sums, n, A.qc, acc.p);
    }
    NZ_TRY(hipGetLastError());
    int h_flag = 0;
    NZ_TRY(hipMemcpyAsync(&h_flag, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    NZ_TRY(hipStreamSynchronize(s));
    if (h_flag) return 1;
    k_norm_factor<<<1, 64, 0, s>>>(acc.p, A.method, nf.p);
    if (A.is_csr) {
        if (n > 0) k_log1p_csr<T, IP><<<grid_rows(n), 256, 0, s>>>(data, indptr, n, A.row_sums, nf.p, A.out);
    } else if (A.nnz > 0) {
        k_log1p_csc<T, IX><<<grid_elems(A.nnz), 256, 0, s>>>(data, indices, A.nnz, A.row_sums, nf.p, A.out);
    }
    NZ_TRY(hipGetLastError());
    if (A.method == 2) {
        Dev<float> sums;
        NZ_TRY(sums.alloc(n));
        if (A.is_csr) {
            k_row_seq_sum<IP><<<(unsigned)std::max<int64_t>(1, ceil_div(n, 256)), 256, 0, s>>>(A.out, indptr, n, sums.p);
        } else {
            // Each cell's entries in column order: the rows of the transpose.
            Dev<float> td;
            Dev<int32_t> ti;
            Dev<int64_t> tp;
            NZ_TRY(td.alloc(A.nnz));
            NZ_TRY(ti.alloc(A.nnz));
            NZ_TRY(tp.alloc(n + 1));
            const int rc = run_csr_transpose(A.out, A.indices, sizeof(IX), A.indptr, sizeof(IP),
                                             A.num_genes, n, A.nnz, td.p, ti.p, tp.p, herr, s);
            if (rc) return rc;
            k_row_seq_sum<int64_t><<<(unsigned)std::max<int64_t>(1, ceil_div(n, 256)), 256, 0, s>>>(td.p, tp.p, n, sums.p);
            NZ_TRY(hipGetLastError());
            NZ_TRY(hipStreamSynchronize(s));
        }
        k_pf_factor<<<1, 256, 0, s>>>(sums.p, n, A.qc, acc.p, nf.p + 1);
        if (A.is_csr) {
            if (n > 0) k_pf_scale_csr<IP><<<grid_rows(n), 256, 0, s>>>(indptr, n, sums.p, nf.p + 1, A.out);
        } else if (A.nnz > 0) {
            k_pf_scale_csc<IX><<<grid_elems(A.nnz), 256, 0, s>>>(indices, A.nnz, sums.p, nf.p + 1, A.out);
        }
        NZ_TRY(hipGetLastError());
        NZ_TRY(hipStreamSynchronize(s));
    }
    NZ_TRY(hipStreamSynchronize(s));
    return 0;
}

template <typename T>
int normalize_index(const NormalizeArgs &A, hipError_t *herr, hipStream_t s) {
    if (A.index_width == 8 && A.indptr_width == 8) return normalize_typed<T, int64_t, int64_t>(A, herr, s);
    if (A.index_width == 8) return normalize_typed<T, int64_t, int32_t>(A, herr, s);
    if (A.indptr_width == 8) return normalize_typed<T, int32_t, int64_t>(A, herr, s);
    return normalize_typed<T, int32_t, int32_t>(A, herr, s);
}

__global__ __launch_bounds__(256) void k_log1pf_eval(const float *__restrict__ x, int64_t n,
                                                     float *__restrict__ y) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) y[i] = log1pf_ref(x[i]);
}

}  // namespace

int run_normalize(const NormalizeArgs &A, hipError_t *herr, hipStream_t s) {
    switch (A.data_type) {
        case 0: return normalize_index<int32_t>(A, herr, s);
        case 1: return normalize_index<uint32_t>(A, herr, s);
        case 2: return normalize_index<int64_t>(A, herr, s);
        case 3: return normalize_index<uint64_t>(A, herr, s);
        case 4: return normalize_index<float>(A, herr, s);
        default: return normalize_index<double>(A, herr, s);
    }
}

hipError_t launch_log1pf_eval(const float *x, int64_t n, float *y, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_log1pf_eval<<<grid_elems(n), 256, 0, s>>>(x, n, y);
    return hipGetLastError();
}

}  // namespace bpca
