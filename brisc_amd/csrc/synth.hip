// Device generator of the synthetic negative-binomial count workloads and
// the log1pPF normalisation of brisc/normalize.pyx:9-128.
//
// Per entry (cell i, gene j) the count is a pure function of
// (seed, i, j, m = type_gene_mean[type[i], j] * cell_factor[i]):
//   h = SplitMix64 finaliser(key + (i*g + j) * 0x9E3779B97F4A7C15),
//   u = (h >> 11) * 2^-53, and the count is the smallest k with
//   u < CDF of NB(mean m, size 5), walked from P(0) = (5/(5+m))^5 with
//   P(k+1) = P(k) * p * (k+5) / (k+1), p = m / (5+m).
// This file is compiled with -ffp-contract=off so each multiply and add is
// rounded separately, which makes the counts independent of where they are
// computed.
#include "kernels.h"
#include "log1pf.h"

namespace bpca {

namespace {

constexpr double kNbSize = 5.0;
constexpr int kNbMaxK = 1000000;

__device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__host__ __device__ inline uint64_t seed_key(uint64_t seed) {
    uint64_t z = seed + 0x632BE59BD9B4E019ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__device__ inline int nb_draw(uint64_t key, uint64_t cell, uint32_t gene,
                              uint32_t g, double m) {
    const uint64_t h = mix64(key + (cell * (uint64_t)g + gene) * 0x9E3779B97F4A7C15ULL);
    const double u = (double)(h >> 11) * 0x1.0p-53;
    const double q = kNbSize / (kNbSize + m);
    const double q2 = q * q;
    const double p0 = q2 * q2 * q;
    if (u < p0) return 0;
    const double p = m / (kNbSize + m);
    double pk = p0, cdf = p0;
    int k = 0;
    while (k < kNbMaxK) {
        pk = pk * p * ((double)k + kNbSize) / ((double)k + 1.0);
        k = k + 1;
        cdf = cdf + pk;
        if (u < cdf) break;
    }
    return k;
}

__device__ inline int wave_id() {
    return __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
}

}  // namespace

// One wave per row; lanes walk the genes in steps of 64.
__global__ __launch_bounds__(256) void synth_row_nnz_kernel(
    uint64_t key, uint64_t row0, int64_t rows, int g,
    const double *__restrict__ tgm, const int32_t *__restrict__ type,
    const double *__restrict__ factor, int64_t *__restrict__ row_nnz) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t r = (int64_t)blockIdx.x * 4 + wave_id();
    if (r >= rows) return;
    const double *base = tgm + (int64_t)type[r] * g;
    const double s = factor[r];
    const uint64_t cell = row0 + (uint64_t)r;
    int count = 0;
    for (int j = lane; j < g; j += kWave) count += nb_draw(key, cell, j, g, base[j] * s) != 0;
    for (int off = 32; off > 0; off >>= 1) count += __shfl_xor(count, off);
    if (lane == 0) row_nnz[r] = count;
}

// One wave per row; nonzeros of each 64-gene chunk are compacted with a
// ballot so column indices come out ascending.
__global__ __launch_bounds__(256) void synth_fill_kernel(
    uint64_t key, uint64_t row0, int64_t rows, int g,
    const double *__restrict__ tgm, const int32_t *__restrict__ type,
    const double *__restrict__ factor, const int64_t *__restrict__ indptr,
    int32_t *__restrict__ indices, int32_t *__restrict__ counts) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t r = (int64_t)blockIdx.x * 4 + wave_id();
    if (r >= rows) return;
    const double *base = tgm + (int64_t)type[r] * g;
    const double s = factor[r];
    const uint64_t cell = row0 + (uint64_t)r;
    int64_t pos = indptr[r];
    for (int j0 = 0; j0 < g; j0 += kWave) {
        const int j = j0 + lane;
        const int x = j < g ? nb_draw(key, cell, j, g, base[j] * s) : 0;
        const uint64_t mask = __ballot(x != 0);
        const int before = __popcll(mask & ((1ULL << lane) - 1ULL));
        if (x != 0) {
            indices[pos + before] = j;
            counts[pos + before] = x;
        }
        pos += __popcll(mask);
    }
}

hipError_t launch_synth_row_nnz(uint64_t seed, uint64_t row0, int64_t rows,
                                int g, const double *tgm, const int32_t *type,
                                const double *factor, int64_t *row_nnz,
                                hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    synth_row_nnz_kernel<<<(unsigned)ceil_div(rows, 4), 256, 0, s>>>(
        seed_key(seed), row0, rows, g, tgm, type, factor, row_nnz);
    return hipGetLastError();
}

hipError_t launch_synth_fill(uint64_t seed, uint64_t row0, int64_t rows, int g,
                             const double *tgm, const int32_t *type,
                             const double *factor, const int64_t *indptr,
                             int32_t *indices, int32_t *counts, hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    synth_fill_kernel<<<(unsigned)ceil_div(rows, 4), 256, 0, s>>>(
        seed_key(seed), row0, rows, g, tgm, type, factor, indptr, indices,
        counts);
    return hipGetLastError();
}

// ------------------------------------------------------------ log1pPF

__global__ __launch_bounds__(256) void row_sums_kernel(
    const int32_t *__restrict__ counts, const int64_t *__restrict__ indptr,
    int64_t n, uint64_t *__restrict__ row_sums) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t r = (int64_t)blockIdx.x * 4 + wave_id();
    if (r >= n) return;
    uint64_t acc = 0;
    for (int64_t p = indptr[r] + lane; p < indptr[r + 1]; p += kWave)
        acc += (uint64_t)counts[p];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) row_sums[r] = acc;
}

hipError_t launch_row_sums(const int32_t *counts, const int64_t *indptr,
                           int64_t n, uint64_t *row_sums, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    row_sums_kernel<<<(unsigned)ceil_div(n, 4), 256, 0, s>>>(counts, indptr, n,
                                                            row_sums);
    return hipGetLastError();
}

// Integer sum: exact, so the order does not matter.
__global__ __launch_bounds__(1024) void sum_u64_kernel(
    const uint64_t *__restrict__ x, int64_t n, uint64_t *__restrict__ out) {
    __shared__ uint64_t part[16];
    uint64_t acc = 0;
    for (int64_t i = threadIdx.x; i < n; i += 1024) acc += x[i];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int i = 0; i < 16; i++) t += part[i];
        *out = t;
    }
}

hipError_t launch_sum_u64(const uint64_t *x, int64_t n, uint64_t *out,
                          hipStream_t s) {
    sum_u64_kernel<<<1, 1024, 0, s>>>(x, n, out);
    return hipGetLastError();
}

// factor = (float)total / (float)n; isf = factor / (float)row_sum;
// x' = log1pf((float)count * isf)   (normalize.pyx:70, 195-209)
__global__ __launch_bounds__(256) void log1ppf_kernel(
    const int32_t *__restrict__ counts, const int64_t *__restrict__ indptr,
    int64_t n, const uint64_t *__restrict__ row_sums,
    const uint64_t *__restrict__ total, float *__restrict__ out) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t r = (int64_t)blockIdx.x * 4 + wave_id();
    if (r >= n) return;
    const float factor = (float)(*total) / (float)n;
    const float isf = factor / (float)row_sums[r];
    for (int64_t p = indptr[r] + lane; p < indptr[r + 1]; p += kWave)
        out[p] = log1pf_ref((float)counts[p] * isf);
}

hipError_t launch_log1ppf(const int32_t *counts, const int64_t *indptr,
                          int64_t n, const uint64_t *row_sums,
                          const uint64_t *total, float *out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    log1ppf_kernel<<<(unsigned)ceil_div(n, 4), 256, 0, s>>>(counts, indptr, n,
                                                           row_sums, total, out);
    return hipGetLastError();
}

}  // namespace bpca
