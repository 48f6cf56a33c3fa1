// Device-side operand preparation: the QC x HVG subset of a CSR matrix.
//
// Reference: SingleCell.pca() takes X[np.ix_(QC rows, HVG columns)]
// (brisc/single_cell.py:11853-11908, through the CSR subset routines of
// brisc/cysparse.pyx:926-1029) and hands the result to irlba as a canonical
// CSR.  Here the subset is formed on the device in two launches:
//   k_subset_count  kept entries of every output row (one wave per row)
//   k_subset_fill   the kept entries, column indices remapped, written in
//                   ascending new column order
// between which the caller turns the counts into row offsets.  Output row r
// is input row rows[r]; input column c becomes col_map[c] (or is dropped when
// col_map[c] < 0).  When col_map is increasing over the kept columns (one
// dataset: the HVG positions in order) the kept entries of a canonical row
// are already in order and are compacted with a ballot prefix; otherwise
// (several datasets aligned on gene names) each row's kept entries are sorted
// by new index in LDS.  Integer work only: the result is exact.
#include "kernels.h"

namespace bpca {

namespace {

__device__ inline int64_t row_of(const int64_t *rows, int64_t r) {
    return rows ? rows[r] : r;
}

__device__ inline int lane_id() { return threadIdx.x & (kWave - 1); }

}  // namespace

// Also validates what the fill will read: selected rows in [0, num_rows),
// row offsets non-decreasing within [0, nnz], column indices in
// [0, num_cols); a violation sets *bad and the row counts as empty.
template <typename IP, typename IX>
__global__ __launch_bounds__(256) void k_subset_count(
    const IP *__restrict__ indptr, const IX *__restrict__ indices,
    int64_t num_rows, int64_t nnz, int64_t num_cols,
    const int64_t *__restrict__ rows, int64_t num_out_rows,
    const int32_t *__restrict__ col_map, int64_t *__restrict__ row_nnz,
    int *bad) {
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= num_out_rows) return;
    const int lane = lane_id();
    const int64_t src = row_of(rows, r);
    int err = src < 0 || src >= num_rows;
    int64_t b = 0, e = 0;
    if (!err) {
        b = indptr[src];
        e = indptr[src + 1];
        err = b < 0 || e < b || e > nnz;
    }
    int64_t cnt = 0;
    if (!err) {
        for (int64_t p = b + lane; p < e; p += kWave) {
            const int64_t c = (int64_t)indices[p];
            if (c < 0 || c >= num_cols) {
                err = 1;
                continue;
            }
            cnt += col_map ? (col_map[c] >= 0) : 1;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
    const uint64_t any_err = __ballot(err != 0);
    if (lane == 0) {
        row_nnz[r] = any_err ? 0 : cnt;
        if (any_err) atomicOr(bad, 1);
    }
}

// Increasing map (or none): order-preserving compaction, 64 entries at a
// time, positions from a ballot prefix.
template <typename IP, typename IX>
__global__ __launch_bounds__(256) void k_subset_fill_ordered(
    const float *__restrict__ data, const IP *__restrict__ indptr,
    const IX *__restrict__ indices, const int64_t *__restrict__ rows,
    int64_t num_out_rows, const int32_t *__restrict__ col_map,
    const int64_t *__restrict__ out_indptr, float *__restrict__ out_data,
    int32_t *__restrict__ out_indices) {
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= num_out_rows) return;
    const int lane = lane_id();
    const int64_t src = row_of(rows, r);
    const int64_t b = indptr[src], e = indptr[src + 1];
    int64_t o = out_indptr[r];
    for (int64_t base = b; base < e; base += kWave) {
        const int64_t p = base + lane;
        int32_t nc = -1;
        float x = 0.0f;
        if (p < e) {
            const int64_t c = (int64_t)indices[p];
            nc = col_map ? col_map[c] : (int32_t)c;
            x = data[p];
        }
        const uint64_t keep = __ballot(nc >= 0);
        const int before = __popcll(keep & ((1ull << lane) - 1ull));
        if (nc >= 0) {
            out_data[o + before] = x;
            out_indices[o + before] = nc;
        }
        o += __popcll(keep);
    }
}

// General map: the row's kept (new index, value) pairs are gathered into
// LDS and sorted by new index (bitonic network over the next power of two,
// one workgroup per row); rows keep at most kSubsetSortMax entries.
constexpr int kSubsetSortMax = 4096;

template <typename IP, typename IX>
__global__ __launch_bounds__(256) void k_subset_fill_sorted(
    const float *__restrict__ data, const IP *__restrict__ indptr,
    const IX *__restrict__ indices, const int64_t *__restrict__ rows,
    int64_t num_out_rows, const int32_t *__restrict__ col_map,
    const int64_t *__restrict__ out_indptr, float *__restrict__ out_data,
    int32_t *__restrict__ out_indices) {
    __shared__ int32_t key[kSubsetSortMax];
    __shared__ float val[kSubsetSortMax];
    __shared__ int32_t count;
    const int64_t r = blockIdx.x;
    if (r >= num_out_rows) return;
    const int64_t src = row_of(rows, r);
    const int64_t b = indptr[src], e = indptr[src + 1];
    const int64_t o = out_indptr[r];
    const int kept = (int)(out_indptr[r + 1] - o);  // <= kSubsetSortMax (checked on the host)
    if (threadIdx.x == 0) count = 0;
    int pow2 = 1;
    while (pow2 < kept) pow2 <<= 1;
    for (int i = threadIdx.x; i < pow2; i += 256) {
        key[i] = INT32_MAX;
        val[i] = 0.0f;
    }
    __syncthreads();
    // Gather in 256-entry chunks; slots come from a block-wide ballot prefix
    // so the gather itself is deterministic (the sort then orders by key,
    // and keys are unique within a row).
    __shared__ int32_t wave_tot[4];
    for (int64_t base = b; base < e; base += 256) {
        const int64_t p = base + threadIdx.x;
        int32_t nc = -1;
        float x = 0.0f;
        if (p < e) {
            nc = col_map[(int64_t)indices[p]];
            x = data[p];
        }
        const uint64_t keep = __ballot(nc >= 0);
        const int w = threadIdx.x >> 6;
        if (lane_id() == 0) wave_tot[w] = __popcll(keep);
        __syncthreads();
        int before = __popcll(keep & ((1ull << lane_id()) - 1ull));
        for (int v = 0; v < w; v++) before += wave_tot[v];
        const int start = count;
        if (nc >= 0) {
            key[start + before] = nc;
            val[start + before] = x;
        }
        __syncthreads();
        if (threadIdx.x == 0) count = start + wave_tot[0] + wave_tot[1] + wave_tot[2] + wave_tot[3];
        __syncthreads();
    }
    for (int size = 2; size <= pow2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < pow2; i += 256) {
                const int j = i ^ stride;
                if (j > i) {
                    const bool up = (i & size) == 0;
                    const int32_t ki = key[i], kj = key[j];
                    if ((ki > kj) == up) {
                        key[i] = kj;
                        key[j] = ki;
                        const float t = val[i];
                        val[i] = val[j];
                        val[j] = t;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < kept; i += 256) {
        out_data[o + i] = val[i];
        out_indices[o + i] = key[i];
    }
}

namespace {

template <typename IP, typename IX>
hipError_t subset_count(const void *indptr, const void *indices,
                        int64_t num_rows, int64_t nnz, int64_t num_cols,
                        const int64_t *rows, int64_t num_out_rows,
                        const int32_t *col_map, int64_t *row_nnz, int *bad,
                        hipStream_t s) {
    const unsigned grid = (unsigned)ceil_div(num_out_rows, 4);
    k_subset_count<IP, IX><<<grid, 256, 0, s>>>((const IP *)indptr, (const IX *)indices,
                                                num_rows, nnz, num_cols, rows,
                                                num_out_rows, col_map, row_nnz, bad);
    return hipGetLastError();
}

template <typename IP, typename IX>
hipError_t subset_fill(const float *data, const void *indptr,
                       const void *indices, const int64_t *rows,
                       int64_t num_out_rows, const int32_t *col_map,
                       bool ordered, const int64_t *out_indptr,
                       float *out_data, int32_t *out_indices, hipStream_t s) {
    if (ordered) {
        const unsigned grid = (unsigned)ceil_div(num_out_rows, 4);
        k_subset_fill_ordered<IP, IX><<<grid, 256, 0, s>>>(
            data, (const IP *)indptr, (const IX *)indices, rows, num_out_rows,
            col_map, out_indptr, out_data, out_indices);
    } else {
        k_subset_fill_sorted<IP, IX><<<(unsigned)num_out_rows, 256, 0, s>>>(
            data, (const IP *)indptr, (const IX *)indices, rows, num_out_rows,
            col_map, out_indptr, out_data, out_indices);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_subset_count(const void *indptr, int indptr_width,
                               const void *indices, int index_width,
                               int64_t num_rows, int64_t nnz, int64_t num_cols,
                               const int64_t *rows, int64_t num_out_rows,
                               const int32_t *col_map, int64_t *row_nnz,
                               int *bad, hipStream_t s) {
    if (num_out_rows <= 0) return hipSuccess;
    if (indptr_width == 8 && index_width == 8)
        return subset_count<int64_t, int64_t>(indptr, indices, num_rows, nnz, num_cols, rows, num_out_rows, col_map, row_nnz, bad, s);
    if (indptr_width == 8)
        return subset_count<int64_t, int32_t>(indptr, indices, num_rows, nnz, num_cols, rows, num_out_rows, col_map, row_nnz, bad, s);
    if (index_width == 8)
        return subset_count<int32_t, int64_t>(indptr, indices, num_rows, nnz, num_cols, rows, num_out_rows, col_map, row_nnz, bad, s);
    return subset_count<int32_t, int32_t>(indptr, indices, num_rows, nnz, num_cols, rows, num_out_rows, col_map, row_nnz, bad, s);
}

hipError_t launch_subset_fill(const float *data, const void *indptr,
                              int indptr_width, const void *indices,
                              int index_width, const int64_t *rows,
                              int64_t num_out_rows, const int32_t *col_map,
                              int ordered, const int64_t *out_indptr,
                              float *out_data, int32_t *out_indices,
                              hipStream_t s) {
    if (num_out_rows <= 0) return hipSuccess;
    const bool o = ordered != 0;
    if (indptr_width == 8 && index_width == 8)
        return subset_fill<int64_t, int64_t>(data, indptr, indices, rows, num_out_rows, col_map, o, out_indptr, out_data, out_indices, s);
    if (indptr_width == 8)
        return subset_fill<int64_t, int32_t>(data, indptr, indices, rows, num_out_rows, col_map, o, out_indptr, out_data, out_indices, s);
    if (index_width == 8)
        return subset_fill<int32_t, int64_t>(data, indptr, indices, rows, num_out_rows, col_map, o, out_indptr, out_data, out_indices, s);
    return subset_fill<int32_t, int32_t>(data, indptr, indices, rows, num_out_rows, col_map, o, out_indptr, out_data, out_indices, s);
}

int subset_sort_max() { return kSubsetSortMax; }

}  // namespace bpca
