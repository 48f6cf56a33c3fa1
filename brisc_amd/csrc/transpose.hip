// Deterministic CSR <-> CSC transpose on the device (the role of
// brisc/cysparse.pyx:1179-1295, csr_tocsc, and of scipy's tocsr/tocsc, whose
// outputs list each output row's entries in ascending input-row order).
//
// Every stored entry becomes a (key = its column, value = its row and value
// bits) pair; a stable LSD radix sort by key (rocPRIM) keeps equal keys in
// input order, which is ascending row order, so output row c lists the rows
// holding column c in ascending order -- the same arrays as the reference's
// stable scatter, with no atomics.  Output row offsets come from a binary
// search of the sorted keys.
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>

#include "kernels.h"

namespace bpca {

namespace {

// One wave per input row: keys[e] = column, vals[e] = (row << 32) | bits of
// data[e].  Flags offsets that decrease or leave [0, nnz] and columns outside
// [0, num_cols).
template <typename IP, typename IC>
__global__ __launch_bounds__(256) void k_transpose_expand(
    const float *__restrict__ data, const IC *__restrict__ indices,
    const IP *__restrict__ indptr, int64_t num_rows, int64_t nnz,
    int64_t num_cols, uint32_t *__restrict__ keys,
    uint64_t *__restrict__ vals, int *flag) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t row = (int64_t)blockIdx.x * 4 + threadIdx.x / kWave;
    if (row >= num_rows) return;
    const int64_t b = indptr[row], e = indptr[row + 1];
    int bad = (row == 0 && b != 0) || e < b || b < 0 || e > nnz;
    if (!bad) {
        for (int64_t p = b + lane; p < e; p += kWave) {
            const int64_t c = (int64_t)indices[p];
            if (c < 0 || c >= num_cols) {
                bad = 1;
                continue;
            }
            keys[p] = (uint32_t)c;
            vals[p] = ((uint64_t)row << 32) | (uint64_t)__float_as_uint(data[p]);
        }
    }
    if (bad) atomicOr(flag, 1);
}

// out_indptr[c] = first position of key c in the sorted keys (c = 0 ..
// num_cols), by binary search.
__global__ __launch_bounds__(256) void k_transpose_offsets(
    const uint32_t *__restrict__ keys, int64_t nnz, int64_t num_cols,
    int64_t *__restrict__ out_indptr) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c > num_cols) return;
    int64_t lo = 0, hi = nnz;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if ((int64_t)keys[mid] < c)
            lo = mid + 1;
        else
            hi = mid;
    }
    out_indptr[c] = lo;
}

__global__ __launch_bounds__(256) void k_transpose_unpack(
    const uint64_t *__restrict__ vals, int64_t nnz, float *__restrict__ out_data,
    int32_t *__restrict__ out_indices) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nnz; i += stride) {
        const uint64_t v = vals[i];
        out_indices[i] = (int32_t)(v >> 32);
        out_data[i] = __uint_as_float((uint32_t)v);
    }
}

template <typename T>
struct DevBuf {
    T *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t count) { return hipMalloc((void **)&p, sizeof(T) * std::max<size_t>(count, 1)); }
};

}  // namespace

int run_csr_transpose(const float *data, const void *indices, int index_width,
                      const void *indptr, int indptr_width, int64_t num_rows,
                      int64_t num_cols, int64_t nnz, float *out_data,
                      int32_t *out_indices, int64_t *out_indptr,
                      hipError_t *herr, hipStream_t s) {
#define TR_TRY(expr)                 \
    do {                             \
        hipError_t _e = (expr);      \
        if (_e != hipSuccess) {      \
            *herr = _e;              \
            return 2;                \
        }                            \
    } while (0)
    DevBuf<uint32_t> keys, keys_sorted;
    DevBuf<uint64_t> vals, vals_sorted;
    DevBuf<int> flag;
    TR_TRY(keys.alloc(nnz));
    TR_TRY(keys_sorted.alloc(nnz));
    TR_TRY(vals.alloc(nnz));
    TR_TRY(vals_sorted.alloc(nnz));
    TR_TRY(flag.alloc(1));
    TR_TRY(launch_zero(flag.p, sizeof(int), s));
    const unsigned grid = (unsigned)ceil_div(std::max<int64_t>(num_rows, 1), 4);
    if (num_rows > 0) {
        if (indptr_width == 8 && index_width == 8)
            k_transpose_expand<int64_t, int64_t><<<grid, 256, 0, s>>>(data, (const int64_t *)indices, (const int64_t *)indptr, num_rows, nnz, num_cols, keys.p, vals.p, flag.p);
        else if (indptr_width == 8)
            k_transpose_expand<int64_t, int32_t><<<grid, 256, 0, s>>>(data, (const int32_t *)indices, (const int64_t *)indptr, num_rows, nnz, num_cols, keys.p, vals.p, flag.p);
        else if (index_width == 8)
            k_transpose_expand<int32_t, int64_t><<<grid, 256, 0, s>>>(data, (const int64_t *)indices, (const int32_t *)indptr, num_rows, nnz, num_cols, keys.p, vals.p, flag.p);
        else
            k_transpose_expand<int32_t, int32_t><<<grid, 256, 0, s>>>(data, (const int32_t *)indices, (const int32_t *)indptr, num_rows, nnz, num_cols, keys.p, vals.p, flag.p);
        TR_TRY(hipGetLastError());
    }
    int h_flag = 0;
    TR_TRY(hipMemcpyAsync(&h_flag, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    TR_TRY(hipStreamSynchronize(s));
    if (h_flag) return 1;
    int bits = 1;
    while (bits < 32 && ((int64_t)1 << bits) < num_cols) bits++;
    if (nnz > 0) {
        size_t temp = 0;
        TR_TRY(rocprim::radix_sort_pairs(nullptr, temp, keys.p, keys_sorted.p, vals.p, vals_sorted.p,
                                         (size_t)nnz, 0, bits, s));
        DevBuf<char> tmp;
        TR_TRY(tmp.alloc(temp));
        TR_TRY(rocprim::radix_sort_pairs(tmp.p, temp, keys.p, keys_sorted.p, vals.p, vals_sorted.p,
                                         (size_t)nnz, 0, bits, s));
        const unsigned ug = (unsigned)std::min<int64_t>(ceil_div(nnz, 256), 65536);
        k_transpose_unpack<<<ug, 256, 0, s>>>(vals_sorted.p, nnz, out_data, out_indices);
        TR_TRY(hipGetLastError());
    }
    k_transpose_offsets<<<(unsigned)ceil_div(num_cols + 1, 256), 256, 0, s>>>(keys_sorted.p, nnz, num_cols, out_indptr);
    TR_TRY(hipGetLastError());
    TR_TRY(hipStreamSynchronize(s));
    return 0;
#undef TR_TRY
}

}  // namespace bpca
