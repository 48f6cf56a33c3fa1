// k nearest neighbours of every cell among all cells, on device-resident PCs
// (the exhaustive form of SingleCell.neighbors()).
//
// Reference: brisc/knn.pyx:13-516 (knn_self) searches each cell's nearest
// clusters; with every cluster searched it returns the exact k nearest
// neighbours.  Distances are squared Euclidean in FP32, formed as the
// reference forms them (knn.pyx:86-94 with cyutils.pyx:1121-1160):
// ||q||^2 + (||c||^2 - 2 q.c); a cell is not its own neighbour; each row of
// the output is in ascending distance (ties in ascending index).
//
// One workgroup of 256 threads takes 64 queries and streams the cells in
// tiles of 64 through LDS (transposed, so a thread reads 4 consecutive
// queries or cells with one 16-byte load).  Thread (ty, tx) forms the 4x4
// block of dot products of queries 4ty..4ty+3 with cells 4tx..4tx+3 (FP32,
// dimensions in order).  Wave w keeps, per query (lane), a sorted list of
// the KMAX best among tile columns 16w..16w+15; the four lists are merged at
// the end.  Integer bookkeeping is exact; every tie is broken by index, so
// the result does not depend on the launch.
#include "kernels.h"

namespace bpca {

namespace {

constexpr int kQueries = 64;  // per workgroup
constexpr int kCells = 64;    // per streamed tile

// Sorted (ascending distance, then index) insertion into a KMAX-long list
// held in registers; a new element equal to an existing one goes after it.
template <int KMAX>
__device__ __forceinline__ void list_insert(float (&D)[KMAX], uint32_t (&I)[KMAX],
                                            float d, uint32_t ix) {
    if (!(d < D[KMAX - 1])) return;
    D[KMAX - 1] = d;
    I[KMAX - 1] = ix;
#pragma unroll
    for (int j = KMAX - 1; j > 0; j--) {
        const bool sw = D[j] < D[j - 1];
        const float a = D[j - 1], b = D[j];
        const uint32_t ia = I[j - 1], ib = I[j];
        D[j - 1] = sw ? b : a;
        D[j] = sw ? a : b;
        I[j - 1] = sw ? ib : ia;
        I[j] = sw ? ia : ib;
    }
}

__device__ __forceinline__ bool before(float da, uint32_t ia, float db, uint32_t ib) {
    return da < db || (da == db && ia < ib);
}

}  // namespace

// Squared norms ||x||^2 of the rows of X (n x d), dimensions in order.
__global__ __launch_bounds__(256) void k_row_norms(const float *__restrict__ X,
                                                   int64_t n, int d,
                                                   float *__restrict__ norms) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float *x = X + i * d;
    float s = x[0] * x[0];
    for (int j = 1; j < d; j++) s = fmaf(x[j], x[j], s);
    norms[i] = s;
}

template <int KMAX>
__global__ __launch_bounds__(256) void k_knn_exact(
    const float *__restrict__ X, const float *__restrict__ norms, int64_t n,
    int d, int k, uint32_t *__restrict__ out_idx, float *__restrict__ out_dist) {
    extern __shared__ float lds[];
    const int dp = (d + 3) & ~3;
    float *qT = lds;                          // dp x 64
    float *cT = qT + dp * kQueries;           // dp x 64
    float *dist = cT + dp * kCells;           // 64 x 65
    float *qn = dist + kQueries * (kCells + 1);
    float *cn = qn + kQueries;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wv = tid >> 6;
    const int64_t q0 = (int64_t)blockIdx.x * kQueries;
    for (int idx = tid; idx < dp * kQueries; idx += 256) {
        const int q = idx & (kQueries - 1), j = idx >> 6;
        qT[idx] = (j < d && q0 + q < n) ? X[(q0 + q) * d + j] : 0.0f;
    }
    if (tid < kQueries) qn[tid] = q0 + tid < n ? norms[q0 + tid] : 0.0f;
    float D[KMAX];
    uint32_t I[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
        D[j] = INFINITY;
        I[j] = 0xffffffffu;
    }
    const int ty = tid >> 4, tx = tid & 15;
    for (int64_t c0 = 0; c0 < n; c0 += kCells) {
        __syncthreads();  // previous tile's distances consumed
        for (int idx = tid; idx < dp * kCells; idx += 256) {
            const int c = idx & (kCells - 1), j = idx >> 6;
            cT[idx] = (j < d && c0 + c < n) ? X[(c0 + c) * d + j] : 0.0f;
        }
        if (tid < kCells) cn[tid] = c0 + tid < n ? norms[c0 + tid] : 0.0f;
        __syncthreads();
        float acc[4][4] = {};
        for (int j = 0; j < d; j++) {
            const float4 a = *(const float4 *)(qT + j * kQueries + 4 * ty);
            const float4 b = *(const float4 *)(cT + j * kCells + 4 * tx);
            const float av[4] = {a.x, a.y, a.z, a.w}, bv[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) acc[u][v] = fmaf(av[u], bv[v], acc[u][v]);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int q = 4 * ty + u, c = 4 * tx + v;
                const int64_t cell = c0 + c;
                float dd = qn[q] + (cn[c] - 2.0f * acc[u][v]);
                if (cell >= n || cell == q0 + q) dd = INFINITY;
                dist[q * (kCells + 1) + c] = dd;
            }
        }
        __syncthreads();
        // Wave wv scans tile columns 16wv..16wv+15 for query `lane`.
        const float *row = dist + lane * (kCells + 1) + 16 * wv;
#pragma unroll 4
        for (int c = 0; c < 16; c++)
            list_insert<KMAX>(D, I, row[c], (uint32_t)(c0 + 16 * wv + c));
    }
    __syncthreads();
    // Merge the four wave lists of each query (LDS reused).
    float *md = lds;
    uint32_t *mi = (uint32_t *)(lds + 4 * kQueries * KMAX);
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
        md[(wv * kQueries + lane) * KMAX + j] = D[j];
        mi[(wv * kQueries + lane) * KMAX + j] = I[j];
    }
    __syncthreads();
    if (wv == 0 && q0 + lane < n) {
        int h[4] = {0, 0, 0, 0};
        const int64_t q = q0 + lane;
        for (int t = 0; t < k; t++) {
            int best = 0;
            float bd = md[(0 * kQueries + lane) * KMAX + h[0]];
            uint32_t bi = mi[(0 * kQueries + lane) * KMAX + h[0]];
            for (int w = 1; w < 4; w++) {
                const float dw = h[w] < KMAX ? md[(w * kQueries + lane) * KMAX + h[w]] : INFINITY;
                const uint32_t iw = h[w] < KMAX ? mi[(w * kQueries + lane) * KMAX + h[w]] : 0xffffffffu;
                if (before(dw, iw, bd, bi)) {
                    best = w;
                    bd = dw;
                    bi = iw;
                }
            }
            h[best]++;
            out_dist[q * k + t] = bd;
            out_idx[q * k + t] = bi;
        }
    }
}

size_t knn_lds_bytes(int d, int kmax) {
    const size_t dp = (size_t)((d + 3) & ~3);
    const size_t tiles = sizeof(float) * (dp * kQueries + dp * kCells + kQueries * (kCells + 1) + kQueries + kCells);
    const size_t merge = (sizeof(float) + sizeof(uint32_t)) * 4 * kQueries * (size_t)kmax;
    return tiles > merge ? tiles : merge;
}

hipError_t launch_row_norms(const float *X, int64_t n, int d, float *norms,
                            hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_row_norms<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(X, n, d, norms);
    return hipGetLastError();
}

int knn_kmax(int k) { return k <= 16 ? 16 : k <= 32 ? 32 : k <= 64 ? 64 : 0; }

hipError_t launch_knn_exact(const float *X, const float *norms, int64_t n,
                            int d, int k, uint32_t *out_idx, float *out_dist,
                            hipStream_t s) {
    const int kmax = knn_kmax(k);
    if (kmax == 0 || n <= 0) return hipErrorInvalidValue;
    const unsigned grid = (unsigned)ceil_div(n, kQueries);
    const size_t lds = knn_lds_bytes(d, kmax);
    switch (kmax) {
        case 16: k_knn_exact<16><<<grid, 256, lds, s>>>(X, norms, n, d, k, out_idx, out_dist); break;
        case 32: k_knn_exact<32><<<grid, 256, lds, s>>>(X, norms, n, d, k, out_idx, out_dist); break;
        default: k_knn_exact<64><<<grid, 256, lds, s>>>(X, norms, n, d, k, out_idx, out_dist); break;
    }
    return hipGetLastError();
}

}  // namespace bpca
