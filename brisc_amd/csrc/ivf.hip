// Inverted-file k nearest neighbours: SingleCell.neighbors() with its
// default search, on device-resident PCs.
//
// Reference: brisc/single_cell.py:12452-12650 drives brisc/kmeans.pyx
// (kmeans: random initialisation 13-39, Lloyd iterations 889-1100, empty
// cluster relocation 747-820) and brisc/knn.pyx:13-516 (knn_self).  The
// steps, each on the device unless noted:
//   1. k-means: centroids = num_clusters distinct random cells, drawn on the
//      host with the reference's PCG32 / Lemire generator (cyutils.pyx
//      1153-1232) so the same cells are drawn; then per iteration an E-step
//      (each cell's nearest centroid by the partial distance ||c||^2 - 2 x.c,
//      ties to the lower index: the tile engine with k = 1) and an M-step
//      whose sums follow the reference's order exactly (per chunk of
//      chunk_size cells, cells in order, then the chunk partials in order),
//      the relocation of empty clusters to the farthest cells (host, rare)
//      and the division by the count, 1 / count rounded from double; after
//      the last iteration a final E-step gives the labels and counts.
//   2. Empty clusters dropped; clusters renumbered by the first coordinate of
//      their centroid (host); cells sorted by cluster (stable radix sort);
//      cluster row ranges from the counts.
//   3. Each cell's num_clusters_searched nearest centroids (partial
//      distances; the tile engine with k = num_clusters_searched).
//   4. Per 64 consecutive sorted cells, the union of their clusters with a
//      64-bit mask of the cells that search each (LDS bitonic sort).
//   5. The search: the tile engine in inverted-file mode over the sorted
//      cells, a cell never its own neighbour, ids = original indices.
// The E-step's dot products use the FP32 matrix cores, not the reference's
// sgemm blocking, so a cell equidistant (to FP32 rounding) from two
// centroids may be assigned differently; everything else is exact integer
// work or follows the reference's FP32 operation order.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "kernels.h"
#include "rng.h"

namespace bpca {

namespace {

constexpr int kIvfQueries = 64;           // queries per search workgroup
constexpr int kIvfMaxKeys = 64 * 64;      // (cluster, query) pairs per workgroup

__global__ __launch_bounds__(256) void k_iota(uint32_t *__restrict__ a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) a[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void k_gather_rows(const float *__restrict__ X,
                                                     int d,
                                                     const uint32_t *__restrict__ rows,
                                                     int64_t num, float *__restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= num * d) return;
    const int64_t r = e / d;
    out[e] = X[(int64_t)rows[r] * d + (e - r * d)];
}

// Block c of out (64 padded rows of stride st): rows near[c * kp + r] of the
// num padded centroids P (d dimensions) for r < kp.  Other rows, and any
// id that is not a centroid, are at infinite distance ([0, 0, inf, 0 ...]).
__global__ __launch_bounds__(256) void k_gather_blocks(const float *__restrict__ P, int d,
                                                       int st,
                                                       const uint32_t *__restrict__ near,
                                                       int64_t num, int kp,
                                                       float *__restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= num * 64 * st) return;
    const int64_t row = e / st;
    const int j = (int)(e - row * st);
    const int64_t c = row / 64;
    const int r = (int)(row - c * 64);
    const uint32_t id = r < kp ? near[c * kp + r] : 0xffffffffu;
    out[e] = id < num ? P[(int64_t)id * st + j] : (j == d + 1 ? INFINITY : 0.0f);
}

__global__ __launch_bounds__(256) void k_count_labels(const uint32_t *__restrict__ labels,
                                                      int64_t n, uint32_t *__restrict__ counts) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) atomicAdd(&counts[labels[i]], 1u);
}

__global__ __launch_bounds__(256) void k_remap_labels(uint32_t *__restrict__ labels,
                                                      int64_t n,
                                                      const uint32_t *__restrict__ map) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) labels[i] = map[labels[i]];
}

// Cluster sums in the reference's order (kmeans.pyx:1040-1062): one wave per
// cluster, a lane per dimension; the cluster's cells in ascending index
// (order[off[j] .. off[j + 1]) from the stable sort), a partial sum per chunk
// of `chunk` cells started from zero, the partials added in chunk order.
__global__ __launch_bounds__(256) void k_cluster_sums(const float *__restrict__ X,
                                                      int d,
                                                      const uint32_t *__restrict__ order,
                                                      const int64_t *__restrict__ off,
                                                      int num_clusters, int64_t chunk,
                                                      float *__restrict__ sums) {
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= num_clusters) return;
    const int lane = threadIdx.x & 63;
    const int64_t b = off[j], e = off[j + 1];
    for (int d0 = 0; d0 < d; d0 += 64) {
        const int dim = d0 + lane;
        const bool on = dim < d;
        float total = 0.0f, part = 0.0f;
        int64_t end = -1;  // first cell past the current chunk (cells ascend)
        bool started = false;
        // Eight cells' values are loaded ahead; the adds stay in cell order.
        for (int64_t p = b; p < e; p += 8) {
            int64_t cell[8];
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                cell[u] = p + u < e ? (int64_t)order[p + u] : 0;
                v[u] = on && p + u < e ? X[cell[u] * d + dim] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                if (p + u >= e) break;
                if (cell[u] >= end) {
                    if (started) total += part;
                    part = 0.0f;
                    started = true;
                    end = (cell[u] / chunk + 1) * chunk;
                }
                part += v[u];
            }
        }
        if (started) total += part;
        if (on) sums[(int64_t)j * d + dim] = total;
    }
}

// centroid = sum * (float)(1.0 / count) for non-empty clusters
// (kmeans.pyx:1068-1073: a double quotient stored in a float).
__global__ __launch_bounds__(256) void k_cluster_normalize(float *__restrict__ sums,
                                                           int d, int num_clusters,
                                                           const uint32_t *__restrict__ counts) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)num_clusters * d) return;
    const uint32_t c = counts[e / d];
    if (c > 0) sums[e] *= (float)(1.0 / (double)c);
}

// Squared distance of each cell to its assigned centroid, dimensions in
// order (kmeans.pyx:779-787), for the relocation of empty clusters.
__global__ __launch_bounds__(256) void k_dist_to_centroid(const float *__restrict__ X,
                                                          int64_t n, int d,
                                                          const uint32_t *__restrict__ labels,
                                                          const float *__restrict__ cent,
                                                          float *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float *x = X + i * d;
    const float *c = cent + (int64_t)labels[i] * d;
    float diff = x[0] - c[0];
    float dist = diff * diff;
    for (int j = 1; j < d; j++) {
        diff = x[j] - c[j];
        dist = __builtin_fmaf(diff, diff, dist);
    }
    out[i] = dist;
}

// Ascending bitonic sort of key[0, pow2) in LDS by a workgroup of 256
// (pow2 a power of two, at most 4096).  At each step pair p compares
// positions i = p + (p & ~(stride - 1)) and i + stride; a thread loads all
// its pairs before it compares and stores any, so the step costs one LDS
// round trip.
__device__ __forceinline__ void lds_bitonic_sort(uint32_t *key, int pow2, int tid) {
    const int pairs = pow2 >> 1;
    for (int size = 2; size <= pow2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            uint32_t a[8], b[8];
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const int p = tid + 256 * r;
                if (p < pairs) {
                    const int i = p + (p & ~(stride - 1));
                    a[r] = key[i];
                    b[r] = key[i + stride];
                }
            }
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const int p = tid + 256 * r;
                if (p < pairs) {
                    const int i = p + (p & ~(stride - 1));
                    if ((a[r] > b[r]) == ((i & size) == 0)) {
                        key[i] = b[r];
                        key[i + stride] = a[r];
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Per workgroup of 64 consecutive sorted cells: the distinct clusters they
// search, with the mask of the cells searching each.  Keys (cluster << 6 |
// cell - 64 b) are bitonically sorted in LDS; a run of equal clusters is one
// entry.  WRITE = false counts the entries (counts[b]); WRITE = true writes
// them from wg_off[b] on (order below), each as four words: the cluster's
// rows [cl_off[c], cl_off[c + 1]) and the 64-bit mask of its cells.
template <bool WRITE>
__global__ __launch_bounds__(256) void k_ivf_entries(const uint32_t *__restrict__ near,
                                                     int64_t n, int searched,
                                                     int64_t *__restrict__ counts,
                                                     const int64_t *__restrict__ wg_off,
                                                     const int64_t *__restrict__ cl_off,
                                                     uint32_t *__restrict__ ent) {
    __shared__ uint32_t key[kIvfMaxKeys];
    __shared__ int part[256];
    const int tid = threadIdx.x;
    const int64_t q0 = (int64_t)blockIdx.x * kIvfQueries;
    const int nq = (int)(n - q0 < kIvfQueries ? n - q0 : kIvfQueries);
    const int total = kIvfQueries * searched;
    int pow2 = 1;
    while (pow2 < total) pow2 <<= 1;
    for (int i = tid; i < pow2; i += 256) {
        const int q = i / searched;
        key[i] = (i < total && q < nq)
                     ? (near[(q0 + q) * searched + (i - q * searched)] << 6) | (uint32_t)q
                     : 0xffffffffu;
    }
    __syncthreads();
    lds_bitonic_sort(key, pow2, tid);
    // Run starts; thread t owns positions [t * per, (t + 1) * per).
    const int per = (pow2 + 255) / 256;
    int mine = 0;
    for (int i = tid * per; i < (tid + 1) * per && i < pow2; i++) {
        const uint32_t k = key[i];
        if (k != 0xffffffffu && (i == 0 || (key[i - 1] >> 6) != (k >> 6))) mine++;
    }
    part[tid] = mine;
    __syncthreads();
    if (!WRITE) {
        if (tid == 0) {
            int64_t s = 0;
            for (int t = 0; t < 256; t++) s += part[t];
            counts[blockIdx.x] = s;
        }
        return;
    }
    // Entries in cluster order into LDS: the run starts before each
    // thread's range (an inclusive scan of part), then each key ORs its
    // cell's bit into its entry's mask.
    __shared__ uint32_t ecl[kIvfMaxKeys];
    __shared__ uint64_t emask[kIvfMaxKeys];
    for (int off = 1; off < 256; off <<= 1) {
        const int v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    const int before = part[tid] - mine;
    const int U = part[255];
    for (int x = tid; x < U; x += 256) emask[x] = 0;
    __syncthreads();
    int e = before - 1;
    for (int i = tid * per; i < (tid + 1) * per && i < pow2; i++) {
        const uint32_t k = key[i];
        if (k == 0xffffffffu) break;  // the invalid keys sort last
        if (i == 0 || (key[i - 1] >> 6) != (k >> 6)) ecl[++e] = k >> 6;
        atomicOr((unsigned long long *)&emask[e], 1ull << (k & 63u));
    }
    __syncthreads();
    // ... then ordered nearest first, as the block's first cell ranks the
    // clusters (its own cluster first; clusters it does not search after
    // those it does), then by the number of cells searching each, most
    // first, then in cluster order.  The neighbours of the block's cells are
    // found in the first tiles, so the thresholds are tight early.
    __syncthreads();  // part is read above; it now holds that ranking
    uint32_t *rep = (uint32_t *)part;
    if (tid < searched) rep[tid] = near[q0 * searched + tid];
    __syncthreads();
    int p2 = 1;
    while (p2 < U) p2 <<= 1;
    for (int i = tid; i < p2; i += 256) {
        uint32_t rank = 64;
        if (i < U)
            for (int j = 0; j < searched; j++)
                if (rep[j] == ecl[i]) {
                    rank = (uint32_t)j;
                    break;
                }
        key[i] = i < U ? (rank << 19) | ((uint32_t)(64 - __popcll(emask[i])) << 12) | (uint32_t)i
                       : 0xffffffffu;
    }
    __syncthreads();
    lds_bitonic_sort(key, p2, tid);
    const int64_t out = wg_off[blockIdx.x];
    for (int i = tid; i < U; i += 256) {
        const uint32_t src = key[i] & 0xfffu;
        const uint32_t cl = ecl[src];
        const uint64_t mk = emask[src];
        uint32_t *rec = ent + 4 * (out + i);
        rec[0] = (uint32_t)cl_off[cl];
        rec[1] = (uint32_t)cl_off[cl + 1];
        rec[2] = (uint32_t)mk;
        rec[3] = (uint32_t)(mk >> 32);
    }
}

// ---------------------------------------------------------------------------
// k-means|| initialisation (kmeans_barbar_init, brisc/kmeans.pyx:41-360; the
// parallel version 362-720 selects the same cells in the same order).
// Selection rounds: each cell's squared distance to its nearest candidate
// (min_distances), the cost (chunk sums in cell order, then the chunk sums
// in order), and a draw per cell from srand(seed + cell * round) taking it
// with probability l * distance / cost.  Candidate updates rank the round's
// new candidates with the tile engine (k = 1, partial distances, ties to the
// lower index).  Then a weighted k-means++ pass over the candidates: per
// draw one launch updates every candidate's weighted distance and its
// 256-candidate halving-tree sum, and a one-workgroup launch adds the tree
// sums in order and walks the reference's sequential probability chain.

// Distances to the first random cell, dimensions in order, the first square
// a product and each later one fused (kmeans.pyx:84-97).
__global__ __launch_bounds__(256) void k_kb_first(const float *__restrict__ X, int64_t n,
                                                  int d, int64_t rc,
                                                  float *__restrict__ md) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float *x = X + i * d, *r = X + rc * d;
    float diff = x[0] - r[0];
    float sum = diff * diff;
    for (int j = 1; j < d; j++) {
        diff = x[j] - r[j];
        sum = fmaf(diff, diff, sum);
    }
    md[i] = sum;
}

// Sum of each chunk of `chunk` cells, cells in order (kmeans.pyx:215-224).
__global__ __launch_bounds__(256) void k_kb_chunk_sums(const float *__restrict__ md,
                                                       int64_t n, int64_t chunk,
                                                       int64_t nchunks,
                                                       float *__restrict__ out) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= nchunks) return;
    const int64_t i0 = c * chunk, i1 = i0 + chunk < n ? i0 + chunk : n;
    float sum = 0.0f;
#pragma unroll 16
    for (int64_t i = i0; i < i1; i++) sum += md[i];
    out[c] = sum;
}

// One selection round's draws (kmeans.pyx:107-112, 227-240): cell i draws
// random_uniform from srand(seed + i * round), the product in 32 bits; a
// taken cell's distance becomes 0.
__global__ __launch_bounds__(256) void k_kb_sample(float *__restrict__ md, int64_t n,
                                                   uint64_t seed, uint32_t round,
                                                   float l_over_cost,
                                                   uint8_t *__restrict__ flags) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    Pcg32 g(seed + (uint64_t)((uint32_t)i * round));
    const float u = g.uniform();
    const bool take = l_over_cost * md[i] >= u;
    flags[i] = take ? 1 : 0;
    if (take) md[i] = 0.0f;
}

// Each new candidate is its own cell's nearest (kmeans.pyx:127-130,
// 247-252); a cell taken twice keeps its later position, as the serial
// loop's last write does.
__global__ __launch_bounds__(256) void k_kb_set_best(const uint32_t *__restrict__ sel,
                                                     int64_t p0, int64_t p1,
                                                     uint32_t *__restrict__ best) {
    const int64_t p = p0 + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p < p1) atomicMax(&best[sel[p]], (uint32_t)p);
}

// Candidate update (kmeans.pyx:196-214): the round's best partial distance
// ||c||^2 - 2 x.c against min_distance - ||x||^2, strictly smaller wins;
// min_distance is stored back with ||x||^2 added.
__global__ __launch_bounds__(256) void k_kb_update(float *__restrict__ md,
                                                   const float *__restrict__ norms,
                                                   uint32_t *__restrict__ best,
                                                   const float *__restrict__ pd,
                                                   const uint32_t *__restrict__ pj,
                                                   uint32_t prev, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float xn = norms[i];
    float m = md[i] - xn;
    const float v = pd[i];
    if (v < m) {
        m = v;
        best[i] = prev + pj[i];
    }
    md[i] = m + xn;
}

// Candidate weights: cells nearest each candidate (kmeans.pyx:276-280).
__global__ __launch_bounds__(256) void k_kb_weights(const uint32_t *__restrict__ best,
                                                    int64_t n, uint32_t *__restrict__ w) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) atomicAdd(&w[best[i]], 1u);
}

// Candidate coordinates by dimension: candT[j S + i] = X[sel[i], j].
__global__ __launch_bounds__(256) void k_kb_gather_t(const float *__restrict__ X, int d,
                                                     const uint32_t *__restrict__ sel,
                                                     int64_t S, float *__restrict__ candT) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= S * d) return;
    const int64_t j = e / S, i = e - j * S;
    candT[e] = X[(int64_t)sel[i] * d + j];
}

// k-means++ update (kmeans.pyx:286-300, 331-348): candidate i's weighted
// squared distance to the centroid at candidate position *pos, kept if
// smaller (first: taken as is), and this workgroup's halving-tree sum of
// the 256 results.
__global__ __launch_bounds__(256) void k_kpp_update(const float *__restrict__ candT,
                                                    int64_t S, int d,
                                                    const uint32_t *__restrict__ w,
                                                    float *__restrict__ m2,
                                                    const uint32_t *__restrict__ pos,
                                                    int first, float *__restrict__ partials) {
    __shared__ float red[256];
    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * 256 + tid;
    const int64_t c = *pos;
    float m = 0.0f;
    if (i < S) {
        float diff = candT[i] - candT[c];
        float dist = diff * diff;
#pragma unroll 8
        for (int j = 1; j < d; j++) {
            diff = candT[(int64_t)j * S + i] - candT[(int64_t)j * S + c];
            dist = fmaf(diff, diff, dist);
        }
        dist = dist * (float)w[i];
        if (first) {
            m = dist;
        } else {
            m = m2[i];
            if (dist < m) m = dist;
        }
        m2[i] = m;
    }
    red[tid] = m;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) red[tid] += red[tid + h];
        __syncthreads();
    }
    if (tid == 0) partials[blockIdx.x] = red[0];
}

constexpr int kKppSeg = 4096;  // candidates staged in LDS per segment

// One k-means++ draw (kmeans.pyx:307-321): cost = the tree sums in order,
// probability = u, then probability -= m2[i] / cost (times 1 / cost, fused)
// for i in order until it goes negative; the last candidate if it never
// does.  The chosen candidate's m2 becomes 0.  The first wave walks the
// chain over a segment of candidates in LDS, 64 at a time (the chain never
// increases, so a block whose end stays non-negative holds no crossing),
// while the other three stage the next segment.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_kpp_draw(float *__restrict__ m2, int64_t S,
                                                  const float *__restrict__ partials,
                                                  int G, float u, uint32_t *__restrict__ pos,
                                                  const uint32_t *__restrict__ sel,
                                                  uint32_t *__restrict__ cell_out) {
    __shared__ __align__(16) float seg[2][kKppSeg];
    __shared__ float sh_p, sh_inv;
    __shared__ int64_t sh_pick[2];  // by segment parity: read after a barrier, rewritten two later
    const int tid = threadIdx.x;
    const int wv = tid >> 6, lane = tid & (kWave - 1);
    // The tree sums, 1024 at a time through LDS (loaded by every thread),
    // added in order by the first wave.
    float cost = 0.0f;
    for (int b0 = 0; b0 < G; b0 += 1024) {
        for (int k = tid; k < 1024; k += 256) seg[0][k] = b0 + k < G ? partials[b0 + k] : 0.0f;
        __syncthreads();
        if (wv == 0) {
            const float4 *v = (const float4 *)seg[0];
            const int nv = (G - b0 < 1024 ? G - b0 + 3 : 1024) / 4;
            for (int t = 0; t < nv; t++) {
                const float4 a = v[t];
                cost += a.x;
                cost += a.y;
                cost += a.z;
                cost += a.w;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        sh_inv = 1.0f / cost;
        sh_p = u;
    }
    const int nseg = (int)((S + kKppSeg - 1) / kKppSeg);
    auto stage = [&](int sidx, int t0, int nt) {
        const int64_t s0 = (int64_t)sidx * kKppSeg;
        const int len = (int)(S - s0 < kKppSeg ? S - s0 : kKppSeg);
        float *dst = seg[sidx & 1];
        for (int k = t0; k < kKppSeg; k += nt) dst[k] = k < len ? m2[s0 + k] : 0.0f;
    };
    stage(0, tid, 256);
    __syncthreads();
    const float inv = sh_inv;
    int64_t chosen = -1;
    for (int sidx = 0; sidx < nseg; sidx++) {
        if (wv == 0) {
            const int64_t s0 = (int64_t)sidx * kKppSeg;
            const int len = (int)(S - s0 < kKppSeg ? S - s0 : kKppSeg);
            const float *cur = seg[sidx & 1];
            const float4 *v = (const float4 *)cur;
            float p = sh_p;
            int64_t pick = -1;
            const int nblk = (len + 63) / 64;
            // Two register sets in turn: block b + 1's loads are in flight
            // while block b's chain runs.  Returns true at a crossing.
            auto walk = [&](const float4 *x, int b) -> bool {
                float q = p;
#pragma unroll
                for (int t = 0; t < 16; t++) {
                    q = fmaf(-x[t].x, inv, q);
                    q = fmaf(-x[t].y, inv, q);
                    q = fmaf(-x[t].z, inv, q);
                    q = fmaf(-x[t].w, inv, q);
                }
                if (q < 0.0f) {
                    q = p;
                    for (int t = 0; t < 64; t++) {
                        q = fmaf(-cur[64 * b + t], inv, q);
                        if (q < 0.0f) {
                            pick = (int64_t)sidx * kKppSeg + 64 * b + t;
                            break;
                        }
                    }
                    return true;
                }
                p = q;
                return false;
            };
            float4 A[16], B[16];
#pragma unroll
            for (int t = 0; t < 16; t++) A[t] = v[t];
            for (int b = 0; b < nblk; b += 2) {
                if (b + 1 < nblk) {
#pragma unroll
                    for (int t = 0; t < 16; t++) B[t] = v[16 * (b + 1) + t];
                }
                if (walk(A, b) || b + 1 >= nblk) break;
                if (b + 2 < nblk) {
#pragma unroll
                    for (int t = 0; t < 16; t++) A[t] = v[16 * (b + 2) + t];
                }
                if (walk(B, b + 1)) break;
            }
            if (lane == 0) {
                sh_p = p;
                sh_pick[sidx & 1] = pick;
            }
        } else if (sidx + 1 < nseg) {
            stage(sidx + 1, tid - kWave, 256 - kWave);
        }
        __syncthreads();
        chosen = sh_pick[sidx & 1];
        if (chosen >= 0) break;
    }
    if (tid == 0) {
        const int64_t pick = chosen >= 0 ? chosen : S - 1;
        m2[pick] = 0.0f;
        *pos = (uint32_t)pick;
        *cell_out = sel[pick];
    }
}

template <typename T>
struct Buf {
    T *p = nullptr;
    size_t count = 0;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t c) {
        count = c;
        return hipMalloc((void **)&p, sizeof(T) * std::max<size_t>(c, 1));
    }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
};

#define IVF_TRY(expr)                        \
    do {                                     \
        hipError_t _e = (expr);              \
        if (_e != hipSuccess) {              \
            *herr = _e;                      \
            return 2;                        \
        }                                    \
    } while (0)

unsigned grid_of(int64_t n) { return (unsigned)ceil_div(n, 256); }

}  // namespace

std::vector<uint32_t> kmeans_random_init_cells(int64_t n, int num_clusters, uint64_t seed) {
    Pcg32 rng(seed);
    std::vector<uint64_t> taken((size_t)(n + 63) / 64, 0);
    std::vector<uint32_t> out;
    out.reserve(num_clusters);
    for (int i = 0; i < num_clusters; i++) {
        for (;;) {
            const uint32_t j = rng.bounded((uint32_t)n);
            if (!(taken[j >> 6] & (1ull << (j & 63)))) {
                taken[j >> 6] |= 1ull << (j & 63);
                out.push_back(j);
                break;
            }
        }
    }
    return out;
}

int kmeans_barbar_init_cells(const float *X, const float *norms_in, const float *Xp_in,
                             const IvfParams &P, std::vector<uint32_t> *cells,
                             int64_t *candidates, std::string *msg,
                             hipError_t *herr, hipStream_t s) {
    const int64_t n = P.n;
    const int d = P.d, C = P.num_clusters;
    const int st = knn_mfma_stride(d);
    const int64_t pad = kKnnPadRows;
    const int64_t chunk = P.chunk, nchunks = ceil_div(n, chunk);
    const float l = P.oversampling * (float)(uint32_t)C;
    Buf<float> norms_own, Xp_own;
    const float *norms = norms_in, *Xp = Xp_in;
    if (!norms || !Xp) {
        IVF_TRY(norms_own.alloc(n + pad));
        IVF_TRY(Xp_own.alloc((n + pad) * st));
        IVF_TRY(launch_row_norms(X, n, d, norms_own.p, s));
        IVF_TRY(launch_knn_pack(X, nullptr, n, d, norms_own.p, Xp_own.p, s));
        norms = norms_own.p;
        Xp = Xp_own.p;
    }
    Buf<float> md, pd, csum;
    Buf<uint32_t> best, sel, fresh, pj, nsel;
    Buf<uint8_t> flags;
    IVF_TRY(md.alloc(n));
    IVF_TRY(pd.alloc(n));
    IVF_TRY(csum.alloc(nchunks));
    IVF_TRY(best.alloc(n));
    IVF_TRY(sel.alloc(n));
    IVF_TRY(fresh.alloc(n));
    IVF_TRY(pj.alloc(n));
    IVF_TRY(nsel.alloc(1));
    IVF_TRY(flags.alloc(n));
    IVF_TRY(hipMemsetAsync(best.p, 0, sizeof(uint32_t) * n, s));
    size_t sel_bytes = 0;
    IVF_TRY(rocprim::select(nullptr, sel_bytes, rocprim::counting_iterator<uint32_t>(0), flags.p,
                            fresh.p, nsel.p, (size_t)n, s));
    Buf<char> sel_tmp;
    IVF_TRY(sel_tmp.alloc(sel_bytes));
    std::vector<float> h_csum((size_t)nchunks);
    // The cost: chunk sums in cell order, then the chunk sums in order.
    auto cost_of = [&](float *out) -> hipError_t {
        k_kb_chunk_sums<<<grid_of(nchunks), 256, 0, s>>>(md.p, n, chunk, nchunks, csum.p);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(h_csum.data(), csum.p, sizeof(float) * nchunks, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        float c = 0.0f;
        for (int64_t i = 0; i < nchunks; i++) c += h_csum[i];
        *out = c;
        return hipSuccess;
    };
    int64_t S = 0;
    // A round's draws; the taken cells (in cell order) are appended to sel.
    auto sample = [&](uint32_t round, float l_over_cost, int64_t *taken) -> int {
        k_kb_sample<<<grid_of(n), 256, 0, s>>>(md.p, n, P.seed, round, l_over_cost, flags.p);
        IVF_TRY(hipGetLastError());
        size_t bytes = sel_bytes;
        IVF_TRY(rocprim::select(sel_tmp.p, bytes, rocprim::counting_iterator<uint32_t>(0), flags.p,
                                fresh.p, nsel.p, (size_t)n, s));
        uint32_t h_count = 0;
        IVF_TRY(hipMemcpyAsync(&h_count, nsel.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        IVF_TRY(hipStreamSynchronize(s));
        if (S + (int64_t)h_count > n) {
            *msg = "k-means||: more candidates than cells";
            return 1;
        }
        if (h_count)
            IVF_TRY(hipMemcpyAsync(sel.p + S, fresh.p, sizeof(uint32_t) * h_count, hipMemcpyDeviceToDevice, s));
        *taken = h_count;
        return 0;
    };

    // The first candidate and every cell's distance to it.
    Pcg32 g0(P.seed - 1);
    const uint32_t rc = g0.bounded((uint32_t)n);
    IVF_TRY(hipMemcpyAsync(sel.p, &rc, sizeof(uint32_t), hipMemcpyHostToDevice, s));
    S = 1;
    k_kb_first<<<grid_of(n), 256, 0, s>>>(X, n, d, rc, md.p);
    IVF_TRY(hipGetLastError());
    float cost = 0.0f;
    IVF_TRY(cost_of(&cost));
    if (cost == 0.0f) {
        *msg = "all cells have the same principal component loadings";
        return 1;
    }
    int64_t taken = 0;
    if (int r = sample(1, l / cost, &taken)) return r;
    int64_t prev = 1;
    S += taken;
    int64_t fresh_n = taken;
    if (fresh_n) {
        k_kb_set_best<<<grid_of(fresh_n), 256, 0, s>>>(sel.p, prev, S, best.p);
        IVF_TRY(hipGetLastError());
    }

    Buf<float> candp;
    KnnTiles E{};
    E.Q = Xp;
    E.qnorm = 0;
    E.d = d;
    E.nq = n;
    E.st = st;
    E.ksteps = knn_mfma_ksteps(d);
    E.k = 1;
    E.self = 0;
    E.out_idx = pj.p;
    E.out_dist = pd.p;
    uint32_t round = 1;
    for (;;) {
        if (fresh_n > 0) {
            if ((int64_t)candp.count < (fresh_n + pad) * st) {
                candp.reset();
                IVF_TRY(candp.alloc((fresh_n + pad) * st));
            }
            IVF_TRY(launch_knn_pack(X, sel.p + prev, fresh_n, d, norms, candp.p, s));
            E.C = candp.p;
            E.nc = fresh_n;
            IVF_TRY(launch_knn_tiles(E, false, s));
            k_kb_update<<<grid_of(n), 256, 0, s>>>(md.p, norms, best.p, pd.p, pj.p, (uint32_t)prev, n);
            IVF_TRY(hipGetLastError());
            IVF_TRY(cost_of(&cost));
        }
        if (int r = sample(round, l / cost, &taken)) return r;
        prev = S;
        S += taken;
        fresh_n = taken;
        if (fresh_n) {
            k_kb_set_best<<<grid_of(fresh_n), 256, 0, s>>>(sel.p, prev, S, best.p);
            IVF_TRY(hipGetLastError());
        }
        if (fresh_n == 0 && cost == 0.0f) {
            *msg = "num_clusters (" + std::to_string(C) +
                   ") is greater than the number of cells with distinct principal component "
                   "loadings; decrease num_clusters";
            return 1;
        }
        round++;
        if ((round == (uint32_t)P.init_iterations && S >= C) || S == n) break;
    }
    *candidates = S;

    // Weighted k-means++ over the S candidates.
    Buf<uint32_t> w, pos, cells_d;
    Buf<float> candT, m2, parts;
    IVF_TRY(w.alloc(S));
    IVF_TRY(hipMemsetAsync(w.p, 0, sizeof(uint32_t) * S, s));
    k_kb_weights<<<grid_of(n), 256, 0, s>>>(best.p, n, w.p);
    IVF_TRY(hipGetLastError());
    Pcg32 g2(P.seed - 2);
    const uint32_t p0 = g2.bounded((uint32_t)S);
    uint32_t first_cell = 0;
    IVF_TRY(hipMemcpyAsync(&first_cell, sel.p + p0, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    IVF_TRY(hipStreamSynchronize(s));
    cells->assign(C, 0);
    (*cells)[0] = first_cell;
    if (C == 1) return 0;
    const int64_t G = ceil_div(S, 256);
    IVF_TRY(candT.alloc(S * d));
    IVF_TRY(m2.alloc(S));
    IVF_TRY(parts.alloc(G));
    IVF_TRY(pos.alloc(1));
    IVF_TRY(cells_d.alloc(C));
    k_kb_gather_t<<<grid_of(S * d), 256, 0, s>>>(X, d, sel.p, S, candT.p);
    IVF_TRY(hipGetLastError());
    IVF_TRY(hipMemcpyAsync(pos.p, &p0, sizeof(uint32_t), hipMemcpyHostToDevice, s));
    for (int c = 1; c < C; c++) {
        const float u = g2.uniform();
        k_kpp_update<<<(unsigned)G, 256, 0, s>>>(candT.p, S, d, w.p, m2.p, pos.p, c == 1, parts.p);
        k_kpp_draw<<<1, 256, 0, s>>>(m2.p, S, parts.p, (int)G, u, pos.p, sel.p, cells_d.p + c);
    }
    IVF_TRY(hipGetLastError());
    IVF_TRY(hipMemcpyAsync(cells->data() + 1, cells_d.p + 1, sizeof(uint32_t) * (C - 1), hipMemcpyDeviceToHost, s));
    IVF_TRY(hipStreamSynchronize(s));
    return 0;
}

int run_knn_ivf(const float *X, const IvfParams &P, uint32_t *out_idx,
                float *out_dist, IvfInfo *info, std::string *msg,
                hipError_t *herr, hipStream_t s) {
    const int64_t n = P.n;
    const int d = P.d, k = P.k;
    int C = P.num_clusters;
    const int st = knn_mfma_stride(d);
    const int ksteps = knn_mfma_ksteps(d);
    const int64_t pad = kKnnPadRows;

    // Cell norms (kept for the search), the padded cells.
    Buf<float> norms, Xp;
    IVF_TRY(norms.alloc(n + pad));
    IVF_TRY(Xp.alloc((n + pad) * st));
    IVF_TRY(launch_row_norms(X, n, d, norms.p, s));
    IVF_TRY(launch_knn_pack(X, nullptr, n, d, norms.p, Xp.p, s));

    // 1. k-means.
    std::vector<uint32_t> init;
    info->candidates = 0;
    if (P.barbar) {
        if (int r = kmeans_barbar_init_cells(X, norms.p, Xp.p, P, &init, &info->candidates, msg, herr, s))
            return r;
    } else {
        init = kmeans_random_init_cells(n, C, P.seed);
    }
    Buf<uint32_t> rows, labels, counts, order, order_tmp, labels_sorted;
    Buf<float> cent, centp, cnorm, mind, sums;
    Buf<int64_t> off;
    IVF_TRY(rows.alloc(C));
    IVF_TRY(labels.alloc(n));
    IVF_TRY(labels_sorted.alloc(n));
    IVF_TRY(order.alloc(n));
    IVF_TRY(order_tmp.alloc(n));
    IVF_TRY(counts.alloc(C));
    IVF_TRY(cent.alloc((size_t)C * d));
    IVF_TRY(sums.alloc((size_t)C * d));
    IVF_TRY(centp.alloc((size_t)(C + pad) * st));
    IVF_TRY(cnorm.alloc(C + pad));
    IVF_TRY(mind.alloc(n));
    IVF_TRY(off.alloc(C + 1));
    IVF_TRY(hipMemcpyAsync(rows.p, init.data(), sizeof(uint32_t) * C, hipMemcpyHostToDevice, s));
    IVF_TRY(hipStreamSynchronize(s));
    k_gather_rows<<<grid_of((int64_t)C * d), 256, 0, s>>>(X, d, rows.p, C, cent.p);
    IVF_TRY(hipGetLastError());

    // Radix sort scratch (labels < 2^bits).
    int bits = 1;
    while ((1ll << bits) < (int64_t)C) bits++;
    size_t sort_bytes = 0;
    IVF_TRY(rocprim::radix_sort_pairs(nullptr, sort_bytes, labels.p, labels_sorted.p,
                                      order_tmp.p, order.p, (size_t)n, 0, bits, s));
    Buf<char> sort_tmp;
    IVF_TRY(sort_tmp.alloc(sort_bytes));

    auto sort_by_label = [&]() -> hipError_t {
        k_iota<<<grid_of(n), 256, 0, s>>>(order_tmp.p, n);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        size_t bytes = sort_bytes;
        return rocprim::radix_sort_pairs(sort_tmp.p, bytes, labels.p, labels_sorted.p,
                                         order_tmp.p, order.p, (size_t)n, 0, bits, s);
    };

    KnnTiles E{};
    E.Q = Xp.p;
    E.C = centp.p;
    E.qnorm = 0;
    E.d = d;
    E.nq = n;
    E.nc = C;
    E.st = st;
    E.ksteps = ksteps;
    E.k = 1;
    E.self = 0;
    E.out_idx = labels.p;
    E.out_dist = mind.p;

    std::vector<uint32_t> h_counts(C);
    std::vector<int64_t> h_off(C + 1);
    float old_total = 0.0f;
    float sum_sq_norms = 0.0f;
    if (P.iterations > 2) {
        std::vector<float> h_norms(n);
        IVF_TRY(hipMemcpyAsync(h_norms.data(), norms.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        IVF_TRY(hipStreamSynchronize(s));
        for (int64_t i = 0; i < n; i++) sum_sq_norms += h_norms[i];
    }
    int iteration = 0;
    for (;;) {
        // E-step against the current centroids.
        IVF_TRY(launch_row_norms(cent.p, C, d, cnorm.p, s));
        IVF_TRY(launch_knn_pack(cent.p, nullptr, C, d, cnorm.p, centp.p, s));
        IVF_TRY(launch_knn_tiles(E, false, s));
        // M-step: counts, the stable order by label, sums in the reference's order.
        IVF_TRY(hipMemsetAsync(counts.p, 0, sizeof(uint32_t) * C, s));
        k_count_labels<<<grid_of(n), 256, 0, s>>>(labels.p, n, counts.p);
        IVF_TRY(hipGetLastError());
        IVF_TRY(sort_by_label());
        IVF_TRY(hipMemcpyAsync(h_counts.data(), counts.p, sizeof(uint32_t) * C, hipMemcpyDeviceToHost, s));
        IVF_TRY(hipStreamSynchronize(s));
        h_off[0] = 0;
        for (int j = 0; j < C; j++) h_off[j + 1] = h_off[j] + h_counts[j];
        IVF_TRY(hipMemcpyAsync(off.p, h_off.data(), sizeof(int64_t) * (C + 1), hipMemcpyHostToDevice, s));
        k_cluster_sums<<<(unsigned)ceil_div(C, 4), 256, 0, s>>>(X, d, order.p, off.p, C, P.chunk, sums.p);
        IVF_TRY(hipGetLastError());
        // Empty clusters take the farthest cells (kmeans.pyx:747-820).
        std::vector<int> empty;
        for (int j = 0; j < C; j++)
            if (h_counts[j] == 0) empty.push_back(j);
        if (!empty.empty()) {
            Buf<float> far;
            IVF_TRY(far.alloc(n));
            k_dist_to_centroid<<<grid_of(n), 256, 0, s>>>(X, n, d, labels.p, cent.p, far.p);
            IVF_TRY(hipGetLastError());
            std::vector<float> h_far(n), h_sums((size_t)C * d);
            std::vector<uint32_t> h_labels(n);
            IVF_TRY(hipMemcpyAsync(h_far.data(), far.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
            IVF_TRY(hipMemcpyAsync(h_labels.data(), labels.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
            IVF_TRY(hipMemcpyAsync(h_sums.data(), sums.p, sizeof(float) * C * d, hipMemcpyDeviceToHost, s));
            IVF_TRY(hipStreamSynchronize(s));
            std::vector<int64_t> idx(n);
            std::iota(idx.begin(), idx.end(), 0);
            const size_t ne = empty.size();
            std::partial_sort(idx.begin(), idx.begin() + ne, idx.end(), [&](int64_t a, int64_t b) {
                return h_far[a] > h_far[b] || (h_far[a] == h_far[b] && a < b);
            });
            if (h_far[idx[0]] == 0.0f) {
                *msg = "num_clusters (" + std::to_string(C) +
                       ") is greater than the number of cells with distinct principal component loadings (" +
                       std::to_string(n - (int64_t)ne) + "); decrease num_clusters";
                return 1;
            }
            std::vector<float> h_x((size_t)d);
            for (size_t i = 0; i < ne; i++) {
                const int nl = empty[i];
                const int64_t cell = idx[i];
                const uint32_t ol = h_labels[cell];
                IVF_TRY(hipMemcpyAsync(h_x.data(), X + cell * d, sizeof(float) * d, hipMemcpyDeviceToHost, s));
                IVF_TRY(hipStreamSynchronize(s));
                for (int j = 0; j < d; j++) {
                    h_sums[(size_t)ol * d + j] -= h_x[j];
                    h_sums[(size_t)nl * d + j] = h_x[j];
                }
                h_counts[ol] -= 1;
                h_counts[nl] += 1;
            }
            IVF_TRY(hipMemcpyAsync(sums.p, h_sums.data(), sizeof(float) * C * d, hipMemcpyHostToDevice, s));
            IVF_TRY(hipMemcpyAsync(counts.p, h_counts.data(), sizeof(uint32_t) * C, hipMemcpyHostToDevice, s));
            IVF_TRY(hipStreamSynchronize(s));
        }
        k_cluster_normalize<<<grid_of((int64_t)C * d), 256, 0, s>>>(sums.p, d, C, counts.p);
        IVF_TRY(hipGetLastError());
        std::swap(cent.p, sums.p);
        iteration++;
        if (iteration == P.iterations) break;
        // Convergence on the relative change of the inertia (kmeans.pyx
        // 1088-1100), summed in FP32 in cell order as the reference does.
        std::vector<float> h_mind(n);
        IVF_TRY(hipMemcpyAsync(h_mind.data(), mind.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        IVF_TRY(hipStreamSynchronize(s));
        float total = sum_sq_norms;
        for (int64_t i = 0; i < n; i++) total += h_mind[i];
        if (iteration > 1 && old_total - total < P.tolerance * old_total) break;
        old_total = total;
    }
    info->iterations = iteration;
    // A last E-step against the final centroids gives the labels and the
    // counts (kmeans.pyx:1357-1432).
    IVF_TRY(launch_row_norms(cent.p, C, d, cnorm.p, s));
    IVF_TRY(launch_knn_pack(cent.p, nullptr, C, d, cnorm.p, centp.p, s));
    IVF_TRY(launch_knn_tiles(E, false, s));
    IVF_TRY(hipMemsetAsync(counts.p, 0, sizeof(uint32_t) * C, s));
    k_count_labels<<<grid_of(n), 256, 0, s>>>(labels.p, n, counts.p);
    IVF_TRY(hipGetLastError());
    IVF_TRY(hipMemcpyAsync(h_counts.data(), counts.p, sizeof(uint32_t) * C, hipMemcpyDeviceToHost, s));
    IVF_TRY(hipStreamSynchronize(s));
    if (P.kmeans_labels) {
        IVF_TRY(hipMemcpyAsync(P.kmeans_labels, labels.p, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
        IVF_TRY(hipMemcpyAsync(P.kmeans_centroids, cent.p, sizeof(float) * C * d, hipMemcpyDeviceToDevice, s));
        IVF_TRY(hipMemcpyAsync(P.kmeans_counts, h_counts.data(), sizeof(uint32_t) * C, hipMemcpyHostToDevice, s));
        IVF_TRY(hipStreamSynchronize(s));
        info->clusters = C;
        return 0;
    }

    // 2. Drop empty clusters, renumber by the first centroid coordinate.
    std::vector<float> h_cent((size_t)C * d);
    IVF_TRY(hipMemcpyAsync(h_cent.data(), cent.p, sizeof(float) * C * d, hipMemcpyDeviceToHost, s));
    IVF_TRY(hipStreamSynchronize(s));
    std::vector<uint32_t> kept;
    for (int j = 0; j < C; j++)
        if (h_counts[j] > 0) kept.push_back((uint32_t)j);
    std::stable_sort(kept.begin(), kept.end(), [&](uint32_t a, uint32_t b) {
        return h_cent[(size_t)a * d] < h_cent[(size_t)b * d];
    });
    const int Ck = (int)kept.size();
    std::vector<uint32_t> map(C, 0);
    std::vector<int64_t> h_off2(Ck + 1);
    h_off2[0] = 0;
    for (int j = 0; j < Ck; j++) {
        map[kept[j]] = (uint32_t)j;
        h_off2[j + 1] = h_off2[j] + h_counts[kept[j]];
    }
    info->clusters = Ck;
    // Enough clusters for k + 1 candidates even if the nearest are the
    // smallest (single_cell.py:12555-12565).
    std::vector<uint32_t> sizes(Ck);
    for (int j = 0; j < Ck; j++) sizes[j] = h_counts[kept[j]];
    std::sort(sizes.begin(), sizes.end());
    int min_searched = Ck;
    uint64_t run = 0;
    for (int j = 0; j < Ck; j++) {
        run += sizes[j];
        if (run >= (uint64_t)k + 1) {
            min_searched = j + 1;
            break;
        }
    }
    int searched = std::min(std::max(P.searched, min_searched), Ck);
    info->searched = searched;
    if (searched == Ck || searched > kWave) {
        // Every cluster searched (or more than the lists hold): the
        // exhaustive search, which is what searching every cluster returns.
        info->exhaustive = 1;
        KnnTiles T{};
        T.Q = T.C = Xp.p;
        T.qnorm = 1;
        T.d = d;
        T.nq = T.nc = n;
        T.st = st;
        T.ksteps = ksteps;
        T.self = 1;
        T.out_idx = out_idx;
        T.out_dist = out_dist;
        IVF_TRY(launch_knn_tiles_k(T, false, k, s));
        IVF_TRY(hipStreamSynchronize(s));
        return 0;
    }
    info->exhaustive = 0;
    Buf<uint32_t> dmap;
    IVF_TRY(dmap.alloc(C));
    IVF_TRY(hipMemcpyAsync(dmap.p, map.data(), sizeof(uint32_t) * C, hipMemcpyHostToDevice, s));
    IVF_TRY(hipMemcpyAsync(off.p, h_off2.data(), sizeof(int64_t) * (Ck + 1), hipMemcpyHostToDevice, s));
    IVF_TRY(hipMemcpyAsync(rows.p, kept.data(), sizeof(uint32_t) * Ck, hipMemcpyHostToDevice, s));
    k_remap_labels<<<grid_of(n), 256, 0, s>>>(labels.p, n, dmap.p);
    IVF_TRY(hipGetLastError());
    k_gather_rows<<<grid_of((int64_t)Ck * d), 256, 0, s>>>(cent.p, d, rows.p, Ck, sums.p);
    IVF_TRY(hipGetLastError());
    IVF_TRY(launch_row_norms(sums.p, Ck, d, cnorm.p, s));
    IVF_TRY(launch_knn_pack(sums.p, nullptr, Ck, d, cnorm.p, centp.p, s));
    IVF_TRY(sort_by_label());

    // Cells in cluster order, padded.
    Buf<float> Xs;
    IVF_TRY(Xs.alloc((n + pad) * st));
    IVF_TRY(launch_knn_pack(X, order.p, n, d, norms.p, Xs.p, s));
    Xp.reset();  // the unsorted copy is no longer needed

    // 3. Each cell's nearest clusters (by partial distance).  A first tile
    // per query block, the kp clusters nearest its first cell's own cluster,
    // caps the thresholds of the scan (the same ranking, found sooner).
    const int kp = std::min(kWave, Ck);
    Buf<uint32_t> cc;
    Buf<float> cc_d, boot;
    IVF_TRY(cc.alloc((size_t)Ck * kp));
    IVF_TRY(cc_d.alloc((size_t)Ck * kp));
    IVF_TRY(boot.alloc((size_t)Ck * kKnnPadRows * st));
    {
        KnnTiles Cc{};
        Cc.Q = Cc.C = centp.p;
        Cc.qnorm = 1;
        Cc.d = d;
        Cc.nq = Cc.nc = Ck;
        Cc.st = st;
        Cc.ksteps = ksteps;
        Cc.k = kp;
        Cc.self = 0;
        Cc.out_idx = cc.p;
        Cc.out_dist = cc_d.p;
        IVF_TRY(launch_knn_tiles(Cc, false, s));
        k_gather_blocks<<<grid_of((int64_t)Ck * kKnnPadRows * st), 256, 0, s>>>(
            centp.p, d, st, cc.p, Ck, kp, boot.p);
        IVF_TRY(hipGetLastError());
    }
    Buf<uint32_t> near;
    Buf<float> near_d;
    IVF_TRY(near.alloc((size_t)n * searched));
    IVF_TRY(near_d.alloc((size_t)n * searched));
    {
        KnnTiles R{};
        R.Q = Xs.p;
        R.C = centp.p;
        R.qnorm = 0;
        R.d = d;
        R.nq = n;
        R.nc = Ck;
        R.st = st;
        R.ksteps = ksteps;
        R.k = searched;
        R.self = 0;
        R.start_hint = labels_sorted.p;  // the block's own cluster first
        R.boot = boot.p;
        R.boot_rows = kp;
        R.out_idx = near.p;
        R.out_dist = near_d.p;
        IVF_TRY(launch_knn_tiles(R, false, s));
    }
    near_d.reset();
    boot.reset();

    // 4. Per-workgroup cluster lists.
    const int64_t nwg = ceil_div(n, kIvfQueries);
    Buf<int64_t> wg_cnt, wg_off;
    IVF_TRY(wg_cnt.alloc(nwg));
    IVF_TRY(wg_off.alloc(nwg + 1));
    k_ivf_entries<false><<<(unsigned)nwg, 256, 0, s>>>(near.p, n, searched, wg_cnt.p, nullptr, nullptr, nullptr);
    IVF_TRY(hipGetLastError());
    std::vector<int64_t> h_wg((size_t)nwg + 1);
    IVF_TRY(hipMemcpyAsync(h_wg.data() + 1, wg_cnt.p, sizeof(int64_t) * nwg, hipMemcpyDeviceToHost, s));
    IVF_TRY(hipStreamSynchronize(s));
    h_wg[0] = 0;
    for (int64_t b = 0; b < nwg; b++) h_wg[b + 1] += h_wg[b];
    const int64_t num_entries = h_wg[nwg];
    IVF_TRY(hipMemcpyAsync(wg_off.p, h_wg.data(), sizeof(int64_t) * (nwg + 1), hipMemcpyHostToDevice, s));
    Buf<uint32_t> ent;
    IVF_TRY(ent.alloc(4 * num_entries));
    k_ivf_entries<true><<<(unsigned)nwg, 256, 0, s>>>(near.p, n, searched, nullptr, wg_off.p, off.p, ent.p);
    IVF_TRY(hipGetLastError());
    info->entries = num_entries;

    // 5. The search.
    KnnTiles T{};
    T.Q = T.C = Xs.p;
    T.qnorm = 1;
    T.d = d;
    T.nq = T.nc = n;
    T.st = st;
    T.ksteps = ksteps;
    T.self = 1;
    T.wg_off = wg_off.p;
    T.ent = ent.p;
    T.ids = order.p;
    T.out_idx = out_idx;
    T.out_dist = out_dist;
    IVF_TRY(launch_knn_tiles_k(T, true, k, s));
    IVF_TRY(hipStreamSynchronize(s));
    return 0;
}

}  // namespace bpca
