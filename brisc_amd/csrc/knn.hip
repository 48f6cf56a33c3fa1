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

// ---------------------------------------------------------------------------
// Matrix-core path (num_dims <= 128): distances from FP32 MFMA tiles,
// candidates buffered per query in LDS, lists merged with a bitonic network.
//
// The points are first copied to a padded row-major layout, row c being
// [x_c, 1, ||x_c||^2, 0 ...] (kdim = round_up(d + 2, 4) columns used, row
// stride st = kdim + 4 floats) so that a 64-cell tile is a run of 16-byte
// loads.  A workgroup of four waves takes 64 queries (wave w owns rows
// 16w..16w+15) and streams the cells in tiles of 64 through LDS (double
// buffered).  In LDS a query row becomes [-2 x_q, ||x_q||^2, 1, 0 ...] (or
// [-2 x_q, 0, 1, 0 ...] for the partial distance ||c||^2 - 2 q.c), so one
// product gives the distance: ||q||^2 + ||c||^2 - 2 q.c with all terms
// summed in the matrix core.  Wave w forms its 16 x 64 block with
// v_mfma_f32_16x16x4f32 (four 16 x 16 blocks; lane l holds rows
// 4(l/16)..4(l/16)+3 of column l%16 of each block); within a group of 16
// columns lane group g supplies columns 4g..4g+3, one per MFMA, to both
// operands, so every column enters each product once.
//
// Selection: each query row keeps its 64 best (distance, index) pairs, sorted,
// one per lane, in registers; a candidate that beats the row's k-th entry is
// appended to the row's LDS buffer (64 slots).  A row whose buffer would
// overflow, and every row at the end, is merged: the buffer is bitonically
// sorted descending across the lanes, the elementwise minimum with the
// ascending list holds the 64 smallest of both (a bitonic sequence), and a
// bitonic merge sorts it.  Keys compare by (distance, index), so the result
// does not depend on arrival order.
// ---------------------------------------------------------------------------
namespace {

constexpr int kTile = 64;  // cells per tile; queries per workgroup
constexpr int kBuf = 64;   // buffered candidates per query row
constexpr int kKsWidest = 65;     // k-steps of 256 dimensions: ceil((256 + 2) / 4)
constexpr int kKsRotBuf = 33;     // widest k-steps whose ranking keeps 64-slot rows
// Row buffer slots of a launch: 64 for the ranking (k = 64 there: half the
// merges) while its tiles leave room, 32 otherwise.
__host__ __device__ constexpr int rot_buffers(bool rot, int ks) {
    return rot && ks <= kKsRotBuf ? kBuf : 32;
}
typedef float knn_f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bool key_lt(float da, uint32_t ia, float db, uint32_t ib) {
    return da < db || (da == db && ia < ib);
}

// Compare-exchange with lane ^ stride; keep the smaller key when keep_min.
__device__ __forceinline__ void cmp_xchg(float &d, uint32_t &i, int stride, bool keep_min) {
    const float od = __int_as_float(xor_lane(__float_as_int(d), stride));
    const uint32_t oi = (uint32_t)xor_lane((int)i, stride);
    const bool take = keep_min ? key_lt(od, oi, d, i) : key_lt(d, i, od, oi);
    d = take ? od : d;
    i = take ? oi : i;
}

// a[r] for a wave-uniform r, as a chain of selects.  The empty asm keeps
// each element an opaque register value: otherwise the selects of loads are
// folded into one load at a variable index, which puts the array in scratch.
template <int N, typename T>
__device__ __forceinline__ T pick(const T (&a)[N], int r) {
    T v = a[0];
    asm("" : "+v"(v));
#pragma unroll
    for (int j = 1; j < N; j++) {
        T x = a[j];
        asm("" : "+v"(x));
        v = r == j ? x : v;
    }
    return v;
}

template <int N, typename T>
__device__ __forceinline__ void put(T (&a)[N], int r, T v) {
#pragma unroll
    for (int j = 0; j < N; j++) a[j] = r == j ? v : a[j];
}

// Merge row r (wave-uniform) of this wave: list (Ld, Li lane = entry) with
// the row's cnt buffered candidates; refresh the row's threshold (entry k-1,
// or the row's cap if that is smaller).
__device__ __forceinline__ void merge_row(float (&Ld)[16], uint32_t (&Li)[16],
                                          float (&tau_d)[4], uint32_t (&tau_i)[4],
                                          int (&cnt)[4], const float (&cap)[4],
                                          const float *bd, const uint32_t *bi, int r,
                                          int k, int lane) {
    const int g = r >> 2, i = r & 3;
    const int c = __builtin_amdgcn_readlane(pick(cnt, i), 16 * g);  // r is wave-uniform
    float B = INFINITY;
    uint32_t BI = 0xffffffffu;
    if (lane < c) {
        B = bd[lane];
        BI = bi[lane];
    }
    // Bitonic sort, descending overall.
#pragma unroll
    for (int size = 2; size <= kWave; size <<= 1) {
        const bool asc = (lane & size) == 0;
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const bool lower = (lane & stride) == 0;
            cmp_xchg(B, BI, stride, lower != asc);
        }
    }
    float L = pick(Ld, r);
    uint32_t LI = pick(Li, r);
    if (key_lt(B, BI, L, LI)) {
        L = B;
        LI = BI;
    }
#pragma unroll
    for (int stride = kWave >> 1; stride > 0; stride >>= 1)
        cmp_xchg(L, LI, stride, (lane & stride) == 0);
    put(Ld, r, L);
    put(Li, r, LI);
    const float td = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(L), k - 1));
    const uint32_t ti = (uint32_t)__builtin_amdgcn_readlane((int)LI, k - 1);
    if ((lane >> 4) == g) {
        // A row's cap (a distance no greater than its k-th nearest) bounds
        // the threshold while the list holds fewer than k close entries.
        const float c = pick(cap, i);
        const bool capped = c < td;
        put(tau_d, i, capped ? c : td);
        put(tau_i, i, capped ? 0xffffffffu : ti);
        put(cnt, i, 0);
    }
}

}  // namespace

// Padded gather: row c of Xp is [x, 1, norms[src], 0 ...] for source row
// src = order ? order[c] : c of X (n x d), with kTile zero rows past n.
__global__ __launch_bounds__(256) void k_knn_pack(const float *__restrict__ X,
                                                  const uint32_t *__restrict__ order,
                                                  int64_t n, int d, int st,
                                                  const float *__restrict__ norms,
                                                  float *__restrict__ Xp) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (n + kTile) * st) return;
    const int64_t c = e / st;
    const int j = (int)(e - c * st);
    float v = 0.0f;
    if (c < n) {
        const int64_t src = order ? (int64_t)order[c] : c;
        v = j < d ? X[src * d + j] : j == d ? 1.0f : j == d + 1 ? norms[src] : 0.0f;
    }
    Xp[e] = v;
}

typedef __attribute__((address_space(3))) void lds_void;

// The 16-byte vectors [0, 64 chunks) of a run starting at src, copied into
// LDS at dst by the LDS DMA path (global_load_lds_dwordx4: lane l of a
// 64-vector chunk lands 16 l bytes past the chunk's base), the chunks dealt
// to the four waves in turn.  Completion is waited for with vmcnt.
__device__ __forceinline__ void dma_run(const float *src, float *dst, int chunks,
                                        int w, int lane) {
    const char *base = (const char *)src;  // wave-uniform; lanes add a 32-bit offset
    for (int c = w; c < chunks; c += 4)
        __builtin_amdgcn_global_load_lds(base + (uint32_t)(16 * (kWave * c + lane)),
                                         (lds_void *)(dst + 4 * kWave * c), 16, 0, 0);
}

// ARGMIN (k = 1, exhaustive scans: the k-means E-step): each lane keeps the
// best (distance, id) of its own columns in registers; the 16 lanes of a
// row group are reduced at the end.  No buffers, no merges.
// ROT (exhaustive scans with a start hint: the centroid ranking): the scan
// starts at the tile holding the block's own cluster and wraps, so the
// thresholds are tight from the first tiles; ids then arrive out of order.
// With T.boot, a first tile holds the clusters nearest the block's own
// (T.boot_rows of them): its k-th distance per row caps that row's threshold
// (the k-th nearest over all candidates is no farther), and the lists then
// restart empty, so every candidate is still seen once.
template <bool IVF, bool ARGMIN, bool ROT, int KS>
__global__ __launch_bounds__(256) void k_knn_mfma(const KnnTiles T) {
    extern __shared__ __align__(16) float lds_f[];
    // The queries' operands are held in registers: the queries pass through
    // candidate buffer 1 on the way (LDS: two tiles, ids, row buffers of 32
    // slots, 64 for the ranking, where k = 64).
    static_assert(KS >= 1 && KS <= kKsWidest, "k-steps");
    constexpr int kB = rot_buffers(ROT, KS);
    const int st = T.st;
    const int chunks = st >> 2;                   // 64 rows x st floats
    float *const cs0 = lds_f;                     // 2 x (64 x st) candidates
    float *const qs = cs0 + kTile * st;           // 64 x st queries, at first
    uint32_t *cid0 = (uint32_t *)(cs0 + 2 * kTile * st);  // 2 x 64 candidate ids
    float *bufd = (float *)(cid0 + 2 * kTile);    // 64 rows x kB
    uint32_t *bufi = (uint32_t *)(bufd + kTile * kB);
    const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid >> 6;
    const int g = lane >> 4, m = lane & 15;
    const int q0 = (int)blockIdx.x * kTile;
    const int nq = (int)T.nq, nc = (int)T.nc;
    const int k = T.k;

    // The candidate tiles, in scan order (wave-uniform state).  Exhaustive:
    // rows 0, 64, ... of the candidates.  Inverted file: the rows of each of
    // the workgroup's (cluster, query mask) entries in turn.
    int pos = 0, lim = nc;
    int64_t e = 0, e_end = 0;
    uint64_t mask = ~0ull;
    bool valid = nc > 0;
    int start = 0;
    bool wrapped = false;
    bool boot = false;
    const float *boot_src = nullptr;
    if (ROT) {
        const uint32_t own = T.start_hint[min(q0, nq - 1)];
        start = min((int)(own & ~(uint32_t)(kTile - 1)), (nc - 1) & ~(kTile - 1));
        pos = start;
        if (T.boot) {
            boot = true;
            boot_src = T.boot + (size_t)own * kTile * st;
            pos = 0;
            lim = T.boot_rows;
        }
    }
    // Inverted file: an entry's four words are read by lanes 0..3 (a vector
    // load) and taken with readlane; the next entry's are loaded when an
    // entry starts, so they have landed (the end-of-tile wait) before use.
    uint32_t nrec = 0;
    auto fetch = [&](int64_t x) { return T.ent[4 * x + (lane & 3)]; };
    auto unpack = [&](uint32_t r, int &p, int &l, uint64_t &mk) {
        p = __builtin_amdgcn_readlane((int)r, 0);
        l = __builtin_amdgcn_readlane((int)r, 1);
        mk = (uint32_t)__builtin_amdgcn_readlane((int)r, 2) |
             ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)r, 3) << 32);
    };
    if (IVF) {
        e = T.wg_off[blockIdx.x];
        e_end = T.wg_off[blockIdx.x + 1];
        valid = e < e_end;
        if (valid) {
            unpack(fetch(e), pos, lim, mask);
            if (e + 1 < e_end) nrec = fetch(e + 1);
        }
    }
    // Tile rows p .. p + 63 (pad rows keep every read in bounds; ids are
    // clamped, and their rows are masked) into buffer b.
    auto stage = [&](const float *src, int p, int b) {
        dma_run(src + (size_t)p * st, cs0 + b * (kTile * st), chunks, w, lane);
        if (IVF && w == 0)
            __builtin_amdgcn_global_load_lds(T.ids + min(p + lane, nc - 1),
                                             (lds_void *)(cid0 + b * kTile), 4, 0, 0);
    };
    dma_run(T.Q + (size_t)q0 * st, qs, chunks, w, lane);
    if (valid) stage(ROT && boot ? boot_src : T.C, pos, 0);

    int qid[4];
#pragma unroll
    for (int i = 0; i < 4; i++) qid[i] = q0 + 16 * w + 4 * g + i;
    float fl_d[4] = {0.f, 0.f, 0.f, 0.f};
    uint32_t fl_i[4] = {0u, 0u, 0u, 0u};
    if (T.floor_d) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int q = qid[i] < nq ? qid[i] : nq - 1;
            const size_t row = (size_t)(IVF ? T.ids[q] : (uint32_t)q) * T.floor_ld;
            fl_d[i] = T.floor_d[row];
            fl_i[i] = T.floor_i[row];
        }
    }
    float Ld[16];
    uint32_t Li[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        Ld[r] = INFINITY;
        Li[r] = 0xffffffffu;
    }
    float tau_d[4], cap[4];
    uint32_t tau_i[4];
    int cnt[4];
    float bestd[4];
    uint32_t besti[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        tau_d[i] = INFINITY;
        cap[i] = INFINITY;
        tau_i[i] = 0xffffffffu;
        cnt[i] = 0;
        bestd[i] = INFINITY;
        besti[i] = 0xffffffffu;
    }
    float *const myb_d = bufd + (16 * w) * kB;
    uint32_t *const myb_i = bufi + (16 * w) * kB;
    const uint32_t below = (1u << m) - 1u;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    {
        // The wave's 16 query rows [x, 1, n, 0 ...] become [-2 x, n or 0, 1,
        // 0 ...].  Each lane reads its row's norm before any lane writes.
        const int d = T.d;
        float *myq = qs + 16 * w * st;
        const float nrm = lane < 16 ? myq[lane * st + d + 1] : 0.0f;
        for (int x = lane; x < 16 * st; x += kWave) {
            const int r = x / st, j = x - r * st;
            const float v = myq[x];
            const float nr = __shfl(nrm, r);
            myq[x] = j < d ? -2.0f * v : j == d ? (T.qnorm ? nr : 0.0f) : j == d + 1 ? 1.0f : 0.0f;
        }
    }
    __syncthreads();
    constexpr int QG = KS / 4 > 0 ? KS / 4 : 1, QT = KS % 4 ? KS % 4 : 1;
    knn_f4 qa[QG];  // this lane's query operands: columns 16 gg + 4 g .. + 3
    float qt[QT];   // and the tail steps' columns 16 (KS / 4) + 4 s + g
    {
        const float *qrow = qs + (16 * w + m) * st + 4 * g;
#pragma unroll
        for (int gg = 0; gg < KS / 4; gg++) qa[gg] = *(const knn_f4 *)(qrow + 16 * gg);
#pragma unroll
        for (int s = 0; s < KS % 4; s++) qt[s] = qrow[16 * (KS / 4) + 4 * s - 3 * g];
        __syncthreads();  // buffer 1 is free for the tiles
    }

    int t = 0;
    while (valid) {
        const float *cs = cs0 + t * (kTile * st);
        const uint32_t *cid = cid0 + t * kTile;
        int npos = pos + kTile, nlim = lim;
        int64_t ne = e;
        uint64_t nmask = mask;
        bool nvalid = npos < lim;
        bool nwrapped = wrapped;
        if (ROT && boot) {
            // The first tile of the scan proper follows the first tile.
            npos = start;
            nlim = nc;
            nvalid = true;
        } else if (ROT && !nvalid && !wrapped && start > 0) {
            npos = 0;
            nlim = start;
            nwrapped = true;
            nvalid = true;
        }
        if (IVF && !nvalid) {
            ne = e + 1;
            nvalid = ne < e_end;
            if (nvalid) {
                unpack(nrec, npos, nlim, nmask);
                if (ne + 1 < e_end) nrec = fetch(ne + 1);
            }
        }
        // The next tile streams into the other buffer while this one is used.
        if (nvalid) stage(T.C, npos, t ^ 1);

        // Inverted file: the wave's queries that search this cluster.  A wave
        // with none skips the tile (it still stages and meets the barrier).
        const uint32_t wmask = IVF ? (uint32_t)((mask >> (16 * w)) & 0xffffull) : 0xffffu;
        if (wmask != 0) {
        // The wave's 16 x 64 block of distances.
        knn_f4 acc[4];
        {
            // Straight-line code in which the next group's (or the tail's)
            // candidate operands are read from LDS before the current
            // group's products are issued; the scheduling barriers keep
            // that order, so the reads complete under the products.  Two
            // operand sets alternate; the query operands are in registers.
            constexpr int G = KS / 4, TL = KS % 4;
            const float *crow = cs + m * st + 4 * g;
            knn_f4 sb[2][4];
            float tb[TL > 0 ? TL : 1][4];
            auto load_group = [&](int gg, int set) {
#pragma unroll
                for (int b = 0; b < 4; b++)
                    sb[set][b] = *(const knn_f4 *)(crow + 16 * b * st + 16 * gg);
            };
            auto load_tail = [&]() {
#pragma unroll
                for (int s = 0; s < TL; s++) {
                    const int off = 16 * G + 4 * s - 3 * g;
#pragma unroll
                    for (int b = 0; b < 4; b++) tb[s][b] = crow[16 * b * st + off];
                }
            };
            const knn_f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
            if constexpr (G > 0) load_group(0, 0);
            else load_tail();
#pragma unroll
            for (int gg = 0; gg < G; gg++) {
                if (gg + 1 < G) load_group(gg + 1, (gg + 1) & 1);
                else load_tail();
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int s = 0; s < 4; s++)
#pragma unroll
                    for (int b = 0; b < 4; b++)
                        acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                            qa[gg][s], sb[gg & 1][b][s], gg == 0 && s == 0 ? zero : acc[b], 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int s = 0; s < TL; s++)
#pragma unroll
                for (int b = 0; b < 4; b++)
                    acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                        qt[s], tb[s][b], G == 0 && s == 0 ? zero : acc[b], 0, 0, 0);
        }
        // Masks: candidates past the tile's end (the last tile of a run) and
        // the query itself (the tile holding it) become NaN, which no
        // comparison below accepts.  Inverted file: rows whose query does not
        // search this cluster get a NaN threshold instead (below).
        const bool selftile = T.self && pos < q0 + 16 * w + 16 && pos + kTile > q0 + 16 * w;
        if (pos + kTile > lim || selftile) {
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const bool ok = pos + 16 * b + m < lim;
#pragma unroll
                for (int i = 0; i < 4; i++) acc[b][i] = ok ? acc[b][i] : __builtin_nanf("");
            }
            if (selftile) {
#pragma unroll
                for (int b = 0; b < 4; b++)
#pragma unroll
                    for (int i = 0; i < 4; i++)
                        acc[b][i] = pos + 16 * b + m == qid[i] ? __builtin_nanf("") : acc[b][i];
            }
        }
        if (T.floor_d) {
            // Passes after the first: a candidate not after its row's floor in
            // (distance, id) order was returned by an earlier pass.
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const uint32_t idb = IVF ? cid[16 * b + m] : (uint32_t)(pos + 16 * b + m);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const float v = acc[b][i];
                    const bool after = v > fl_d[i] || (v == fl_d[i] && idb > fl_i[i]);
                    acc[b][i] = after ? v : __builtin_nanf("");
                }
            }
        }
        if constexpr (ARGMIN) {
            // Columns arrive in increasing id per lane: a tie keeps the first.
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const uint32_t idv = (uint32_t)(pos + 16 * b + m);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const bool better = acc[b][i] < bestd[i];
                    bestd[i] = better ? acc[b][i] : bestd[i];
                    besti[i] = better ? idv : besti[i];
                }
            }
        } else {
        // A candidate is kept if it does not exceed its row's k-th entry
        // (exhaustive scans in increasing id: if it is below it, as a tie
        // loses).  Ties kept here are ordered by (distance, id) when merged,
        // so the k best are the same whichever order the tiles come in.
        float te[4];  // the thresholds, NaN for rows that skip this tile
        auto set_te = [&]() {
#pragma unroll
            for (int i = 0; i < 4; i++)
                te[i] = (!IVF || ((wmask >> (4 * g + i)) & 1u)) ? tau_d[i] : __builtin_nanf("");
        };
        set_te();
        auto passes = [&](float v, int i) {
            return (IVF || ROT) ? v <= te[i] : v < te[i];
        };
        uint64_t any = 0;
#pragma unroll
        for (int b = 0; b < 4; b++)
#pragma unroll
            for (int i = 0; i < 4; i++) any |= __ballot(passes(acc[b][i], i));
        if (any) {
            uint32_t idc[4];
#pragma unroll
            for (int b = 0; b < 4; b++)
                idc[b] = IVF ? cid[16 * b + m] : (uint32_t)(pos + 16 * b + m);
            // Two forms of the same appends (either keeps the k best, and
            // the merges order by (distance, id)); each is the faster one,
            // as measured, for its kind of scan.  The counts-first form
            // needs room for a whole tile (64 per row) after a merge, so
            // the 32-slot buffers take the per-block form.
            if constexpr (IVF || kB < kBuf) {
                // One 16-column block at a time (at most 16 per row): rows
                // whose buffer cannot take the block's candidates are merged
                // first, and the block is then tested against their new
                // thresholds.
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    // Late in a scan a tile has a candidate or two: blocks
                    // and rows with none are skipped (wave-uniform tests).
                    uint64_t bal[4], bany = 0;
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        bal[i] = __ballot(passes(acc[b][i], i));
                        bany |= bal[i];
                    }
                    if (!bany) continue;
                    uint32_t gb[4];
                    uint64_t need[4], need_any = 0;
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        gb[i] = (uint32_t)((bal[i] >> (16 * g)) & 0xffffull);
                        need[i] = __ballot(cnt[i] + __popc(gb[i]) > kB);
                        need_any |= need[i];
                    }
                    if (need_any) {
                        uint32_t rmask = 0;
#pragma unroll
                        for (int i = 0; i < 4; i++)
#pragma unroll
                            for (int gq = 0; gq < 4; gq++)
                                rmask |= (uint32_t)((need[i] >> (16 * gq)) & 1ull) << (4 * gq + i);
                        while (rmask) {
                            const int r = __builtin_ctz(rmask);
                            rmask &= rmask - 1;
                            merge_row(Ld, Li, tau_d, tau_i, cnt, cap, myb_d + r * kB, myb_i + r * kB, r, k, lane);
                        }
                        set_te();
#pragma unroll
                        for (int i = 0; i < 4; i++) {
                            bal[i] = __ballot(passes(acc[b][i], i));
                            gb[i] = (uint32_t)((bal[i] >> (16 * g)) & 0xffffull);
                        }
                    }
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        if (!bal[i]) continue;
                        if ((gb[i] >> m) & 1u) {
                            const int slot = (4 * g + i) * kB + cnt[i] + __popc(gb[i] & below);
                            myb_d[slot] = acc[b][i];
                            myb_i[slot] = idc[b];
                        }
                        cnt[i] += __popc(gb[i]);
                    }
                }
            } else {
                // Per-row counts first.  (The thresholds are passed through an
                // opaque copy so that the compares above are not kept live.)
                float th[4];
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    th[i] = te[i];
                    asm("" : "+v"(th[i]));
                }
                int arr[4] = {0, 0, 0, 0};
#pragma unroll
                for (int b = 0; b < 4; b++) {
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const bool p = ROT ? acc[b][i] <= th[i] : acc[b][i] < th[i];
                        const uint64_t bal = __ballot(p);
                        arr[i] += __popcll((bal >> (16 * g)) & 0xffffull);
                    }
                }
                // Rows whose buffer would overflow are merged first.  (The
                // appends below test against the thresholds as they are then:
                // a candidate a merge has put out of reach is not kept, so
                // arr bounds the appends.)
                uint32_t rmask = 0;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const uint64_t need = __ballot(cnt[i] + arr[i] > kB);
#pragma unroll
                    for (int gq = 0; gq < 4; gq++)
                        rmask |= (uint32_t)((need >> (16 * gq)) & 1ull) << (4 * gq + i);
                }
                while (rmask) {
                    const int r = __builtin_ctz(rmask);
                    rmask &= rmask - 1;
                    merge_row(Ld, Li, tau_d, tau_i, cnt, cap, myb_d + r * kB, myb_i + r * kB, r, k, lane);
                }
                set_te();
#pragma unroll
                for (int b = 0; b < 4; b++) {
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const bool p = passes(acc[b][i], i);
                        const uint32_t gb = (uint32_t)((__ballot(p) >> (16 * g)) & 0xffffull);
                        if (p) {
                            const int slot = (4 * g + i) * kB + cnt[i] + __popc(gb & below);
                            myb_d[slot] = acc[b][i];
                            myb_i[slot] = idc[b];
                        }
                        cnt[i] += __popc(gb);
                    }
                }
            }
        }
        }  // selection
        }  // wmask
        if (ROT && boot) {
            // The first tile's k-th distances become the caps; the lists
            // restart empty (the scan proper sees these candidates again).
            uint32_t rmask = 0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint64_t need = __ballot(cnt[i] > 0);
#pragma unroll
                for (int gq = 0; gq < 4; gq++)
                    rmask |= (uint32_t)((need >> (16 * gq)) & 1ull) << (4 * gq + i);
            }
            while (rmask) {
                const int r = __builtin_ctz(rmask);
                rmask &= rmask - 1;
                merge_row(Ld, Li, tau_d, tau_i, cnt, cap, myb_d + r * kB, myb_i + r * kB, r, k, lane);
            }
#pragma unroll
            for (int i = 0; i < 4; i++) {
                cap[i] = tau_d[i];
                tau_i[i] = 0xffffffffu;
            }
#pragma unroll
            for (int r = 0; r < 16; r++) {
                Ld[r] = INFINITY;
                Li[r] = 0xffffffffu;
            }
            boot = false;
        }
        // The next tile has landed (this wave's share), then every wave's.
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        pos = npos;
        lim = nlim;
        e = ne;
        mask = nmask;
        valid = nvalid;
        wrapped = nwrapped;
        t ^= 1;
    }
    if constexpr (ARGMIN) {
        // Reduce each row over the 16 lanes of its group.
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float d = bestd[i];
            uint32_t id = besti[i];
#pragma unroll
            for (int x = 1; x < 16; x <<= 1) {
                const float od = __shfl_xor(d, x);
                const uint32_t oi = (uint32_t)__shfl_xor((int)id, x);
                const bool take = key_lt(od, oi, d, id);
                d = take ? od : d;
                id = take ? oi : id;
            }
            const int q = qid[i];
            if (m == 0 && q < nq) {
                const size_t row = (size_t)q;
                T.out_idx[row] = id;
                T.out_dist[row] = d;
            }
        }
        return;
    }
    // Final merges.
    uint32_t rmask = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint64_t need = __ballot(cnt[i] > 0);
#pragma unroll
        for (int gq = 0; gq < 4; gq++)
            rmask |= (uint32_t)((need >> (16 * gq)) & 1ull) << (4 * gq + i);
    }
    while (rmask) {
        const int r = __builtin_ctz(rmask);
        rmask &= rmask - 1;
        merge_row(Ld, Li, tau_d, tau_i, cnt, cap, myb_d + r * kB, myb_i + r * kB, r, k, lane);
    }
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int q = q0 + 16 * w + r;
        if (q < nq && lane < k) {
            const size_t row = IVF ? (size_t)T.ids[q] : (size_t)q;
            T.out_idx[row * k + lane] = Li[r];
            T.out_dist[row * k + lane] = Ld[r];
        }
    }
}

// The tile kernels are compiled for these k-step counts; a width runs with
// the smallest that covers its ceil((d + 2) / 4) columns of four (the extra
// columns are zero in every padded row, so they add nothing).
static constexpr int kKsBuckets[] = {2, 3, 4, 5, 6, 8, 10, 13, 16, 20, 26, 33, 40, 48, 56, 65};

int knn_mfma_ksteps(int d) {
    const int need = (d + 2 + 3) / 4;
    for (int b : kKsBuckets)
        if (b >= need) return b;
    return need;  // past the widest: rejected by the launch
}

int knn_mfma_stride(int d) { return 4 * knn_mfma_ksteps(d) + 4; }

// LDS of a launch: two candidate tiles (the first holds the queries at the
// start), ids, and (unless k = 1) the row buffers: 32 slots per row, 64 for
// the ranking.
static size_t tile_lds_bytes(int st, bool buffers, bool rot) {
    const size_t tiles = 2 * (size_t)kTile * st + 2 * kTile;
    const size_t bufs = buffers ? 2 * (size_t)kTile * rot_buffers(rot, (st - 4) / 4) : 0;
    return sizeof(float) * (tiles + bufs);
}

size_t knn_mfma_lds_bytes(int d) { return tile_lds_bytes(knn_mfma_stride(d), true, true); }

template <int KS>
static void launch_ks(const KnnTiles &T, bool ivf, unsigned grid, hipStream_t s) {
    if (ivf)
        k_knn_mfma<true, false, false, KS><<<grid, 256, tile_lds_bytes(T.st, true, false), s>>>(T);
    else if (T.k == 1)
        k_knn_mfma<false, true, false, KS><<<grid, 256, tile_lds_bytes(T.st, false, false), s>>>(T);
    else if (T.start_hint)
        k_knn_mfma<false, false, true, KS><<<grid, 256, tile_lds_bytes(T.st, true, true), s>>>(T);
    else
        k_knn_mfma<false, false, false, KS><<<grid, 256, tile_lds_bytes(T.st, true, false), s>>>(T);
}

hipError_t launch_knn_pack(const float *X, const uint32_t *order, int64_t n,
                           int d, const float *norms, float *Xp, hipStream_t s) {
    const int st = knn_mfma_stride(d);
    const int64_t total = (n + kTile) * st;
    k_knn_pack<<<(unsigned)ceil_div(total, 256), 256, 0, s>>>(X, order, n, d, st, norms, Xp);
    return hipGetLastError();
}

hipError_t launch_knn_tiles(const KnnTiles &T, bool ivf, hipStream_t s) {
    if (T.nq <= 0 || T.d < 1 || T.d > kKnnMfmaMaxDims || T.k < 1 || T.k > kWave ||
        T.nq > kKnnMaxRows || T.nc > kKnnMaxRows || T.st != knn_mfma_stride(T.d) ||
        T.ksteps != knn_mfma_ksteps(T.d))
        return hipErrorInvalidValue;
    const unsigned grid = (unsigned)ceil_div(T.nq, kTile);
    switch (T.ksteps) {
        case 2: launch_ks<2>(T, ivf, grid, s); break;
        case 3: launch_ks<3>(T, ivf, grid, s); break;
        case 4: launch_ks<4>(T, ivf, grid, s); break;
        case 5: launch_ks<5>(T, ivf, grid, s); break;
        case 6: launch_ks<6>(T, ivf, grid, s); break;
        case 8: launch_ks<8>(T, ivf, grid, s); break;
        case 10: launch_ks<10>(T, ivf, grid, s); break;
        case 13: launch_ks<13>(T, ivf, grid, s); break;
        case 16: launch_ks<16>(T, ivf, grid, s); break;
        case 20: launch_ks<20>(T, ivf, grid, s); break;
        case 26: launch_ks<26>(T, ivf, grid, s); break;
        case 33: launch_ks<33>(T, ivf, grid, s); break;
        case 40: launch_ks<40>(T, ivf, grid, s); break;
        case 48: launch_ks<48>(T, ivf, grid, s); break;
        case 56: launch_ks<56>(T, ivf, grid, s); break;
        case 65: launch_ks<65>(T, ivf, grid, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_knn_tiles_k(const KnnTiles &T0, bool ivf, int k, hipStream_t s) {
    if (k < 1) return hipErrorInvalidValue;
    KnnTiles T = T0;
    if (k <= kWave) {
        T.k = k;
        return launch_knn_tiles(T, ivf, s);
    }
    const size_t rows = (size_t)T.nq;
    uint32_t *pi = nullptr;
    float *pd = nullptr;
    hipError_t e = hipMalloc((void **)&pi, sizeof(uint32_t) * 2 * rows * kWave);
    if (e == hipSuccess) e = hipMalloc((void **)&pd, sizeof(float) * 2 * rows * kWave);
    int prev_k = 0;
    for (int p = 0; e == hipSuccess && kWave * p < k; p++) {
        const int kp = k - kWave * p < kWave ? k - kWave * p : kWave;
        uint32_t *oi = pi + (size_t)(p & 1) * rows * kWave;
        float *od = pd + (size_t)(p & 1) * rows * kWave;
        T.k = kp;
        T.out_idx = oi;
        T.out_dist = od;
        if (p > 0) {
            const size_t prev = (size_t)((p - 1) & 1) * rows * kWave;
            T.floor_d = pd + prev + (prev_k - 1);
            T.floor_i = pi + prev + (prev_k - 1);
            T.floor_ld = prev_k;
        }
        e = launch_knn_tiles(T, ivf, s);
        if (e == hipSuccess)
            e = hipMemcpy2DAsync(T0.out_idx + kWave * p, sizeof(uint32_t) * k, oi, sizeof(uint32_t) * kp,
                                 sizeof(uint32_t) * kp, rows, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpy2DAsync(T0.out_dist + kWave * p, sizeof(float) * k, od, sizeof(float) * kp,
                                 sizeof(float) * kp, rows, hipMemcpyDeviceToDevice, s);
        prev_k = kp;
    }
    // The buffers are freed once the passes are done.
    const hipError_t es = hipStreamSynchronize(s);
    if (e == hipSuccess) e = es;
    if (pi) (void)hipFree(pi);
    if (pd) (void)hipFree(pd);
    return e;
}

hipError_t launch_knn_mfma(const float *X, const float *norms, int64_t n,
                           int d, int k, float *Xp, uint32_t *out_idx,
                           float *out_dist, hipStream_t s) {
    if (n <= 0 || d < 1 || d > kKnnMfmaMaxDims || k < 1 || k >= n)
        return hipErrorInvalidValue;
    hipError_t err = launch_knn_pack(X, nullptr, n, d, norms, Xp, s);
    if (err != hipSuccess) return err;
    KnnTiles T{};
    T.Q = T.C = Xp;
    T.qnorm = 1;
    T.nq = T.nc = n;
    T.d = d;
    T.st = knn_mfma_stride(d);
    T.ksteps = knn_mfma_ksteps(d);
    T.self = 1;
    T.out_idx = out_idx;
    T.out_dist = out_dist;
    return launch_knn_tiles_k(T, false, k, s);
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
