// Host side of the IRLBA engine and the C ABI declared in include/bpca.h.
//
// The IRLBA of brisc/pca.pyx:973-1254 runs entirely on the device.  The host
// admits problems (workspace, start vector, set-up kernels) and then replays
// one captured "macro step" graph: every active problem advances one Lanczos
// step per replay, and a problem whose sweep ends gets its small SVD,
// convergence test and restart (or its final PCs) inside the same replay.
// The host reads a done-counter every few replays.  A single solve is a batch
// of one problem, so single, batched and served solves share every kernel.
#include "bpca.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "kernels.h"
#include "rng.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(e == hipErrorOutOfMemory ? BPCA_ERR_OUT_OF_MEMORY : BPCA_ERR_DEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

// The k-means|| arguments (brisc/single_cell.py:12389-12401), or null.
const char *barbar_arg_error(int32_t barbar, int32_t init_iterations, float oversampling) {
    if (barbar != 0 && barbar != 1) return "kmeans_barbar must be 0 or 1";
    if (init_iterations < 1) return "num_init_iterations must be a positive integer";
    if (!(oversampling > 0.0f) || !std::isfinite(oversampling))
        return "oversampling_factor must be a positive number";
    return nullptr;
}

#define HIP_TRY(expr)                                      \
    do {                                                   \
        hipError_t _e = (expr);                            \
        if (_e != hipSuccess) return hip_fail(_e, #expr);  \
    } while (0)

#define BPCA_TRY(expr)                  \
    do {                                \
        int _rc = (expr);               \
        if (_rc != BPCA_OK) return _rc; \
    } while (0)

constexpr size_t kAlign = 256;
inline size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Per-problem workspace: one allocation carved into these buffers.
struct WsLayout {
    size_t U0, U1, V0, V1, R0, R1, u32, P, red, sq, tpart, grp, B, sd, colsum, rem, vs, W, Jv,
        Ub, Sb, vw, vdot, vsq, indptr64, idx, data4, segbase, segs, total;
    size_t idx_bytes;  // bytes per packed column index (2, or 4 when g > 65535)
    int64_t nnz4;  // nnz rounded up to whole quads of 4 elements
    // Gene-sorted copy for X^T u (tsorted): values, row bytes, run-start
    // bits (thb_words 16-bit words) and per-sub-block gene counts (row
    // stride tgl).
    bool tsorted;
    size_t tval, trow, thb, tlens;
    int64_t thb_words;
    int tgl;
};

// Whether a problem with g genes takes the gene-sorted X^T u.
bool uses_sorted(int g, bool enabled) { return enabled && g <= bpca::kMaxSortedGenes; }

// Whether a problem reorthogonalises in one pass over U per Lanczos step
// (kernels.hip, k_reorth_onepass).  Its gene-space X^T U carries rounding
// forward by about sqrt(g / n) per step, so it needs n >= 8 g; passes = 0
// (automatic) also needs n >= kOnePassMinCells, below which U is small and
// the two-pass kernels cost little; passes = 1: whenever n >= 8 g; 2: never.
bool uses_onepass(int64_t n, int g, int passes) {
    if (passes == 2 || n < 8 * (int64_t)g) return false;
    return passes == 1 || n >= bpca::kOnePassMinCells;
}

// Whether a one-pass problem forms X v and X^T u in one read of X
// (kernels.hip, k_spmv_pair): 16-bit indices with v in LDS and g <=
// kMaxPairGenes.  Such a problem needs no gene-sorted copy.
bool uses_pair(int64_t n, int g, int passes, bool enabled) {
    return enabled && g <= bpca::kMaxPairGenes && uses_onepass(n, g, passes);
}

WsLayout ws_layout(int64_t n, int g, int m, bool widen, int64_t nnz, bool sorted, bool pair) {
    const int64_t nb = bpca::ceil_div(n, bpca::kRowBlock);
    const int64_t ng = bpca::ceil_div(nb, bpca::kGroup);
    const int64_t mm = m + (m & 1);
    WsLayout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t at = off;
        off += align_up(std::max<size_t>(bytes, 8));
        return at;
    };
    L.U0 = take(sizeof(double) * n * m);
    L.U1 = take(sizeof(double) * n * m);
    L.V0 = take(sizeof(double) * g * (m + 1));
    L.V1 = take(sizeof(double) * g * (m + 1));
    L.R0 = take(sizeof(double) * g * (m + 1));
    L.R1 = take(sizeof(double) * g * (m + 1));
    // Two step vectors (current and previous), each with a sink slot for
    // masked lanes at [n].
    L.u32 = take(sizeof(float) * 2 * (n + bpca::kWave));
    L.P = take(sizeof(double) * (m + 1) * nb);
    L.red = take(sizeof(double) * (m + 1));
    L.sq = take(sizeof(double) * 2 * nb);
    L.tpart = take(sizeof(float) * nb * 2 * g);
    L.grp = take(sizeof(double) * ng * 2 * g);
    L.B = take(sizeof(double) * m * (m + 1));
    L.sd = take(sizeof(double) * g);
    L.colsum = take(sizeof(double) * g);
    L.rem = take(sizeof(double) * g);
    L.vs = take(sizeof(float) * g);
    L.W = take(sizeof(double) * m * mm);
    L.Jv = take(sizeof(double) * (mm * mm + (int64_t)m * m));
    L.Ub = take(sizeof(double) * m * m);
    L.Sb = take(sizeof(double) * m);
    L.vw = take(sizeof(double) * g);
    L.vdot = take(sizeof(double) * (m + 1));
    L.vsq = take(sizeof(double) * bpca::ceil_div(g, 256));
    // Pair problems: the padded row offsets; others: the widened caller's.
    L.indptr64 = widen || pair ? take(sizeof(int64_t) * (n + 1)) : 0;
    L.nnz4 = pair ? bpca::pad_packed_bound(nnz, n) : (nnz + 3) / 4 * 4;
    L.idx_bytes = g > bpca::kMaxPacked16 ? 4 : 2;
    L.idx = take(L.idx_bytes * L.nnz4);
    L.data4 = take(sizeof(float) * L.nnz4);
    if (pair) {
        // The pair kernel's segment table (bound from the padded size).
        L.segbase = take(sizeof(int64_t) * (bpca::ceil_div(n, bpca::kSubRows) + 1));
        L.segs = take(sizeof(uint16_t) * bpca::pair_segments_bound(n, L.nnz4));
    }
    L.tsorted = sorted;
    if (sorted) {
        // Lane chunks of kTChunk entries are loaded whole up to the chunk
        // holding the last entry.
        const int64_t tlen = (nnz + 2 * bpca::kTChunk + 63) / 64 * 64;
        const int64_t nsub = bpca::ceil_div(n, bpca::kSubRows);
        L.tgl = (g + 7) / 8 * 8;
        L.thb_words = (tlen / 16 + 1) / 2 * 2;
        L.tval = take(sizeof(float) * tlen);
        L.trow = take(tlen);
        L.thb = take(sizeof(uint16_t) * L.thb_words);
        L.tlens = take(sizeof(uint16_t) * (size_t)nsub * L.tgl);
    }
    L.total = off;
    return L;
}

// The start vector of brisc/pca.pyx:801-806: g FP32 normal draws of the
// reference's generator (rng.h), normalised here in FP64.  `draws` (may be
// null) receives the raw FP32 draws.
std::vector<double> start_vector(uint64_t seed, int g, float *draws = nullptr) {
    bpca::Pcg32 rng(seed);
    std::vector<double> v(g);
    double nrm = 0.0;
    for (int i = 0; i < g; i++) {
        const float x = rng.normal();
        if (draws) draws[i] = x;
        v[i] = (double)x;
        nrm += v[i] * v[i];
    }
    nrm = std::sqrt(nrm);
    for (int i = 0; i < g; i++) v[i] /= nrm;
    return v;
}

int check_shape(int64_t n, int64_t g, int64_t k, int64_t m, float tol,
                int64_t maxit) {
    if (n < 2) return fail(BPCA_ERR_INVALID_ARGUMENT, "there is only one cell, so principal components cannot be calculated");
    if (g < 2) return fail(BPCA_ERR_INVALID_ARGUMENT, "there is only one gene, so principal components cannot be calculated");
    if (k < 1) return fail(BPCA_ERR_INVALID_ARGUMENT, "k must be a positive integer");
    if (m < k) return fail(BPCA_ERR_INVALID_ARGUMENT, "subspace_size must be >= k");
    if (m > n || m > g) return fail(BPCA_ERR_INVALID_ARGUMENT, "subspace_size must be <= min(num_cells, num_genes)");
    if (!(tol > 0.0f)) return fail(BPCA_ERR_INVALID_ARGUMENT, "tolerance must be positive");
    if (maxit < 1) return fail(BPCA_ERR_INVALID_ARGUMENT, "max_iterations must be a positive integer");
    if (g > INT32_MAX / 2) return fail(BPCA_ERR_INVALID_ARGUMENT, "num_genes too large");
    if (m > bpca::kMaxSubspace) return fail(BPCA_ERR_INVALID_ARGUMENT, "subspace_size above 1024 is not supported by the device path");
    if (bpca::ceil_div(n, bpca::kRowBlock) > INT32_MAX) return fail(BPCA_ERR_INVALID_ARGUMENT, "num_cells too large");
    return BPCA_OK;
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// Scratch of the allocation-free solve (bpca_irlba_device_noalloc): a
// file-scope device array, one copy per device, held until the process
// ends.  Host code never uses the array's name as an address: the device
// address of each device's copy is looked up once, when the first context on
// that device is created, and kept in g_scratch_ptr.
constexpr size_t kScratchBytes = (size_t)256 << 20;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_lock;
void *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_use[kMaxDevices];  // one allocation-free solve per device at a time

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Scratch bytes of an allocation-free solve: descriptor, block and group
// lists, counters, workspace.
uint64_t noalloc_bytes(int64_t n, int g, int m, int64_t nnz, bool widen, bool sorted, bool pair) {
    const int64_t nb = bpca::ceil_div(n, bpca::kRowBlock), ng = bpca::ceil_div(nb, bpca::kGroup);
    return align_up(sizeof(bpca::Prob)) + align_up(sizeof(int2) * nb) + align_up(sizeof(int2) * ng) +
           kAlign + ws_layout(n, g, m, widen, nnz, sorted, pair).total;
}

// With `device` current: record the device address of its scratch copy.
int scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "device index above the library's limit of 64");
    std::lock_guard<std::mutex> guard(g_scratch_lock);
    if (g_scratch_ptr[device]) return BPCA_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(BPCA_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[device] = p;
    return BPCA_OK;
}

}  // namespace

// ------------------------------------------------------------ context

struct bpca_context {
    int device = 0;
    hipStream_t stream = nullptr;       // the engine's own stream
    hipStream_t side = nullptr;         // small SVDs, forked from `stream`
    hipEvent_t fork_event = nullptr, join_event = nullptr;
    hipStream_t user_stream = nullptr;  // caller's stream to order against
    hipEvent_t order_event = nullptr;
    uint64_t budget = 0;                // 0 = unlimited
    uint64_t reserved = 0;              // bytes charged by bpca_context_reserve
    size_t max_lds = 65536;
    // Incremented by bpca_context_cancel; a call that advances solves
    // stops when it differs from the value it saw on entry.
    std::atomic<uint64_t> cancel_gen{0};
    std::recursive_mutex lock;
    // Device memory pool of whole workspaces, reused across problems.
    struct Block { void *p; size_t bytes; };
    std::vector<Block> free_blocks;
    std::map<void *, size_t> live_blocks;
    uint64_t live_bytes = 0, cached_bytes = 0;
    // Kernel timing: with profiling on, macro steps are launched directly
    // with events around each kernel class on the engine's stream.
    bool profiling = false;
    uint32_t prof_classes = ~0u;   // classes timed while profiling (bit per class)
    std::vector<hipEvent_t> events;
    size_t next_event = 0;
    struct Pending { int kind, a, b; };
    std::vector<Pending> pending;
    int open_event[BPCA_NUM_KERNEL_CLASSES] = {};
    double k_ms[BPCA_NUM_KERNEL_CLASSES] = {};
    int64_t k_count[BPCA_NUM_KERNEL_CLASSES] = {};
    bool use_graph = true;
    int spmv_group = 4, spmvt_group = 4;   // autotuned load-group sizes
    int spmv_wg = 0, spmvt_wg = 0;         // autotuned workgroups-per-CU caps
    int64_t lds_per_cu = 0;                // LDS bytes of one compute unit
    int num_cus = 0;                       // compute units of the device
    bool sorted_xtu = true;                // gene-sorted X^T u for g <= kMaxSortedGenes
    bool pair_products = true;             // X v and X^T u in one pass (one-pass problems, g <= kMaxPairGenes)
    int reorth_passes = 0;                 // 0 automatic, 1 one pass (n >= 8 g), 2 two passes
};

namespace {

// Whether a problem of this context takes the gene-sorted copy for X^T u:
// enabled, g <= kMaxSortedGenes, and X^T u not formed in the pass of X v.
// Whether a problem of this context forms X v and X^T u in one pass (its
// padded packed copy must stay below 2^31 quads).
bool problem_pair(const bpca_context *ctx, int64_t n, int g, int64_t nnz) {
    return uses_pair(n, g, ctx->reorth_passes, ctx->pair_products) &&
           bpca::pad_packed_bound(nnz, n) / 4 < (int64_t)INT32_MAX;
}

bool layout_sorted(const bpca_context *ctx, int64_t n, int g, int64_t nnz) {
    return uses_sorted(g, ctx->sorted_xtu) && !problem_pair(ctx, n, g, nnz);
}

int pool_alloc(bpca_context *ctx, size_t bytes, void **out) {
    bytes = align_up(bytes);
    int best = -1;  // best fit among cached blocks at most twice the request
    for (size_t i = 0; i < ctx->free_blocks.size(); i++) {
        const auto &b = ctx->free_blocks[i];
        if (b.bytes >= bytes && b.bytes <= 2 * bytes &&
            (best < 0 || b.bytes < ctx->free_blocks[best].bytes))
            best = (int)i;
    }
    if (best >= 0) {
        auto b = ctx->free_blocks[best];
        ctx->free_blocks.erase(ctx->free_blocks.begin() + best);
        ctx->cached_bytes -= b.bytes;
        ctx->live_blocks[b.p] = b.bytes;
        ctx->live_bytes += b.bytes;
        *out = b.p;
        return BPCA_OK;
    }
    if (ctx->budget && ctx->live_bytes + ctx->reserved + bytes > ctx->budget)
        return fail(BPCA_ERR_BUDGET, "request does not fit the context's HBM budget now");
    while (ctx->budget && !ctx->free_blocks.empty() &&
           ctx->live_bytes + ctx->reserved + ctx->cached_bytes + bytes > ctx->budget) {
        auto b = ctx->free_blocks.back();
        ctx->free_blocks.pop_back();
        ctx->cached_bytes -= b.bytes;
        HIP_TRY(hipFree(b.p));
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipErrorOutOfMemory && !ctx->free_blocks.empty()) {
        (void)hipGetLastError();
        for (auto &b : ctx->free_blocks) (void)hipFree(b.p);
        ctx->free_blocks.clear();
        ctx->cached_bytes = 0;
        e = hipMalloc(&p, bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return hip_fail(e, "hipMalloc(workspace)");
    }
    ctx->live_blocks[p] = bytes;
    ctx->live_bytes += bytes;
    *out = p;
    return BPCA_OK;
}

void pool_free(bpca_context *ctx, void *p) {
    auto it = ctx->live_blocks.find(p);
    if (it == ctx->live_blocks.end()) return;
    ctx->free_blocks.push_back({p, it->second});
    ctx->cached_bytes += it->second;
    ctx->live_bytes -= it->second;
    ctx->live_blocks.erase(it);
}

int prof_record(bpca_context *ctx) {
    if (ctx->next_event == ctx->events.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        ctx->events.push_back(e);
    }
    int i = (int)ctx->next_event++;
    if (hipEventRecord(ctx->events[i], ctx->stream) != hipSuccess) return -1;
    return i;
}

void prof_mark(void *user, int kind, int end) {
    bpca_context *ctx = (bpca_context *)user;
    if (!((ctx->prof_classes >> kind) & 1u)) return;
    if (!end) {
        ctx->open_event[kind] = prof_record(ctx);
    } else {
        int a = ctx->open_event[kind];
        int b = prof_record(ctx);
        if (a >= 0 && b >= 0) ctx->pending.push_back({kind, a, b});
    }
}

void prof_flush(bpca_context *ctx) {
    for (const auto &p : ctx->pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, ctx->events[p.a], ctx->events[p.b]) == hipSuccess) {
            ctx->k_ms[p.kind] += ms;
            ctx->k_count[p.kind] += 1;
        }
    }
    ctx->pending.clear();
    ctx->next_event = 0;
}

void begin_call(bpca_context *ctx) {
    // Order the engine's stream after work already queued on the caller's.
    if (ctx->order_event) {
        (void)hipEventRecord(ctx->order_event, ctx->user_stream);
        (void)hipStreamWaitEvent(ctx->stream, ctx->order_event, 0);
    }
}

// The descriptor of problem `pb` whose workspace (layout L) starts at base.
bpca::Prob make_prob(const bpca_problem *pb, const WsLayout &L, char *base,
                     bool widen, int id, int passes, bool pair) {
    // pair: the problem forms X v and X^T u in one pass (problem_pair), and L
    // holds its padded packed copy.
    const int64_t n = (int64_t)pb->num_cells;
    bpca::Prob d{};
    d.data = (const float *)(base + L.data4);
    d.idx = (const void *)(base + L.idx);
    d.indptr = widen || pair ? (const int64_t *)(base + L.indptr64) : (const int64_t *)pb->indptr;
    d.PCs = pb->PCs;
    d.n = n;
    d.nb = bpca::ceil_div(n, bpca::kRowBlock);
    d.ngrp = bpca::ceil_div(d.nb, bpca::kGroup);
    d.g = (int)pb->num_genes;
    d.mode = d.g > bpca::kMaxPacked16 ? bpca::kModeWide
             : d.g > bpca::kMaxLdsV   ? bpca::kModeGlobalV
                                      : bpca::kModeLds;
    d.k = (int)pb->k;
    d.m = (int)pb->subspace_size;
    d.maxit = (int)pb->max_iterations;
    d.tol = pb->tolerance;
    d.id = id;
    d.U[0] = (double *)(base + L.U0);
    d.U[1] = (double *)(base + L.U1);
    d.V[0] = (double *)(base + L.V0);
    d.V[1] = (double *)(base + L.V1);
    d.R[0] = (double *)(base + L.R0);
    d.R[1] = (double *)(base + L.R1);
    d.u32 = (float *)(base + L.u32);
    d.P = (double *)(base + L.P);
    d.red = (double *)(base + L.red);
    d.sq = (double *)(base + L.sq);
    d.tpart = (float *)(base + L.tpart);
    d.grp = (double *)(base + L.grp);
    d.B = (double *)(base + L.B);
    d.sd = (double *)(base + L.sd);
    d.colsum = (double *)(base + L.colsum);
    d.rem = (double *)(base + L.rem);
    d.vs = (float *)(base + L.vs);
    d.W = (double *)(base + L.W);
    d.Jv = (double *)(base + L.Jv);
    d.Ub = (double *)(base + L.Ub);
    d.Sb = (double *)(base + L.Sb);
    d.vw = (double *)(base + L.vw);
    d.vdot = (double *)(base + L.vdot);
    d.vsq = (double *)(base + L.vsq);
    d.onepass = uses_onepass(n, d.g, passes) ? 1 : 0;
    d.pair = pair ? 1 : 0;
    d.idx_shift = pair ? 2 : 0;
    if (pair) {
        d.segs = (const uint16_t *)(base + L.segs);
        d.segbase = (const int64_t *)(base + L.segbase);
    }
    d.tsorted = L.tsorted ? 1 : 0;
    if (L.tsorted) {
        d.tval = (const float *)(base + L.tval);
        d.trow = (const uint8_t *)(base + L.trow);
        d.thb = (const uint16_t *)(base + L.thb);
        d.tlens = (const uint16_t *)(base + L.tlens);
        d.tgl = L.tgl;
    }
    d.phase = bpca::PH_SETUP;
    return d;
}

struct Slot {
    bool used = false;
    bpca::Prob desc{};   // host copy of the descriptor as admitted
    void *ws = nullptr;
    std::chrono::steady_clock::time_point t_admit;
};

}  // namespace

struct bpca_engine {
    bpca_context *ctx = nullptr;
    std::vector<Slot> slots;
    bpca::Prob *d_probs = nullptr;
    int capacity = 0;
    int2 *d_blocks = nullptr, *d_groups = nullptr;
    int64_t blocks_cap = 0, groups_cap = 0;
    int *d_done = nullptr;
    int *h_done = nullptr;   // pinned
    bpca::Tables T{};
    bool dirty = true, setup_pending = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int done_seen = 0;       // finished problems already collected
    int64_t replays = 0;
};

namespace {

void engine_drop_graph(bpca_engine *E) {
    if (E->exec) (void)hipGraphExecDestroy(E->exec);
    if (E->graph) (void)hipGraphDestroy(E->graph);
    E->exec = nullptr;
    E->graph = nullptr;
}

int engine_rebuild_tables(bpca_engine *E) {
    bpca_context *ctx = E->ctx;
    std::vector<int2> blocks, groups;
    int max_g = 2, max_m = 1, modes = 0, vmodes = 0, tmodes = 0, rmodes = 0;
    for (int s = 0; s < (int)E->slots.size(); s++) {
        const Slot &sl = E->slots[s];
        if (!sl.used) continue;
        modes |= 1 << sl.desc.mode;
        if (!sl.desc.pair) vmodes |= 1 << sl.desc.mode;
        tmodes |= sl.desc.pair ? bpca::kPairBit : sl.desc.tsorted ? bpca::kSortedBit : 1 << sl.desc.mode;
        rmodes |= sl.desc.onepass ? bpca::kReorthOneBit : bpca::kReorthTwoBit;
        for (int64_t b = 0; b < sl.desc.nb; b++) blocks.push_back(make_int2(s, (int)b));
        for (int64_t G = 0; G < sl.desc.ngrp; G++) groups.push_back(make_int2(s, (int)G));
        max_g = std::max(max_g, sl.desc.g);
        max_m = std::max(max_m, sl.desc.m);
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if ((int64_t)blocks.size() > E->blocks_cap) {
        if (E->d_blocks) HIP_TRY(hipFree(E->d_blocks));
        E->d_blocks = nullptr;
        E->blocks_cap = std::max<int64_t>(blocks.size(), 2 * E->blocks_cap);
        HIP_TRY(hipMalloc((void **)&E->d_blocks, sizeof(int2) * E->blocks_cap));
    }
    if ((int64_t)groups.size() > E->groups_cap) {
        if (E->d_groups) HIP_TRY(hipFree(E->d_groups));
        E->d_groups = nullptr;
        E->groups_cap = std::max<int64_t>(groups.size(), 2 * E->groups_cap);
        HIP_TRY(hipMalloc((void **)&E->d_groups, sizeof(int2) * E->groups_cap));
    }
    if (!blocks.empty())
        HIP_TRY(hipMemcpyAsync(E->d_blocks, blocks.data(), sizeof(int2) * blocks.size(), hipMemcpyHostToDevice, ctx->stream));
    if (!groups.empty())
        HIP_TRY(hipMemcpyAsync(E->d_groups, groups.data(), sizeof(int2) * groups.size(), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    E->T.probs = E->d_probs;
    E->T.num_probs = (int)E->slots.size();
    E->T.blocks = E->d_blocks;
    E->T.num_blocks = (int64_t)blocks.size();
    E->T.groups = E->d_groups;
    E->T.num_groups = (int64_t)groups.size();
    E->T.max_g = max_g;
    E->T.max_m = max_m;
    E->T.modes = modes;
    E->T.vmodes = vmodes;
    E->T.tmodes = tmodes;
    E->T.rmodes = rmodes;
    E->T.max_lds = (int64_t)ctx->max_lds;
    E->T.spmv_group = ctx->spmv_group;
    E->T.spmvt_group = ctx->spmvt_group;
    E->T.spmv_wg = ctx->spmv_wg;
    E->T.spmvt_wg = ctx->spmvt_wg;
    E->T.lds_per_cu = ctx->lds_per_cu;
    E->T.num_cus = ctx->num_cus;
    engine_drop_graph(E);
    E->dirty = false;
    return BPCA_OK;
}

int engine_launch_steps(bpca_engine *E, int count) {
    bpca_context *ctx = E->ctx;
    hipStream_t s = ctx->stream;
    if (E->T.num_blocks == 0) return BPCA_OK;
    if (ctx->profiling || !ctx->use_graph) {
        // Timed per kernel class: everything on the engine's stream.
        bpca::LaunchHook hook{prof_mark, ctx};
        for (int i = 0; i < count; i++)
            HIP_TRY(bpca::launch_macro_step(E->T, E->d_done, s, ctx->profiling ? &hook : nullptr));
        return BPCA_OK;
    }
    if (!E->exec) {
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        // Fork the small SVDs only when other problems can overlap them.
        const bool fork = E->T.num_probs > 1;
        hipError_t e = bpca::launch_macro_step(E->T, E->d_done, s, nullptr,
                                               fork ? ctx->side : nullptr,
                                               ctx->fork_event, ctx->join_event);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(s, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return hip_fail(e != hipSuccess ? e : e2, "macro-step graph capture");
        }
        E->graph = g;
        HIP_TRY(hipGraphInstantiate(&E->exec, E->graph, nullptr, nullptr, 0));
    }
    for (int i = 0; i < count; i++) HIP_TRY(hipGraphLaunch(E->exec, s));
    E->replays += count;
    return BPCA_OK;
}

int engine_grow(bpca_engine *E, int need) {
    if (need <= E->capacity) return BPCA_OK;
    const int cap = std::max(need, std::max(4, 2 * E->capacity));
    bpca::Prob *np = nullptr;
    HIP_TRY(hipMalloc((void **)&np, sizeof(bpca::Prob) * cap));
    if (E->d_probs) {
        HIP_TRY(hipMemcpyAsync(np, E->d_probs, sizeof(bpca::Prob) * E->capacity, hipMemcpyDeviceToDevice, E->ctx->stream));
        HIP_TRY(hipStreamSynchronize(E->ctx->stream));
        HIP_TRY(hipFree(E->d_probs));
    }
    E->d_probs = np;
    E->capacity = cap;
    E->dirty = true;
    return BPCA_OK;
}

}  // namespace

extern "C" {

// ------------------------------------------------------------ context ABI

const char *bpca_last_error(void) { return g_last_error.c_str(); }

int bpca_version(void) { return 1 * 10000 + 1 * 100 + 0; }

int bpca_context_info(bpca_context *ctx, int64_t *max_lds_per_workgroup,
                      int32_t *num_cus, int32_t *uses_graph) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (max_lds_per_workgroup) *max_lds_per_workgroup = (int64_t)ctx->max_lds;
    if (num_cus) {
        int cus = 0;
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
        *num_cus = cus;
    }
    if (uses_graph) *uses_graph = ctx->use_graph ? 1 : 0;
    return BPCA_OK;
}

int bpca_test_gemm(bpca_context *ctx, int64_t n, int32_t m, int32_t kout,
                   const double *A, const double *C, double *out) {
    if (!ctx || !A || !C || !out || n < 1 || m < 1 || kout < 1)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "bad argument");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    hipStream_t s = ctx->stream;
    double *dA = nullptr, *dC = nullptr, *dO = nullptr;
    hipError_t e = hipMalloc((void **)&dA, sizeof(double) * n * m);
    if (e == hipSuccess) e = hipMalloc((void **)&dC, sizeof(double) * m * kout);
    if (e == hipSuccess) e = hipMalloc((void **)&dO, sizeof(double) * n * kout);
    if (e == hipSuccess) e = hipMemcpyAsync(dA, A, sizeof(double) * n * m, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dC, C, sizeof(double) * m * kout, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = bpca::launch_gemm_mfma_test(dA, n, m, dC, kout, dO, s);
    if (e == hipSuccess) e = hipMemcpyAsync(out, dO, sizeof(double) * n * kout, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (dA) (void)hipFree(dA);
    if (dC) (void)hipFree(dC);
    if (dO) (void)hipFree(dO);
    if (e != hipSuccess) return hip_fail(e, "test gemm");
    return BPCA_OK;
}

int bpca_test_gemm_device(bpca_context *ctx, int64_t n, int32_t m,
                          int32_t kout, const double *d_A, const double *d_C,
                          double *d_out) {
    if (!ctx || !d_A || !d_C || !d_out || n < 1 || m < 1 || kout < 1 || m > 128 || kout > m)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "bad argument");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    begin_call(ctx);
    HIP_TRY(bpca::launch_gemm_mfma_test(d_A, n, m, d_C, kout, d_out, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return BPCA_OK;
}

int bpca_test_fp64_rate(bpca_context *ctx, int32_t which, int32_t reps,
                        double *tflops) {
    if (!ctx || !tflops || reps < 1 || (which != 0 && which != 1))
        return fail(BPCA_ERR_INVALID_ARGUMENT, "bad argument");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const int blocks = cus * 8;   // eight workgroups of four waves per compute unit
    const int iters = which == 0 ? 2000 : 16000;
    double *out = nullptr;
    HIP_TRY(hipMalloc((void **)&out, sizeof(double) * (size_t)blocks * 256));
    hipEvent_t a = nullptr, z = nullptr;
    double flops = 0.0, best = 0.0;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&z);
    // One untimed launch, then the best of `reps` timed ones.
    if (e == hipSuccess) e = bpca::launch_fp64_rate(which, blocks, iters, out, &flops, ctx->stream);
    for (int r = 0; r < reps && e == hipSuccess; r++) {
        e = hipEventRecord(a, ctx->stream);
        if (e == hipSuccess) e = bpca::launch_fp64_rate(which, blocks, iters, out, &flops, ctx->stream);
        if (e == hipSuccess) e = hipEventRecord(z, ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(z);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, z);
        if (e == hipSuccess && ms > 0.0f) best = std::max(best, flops / (ms * 1e-3) / 1e12);
    }
    if (a) (void)hipEventDestroy(a);
    if (z) (void)hipEventDestroy(z);
    (void)hipFree(out);
    if (e != hipSuccess) return hip_fail(e, "fp64 rate");
    *tflops = best;
    return BPCA_OK;
}

int bpca_context_create(int device, uint64_t hbm_budget_bytes,
                        bpca_context **out) {
    if (!out) return fail(BPCA_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "no such device");
    bpca_context *ctx = new (std::nothrow) bpca_context();
    if (!ctx) return fail(BPCA_ERR_OUT_OF_MEMORY, "context allocation failed");
    ctx->device = device;
    ctx->budget = hbm_budget_bytes;
    DeviceGuard dg(device);
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->order_event, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->fork_event, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->join_event, hipEventDisableTiming);
    if (e != hipSuccess) {
        if (ctx->order_event) (void)hipEventDestroy(ctx->order_event);
        if (ctx->fork_event) (void)hipEventDestroy(ctx->fork_event);
        if (ctx->side) (void)hipStreamDestroy(ctx->side);
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return hip_fail(e, "context stream");
    }
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
        ctx->max_lds = (size_t)lds;
    int lds_cu = 0;
    if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) == hipSuccess && lds_cu > 0)
        ctx->lds_per_cu = lds_cu;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        ctx->num_cus = cus;
    const char *env = std::getenv("BPCA_NO_GRAPH");
    ctx->use_graph = !(env && env[0] == '1');
    const int rc = scratch_init(device);
    if (rc != BPCA_OK) {
        (void)hipEventDestroy(ctx->order_event);
        (void)hipEventDestroy(ctx->fork_event);
        (void)hipEventDestroy(ctx->join_event);
        (void)hipStreamDestroy(ctx->side);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return rc;
    }
    *out = ctx;
    return BPCA_OK;
}

int bpca_context_destroy(bpca_context *ctx) {
    if (!ctx) return BPCA_OK;
    {
        DeviceGuard dg(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        for (auto &b : ctx->free_blocks) (void)hipFree(b.p);
        for (auto &kv : ctx->live_blocks) (void)hipFree(kv.first);
        for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
        (void)hipStreamSynchronize(ctx->side);
        if (ctx->order_event) (void)hipEventDestroy(ctx->order_event);
        if (ctx->fork_event) (void)hipEventDestroy(ctx->fork_event);
        if (ctx->join_event) (void)hipEventDestroy(ctx->join_event);
        (void)hipStreamDestroy(ctx->side);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
    return BPCA_OK;
}

int bpca_context_set_stream(bpca_context *ctx, void *hip_stream) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    ctx->user_stream = (hipStream_t)hip_stream;
    return BPCA_OK;
}

int bpca_context_set_kernel_config(bpca_context *ctx, int32_t spmv_group,
                                   int32_t spmvt_group) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    auto ok = [](int v) { return v == 2 || v == 4 || v == 8; };
    if (!ok(spmv_group) || !ok(spmvt_group))
        return fail(BPCA_ERR_INVALID_ARGUMENT, "load-group sizes must be 2, 4 or 8");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    ctx->spmv_group = spmv_group;
    ctx->spmvt_group = spmvt_group;
    return BPCA_OK;
}

int bpca_context_get_kernel_config(bpca_context *ctx, int32_t *spmv_group,
                                   int32_t *spmvt_group) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (spmv_group) *spmv_group = ctx->spmv_group;
    if (spmvt_group) *spmvt_group = ctx->spmvt_group;
    return BPCA_OK;
}

int bpca_context_set_occupancy(bpca_context *ctx, int32_t spmv_wg,
                               int32_t spmvt_wg) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    auto ok = [](int v) { return v >= 0 && v <= 16; };
    if (!ok(spmv_wg) || !ok(spmvt_wg))
        return fail(BPCA_ERR_INVALID_ARGUMENT, "workgroups per CU must be 0 (no cap) to 16");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    ctx->spmv_wg = spmv_wg;
    ctx->spmvt_wg = spmvt_wg;
    return BPCA_OK;
}

int bpca_context_set_sorted_xtu(bpca_context *ctx, int32_t enable) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    ctx->sorted_xtu = enable != 0;
    return BPCA_OK;
}

int bpca_context_get_sorted_xtu(bpca_context *ctx, int32_t *enable) {
    if (!ctx || !enable) return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL argument");
    *enable = ctx->sorted_xtu ? 1 : 0;
    return BPCA_OK;
}

int bpca_context_set_pair_products(bpca_context *ctx, int32_t enable) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    ctx->pair_products = enable != 0;
    return BPCA_OK;
}

int bpca_context_get_pair_products(bpca_context *ctx, int32_t *enable) {
    if (!ctx || !enable) return fail(BPCA_ERR_INVALID_ARGUMENT, "bad argument");
    *enable = ctx->pair_products ? 1 : 0;
    return BPCA_OK;
}

int bpca_context_set_reorth_passes(bpca_context *ctx, int32_t passes) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (passes < 0 || passes > 2)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "passes must be 0 (automatic), 1 or 2");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    ctx->reorth_passes = passes;
    return BPCA_OK;
}

int bpca_context_get_reorth_passes(bpca_context *ctx, int32_t *passes) {
    if (!ctx || !passes) return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL argument");
    *passes = ctx->reorth_passes;
    return BPCA_OK;
}

int bpca_context_get_occupancy(bpca_context *ctx, int32_t *spmv_wg,
                               int32_t *spmvt_wg) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (spmv_wg) *spmv_wg = ctx->spmv_wg;
    if (spmvt_wg) *spmvt_wg = ctx->spmvt_wg;
    return BPCA_OK;
}

int bpca_context_cancel(bpca_context *ctx) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    ctx->cancel_gen.fetch_add(1);
    return BPCA_OK;
}

int bpca_context_reserve(bpca_context *ctx, uint64_t bytes, int force) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    if (!force && ctx->budget && ctx->live_bytes + ctx->reserved + bytes > ctx->budget)
        return fail(BPCA_ERR_BUDGET, "reservation does not fit the context's HBM budget now");
    ctx->reserved += bytes;
    // Cached workspaces give way to what is charged.
    DeviceGuard dg(ctx->device);
    while (ctx->budget && !ctx->free_blocks.empty() &&
           ctx->live_bytes + ctx->reserved + ctx->cached_bytes > ctx->budget) {
        auto b = ctx->free_blocks.back();
        ctx->free_blocks.pop_back();
        ctx->cached_bytes -= b.bytes;
        HIP_TRY(hipFree(b.p));
    }
    return BPCA_OK;
}

int bpca_context_release(bpca_context *ctx, uint64_t bytes) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    if (bytes > ctx->reserved) return fail(BPCA_ERR_INVALID_ARGUMENT, "release exceeds the reserved bytes");
    ctx->reserved -= bytes;
    return BPCA_OK;
}

int bpca_context_budget(bpca_context *ctx, uint64_t *budget, uint64_t *charged,
                        uint64_t *cached) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    if (budget) *budget = ctx->budget;
    if (charged) *charged = ctx->live_bytes + ctx->reserved;
    if (cached) *cached = ctx->cached_bytes;
    return BPCA_OK;
}

int bpca_context_set_profiling(bpca_context *ctx, int enable) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    ctx->profiling = enable != 0;
    return BPCA_OK;
}

int bpca_context_set_profiled_classes(bpca_context *ctx, uint32_t mask) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    ctx->prof_classes = mask;
    return BPCA_OK;
}

int bpca_context_kernel_times(bpca_context *ctx, double *total_ms,
                              int64_t *launches) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    for (int i = 0; i < BPCA_NUM_KERNEL_CLASSES; i++) {
        if (total_ms) total_ms[i] = ctx->k_ms[i];
        if (launches) launches[i] = ctx->k_count[i];
    }
    return BPCA_OK;
}

int bpca_context_reset_kernel_times(bpca_context *ctx) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    for (int i = 0; i < BPCA_NUM_KERNEL_CLASSES; i++) {
        ctx->k_ms[i] = 0.0;
        ctx->k_count[i] = 0;
    }
    return BPCA_OK;
}

uint64_t bpca_workspace_bytes(uint64_t num_cells, uint32_t num_genes,
                              uint32_t k, uint32_t subspace_size) {
    (void)k;
    // Without the CSR's nnz this counts the packed copies at 7% density.
    const int64_t nnz = (int64_t)((double)num_cells * num_genes * 0.07);
    // Default context settings: one-pass problems with few genes take the
    // pair products, other problems with few genes the gene-sorted copy.
    const bool pair = uses_pair((int64_t)num_cells, (int)num_genes, 0, true);
    return ws_layout((int64_t)num_cells, (int)num_genes, (int)subspace_size,
                     true, nnz, uses_sorted((int)num_genes, true) && !pair, pair).total;
}

// ------------------------------------------------------------ engine ABI

int bpca_engine_create(bpca_context *ctx, bpca_engine **out) {
    if (!ctx || !out) return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    auto *E = new (std::nothrow) bpca_engine();
    if (!E) return fail(BPCA_ERR_OUT_OF_MEMORY, "engine allocation failed");
    E->ctx = ctx;
    hipError_t e = hipMalloc((void **)&E->d_done, sizeof(int));
    if (e == hipSuccess) e = hipHostMalloc((void **)&E->h_done, sizeof(int), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMemsetAsync(E->d_done, 0, sizeof(int), ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        if (E->d_done) (void)hipFree(E->d_done);
        if (E->h_done) (void)hipHostFree(E->h_done);
        delete E;
        return hip_fail(e, "engine counters");
    }
    *E->h_done = 0;
    *out = E;
    return BPCA_OK;
}

int bpca_engine_destroy(bpca_engine *E) {
    if (!E) return BPCA_OK;
    bpca_context *ctx = E->ctx;
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    engine_drop_graph(E);
    for (auto &sl : E->slots)
        if (sl.used && sl.ws) pool_free(ctx, sl.ws);
    if (E->d_probs) (void)hipFree(E->d_probs);
    if (E->d_blocks) (void)hipFree(E->d_blocks);
    if (E->d_groups) (void)hipFree(E->d_groups);
    if (E->d_done) (void)hipFree(E->d_done);
    if (E->h_done) (void)hipHostFree(E->h_done);
    delete E;
    return BPCA_OK;
}

int bpca_engine_add(bpca_engine *E, const bpca_problem *pb, int32_t *slot_out) {
    if (!E || !pb) return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL argument");
    bpca_context *ctx = E->ctx;
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    if (!pb->data || !pb->indices || !pb->indptr || !pb->PCs)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL device pointer in problem");
    if (pb->indptr_width != 4 && pb->indptr_width != 8)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "indptr_width must be 4 or 8");
    const int64_t n = (int64_t)pb->num_cells;
    const int g = (int)pb->num_genes, k = (int)pb->k, m = (int)pb->subspace_size;
    BPCA_TRY(check_shape(n, g, k, m, pb->tolerance, pb->max_iterations));
    if (bpca::lds_bytes_needed(g, m) > ctx->max_lds)
        return fail(BPCA_ERR_INVALID_ARGUMENT,
                    "num_genes / subspace_size need " + std::to_string(bpca::lds_bytes_needed(g, m)) +
                        " bytes of LDS per workgroup; the device has " + std::to_string(ctx->max_lds));
    hipStream_t s = ctx->stream;
    begin_call(ctx);
    int64_t nnz = 0;
    if (pb->indptr_width == 8) {
        HIP_TRY(hipMemcpyAsync(&nnz, (const int64_t *)pb->indptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    } else {
        int32_t nnz32 = 0;
        HIP_TRY(hipMemcpyAsync(&nnz32, (const int32_t *)pb->indptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        nnz = nnz32;
    }
    HIP_TRY(hipStreamSynchronize(s));
    if (nnz < 0) return fail(BPCA_ERR_NONCANONICAL, "indptr[num_cells] is negative");
    const bool widen = pb->indptr_width == 4;
    const bool pair = problem_pair(ctx, n, g, nnz);
    const WsLayout L = ws_layout(n, g, m, widen, nnz, layout_sorted(ctx, n, g, nnz), pair);
    void *ws = nullptr;
    BPCA_TRY(pool_alloc(ctx, L.total, &ws));
    {
        // Canonical-CSR check fused with packing the column indices.
        int *flag = nullptr;
        hipError_t e = hipMalloc((void **)&flag, sizeof(int));
        int h_flag = 0;
        if (e == hipSuccess) e = hipMemsetAsync(flag, 0, sizeof(int), s);
        // Zero the quad padding past nnz in both packed copies (pair
        // problems: the padded copy is written after the check).
        if (e == hipSuccess && L.nnz4 > nnz && !pair)
            e = hipMemsetAsync((char *)ws + L.idx + L.idx_bytes * nnz, 0,
                               L.idx_bytes * (L.nnz4 - nnz), s);
        if (e == hipSuccess && L.nnz4 > nnz && !pair)
            e = hipMemsetAsync((char *)ws + L.data4 + sizeof(float) * nnz, 0,
                               sizeof(float) * (L.nnz4 - nnz), s);
        if (e == hipSuccess)
            e = bpca::launch_validate_pack(pb->data, pb->indices, pb->indptr, pb->indptr_width, n, nnz, g,
                                           (float *)((char *)ws + L.data4),
                                           (char *)ws + L.idx, flag, s, pair);
        if (e == hipSuccess) e = hipMemcpyAsync(&h_flag, flag, sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (flag) (void)hipFree(flag);
        if (e != hipSuccess) {
            pool_free(ctx, ws);
            return hip_fail(e, "CSR validation");
        }
        if (h_flag) {
            pool_free(ctx, ws);
            return fail(BPCA_ERR_NONCANONICAL,
                        "CSR is not canonical: indptr must start at 0 and be non-decreasing, and each row's column indices strictly increasing and below num_genes");
        }
    }
    int slot = -1;
    for (int i = 0; i < (int)E->slots.size(); i++)
        if (!E->slots[i].used) { slot = i; break; }
    if (slot < 0) {
        slot = (int)E->slots.size();
        E->slots.emplace_back();
    }
    int rc = engine_grow(E, (int)E->slots.size());
    if (rc != BPCA_OK) {
        pool_free(ctx, ws);
        return rc;
    }
    const bpca::Prob d = make_prob(pb, L, (char *)ws, widen, slot, ctx->reorth_passes, pair);
    char *base = (char *)ws;
    const std::vector<double> v0 = start_vector(pb->seed, g);
    hipError_t e = hipSuccess;
    if (pair)   // tpart (at least nb + 1 int64) is free until set-up
        e = bpca::launch_pack_padded(pb->data, pb->indices, pb->indptr, pb->indptr_width, n, g,
                                     (int64_t *)(base + L.indptr64), (int64_t *)(base + L.tpart),
                                     (float *)(base + L.data4), (uint16_t *)(base + L.idx), s);
    if (pair && e == hipSuccess)
        e = bpca::launch_pair_segments((const int64_t *)(base + L.indptr64), n, (int64_t *)(base + L.segbase),
                                       (uint16_t *)(base + L.segs), s);
    if (!pair && widen)
        e = bpca::launch_widen_indptr((const int32_t *)pb->indptr, n + 1, (int64_t *)(base + L.indptr64), s);
    if (e == hipSuccess && L.tsorted)
        e = bpca::launch_tsort_build(d.data, (const uint16_t *)d.idx, d.indptr, n, g, L.tgl, (float *)d.tval,
                                     (uint8_t *)d.trow, (uint16_t *)d.thb, L.thb_words, (uint16_t *)d.tlens, (int64_t)ctx->max_lds, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d.V[0], v0.data(), sizeof(double) * g, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(E->d_probs + slot, &d, sizeof(d), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        pool_free(ctx, ws);
        return hip_fail(e, "problem upload");
    }
    Slot &sl = E->slots[slot];
    sl.used = true;
    sl.desc = d;
    sl.ws = ws;
    sl.t_admit = std::chrono::steady_clock::now();
    E->dirty = true;
    E->setup_pending = true;
    if (slot_out) *slot_out = slot;
    return BPCA_OK;
}

namespace {

// bpca_engine_step for a caller that entered at cancel generation gen0.
int engine_step_impl(bpca_engine *E, int32_t max_steps, int32_t *num_done,
                     uint64_t gen0) {
    bpca_context *ctx = E->ctx;
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    hipStream_t s = ctx->stream;
    begin_call(ctx);
    if (E->T.spmv_group != ctx->spmv_group || E->T.spmvt_group != ctx->spmvt_group ||
        E->T.spmv_wg != ctx->spmv_wg || E->T.spmvt_wg != ctx->spmvt_wg)
        E->dirty = true;
    if (E->dirty) BPCA_TRY(engine_rebuild_tables(E));
    if (E->setup_pending) {
        if (ctx->profiling) prof_mark(ctx, BPCA_K_SETUP, 0);
        HIP_TRY(bpca::launch_setup(E->T, s));
        if (ctx->profiling) prof_mark(ctx, BPCA_K_SETUP, 1);
        E->setup_pending = false;
    }
    int remaining = std::max(0, (int)max_steps);
    while (remaining > 0) {
        if (ctx->cancel_gen.load() != gen0) {
            (void)hipStreamSynchronize(s);
            return fail(BPCA_ERR_CANCELLED, "solve cancelled");
        }
        const int batch = std::min(remaining, 8);
        BPCA_TRY(engine_launch_steps(E, batch));
        remaining -= batch;
        HIP_TRY(hipMemcpyAsync(E->h_done, E->d_done, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (ctx->profiling) prof_flush(ctx);
        if (*E->h_done > E->done_seen) break;
    }
    if (num_done) *num_done = *E->h_done - E->done_seen;
    return BPCA_OK;
}

}  // namespace

extern "C" int bpca_engine_step(bpca_engine *E, int32_t max_steps, int32_t *num_done) {
    if (!E) return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL engine");
    return engine_step_impl(E, max_steps, num_done, E->ctx->cancel_gen.load());
}

int bpca_engine_collect(bpca_engine *E, int32_t max_out, int32_t *slots,
                        bpca_solve_stats *stats, double *singular_values,
                        int32_t sv_stride, int32_t *num_out) {
    if (!E) return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL engine");
    bpca_context *ctx = E->ctx;
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    hipStream_t s = ctx->stream;
    int nout = 0;
    const int nslots = (int)E->slots.size();
    if (num_out) *num_out = 0;
    if (nslots == 0) return BPCA_OK;
    std::vector<bpca::Prob> cur(nslots);
    HIP_TRY(hipMemcpyAsync(cur.data(), E->d_probs, sizeof(bpca::Prob) * nslots, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const auto now = std::chrono::steady_clock::now();
    for (int i = 0; i < nslots && nout < max_out; i++) {
        Slot &sl = E->slots[i];
        if (!sl.used || cur[i].phase != bpca::PH_DONE) continue;
        if (singular_values)
            HIP_TRY(hipMemcpyAsync(singular_values + (int64_t)nout * sv_stride, sl.desc.Sb,
                                   sizeof(double) * sl.desc.k, hipMemcpyDeviceToHost, s));
        if (stats) {
            bpca_solve_stats &st = stats[nout];
            st.converged = cur[i].converged;
            st.restarts = cur[i].iter + 1;
            st.lanczos_steps = cur[i].steps;
            st.matvecs = 2 * cur[i].steps;
            st.seconds = std::chrono::duration<double>(now - sl.t_admit).count();
            st.svd_sweeps = cur[i].svd_sweeps;
            st.slot = i;
            st.smax = cur[i].smax;
            st.reorth_passes = cur[i].onepass ? 1 : 2;
            st.pair_products = cur[i].pair;
        }
        if (slots) slots[nout] = i;
        nout++;
        bpca::Prob idle = cur[i];
        idle.phase = bpca::PH_IDLE;
        HIP_TRY(hipMemcpyAsync(E->d_probs + i, &idle, sizeof(idle), hipMemcpyHostToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
        pool_free(ctx, sl.ws);
        sl = Slot();
        E->done_seen++;
        E->dirty = true;
    }
    HIP_TRY(hipStreamSynchronize(s));
    if (num_out) *num_out = nout;
    return BPCA_OK;
}

int bpca_engine_active(bpca_engine *E, int32_t *active) {
    if (!E || !active) return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL argument");
    int a = 0;
    for (const auto &sl : E->slots) a += sl.used;
    *active = a;
    return BPCA_OK;
}

// ------------------------------------------------------------ solves

int bpca_autotune(bpca_context *ctx, const bpca_problem *problem, int32_t reps,
                  double *timings_ms) {
    if (!ctx || !problem || reps < 1)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "bad argument");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    bpca_engine *E = nullptr;
    BPCA_TRY(bpca_engine_create(ctx, &E));
    std::unique_ptr<bpca_engine, int (*)(bpca_engine *)> hold(E, bpca_engine_destroy);
    int32_t slot = -1, done = 0;
    BPCA_TRY(bpca_engine_add(E, problem, &slot));
    const bool prof = ctx->profiling;
    ctx->profiling = false;
    int rc = bpca_engine_step(E, 1, &done);  // set-up and one Lanczos step
    ctx->profiling = prof;
    BPCA_TRY(rc);
    hipStream_t s = ctx->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    hipError_t e = hipSuccess;
    // Average of `reps` launches after one untimed launch.
    auto time_variant = [&](int which, int group, int wg) -> double {
        e = bpca::launch_spmv_variant(E->T, which, group, wg, s);
        if (e == hipSuccess) e = hipEventRecord(e0, s);
        for (int r = 0; r < reps && e == hipSuccess; r++)
            e = bpca::launch_spmv_variant(E->T, which, group, wg, s);
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float t = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        return (double)t / reps;
    };
    // Every pair of group size and cap on the workgroups per CU (0 = none):
    // fewer resident waves can contend less for LDS banks and memory
    // channels, and then deeper load groups may pay.
    const int groups[3] = {2, 4, 8};
    const int caps[4] = {0, 4, 3, 2};
    double ms[2][4][3] = {};
    int best[2][2] = {{0, 1}, {0, 1}};  // (cap index, group index)
    for (int which = 0; which < 2 && e == hipSuccess; which++)
        for (int ci = 0; ci < 4 && e == hipSuccess; ci++)
            for (int gi = 0; gi < 3 && e == hipSuccess; gi++) {
                ms[which][ci][gi] = time_variant(which, groups[gi], caps[ci]);
                if (ms[which][ci][gi] < ms[which][best[which][0]][best[which][1]]) {
                    best[which][0] = ci;
                    best[which][1] = gi;
                }
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return hip_fail(e, "autotune timing");
    ctx->spmv_group = groups[best[0][1]];
    ctx->spmvt_group = groups[best[1][1]];
    ctx->spmv_wg = caps[best[0][0]];
    ctx->spmvt_wg = caps[best[1][0]];
    if (timings_ms)
        for (int which = 0; which < 2; which++)
            for (int gi = 0; gi < 3; gi++) timings_ms[which * 3 + gi] = ms[which][0][gi];
    return BPCA_OK;
}

int bpca_irlba_batched(bpca_context *ctx, int32_t num_problems,
                       const bpca_problem *problems, bpca_solve_stats *stats,
                       double *singular_values, int32_t sv_stride) {
    if (!ctx || num_problems < 0 || (num_problems > 0 && !problems))
        return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL argument");
    const uint64_t gen0 = ctx->cancel_gen.load();
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    if (singular_values)
        for (int i = 0; i < num_problems; i++)
            if ((int)problems[i].k > sv_stride)
                return fail(BPCA_ERR_INVALID_ARGUMENT, "sv_stride is smaller than a problem's k");
    bpca_engine *E = nullptr;
    BPCA_TRY(bpca_engine_create(ctx, &E));
    std::unique_ptr<bpca_engine, int (*)(bpca_engine *)> hold(E, bpca_engine_destroy);
    std::vector<int32_t> slot_of(num_problems);
    for (int i = 0; i < num_problems; i++) BPCA_TRY(bpca_engine_add(E, &problems[i], &slot_of[i]));
    std::vector<int> index_of_slot(E->slots.size(), -1);
    for (int i = 0; i < num_problems; i++) index_of_slot[slot_of[i]] = i;
    int collected = 0;
    std::vector<int32_t> slots(num_problems);
    std::vector<bpca_solve_stats> st(num_problems);
    const int stride = std::max(1, (int)sv_stride);
    std::vector<double> sv((size_t)num_problems * stride);
    while (collected < num_problems) {
        int32_t nd = 0;
        BPCA_TRY(engine_step_impl(E, 1 << 30, &nd, gen0));
        int32_t got = 0;
        BPCA_TRY(bpca_engine_collect(E, num_problems - collected, slots.data() + collected,
                                     st.data() + collected,
                                     singular_values ? sv.data() + (size_t)collected * stride : nullptr,
                                     stride, &got));
        collected += got;
    }
    for (int c = 0; c < num_problems; c++) {
        const int i = index_of_slot[slots[c]];
        if (stats) stats[i] = st[c];
        if (singular_values)
            std::copy(sv.begin() + (size_t)c * stride, sv.begin() + (size_t)c * stride + problems[i].k,
                      singular_values + (size_t)i * sv_stride);
    }
    return BPCA_OK;
}

int bpca_irlba_device(bpca_context *ctx, const float *d_data,
                      const int32_t *d_indices, const void *d_indptr,
                      int indptr_width, uint64_t num_cells, uint32_t num_genes,
                      uint32_t k, uint32_t subspace_size, float tolerance,
                      uint32_t max_iterations, uint64_t seed, int validate,
                      float *d_PCs, double *singular_values,
                      bpca_solve_stats *stats) {
    bpca_problem pb{};
    pb.data = d_data;
    pb.indices = d_indices;
    pb.indptr = d_indptr;
    pb.indptr_width = indptr_width;
    pb.validate = validate;
    pb.num_cells = num_cells;
    pb.num_genes = num_genes;
    pb.k = k;
    pb.subspace_size = subspace_size;
    pb.max_iterations = max_iterations;
    pb.tolerance = tolerance;
    pb.seed = seed;
    pb.PCs = d_PCs;
    return bpca_irlba_batched(ctx, 1, &pb, stats, singular_values, (int32_t)k);
}

int bpca_irlba_device_noalloc(bpca_context *ctx, const float *d_data,
                              const int32_t *d_indices, const void *d_indptr,
                              int indptr_width, uint64_t num_cells,
                              uint32_t num_genes, uint32_t k,
                              uint32_t subspace_size, float tolerance,
                              uint32_t max_iterations, uint64_t seed,
                              float *d_PCs, double *singular_values,
                              bpca_solve_stats *stats) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!d_data || !d_indices || !d_indptr || !d_PCs)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL device pointer");
    if (indptr_width != 4 && indptr_width != 8)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "indptr_width must be 4 or 8");
    const int64_t n = (int64_t)num_cells;
    const int g = (int)num_genes, m = (int)subspace_size;
    BPCA_TRY(check_shape(n, g, k, m, tolerance, max_iterations));
    if (bpca::lds_bytes_needed(g, m) > ctx->max_lds)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "num_genes / subspace_size need more LDS per workgroup than the device has");
    const uint64_t gen0 = ctx->cancel_gen.load();
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    std::lock_guard<std::mutex> use(g_scratch_use[ctx->device]);
    DeviceGuard dg(ctx->device);
    char *const base = (char *)g_scratch_ptr[ctx->device];
    if (!base) return fail(BPCA_ERR_DEVICE, "scratch address of this device was not recorded");
    hipStream_t s = ctx->stream;
    begin_call(ctx);
    int64_t nnz = 0;
    if (indptr_width == 8) {
        HIP_TRY(hipMemcpyAsync(&nnz, (const int64_t *)d_indptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    } else {
        int32_t nnz32 = 0;
        HIP_TRY(hipMemcpyAsync(&nnz32, (const int32_t *)d_indptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        nnz = nnz32;
    }
    if (nnz < 0) return fail(BPCA_ERR_NONCANONICAL, "indptr[num_cells] is negative");
    const bool pair = problem_pair(ctx, n, g, nnz);
    const uint64_t need = noalloc_bytes(n, g, m, nnz, indptr_width == 4, layout_sorted(ctx, n, g, nnz), pair);
    if (need > kScratchBytes)
        return fail(BPCA_ERR_INVALID_ARGUMENT,
                    "the allocation-free solve needs " + std::to_string(need) +
                        " bytes of device scratch for this problem; the library holds " +
                        std::to_string(kScratchBytes) + " per device");
    // Scratch layout: problem descriptor, data-block and group lists, two
    // counters (done, flag), then the workspace.
    const int64_t nb = bpca::ceil_div(n, bpca::kRowBlock), ngrp = bpca::ceil_div(nb, bpca::kGroup);
    size_t off = 0;
    bpca::Prob *d_prob = (bpca::Prob *)(base + off);
    off += align_up(sizeof(bpca::Prob));
    int2 *d_blocks = (int2 *)(base + off);
    off += align_up(sizeof(int2) * nb);
    int2 *d_groups = (int2 *)(base + off);
    off += align_up(sizeof(int2) * ngrp);
    int *d_done = (int *)(base + off);
    int *d_flag = d_done + 1;
    off += kAlign;
    char *ws = base + off;
    const bool widen = indptr_width == 4;
    const WsLayout L = ws_layout(n, g, m, widen, nnz, layout_sorted(ctx, n, g, nnz), pair);
    // The scratch keeps the last call's contents: clear the counters and the
    // quad padding past nnz of both packed copies (pair problems: the padded
    // copy is written whole after the check).
    HIP_TRY(bpca::launch_zero(d_done, kAlign, s));
    if (!pair) {
        HIP_TRY(bpca::launch_zero(ws + L.idx + L.idx_bytes * nnz, L.idx_bytes * (L.nnz4 - nnz), s));
        HIP_TRY(bpca::launch_zero(ws + L.data4 + sizeof(float) * nnz, sizeof(float) * (L.nnz4 - nnz), s));
    }
    HIP_TRY(bpca::launch_validate_pack(d_data, d_indices, d_indptr, indptr_width, n, nnz, g,
                                       (float *)(ws + L.data4), ws + L.idx, d_flag, s, pair));
    int h_flag = 0;
    HIP_TRY(hipMemcpyAsync(&h_flag, d_flag, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (h_flag)
        return fail(BPCA_ERR_NONCANONICAL,
                    "CSR is not canonical: indptr must start at 0 and be non-decreasing, and each row's column indices strictly increasing and below num_genes");
    bpca_problem pb{};
    pb.data = d_data;
    pb.indices = d_indices;
    pb.indptr = d_indptr;
    pb.indptr_width = indptr_width;
    pb.num_cells = num_cells;
    pb.num_genes = num_genes;
    pb.k = k;
    pb.subspace_size = subspace_size;
    pb.max_iterations = max_iterations;
    pb.tolerance = tolerance;
    pb.seed = seed;
    pb.PCs = d_PCs;
    const bpca::Prob d = make_prob(&pb, L, ws, widen, 0, ctx->reorth_passes, pair);
    const std::vector<double> v0 = start_vector(seed, g);
    std::vector<int2> blocks((size_t)nb), groups((size_t)ngrp);
    for (int64_t b = 0; b < nb; b++) blocks[b] = make_int2(0, (int)b);
    for (int64_t G = 0; G < ngrp; G++) groups[G] = make_int2(0, (int)G);
    if (pair)
        HIP_TRY(bpca::launch_pack_padded(d_data, d_indices, d_indptr, indptr_width, n, g,
                                         (int64_t *)(ws + L.indptr64), (int64_t *)(ws + L.tpart),
                                         (float *)(ws + L.data4), (uint16_t *)(ws + L.idx), s));
    if (pair)
        HIP_TRY(bpca::launch_pair_segments((const int64_t *)(ws + L.indptr64), n, (int64_t *)(ws + L.segbase),
                                           (uint16_t *)(ws + L.segs), s));
    if (!pair && widen)
        HIP_TRY(bpca::launch_widen_indptr((const int32_t *)d_indptr, n + 1, (int64_t *)(ws + L.indptr64), s));
    if (L.tsorted)
        HIP_TRY(bpca::launch_tsort_build(d.data, (const uint16_t *)d.idx, d.indptr, n, g, L.tgl, (float *)d.tval,
                                         (uint8_t *)d.trow, (uint16_t *)d.thb, L.thb_words, (uint16_t *)d.tlens, (int64_t)ctx->max_lds, s));
    HIP_TRY(hipMemcpyAsync(d.V[0], v0.data(), sizeof(double) * g, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_prob, &d, sizeof(d), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_blocks, blocks.data(), sizeof(int2) * nb, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_groups, groups.data(), sizeof(int2) * ngrp, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    bpca::Tables T{};
    T.probs = d_prob;
    T.num_probs = 1;
    T.blocks = d_blocks;
    T.num_blocks = nb;
    T.groups = d_groups;
    T.num_groups = ngrp;
    T.max_g = std::max(g, 2);
    T.max_m = std::max(m, 1);
    T.modes = 1 << d.mode;
    T.vmodes = d.pair ? 0 : 1 << d.mode;
    T.tmodes = d.pair ? bpca::kPairBit : d.tsorted ? bpca::kSortedBit : 1 << d.mode;
    T.rmodes = d.onepass ? bpca::kReorthOneBit : bpca::kReorthTwoBit;
    T.max_lds = (int64_t)ctx->max_lds;
    T.spmv_group = ctx->spmv_group;
    T.spmvt_group = ctx->spmvt_group;
    T.spmv_wg = ctx->spmv_wg;
    T.spmvt_wg = ctx->spmvt_wg;
    T.lds_per_cu = ctx->lds_per_cu;
    T.num_cus = ctx->num_cus;
    const auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(bpca::launch_setup(T, s));
    // Every sweep ends within m + 1 macro steps (its SVD, restart or final
    // product run in the step that ends it), so this bound is never reached
    // by a healthy solve.
    const int64_t max_macro = ((int64_t)max_iterations + 1) * (m + 2) + 16;
    int64_t launched = 0;
    int h_done = 0;
    while (!h_done) {
        if (ctx->cancel_gen.load() != gen0) {
            (void)hipStreamSynchronize(s);
            return fail(BPCA_ERR_CANCELLED, "solve cancelled");
        }
        if (launched >= max_macro) return fail(BPCA_ERR_NUMERICAL, "solve did not finish within its step bound");
        for (int i = 0; i < 8; i++) HIP_TRY(bpca::launch_macro_step(T, d_done, s, nullptr));
        launched += 8;
        HIP_TRY(hipMemcpyAsync(&h_done, d_done, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    bpca::Prob fin{};
    HIP_TRY(hipMemcpyAsync(&fin, d_prob, sizeof(fin), hipMemcpyDeviceToHost, s));
    if (singular_values)
        HIP_TRY(hipMemcpyAsync(singular_values, d.Sb, sizeof(double) * k, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (stats) {
        stats->converged = fin.converged;
        stats->restarts = fin.iter + 1;
        stats->lanczos_steps = fin.steps;
        stats->matvecs = 2 * fin.steps;
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        stats->svd_sweeps = fin.svd_sweeps;
        stats->slot = 0;
        stats->smax = fin.smax;
        stats->reorth_passes = fin.onepass ? 1 : 2;
        stats->pair_products = fin.pair;
    }
    return BPCA_OK;
}

}  // extern "C"

namespace {

std::mutex g_default_lock;
bpca_context *g_default = nullptr;

int default_context(bpca_context **out) {
    std::lock_guard<std::mutex> guard(g_default_lock);
    if (!g_default) {
        int dev = 0;
        HIP_TRY(hipGetDevice(&dev));
        BPCA_TRY(bpca_context_create(dev, 0, &g_default));
    }
    *out = g_default;
    return BPCA_OK;
}

template <typename T>
struct DevMem {
    T *p = nullptr;
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t count) {
        return hipMalloc((void **)&p, sizeof(T) * std::max<size_t>(count, 1));
    }
};

}  // namespace

extern "C" {

int bpca_irlba_f32(const float *data_matvec, const void *indices_matvec,
                   const void *indptr_matvec, const float *data_rmatvec,
                   const void *indices_rmatvec, const void *indptr_rmatvec,
                   int index_width, uint64_t num_cells, uint32_t num_genes,
                   uint32_t k, uint32_t subspace_size, float tolerance,
                   uint32_t max_iterations, uint64_t seed, int match_parallel,
                   int is_csr, uint32_t num_threads, uint64_t chunk_size,
                   float *PCs, int *converged) {
    (void)data_rmatvec;
    (void)indices_rmatvec;
    (void)indptr_rmatvec;
    if (!data_matvec || !indices_matvec || !indptr_matvec || !PCs)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL pointer argument");
    if (index_width != 4 && index_width != 8)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "index_width must be 4 or 8");
    if (num_threads < 1) return fail(BPCA_ERR_INVALID_ARGUMENT, "num_threads must be a positive integer");
    if (match_parallel && num_threads != 1)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "match_parallel must be False unless num_threads is 1");
    if (chunk_size < 1) return fail(BPCA_ERR_INVALID_ARGUMENT, "chunk_size must be a positive integer");
    BPCA_TRY(check_shape((int64_t)num_cells, num_genes, k, subspace_size, tolerance, max_iterations));
    bpca_context *ctx = nullptr;
    BPCA_TRY(default_context(&ctx));
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    const int64_t n = (int64_t)num_cells;
    // On the reference's serial CSC path (single_cell.py:12054-12069) the
    // matvec operand is the CSC of X: its offsets run over genes and its
    // indices are cells.  It is transposed to the CSR on the device.
    const bool csc = !is_csr && num_threads == 1 && !match_parallel;
    const int64_t outer = csc ? (int64_t)num_genes : n;
    const int64_t nnz = index_width == 8 ? ((const int64_t *)indptr_matvec)[outer]
                                         : (int64_t)((const int32_t *)indptr_matvec)[outer];
    if (nnz < 0) return fail(BPCA_ERR_NONCANONICAL, "indptr[num_cells] is negative");
    hipStream_t s = ctx->stream;
    if (csc) {
        DevMem<float> c_data, d_data, d_pcs;
        DevMem<char> c_idx, c_ptr;
        DevMem<int32_t> d_idx;
        DevMem<int64_t> d_ptr;
        HIP_TRY(c_data.alloc(nnz));
        HIP_TRY(c_idx.alloc((size_t)nnz * index_width));
        HIP_TRY(c_ptr.alloc((size_t)(outer + 1) * index_width));
        HIP_TRY(d_data.alloc(nnz));
        HIP_TRY(d_idx.alloc(nnz));
        HIP_TRY(d_ptr.alloc(n + 1));
        HIP_TRY(d_pcs.alloc((size_t)n * k));
        HIP_TRY(hipMemcpyAsync(c_data.p, data_matvec, sizeof(float) * nnz, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(c_idx.p, indices_matvec, (size_t)nnz * index_width, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(c_ptr.p, indptr_matvec, (size_t)(outer + 1) * index_width, hipMemcpyHostToDevice, s));
        hipError_t herr = hipSuccess;
        const int rc = bpca::run_csr_transpose(c_data.p, c_idx.p, index_width, c_ptr.p, index_width, outer, n,
                                               nnz, d_data.p, d_idx.p, d_ptr.p, &herr, s);
        if (rc == 1)
            return fail(BPCA_ERR_NONCANONICAL,
                        "CSC is not canonical: indptr must start at 0 and be non-decreasing, and row indices must lie in [0, num_cells)");
        if (rc == 2) return hip_fail(herr, "CSC to CSR transpose");
        bpca_solve_stats st{};
        BPCA_TRY(bpca_irlba_device(ctx, d_data.p, d_idx.p, d_ptr.p, 8, num_cells, num_genes, k,
                                   subspace_size, tolerance, max_iterations, seed, 1, d_pcs.p,
                                   nullptr, &st));
        HIP_TRY(hipMemcpyAsync(PCs, d_pcs.p, sizeof(float) * n * k, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (converged) *converged = st.converged;
        return BPCA_OK;
    }
    DevMem<float> d_data, d_pcs;
    DevMem<int32_t> d_idx;
    DevMem<int64_t> d_idx64;
    DevMem<char> d_ptr;
    DevMem<int> d_flag;
    HIP_TRY(d_data.alloc(nnz));
    HIP_TRY(d_idx.alloc(nnz));
    HIP_TRY(d_ptr.alloc((size_t)(n + 1) * index_width));
    HIP_TRY(d_pcs.alloc((size_t)n * k));
    HIP_TRY(d_flag.alloc(1));
    HIP_TRY(hipMemcpyAsync(d_data.p, data_matvec, sizeof(float) * nnz, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_ptr.p, indptr_matvec, (size_t)(n + 1) * index_width, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemsetAsync(d_flag.p, 0, sizeof(int), s));
    if (index_width == 8) {
        HIP_TRY(d_idx64.alloc(nnz));
        HIP_TRY(hipMemcpyAsync(d_idx64.p, indices_matvec, sizeof(int64_t) * nnz, hipMemcpyHostToDevice, s));
        HIP_TRY(bpca::launch_narrow_indices(d_idx64.p, nnz, (int)num_genes, d_idx.p, d_flag.p, s));
    } else {
        HIP_TRY(hipMemcpyAsync(d_idx.p, indices_matvec, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, s));
    }
    int h_flag = 0;
    HIP_TRY(hipMemcpyAsync(&h_flag, d_flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (h_flag) return fail(BPCA_ERR_NONCANONICAL, "column index outside [0, num_genes)");
    bpca_solve_stats st{};
    BPCA_TRY(bpca_irlba_device(ctx, d_data.p, d_idx.p, d_ptr.p, index_width, num_cells,
                               num_genes, k, subspace_size, tolerance, max_iterations,
                               seed, 1, d_pcs.p, nullptr, &st));
    HIP_TRY(hipMemcpyAsync(PCs, d_pcs.p, sizeof(float) * n * k, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (converged) *converged = st.converged;
    return BPCA_OK;
}

int bpca_test_small_svd(bpca_context *ctx, int32_t m, const double *B,
                        double *Ub, double *Sb, double *Vb, int32_t *sweeps) {
    if (!ctx || !B || !Ub || !Sb || !Vb || m < 1)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "bad argument");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    hipStream_t s = ctx->stream;
    const int64_t mm = m + (m & 1);
    DevMem<double> dB, dW, dJ, dU, dS;
    DevMem<int> dsw;
    HIP_TRY(dB.alloc((size_t)m * (m + 1)));
    HIP_TRY(dW.alloc((size_t)m * mm));
    HIP_TRY(dJ.alloc((size_t)(mm * mm + (int64_t)m * m)));
    HIP_TRY(dU.alloc((size_t)m * m));
    HIP_TRY(dS.alloc(m));
    HIP_TRY(dsw.alloc(1));
    HIP_TRY(hipMemcpyAsync(dB.p, B, sizeof(double) * m * (m + 1), hipMemcpyHostToDevice, s));
    HIP_TRY(bpca::launch_svd_single(dB.p, m, dW.p, dJ.p, dU.p, dS.p, dsw.p, (int64_t)ctx->max_lds, s));
    HIP_TRY(hipMemcpyAsync(Ub, dU.p, sizeof(double) * m * m, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(Sb, dS.p, sizeof(double) * m, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(Vb, dJ.p + mm * mm, sizeof(double) * m * m, hipMemcpyDeviceToHost, s));
    int h_sw = 0;
    HIP_TRY(hipMemcpyAsync(&h_sw, dsw.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (sweeps) *sweeps = h_sw;
    return BPCA_OK;
}

int bpca_synth_row_nnz(bpca_context *ctx, uint64_t seed, uint64_t row0,
                       uint64_t num_rows, uint32_t num_genes,
                       const double *d_type_gene_mean,
                       const int32_t *d_cell_type, const double *d_cell_factor,
                       int64_t *d_row_nnz) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    begin_call(ctx);
    HIP_TRY(bpca::launch_synth_row_nnz(seed, row0, (int64_t)num_rows, (int)num_genes,
                                       d_type_gene_mean, d_cell_type, d_cell_factor,
                                       d_row_nnz, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return BPCA_OK;
}

int bpca_synth_fill(bpca_context *ctx, uint64_t seed, uint64_t row0,
                    uint64_t num_rows, uint32_t num_genes,
                    const double *d_type_gene_mean, const int32_t *d_cell_type,
                    const double *d_cell_factor, const int64_t *d_indptr,
                    int32_t *d_indices, int32_t *d_counts) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    begin_call(ctx);
    HIP_TRY(bpca::launch_synth_fill(seed, row0, (int64_t)num_rows, (int)num_genes,
                                    d_type_gene_mean, d_cell_type, d_cell_factor,
                                    d_indptr, d_indices, d_counts, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return BPCA_OK;
}

int bpca_normalize_log1ppf(bpca_context *ctx, const int32_t *d_counts,
                           const int64_t *d_indptr, uint64_t num_cells,
                           float *d_out) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (num_cells < 1) return fail(BPCA_ERR_INVALID_ARGUMENT, "num_cells must be positive");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    begin_call(ctx);
    hipStream_t s = ctx->stream;
    DevMem<uint64_t> sums;
    HIP_TRY(sums.alloc(num_cells + 1));
    HIP_TRY(bpca::launch_row_sums(d_counts, d_indptr, (int64_t)num_cells, sums.p, s));
    HIP_TRY(bpca::launch_sum_u64(sums.p, (int64_t)num_cells, sums.p + num_cells, s));
    HIP_TRY(bpca::launch_log1ppf(d_counts, d_indptr, (int64_t)num_cells, sums.p,
                                 sums.p + num_cells, d_out, s));
    HIP_TRY(hipStreamSynchronize(s));
    return BPCA_OK;
}

int bpca_csr_subset_count(bpca_context *ctx, const void *d_indptr,
                          int indptr_width, const void *d_indices,
                          int index_width, int64_t num_rows, int64_t nnz,
                          int64_t num_cols, const int64_t *d_rows,
                          int64_t num_out_rows, const int32_t *d_col_map,
                          int64_t *d_row_nnz) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!d_indptr || !d_indices || !d_row_nnz) return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL device pointer");
    if ((indptr_width != 4 && indptr_width != 8) || (index_width != 4 && index_width != 8))
        return fail(BPCA_ERR_INVALID_ARGUMENT, "indptr_width and index_width must be 4 or 8");
    if (num_rows < 0 || nnz < 0 || num_cols < 0 || num_out_rows < 0)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "negative size");
    if (!d_rows && num_out_rows != num_rows)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "without rows, num_out_rows must equal num_rows");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    begin_call(ctx);
    hipStream_t s = ctx->stream;
    DevMem<int> bad;
    HIP_TRY(bad.alloc(1));
    HIP_TRY(hipMemsetAsync(bad.p, 0, sizeof(int), s));
    HIP_TRY(bpca::launch_subset_count(d_indptr, indptr_width, d_indices, index_width, num_rows, nnz,
                                      num_cols, d_rows, num_out_rows, d_col_map, d_row_nnz, bad.p, s));
    int h_bad = 0;
    HIP_TRY(hipMemcpyAsync(&h_bad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (h_bad)
        return fail(BPCA_ERR_NONCANONICAL,
                    "CSR subset: a selected row is out of range, its row offsets are invalid, or a column index is outside [0, num_cols)");
    return BPCA_OK;
}

int bpca_csr_subset_fill(bpca_context *ctx, const float *d_data,
                         const void *d_indptr, int indptr_width,
                         const void *d_indices, int index_width,
                         const int64_t *d_rows, int64_t num_out_rows,
                         const int32_t *d_col_map, int col_map_increasing,
                         const int64_t *d_out_indptr, int64_t max_row_nnz,
                         float *d_out_data, int32_t *d_out_indices) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!d_data || !d_indptr || !d_indices || !d_out_indptr ||
        (num_out_rows > 0 && max_row_nnz > 0 && (!d_out_data || !d_out_indices)))
        return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL device pointer");
    if ((indptr_width != 4 && indptr_width != 8) || (index_width != 4 && index_width != 8))
        return fail(BPCA_ERR_INVALID_ARGUMENT, "indptr_width and index_width must be 4 or 8");
    if (num_out_rows < 0 || max_row_nnz < 0) return fail(BPCA_ERR_INVALID_ARGUMENT, "negative size");
    const bool ordered = col_map_increasing != 0 || d_col_map == nullptr;
    if (!ordered && max_row_nnz > bpca::subset_sort_max())
        return fail(BPCA_ERR_INVALID_ARGUMENT,
                    "CSR subset with a reordering column map: rows may keep at most " +
                        std::to_string(bpca::subset_sort_max()) + " entries");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    begin_call(ctx);
    hipStream_t s = ctx->stream;
    HIP_TRY(bpca::launch_subset_fill(d_data, d_indptr, indptr_width, d_indices, index_width, d_rows,
                                     num_out_rows, d_col_map, ordered ? 1 : 0, d_out_indptr,
                                     d_out_data, d_out_indices, s));
    HIP_TRY(hipStreamSynchronize(s));
    return BPCA_OK;
}

int bpca_knn_exact(bpca_context *ctx, const float *d_X, int64_t num_cells,
                   int32_t num_dims, int32_t k, uint32_t *d_neighbors,
                   float *d_distances) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!d_X || !d_neighbors || !d_distances) return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL device pointer");
    if (num_cells < 2) return fail(BPCA_ERR_INVALID_ARGUMENT, "there is only one cell, so nearest neighbors cannot be calculated");
    if (num_dims < 1 || num_dims > 256) return fail(BPCA_ERR_INVALID_ARGUMENT, "num_dims must be in [1, 256]");
    if (k < 1 || (int64_t)k >= num_cells)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "k must be >= 1 and less than the number of cells");
    if (num_cells > (int64_t)UINT32_MAX) return fail(BPCA_ERR_INVALID_ARGUMENT, "num_cells too large");
    const bool mfma = num_dims <= bpca::kKnnMfmaMaxDims;
    if (!mfma && k > 64)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "more than 64 neighbours need num_dims <= " + std::to_string(bpca::kKnnMfmaMaxDims));
    const size_t lds = mfma ? bpca::knn_mfma_lds_bytes(num_dims)
                            : bpca::knn_lds_bytes(num_dims, bpca::knn_kmax(k));
    if ((int64_t)lds > (int64_t)ctx->max_lds)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "num_dims / k need more LDS than the device has");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    begin_call(ctx);
    hipStream_t s = ctx->stream;
    DevMem<float> norms;
    HIP_TRY(norms.alloc(num_cells + bpca::kKnnPadRows));
    HIP_TRY(bpca::launch_row_norms(d_X, num_cells, num_dims, norms.p, s));
    if (mfma) {
        DevMem<float> padded;
        HIP_TRY(padded.alloc((num_cells + bpca::kKnnPadRows) * bpca::knn_mfma_stride(num_dims)));
        HIP_TRY(bpca::launch_knn_mfma(d_X, norms.p, num_cells, num_dims, k, padded.p, d_neighbors, d_distances, s));
        HIP_TRY(hipStreamSynchronize(s));
        return BPCA_OK;
    }
    HIP_TRY(bpca::launch_knn_exact(d_X, norms.p, num_cells, num_dims, k, d_neighbors, d_distances, s));
    HIP_TRY(hipStreamSynchronize(s));
    return BPCA_OK;
}

int bpca_knn_ivf(bpca_context *ctx, const float *d_X, int64_t num_cells,
                 int32_t num_dims, int32_t k, int32_t num_clusters,
                 int32_t num_clusters_searched, int32_t num_kmeans_iterations,
                 float kmeans_tolerance, int64_t chunk_size_kmeans,
                 uint64_t seed, uint32_t *d_neighbors, float *d_distances,
                 int64_t *info) {
    int64_t info6[6];
    const int rc = bpca_knn_ivf_ex(ctx, d_X, num_cells, num_dims, k, num_clusters,
                                   num_clusters_searched, num_kmeans_iterations,
                                   kmeans_tolerance, chunk_size_kmeans, seed, 0, 5, 1.0f,
                                   d_neighbors, d_distances, info ? info6 : nullptr);
    if (rc == BPCA_OK) std::copy(info6, info6 + 5, info);
    return rc;
}

int bpca_knn_ivf_ex(bpca_context *ctx, const float *d_X, int64_t num_cells,
                    int32_t num_dims, int32_t k, int32_t num_clusters,
                    int32_t num_clusters_searched, int32_t num_kmeans_iterations,
                    float kmeans_tolerance, int64_t chunk_size_kmeans,
                    uint64_t seed, int32_t kmeans_barbar,
                    int32_t num_init_iterations, float oversampling_factor,
                    uint32_t *d_neighbors, float *d_distances, int64_t *info) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!d_X || !d_neighbors || !d_distances || !info) return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL pointer");
    if (num_cells < 2) return fail(BPCA_ERR_INVALID_ARGUMENT, "there is only one cell, so nearest neighbors cannot be calculated");
    if (num_cells > (int64_t)UINT32_MAX) return fail(BPCA_ERR_INVALID_ARGUMENT, "num_cells too large");
    if (num_dims < 1 || num_dims > bpca::kKnnMfmaMaxDims)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "num_dims must be in [1, " + std::to_string(bpca::kKnnMfmaMaxDims) + "]");
    if (k < 1 || (int64_t)k >= num_cells)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "k must be >= 1 and less than the number of cells");
    if (num_clusters < 1 || (int64_t)num_clusters > num_cells || num_clusters >= (1 << 26))
        return fail(BPCA_ERR_INVALID_ARGUMENT, "num_clusters must be >= 1 and <= the number of cells (and < 2^26)");
    if (num_clusters_searched < 1 || num_clusters_searched > num_clusters)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "num_clusters_searched must be >= 1 and <= num_clusters");
    if (num_kmeans_iterations < 1) return fail(BPCA_ERR_INVALID_ARGUMENT, "num_kmeans_iterations must be >= 1");
    if (!(kmeans_tolerance > 0.0f)) return fail(BPCA_ERR_INVALID_ARGUMENT, "kmeans_tolerance must be > 0");
    if (chunk_size_kmeans < 1 || chunk_size_kmeans > num_cells)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "chunk_size_kmeans must be >= 1 and <= the number of cells");
    if ((int64_t)bpca::knn_mfma_lds_bytes(num_dims) > (int64_t)ctx->max_lds)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "num_dims needs more LDS than the device has");
    if (const char *e = barbar_arg_error(kmeans_barbar, num_init_iterations, oversampling_factor))
        return fail(BPCA_ERR_INVALID_ARGUMENT, e);
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    begin_call(ctx);
    bpca::IvfParams P{num_cells, num_dims, k, num_clusters, num_clusters_searched,
                      num_kmeans_iterations, kmeans_tolerance, chunk_size_kmeans, seed};
    P.barbar = kmeans_barbar;
    P.init_iterations = num_init_iterations;
    P.oversampling = oversampling_factor;
    bpca::IvfInfo I{};
    std::string msg;
    hipError_t herr = hipSuccess;
    const int rc = bpca::run_knn_ivf(d_X, P, d_neighbors, d_distances, &I, &msg, &herr, ctx->stream);
    if (rc == 1) return fail(BPCA_ERR_INVALID_ARGUMENT, msg);
    if (rc == 2) return hip_fail(herr, "run_knn_ivf");
    info[0] = I.clusters;
    info[1] = I.searched;
    info[2] = I.iterations;
    info[3] = I.exhaustive;
    info[4] = I.entries;
    info[5] = I.candidates;
    return BPCA_OK;
}

int bpca_kmeans(bpca_context *ctx, const float *d_X, int64_t num_cells,
                int32_t num_dims, int32_t num_clusters,
                int32_t num_kmeans_iterations, float kmeans_tolerance,
                int64_t chunk_size_kmeans, uint64_t seed, uint32_t *d_labels,
                float *d_centroids, uint32_t *d_counts,
                int32_t *iterations_run) {
    return bpca_kmeans_ex(ctx, d_X, num_cells, num_dims, num_clusters, num_kmeans_iterations,
                          kmeans_tolerance, chunk_size_kmeans, seed, 0, 5, 1.0f, d_labels,
                          d_centroids, d_counts, iterations_run);
}

int bpca_kmeans_ex(bpca_context *ctx, const float *d_X, int64_t num_cells,
                   int32_t num_dims, int32_t num_clusters,
                   int32_t num_kmeans_iterations, float kmeans_tolerance,
                   int64_t chunk_size_kmeans, uint64_t seed, int32_t kmeans_barbar,
                   int32_t num_init_iterations, float oversampling_factor,
                   uint32_t *d_labels, float *d_centroids, uint32_t *d_counts,
                   int32_t *iterations_run) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!d_X || !d_labels || !d_centroids || !d_counts || !iterations_run)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL pointer");
    if (num_cells < 1 || num_cells > (int64_t)UINT32_MAX) return fail(BPCA_ERR_INVALID_ARGUMENT, "num_cells out of range");
    if (num_dims < 1 || num_dims > bpca::kKnnMfmaMaxDims)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "num_dims must be in [1, " + std::to_string(bpca::kKnnMfmaMaxDims) + "]");
    if (num_clusters < 1 || (int64_t)num_clusters > num_cells || num_clusters >= (1 << 26))
        return fail(BPCA_ERR_INVALID_ARGUMENT, "num_clusters must be >= 1 and <= the number of cells (and < 2^26)");
    if (num_kmeans_iterations < 1) return fail(BPCA_ERR_INVALID_ARGUMENT, "num_kmeans_iterations must be >= 1");
    if (!(kmeans_tolerance > 0.0f)) return fail(BPCA_ERR_INVALID_ARGUMENT, "kmeans_tolerance must be > 0");
    if (chunk_size_kmeans < 1 || chunk_size_kmeans > num_cells)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "chunk_size_kmeans must be >= 1 and <= the number of cells");
    if (const char *e = barbar_arg_error(kmeans_barbar, num_init_iterations, oversampling_factor))
        return fail(BPCA_ERR_INVALID_ARGUMENT, e);
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    begin_call(ctx);
    bpca::IvfParams P{num_cells, num_dims, 1, num_clusters, 1,
                      num_kmeans_iterations, kmeans_tolerance, chunk_size_kmeans, seed};
    P.barbar = kmeans_barbar;
    P.init_iterations = num_init_iterations;
    P.oversampling = oversampling_factor;
    P.kmeans_labels = d_labels;
    P.kmeans_centroids = d_centroids;
    P.kmeans_counts = d_counts;
    bpca::IvfInfo I{};
    std::string msg;
    hipError_t herr = hipSuccess;
    const int rc = bpca::run_knn_ivf(d_X, P, nullptr, nullptr, &I, &msg, &herr, ctx->stream);
    if (rc == 1) return fail(BPCA_ERR_INVALID_ARGUMENT, msg);
    if (rc == 2) return hip_fail(herr, "run_knn_ivf");
    *iterations_run = I.iterations;
    return BPCA_OK;
}

int bpca_kmeans_barbar_init(bpca_context *ctx, const float *d_X, int64_t num_cells,
                            int32_t num_dims, int32_t num_clusters,
                            int32_t num_init_iterations, float oversampling_factor,
                            int64_t chunk_size_kmeans, uint64_t seed, uint32_t *cells,
                            int64_t *num_candidates) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!d_X || !cells || !num_candidates) return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL pointer");
    if (num_cells < 1 || num_cells > (int64_t)UINT32_MAX) return fail(BPCA_ERR_INVALID_ARGUMENT, "num_cells out of range");
    if (num_dims < 1 || num_dims > bpca::kKnnMfmaMaxDims)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "num_dims must be in [1, " + std::to_string(bpca::kKnnMfmaMaxDims) + "]");
    if (num_clusters < 1 || (int64_t)num_clusters > num_cells)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "num_clusters must be >= 1 and <= the number of cells");
    if (chunk_size_kmeans < 1 || chunk_size_kmeans > num_cells)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "chunk_size_kmeans must be >= 1 and <= the number of cells");
    if (const char *e = barbar_arg_error(1, num_init_iterations, oversampling_factor))
        return fail(BPCA_ERR_INVALID_ARGUMENT, e);
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    begin_call(ctx);
    bpca::IvfParams P{num_cells, num_dims, 1, num_clusters, 1, 1, 1.0f, chunk_size_kmeans, seed};
    P.barbar = 1;
    P.init_iterations = num_init_iterations;
    P.oversampling = oversampling_factor;
    std::vector<uint32_t> v;
    std::string msg;
    hipError_t herr = hipSuccess;
    const int rc = bpca::kmeans_barbar_init_cells(d_X, nullptr, nullptr, P, &v, num_candidates, &msg, &herr,
                                                  ctx->stream);
    if (rc == 1) return fail(BPCA_ERR_INVALID_ARGUMENT, msg);
    if (rc == 2) return hip_fail(herr, "kmeans_barbar_init_cells");
    std::copy(v.begin(), v.end(), cells);
    return BPCA_OK;
}

int bpca_csr_transpose(bpca_context *ctx, const float *d_data,
                       const void *d_indices, int index_width,
                       const void *d_indptr, int indptr_width,
                       int64_t num_rows, int64_t num_cols, float *d_out_data,
                       int32_t *d_out_indices, int64_t *d_out_indptr) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!d_indptr || !d_out_indptr) return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL device pointer");
    if ((index_width != 4 && index_width != 8) || (indptr_width != 4 && indptr_width != 8))
        return fail(BPCA_ERR_INVALID_ARGUMENT, "index_width and indptr_width must be 4 or 8");
    if (num_rows < 0 || num_cols < 0 || num_rows > (int64_t)UINT32_MAX || num_cols > (int64_t)INT32_MAX)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "num_rows must fit 32 bits and num_cols 31 bits");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    begin_call(ctx);
    hipStream_t s = ctx->stream;
    int64_t nnz = 0;
    if (indptr_width == 8) {
        HIP_TRY(hipMemcpyAsync(&nnz, (const int64_t *)d_indptr + num_rows, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    } else {
        int32_t v = 0;
        HIP_TRY(hipMemcpyAsync(&v, (const int32_t *)d_indptr + num_rows, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        nnz = v;
    }
    HIP_TRY(hipStreamSynchronize(s));
    if (nnz < 0) return fail(BPCA_ERR_NONCANONICAL, "indptr[num_rows] is negative");
    if (nnz > 0 && (!d_data || !d_indices || !d_out_data || !d_out_indices))
        return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL device pointer");
    hipError_t herr = hipSuccess;
    const int rc = bpca::run_csr_transpose(d_data, d_indices, index_width, d_indptr, indptr_width, num_rows,
                                           num_cols, nnz, d_out_data, d_out_indices, d_out_indptr, &herr, s);
    if (rc == 1)
        return fail(BPCA_ERR_NONCANONICAL,
                    "transpose: indptr must start at 0 and be non-decreasing, and indices must lie in [0, num_cols)");
    if (rc == 2) return hip_fail(herr, "transpose");
    return BPCA_OK;
}

int bpca_normalize(bpca_context *ctx, const void *d_data, int data_type,
                   const void *d_indices, int index_width, const void *d_indptr,
                   int indptr_width, int is_csr, int64_t num_cells,
                   int64_t num_genes, const uint8_t *d_qc, int method,
                   float *d_normalized, uint64_t *d_row_sums) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!d_indptr || !d_row_sums) return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL device pointer");
    if (data_type < 0 || data_type > 5)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "data_type must be 0 (int32), 1 (uint32), 2 (int64), 3 (uint64), 4 (float32) or 5 (float64)");
    if ((index_width != 4 && index_width != 8) || (indptr_width != 4 && indptr_width != 8))
        return fail(BPCA_ERR_INVALID_ARGUMENT, "index_width and indptr_width must be 4 or 8");
    if (method < 0 || method > 2)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "method must be 0 (logCP10k), 1 (log1pPF) or 2 (PFlog1pPF)");
    if (num_cells < 1 || num_genes < 1) return fail(BPCA_ERR_INVALID_ARGUMENT, "num_cells and num_genes must be positive");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    begin_call(ctx);
    hipStream_t s = ctx->stream;
    const int64_t outer = is_csr ? num_cells : num_genes;
    int64_t nnz = 0;
    if (indptr_width == 8) {
        HIP_TRY(hipMemcpyAsync(&nnz, (const int64_t *)d_indptr + outer, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    } else {
        int32_t v = 0;
        HIP_TRY(hipMemcpyAsync(&v, (const int32_t *)d_indptr + outer, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        nnz = v;
    }
    HIP_TRY(hipStreamSynchronize(s));
    if (nnz < 0) return fail(BPCA_ERR_NONCANONICAL, "indptr's last offset is negative");
    if (nnz > 0 && (!d_data || !d_indices || !d_normalized))
        return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL device pointer");
    bpca::NormalizeArgs A{d_data, data_type, d_indices, index_width, d_indptr, indptr_width,
                          is_csr != 0, num_cells, num_genes, nnz, d_qc, method, d_normalized,
                          d_row_sums};
    hipError_t herr = hipSuccess;
    const int rc = bpca::run_normalize(A, &herr, s);
    if (rc == 1) return fail(BPCA_ERR_NONCANONICAL, "normalize: an index is outside [0, num_cells)");
    if (rc == 2) return hip_fail(herr, "normalize");
    return BPCA_OK;
}

int bpca_test_log1pf(bpca_context *ctx, const float *d_x, int64_t n, float *d_y) {
    if (!ctx || (n > 0 && (!d_x || !d_y))) return fail(BPCA_ERR_INVALID_ARGUMENT, "bad argument");
    std::lock_guard<std::recursive_mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    begin_call(ctx);
    HIP_TRY(bpca::launch_log1pf_eval(d_x, n, d_y, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return BPCA_OK;
}

int bpca_start_vector(uint64_t seed, uint32_t num_genes, float *draws,
                      double *v0) {
    if (!draws && !v0) return fail(BPCA_ERR_INVALID_ARGUMENT, "draws and v0 are both NULL");
    if (num_genes < 1) return fail(BPCA_ERR_INVALID_ARGUMENT, "num_genes must be positive");
    const std::vector<double> v = start_vector(seed, (int)num_genes, draws);
    if (v0) std::copy(v.begin(), v.end(), v0);
    return BPCA_OK;
}

int bpca_kmeans_init_cells(int64_t num_cells, int32_t num_clusters,
                           uint64_t seed, uint32_t *cells) {
    if (!cells) return fail(BPCA_ERR_INVALID_ARGUMENT, "cells is NULL");
    if (num_cells < 1 || num_cells > (int64_t)UINT32_MAX) return fail(BPCA_ERR_INVALID_ARGUMENT, "num_cells out of range");
    if (num_clusters < 1 || (int64_t)num_clusters > num_cells)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "num_clusters must be >= 1 and <= the number of cells");
    const std::vector<uint32_t> v = bpca::kmeans_random_init_cells(num_cells, num_clusters, seed);
    std::copy(v.begin(), v.end(), cells);
    return BPCA_OK;
}

}  // extern "C"
