// Host driver of the IRLBA solve and the C ABI declared in include/bpca.h.
//
// The restart loop follows brisc/pca.pyx:973-1254 (parallel path):
//   per Lanczos step j = k_current .. m-1
//     u = Z v_j            spmv + reorth partials/apply + alpha   (device)
//     w = Z^T u_j - a v_j  spmv_t + group sums + vside (CGS, beta) (device)
//   then B (m x (m+1)) comes to the host once per restart for the FP64 SVD,
//   the residual test |r Ub[m-1, i]| < tol * Smax, k_current update, and the
//   restart U <- U Ub[:, :kc], V <- V Vbt[:kc]^T, V[:, kc] = remainder.
// The sweep itself has no host synchronisation: kernels are queued back to
// back on one stream and the host waits only at the end of each sweep.
#include "bpca.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "kernels.h"
#include "svd.h"

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

#define HIP_TRY(expr)                                    \
    do {                                                 \
        hipError_t _e = (expr);                          \
        if (_e != hipSuccess) return hip_fail(_e, #expr); \
    } while (0)

#define BPCA_TRY(expr)            \
    do {                          \
        int _rc = (expr);         \
        if (_rc != BPCA_OK) return _rc; \
    } while (0)

constexpr size_t kAlign = 256;

inline size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// One device allocation carved into the solve's buffers.
struct Layout {
    size_t U, U2, u32, P, red, sq, tpart, grp, V, V2, B, sd, vs, scal, rem,
        rscal, coef, vbt, flag, total;
};

Layout make_layout(int64_t n, int g, int m, int k) {
    using bpca::ceil_div;
    const int64_t nb = ceil_div(n, bpca::kRowBlock);
    const int64_t ng = ceil_div(nb, bpca::kGroup);
    Layout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t at = off;
        off += align_up(std::max<size_t>(bytes, 1));
        return at;
    };
    L.U = take(sizeof(double) * n * m);
    L.U2 = take(sizeof(double) * n * m);
    L.u32 = take(sizeof(float) * n);
    L.P = take(sizeof(double) * (m + 1) * nb);
    L.red = take(sizeof(double) * (m + 1));
    L.sq = take(sizeof(double) * nb);
    L.tpart = take(sizeof(float) * nb * 2 * g);
    L.grp = take(sizeof(double) * ng * 2 * g);
    L.V = take(sizeof(double) * g * (m + 1));
    L.V2 = take(sizeof(double) * g * (m + 1));
    L.B = take(sizeof(double) * m * (m + 1));
    L.sd = take(sizeof(double) * g);
    L.vs = take(sizeof(float) * g);
    L.scal = take(sizeof(double) * 4);
    L.rem = take(sizeof(double) * g);
    L.rscal = take(sizeof(double) * 4);
    L.coef = take(sizeof(double) * m * std::max(m, k));
    L.vbt = take(sizeof(double) * m * (m + 1));
    L.flag = take(sizeof(int) * 4);
    L.total = off;
    (void)k;
    return L;
}

// PCG-XSH-RR 64/32 (O'Neill 2014) seeded as brisc's srand does
// (cyutils.pyx:1153-1158, 1228-1232), with the FP32 Box-Muller draw of
// cyutils.pyx:1173-1191; gives the start vector of pca.pyx:801-806.
struct Pcg32 {
    uint64_t state;
    explicit Pcg32(uint64_t seed) : state(seed + 1442695040888963407ULL) { next(); }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + 1442695040888963407ULL;
        uint32_t xs = (uint32_t)((old ^ (old >> 18)) >> 27);
        uint32_t rot = (uint32_t)(old >> 59);
        return (xs >> rot) | (xs << ((-rot) & 31));
    }
    float uniform() { return (float)next() / (float)UINT_MAX; }
    float normal() {
        float u1;
        do { u1 = uniform(); } while (u1 == 0.0f);
        float u2 = uniform();
        float r = std::sqrt(-2.0f * std::log(u1));
        float theta = (float)(2.0 * M_PI * (double)u2);
        return r * std::cos(theta);
    }
};

}  // namespace

struct bpca_context {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    size_t max_lds = 65536;  // dynamic LDS per workgroup on this device
    char *ws = nullptr;
    size_t ws_bytes = 0;
    double *host = nullptr;  // pinned staging: B, coef, Vbt, v0, rscal
    size_t host_doubles = 0;
    std::atomic<int> cancel{0};
    std::mutex lock;
    // Kernel timing: events recorded around launch classes, resolved after
    // each host synchronisation.
    bool profiling = false;
    std::vector<hipEvent_t> events;
    size_t next_event = 0;
    struct Pending { int kind, a, b; };
    std::vector<Pending> pending;
    double k_ms[BPCA_NUM_KERNEL_CLASSES] = {};
    int64_t k_count[BPCA_NUM_KERNEL_CLASSES] = {};
};

namespace {

int ensure_workspace(bpca_context *ctx, size_t bytes) {
    if (ctx->ws_bytes >= bytes) return BPCA_OK;
    if (ctx->ws) {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        HIP_TRY(hipFree(ctx->ws));
        ctx->ws = nullptr;
        ctx->ws_bytes = 0;
    }
    hipError_t e = hipMalloc((void **)&ctx->ws, bytes);
    if (e != hipSuccess) {
        ctx->ws = nullptr;
        return hip_fail(e, "hipMalloc(workspace)");
    }
    ctx->ws_bytes = bytes;
    return BPCA_OK;
}

int ensure_host(bpca_context *ctx, size_t doubles) {
    if (ctx->host_doubles >= doubles) return BPCA_OK;
    if (ctx->host) {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        HIP_TRY(hipHostFree(ctx->host));
        ctx->host = nullptr;
    }
    HIP_TRY(hipHostMalloc((void **)&ctx->host, doubles * sizeof(double), hipHostMallocDefault));
    ctx->host_doubles = doubles;
    return BPCA_OK;
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

inline int prof_begin(bpca_context *ctx) {
    return ctx->profiling ? prof_record(ctx) : -1;
}

inline void prof_end(bpca_context *ctx, int kind, int a) {
    if (a < 0) return;
    int b = prof_record(ctx);
    if (b >= 0) ctx->pending.push_back({kind, a, b});
}

// Call after the stream has been synchronised.
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

int check_shape(int64_t n, int64_t g, int64_t k, int64_t m, float tol,
                int64_t maxit) {
    if (n < 2) return fail(BPCA_ERR_INVALID_ARGUMENT, "there is only one cell, so principal components cannot be calculated");
    if (g < 2) return fail(BPCA_ERR_INVALID_ARGUMENT, "there is only one gene, so principal components cannot be calculated");
    if (k < 1) return fail(BPCA_ERR_INVALID_ARGUMENT, "k must be a positive integer");
    if (m < k) return fail(BPCA_ERR_INVALID_ARGUMENT, "subspace_size must be >= k");
    if (m > n || m > g) return fail(BPCA_ERR_INVALID_ARGUMENT, "subspace_size must be <= min(num_cells, num_genes)");
    if (!(tol > 0.0f)) return fail(BPCA_ERR_INVALID_ARGUMENT, "tolerance must be positive");
    if (maxit < 1) return fail(BPCA_ERR_INVALID_ARGUMENT, "max_iterations must be a positive integer");
    if (g > bpca::kMaxGenesTranspose) return fail(BPCA_ERR_INVALID_ARGUMENT, "num_genes above 8192 is not supported by the device path");
    if (n > (int64_t)INT32_MAX * (int64_t)bpca::kRowBlock) return fail(BPCA_ERR_INVALID_ARGUMENT, "num_cells too large");
    return BPCA_OK;
}

// The solve proper, on device-resident CSR.
int solve(bpca_context *ctx, const bpca::Csr &X, int k, int m, float tol,
          int maxit, uint64_t seed, bool validate, float *d_PCs,
          double *sv_out, bpca_solve_stats *st) {
    using namespace bpca;
    const auto t_start = std::chrono::steady_clock::now();
    const int64_t n = X.n;
    const int g = X.g;
    BPCA_TRY(check_shape(n, g, k, m, tol, maxit));
    if (lds_bytes_needed(g, m) > ctx->max_lds)
        return fail(BPCA_ERR_INVALID_ARGUMENT,
                    "num_genes / subspace_size need more LDS per workgroup than this device has (" +
                        std::to_string(lds_bytes_needed(g, m)) + " > " + std::to_string(ctx->max_lds) + " bytes)");
    const int64_t nb = ceil_div(n, kRowBlock);
    const int64_t ngrp = ceil_div(nb, kGroup);
    const Layout L = make_layout(n, g, m, k);
    if (ctx->budget && L.total > ctx->budget)
        return fail(BPCA_ERR_BUDGET, "solve workspace exceeds the context's HBM budget");
    BPCA_TRY(ensure_workspace(ctx, L.total));
    const size_t hB = 0, hCoef = hB + (size_t)m * (m + 1),
                 hVbt = hCoef + (size_t)m * std::max(m, k),
                 hV0 = hVbt + (size_t)m * (m + 1), hR = hV0 + g,
                 hEnd = hR + 4;
    BPCA_TRY(ensure_host(ctx, hEnd));
    hipStream_t s = ctx->stream;
    char *base = ctx->ws;
    double *U = (double *)(base + L.U), *U2 = (double *)(base + L.U2);
    float *u32 = (float *)(base + L.u32);
    double *P = (double *)(base + L.P), *red = (double *)(base + L.red);
    double *sq = (double *)(base + L.sq);
    float *tpart = (float *)(base + L.tpart);
    double *grp = (double *)(base + L.grp);
    double *V = (double *)(base + L.V), *V2 = (double *)(base + L.V2);
    double *B = (double *)(base + L.B), *sd = (double *)(base + L.sd);
    float *vs = (float *)(base + L.vs);
    double *scal = (double *)(base + L.scal), *rem = (double *)(base + L.rem);
    double *rscal = (double *)(base + L.rscal);
    double *coef = (double *)(base + L.coef), *vbt = (double *)(base + L.vbt);
    int *flag = (int *)(base + L.flag);
    double *H = ctx->host;

    ctx->pending.clear();
    ctx->next_event = 0;
    int tp = prof_begin(ctx);
    if (validate) {
        HIP_TRY(hipMemsetAsync(flag, 0, sizeof(int), s));
        HIP_TRY(launch_validate(X, flag, s));
        int h_flag = 0;
        HIP_TRY(hipMemcpyAsync(&h_flag, flag, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (h_flag)
            return fail(BPCA_ERR_NONCANONICAL,
                        "CSR is not canonical: indptr must start at 0 and be non-decreasing, and each row's column indices strictly increasing and below num_genes");
    }

    // Per-gene clipped standard deviation.
    HIP_TRY(launch_colstats(X, tpart, s));
    HIP_TRY(launch_reduce_groups(tpart, nb, 2 * g, grp, s));
    HIP_TRY(launch_finish_stddev(grp, ngrp, g, n, sd, s));
    prof_end(ctx, BPCA_K_SETUP, tp);

    // Start vector: FP32 Box-Muller draws, normalised in FP64.
    {
        Pcg32 rng(seed);
        double nrm = 0.0;
        for (int i = 0; i < g; i++) {
            double x = (double)rng.normal();
            H[hV0 + i] = x;
            nrm += x * x;
        }
        nrm = std::sqrt(nrm);
        for (int i = 0; i < g; i++) H[hV0 + i] /= nrm;
        HIP_TRY(hipMemcpyAsync(V, H + hV0, sizeof(double) * g, hipMemcpyHostToDevice, s));
    }
    HIP_TRY(launch_init_vs(V, sd, g, vs, s));
    HIP_TRY(hipMemsetAsync(B, 0, sizeof(double) * m * (m + 1), s));

    std::vector<double> Ub((size_t)m * m), Sb(m), Vbt((size_t)m * (m + 1));
    double Smax = 2.4221817809573368e-05;  // FP32 eps^(2/3), pca.pyx:688
    int kc = 0, iteration = 0;
    int64_t steps = 0;
    bool converged = false;
    double svd_seconds = 0.0;
    for (iteration = 0; iteration < maxit; iteration++) {
        for (int j = kc; j < m; j++) {
            if (ctx->cancel.load(std::memory_order_relaxed)) {
                (void)hipStreamSynchronize(s);
                ctx->cancel.store(0);
                return fail(BPCA_ERR_CANCELLED, "solve cancelled");
            }
            tp = prof_begin(ctx);
            HIP_TRY(launch_spmv(X, vs, u32, 0, s));
            prof_end(ctx, BPCA_K_SPMV, tp);
            tp = prof_begin(ctx);
            HIP_TRY(launch_reorth_partials(u32, U, n, j, P, s));
            HIP_TRY(launch_reduce_columns(P, nb, j + 1, red, s));
            HIP_TRY(launch_reorth_apply(u32, U, n, j, red, sq, s));
            HIP_TRY(launch_alpha(sq, nb, scal, B, m, j, s));
            prof_end(ctx, BPCA_K_REORTH, tp);
            tp = prof_begin(ctx);
            HIP_TRY(launch_spmv_t(X, U + (int64_t)j * n, scal, tpart, s));
            prof_end(ctx, BPCA_K_SPMV_T, tp);
            tp = prof_begin(ctx);
            HIP_TRY(launch_reduce_groups(tpart, nb, g, grp, s));
            HIP_TRY(launch_vside(grp, ngrp, g, sd, V, j, m, B, scal, rem, rscal, vs, s));
            prof_end(ctx, BPCA_K_VSIDE, tp);
            steps++;
        }
        HIP_TRY(hipMemcpyAsync(H + hB, B, sizeof(double) * m * (m + 1), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(H + hR, rscal, sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        prof_flush(ctx);
        const auto t0 = std::chrono::steady_clock::now();
        if (bidiag_svd(m, H + hB, Ub.data(), Sb.data(), Vbt.data()) != 0)
            return fail(BPCA_ERR_NUMERICAL, "SVD of the projected matrix did not converge");
        svd_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const double rnorm = H[hR];
        Smax = std::max(Sb[0], Smax);
        int nconv = 0;
        for (int i = 0; i < k; i++) {
            double res = rnorm * Ub[(m - 1) + (size_t)i * m];
            nconv += std::fabs(res) < (double)tol * Smax;
        }
        if (nconv == k) {
            converged = true;
            break;
        }
        if (iteration == maxit - 1) break;  // keep U, Ub consistent: no restart
        kc = std::min(std::max(kc, k + nconv), m - 1);
        std::copy(Ub.begin(), Ub.begin() + (size_t)m * kc, H + hCoef);
        std::copy(Vbt.begin(), Vbt.end(), H + hVbt);
        HIP_TRY(hipMemcpyAsync(coef, H + hCoef, sizeof(double) * m * kc, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(vbt, H + hVbt, sizeof(double) * m * (m + 1), hipMemcpyHostToDevice, s));
        tp = prof_begin(ctx);
        HIP_TRY(launch_tall_gemm_f64(U, n, n, m, coef, kc, U2, n, s));
        std::swap(U, U2);
        HIP_TRY(launch_v_restart(V, g, m, vbt, kc, rem, V2, s));
        std::swap(V, V2);
        prof_end(ctx, BPCA_K_RESTART, tp);
        std::fill(H + hB, H + hB + (size_t)m * (m + 1), 0.0);
        for (int i = 0; i < kc; i++) {
            H[hB + i + (size_t)i * m] = Sb[i];
            H[hB + i + (size_t)kc * m] = rnorm * Ub[(m - 1) + (size_t)i * m];
        }
        HIP_TRY(hipMemcpyAsync(B, H + hB, sizeof(double) * m * (m + 1), hipMemcpyHostToDevice, s));
        HIP_TRY(launch_init_vs(V + (int64_t)kc * g, sd, g, vs, s));
    }

    // PCs = U (Ub[:, :k] diag(Sb[:k])), float32 row-major.
    for (int c = 0; c < k; c++)
        for (int l = 0; l < m; l++)
            H[hCoef + l + (size_t)c * m] = Ub[l + (size_t)c * m] * Sb[c];
    HIP_TRY(hipMemcpyAsync(coef, H + hCoef, sizeof(double) * m * k, hipMemcpyHostToDevice, s));
    tp = prof_begin(ctx);
    HIP_TRY(launch_tall_gemm_rowmajor_f32(U, n, n, m, coef, k, d_PCs, s));
    prof_end(ctx, BPCA_K_FINAL, tp);
    HIP_TRY(hipStreamSynchronize(s));
    prof_flush(ctx);
    if (sv_out)
        for (int c = 0; c < k; c++) sv_out[c] = Sb[c];
    if (st) {
        st->converged = converged ? 1 : 0;
        st->restarts = std::min(iteration + 1, maxit);
        st->lanczos_steps = steps;
        st->matvecs = 2 * steps;
        st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
        st->svd_seconds = svd_seconds;
        st->smax = Smax;
    }
    return BPCA_OK;
}

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

const char *bpca_last_error(void) { return g_last_error.c_str(); }

int bpca_version(void) { return 1 * 10000 + 0 * 100 + 0; }

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
    hipError_t e = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return hip_fail(e, "hipStreamCreateWithFlags");
    }
    ctx->stream = ctx->own_stream;
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
        ctx->max_lds = (size_t)lds;
    *out = ctx;
    return BPCA_OK;
}

int bpca_context_destroy(bpca_context *ctx) {
    if (!ctx) return BPCA_OK;
    {
        DeviceGuard dg(ctx->device);
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        if (ctx->ws) (void)hipFree(ctx->ws);
        if (ctx->host) (void)hipHostFree(ctx->host);
        for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
        if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    }
    {
        std::lock_guard<std::mutex> guard(g_default_lock);
        if (ctx == g_default) g_default = nullptr;
    }
    delete ctx;
    return BPCA_OK;
}

int bpca_context_set_stream(bpca_context *ctx, void *hip_stream) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    ctx->stream = hip_stream ? (hipStream_t)hip_stream : ctx->own_stream;
    return BPCA_OK;
}

int bpca_context_cancel(bpca_context *ctx) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    ctx->cancel.store(1);
    return BPCA_OK;
}

int bpca_context_set_profiling(bpca_context *ctx, int enable) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    std::lock_guard<std::mutex> guard(ctx->lock);
    ctx->profiling = enable != 0;
    return BPCA_OK;
}

int bpca_context_kernel_times(bpca_context *ctx, double *total_ms,
                              int64_t *launches) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    std::lock_guard<std::mutex> guard(ctx->lock);
    for (int i = 0; i < BPCA_NUM_KERNEL_CLASSES; i++) {
        if (total_ms) total_ms[i] = ctx->k_ms[i];
        if (launches) launches[i] = ctx->k_count[i];
    }
    return BPCA_OK;
}

int bpca_context_reset_kernel_times(bpca_context *ctx) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    std::lock_guard<std::mutex> guard(ctx->lock);
    for (int i = 0; i < BPCA_NUM_KERNEL_CLASSES; i++) {
        ctx->k_ms[i] = 0.0;
        ctx->k_count[i] = 0;
    }
    return BPCA_OK;
}

uint64_t bpca_workspace_bytes(uint64_t num_cells, uint32_t num_genes,
                              uint32_t k, uint32_t subspace_size) {
    return make_layout((int64_t)num_cells, (int)num_genes, (int)subspace_size,
                       (int)k)
        .total;
}

int bpca_irlba_device(bpca_context *ctx, const float *d_data,
                      const int32_t *d_indices, const void *d_indptr,
                      int indptr_width, uint64_t num_cells, uint32_t num_genes,
                      uint32_t k, uint32_t subspace_size, float tolerance,
                      uint32_t max_iterations, uint64_t seed, int validate,
                      float *d_PCs, double *singular_values,
                      bpca_solve_stats *stats) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!d_data || !d_indices || !d_indptr || !d_PCs)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "NULL device pointer");
    if (indptr_width != 4 && indptr_width != 8)
        return fail(BPCA_ERR_INVALID_ARGUMENT, "indptr_width must be 4 or 8");
    std::lock_guard<std::mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    bpca::Csr X{d_data, d_indices, d_indptr, indptr_width, (int64_t)num_cells,
                (int)num_genes};
    try {
        return solve(ctx, X, (int)k, (int)subspace_size, tolerance,
                     (int)max_iterations, seed, validate != 0, d_PCs,
                     singular_values, stats);
    } catch (const std::bad_alloc &) {
        return fail(BPCA_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

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
    if (!is_csr && num_threads == 1 && !match_parallel)
        return fail(BPCA_ERR_INVALID_ARGUMENT,
                    "the matvec operand must be the CSR of X; on the serial CSC path pass X.tocsr() as the matvec operand");
    BPCA_TRY(check_shape((int64_t)num_cells, num_genes, k, subspace_size, tolerance, max_iterations));
    bpca_context *ctx = nullptr;
    BPCA_TRY(default_context(&ctx));
    std::lock_guard<std::mutex> guard(ctx->lock);
    DeviceGuard dg(ctx->device);
    const int64_t n = (int64_t)num_cells;
    const int64_t nnz = index_width == 8 ? ((const int64_t *)indptr_matvec)[n]
                                         : (int64_t)((const int32_t *)indptr_matvec)[n];
    if (nnz < 0) return fail(BPCA_ERR_NONCANONICAL, "indptr[num_cells] is negative");
    hipStream_t s = ctx->stream;
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
    bpca::Csr X{d_data.p, d_idx.p, d_ptr.p, index_width, n, (int)num_genes};
    bpca_solve_stats st{};
    int rc;
    try {
        rc = solve(ctx, X, (int)k, (int)subspace_size, tolerance,
                   (int)max_iterations, seed, true, d_pcs.p, nullptr, &st);
    } catch (const std::bad_alloc &) {
        return fail(BPCA_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    if (rc != BPCA_OK) return rc;
    HIP_TRY(hipMemcpyAsync(PCs, d_pcs.p, sizeof(float) * n * k, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (converged) *converged = st.converged;
    return BPCA_OK;
}

int bpca_synth_row_nnz(bpca_context *ctx, uint64_t seed, uint64_t row0,
                       uint64_t num_rows, uint32_t num_genes,
                       const double *d_type_gene_mean,
                       const int32_t *d_cell_type, const double *d_cell_factor,
                       int64_t *d_row_nnz) {
    if (!ctx) return fail(BPCA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    DeviceGuard dg(ctx->device);
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
    DeviceGuard dg(ctx->device);
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
    DeviceGuard dg(ctx->device);
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

}  // extern "C"
