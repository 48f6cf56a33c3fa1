// Device kernels of the IRLBA engine for gfx950 (CDNA4, wave64).
//
// Reference code each kernel replaces (brisc/pca.pyx unless noted):
//   k_colstats, k_setup_finish  clipped_stddev_csc                51-103
//   k_spmv                      matvec_csr_parallel_raw           440-475
//   k_reorth_*, k_alpha         fused passes 1, 2 and alpha       1006-1090
//   k_spmv_t                    rmatvec_csc_parallel_raw          478-511
//                               (computed from the CSR, no CSC copy)
//   k_vside                     V update, CGS, beta / remainder   1103-1171
//   k_svd                       sgesdd + convergence + k_current  514-545,
//                                                                 1178-1202
//   k_tall_gemm<false>,         matmul_nn_f_out + V sgemm + B     578-616,
//   k_restart_v, _commit                                          1204-1231
//   k_tall_gemm<true>           matmul_nn_c_out with Ub*Sb        619-658,
//                                                                 1233-1254
//
// Determinism: no floating-point atomics.  Every sum has a fixed order set by
// the data -- the position inside a row, fixed data blocks of kRowBlock rows,
// fixed sub-blocks of kSubRows rows, fixed groups of kGroup blocks -- never by
// grid size, scheduling or the other problems sharing a launch.  A problem's
// results are therefore bit-identical run to run, alone or in a batch.
#include "kernels.h"

#include <type_traits>

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

// wave_sum with the lane exchanges of strides 16 and below done by DPP moves
// and a swizzle (xor_lane_f64) instead of LDS permutes: the same pairs in the
// same order, so the same bits.  Every lane of the wave must be active.
__device__ inline double wave_sum_dpp(double v) {
    v += __shfl_xor(v, 32);
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) v += xor_lane_f64(v, off);
    return v;
}

// Sum over a block of NW waves: xor butterfly inside each wave, then a
// pairwise tree over the wave sums in wave order.  `red` holds NW doubles.
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

// Rows [first, first + count) of data block `lb` owned by wave w when the
// block is split into 4 sub-blocks of kSubRows rows.
__device__ inline void sub_block(int64_t n, int64_t lb, int w, int64_t &first,
                                 int64_t &count) {
    first = lb * kRowBlock + (int64_t)w * kSubRows;
    count = n - first;
    count = count < 0 ? 0 : (count > kSubRows ? kSubRows : count);
}

// colsum . vs / n, the mean of X vs, for the problem's current vs: thread t
// takes genes t, t + 256, ... in order, then block_sum<4>.  Every thread of
// a 256-thread block calls it after vs is written.
__device__ inline double vs_mean(const Prob &pr, double *tmp) {
    double a = 0.0;
    for (int i = threadIdx.x; i < pr.g; i += 256) a = fma(pr.colsum[i], (double)pr.vs[i], a);
    return block_sum<4>(a, tmp) / (double)pr.n;
}

}  // namespace

// ================================================================ set-up

// Column sums of x (kSquare = false) or x^2 (true) for one data block and
// the gene tile [c0, c0 + tw) (tw <= kGeneTile; one tile when g <= kGeneTile):
// wave w accumulates its sub-block in a private LDS vector (each row's
// column indices are unique, so the 64 lanes touch distinct entries); the
// per-gene sum is the FP32 chain over the sub-block's rows in order, and
// the four sub-blocks combine as (w0 + w1) + (w2 + w3), whatever the tile.
// IX is the width of the packed column indices (problems of the other
// width return at once).
template <bool kSquare, typename IX>
__global__ __launch_bounds__(256) void k_colstats(Tables T, int c0) {
    extern __shared__ float lds[];
    const int2 bp = T.blocks[blockIdx.x];
    Prob &pr = T.probs[bp.x];
    if (pr.phase != PH_SETUP || pr.pair || (pr.mode == kModeWide) != (sizeof(IX) == 4)) return;
    const int g = pr.g;
    if (c0 >= g) return;
    const int tw = g - c0 < kGeneTile ? g - c0 : kGeneTile;
    const int gs = tw + 1;   // accumulator stride: a spare slot at [tw]
    const int lane = threadIdx.x & (kWave - 1);
    const int w = wave_id();
    float *acc = lds + (size_t)w * gs;
    for (int i = lane; i < gs; i += kWave) acc[i] = 0.0f;
    int64_t r0, nrows;
    sub_block(pr.n, bp.y, w, r0, nrows);
    const int64_t *indptr = pr.indptr;
    const IX *indices = (const IX *)pr.idx;
    const float *data = pr.data;
    const int shift = pr.idx_shift;
    for (int64_t rr = 0; rr < nrows; rr++) {
        const int64_t b = indptr[r0 + rr], e = indptr[r0 + rr + 1];
        for (int64_t p = b + lane; p < e; p += kWave) {
            // Genes outside the tile (and pair problems' padding, gene g;
            // their indices hold gene << 2) update the spare slot.
            int c = ((int)indices[p] >> shift) - c0;
            c = (unsigned)c < (unsigned)tw ? c : tw;
            const float x = data[p];
            acc[c] = kSquare ? fmaf(x, x, acc[c]) : acc[c] + x;
        }
    }
    __syncthreads();
    float *out = pr.tpart + (int64_t)bp.y * 2 * g + (kSquare ? g : 0) + c0;
    for (int i = threadIdx.x; i < tw; i += 256)
        out[i] = (lds[i] + lds[gs + i]) + (lds[2 * gs + i] + lds[3 * gs + i]);
}

// Column sums of x and x^2 of a pair problem's data block in one read of its
// padded copy (rows of whole quads, padding of gene g): wave w walks its
// sub-block's rows in order, a quad a lane, and adds each element to the
// (sum, sum of squares) pair of its gene in a private LDS vector (float2 a
// gene, the spare pair at [g]); per gene each sum is the FP32 chain over
// the sub-block's rows in order, and the four sub-blocks combine as
// (w0 + w1) + (w2 + w3).
__global__ __launch_bounds__(256) void k_colstats_pair(Tables T) {
    extern __shared__ float lds[];
    const int2 bp = T.blocks[blockIdx.x];
    Prob &pr = T.probs[bp.x];
    if (pr.phase != PH_SETUP || !pr.pair) return;
    const int g = pr.g, gs = g + 1;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = wave_id();
    float2 *acc = (float2 *)lds + (size_t)w * gs;
    for (int i = lane; i < gs; i += kWave) acc[i] = make_float2(0.0f, 0.0f);
    int64_t r0, nrows;
    sub_block(pr.n, bp.y, w, r0, nrows);
    const int64_t *indptr = pr.indptr;
    const float4 *data = (const float4 *)pr.data;
    const uint2 *idx = (const uint2 *)pr.idx;
    for (int64_t rr = 0; rr < nrows; rr++) {
        const int64_t b = indptr[r0 + rr] >> 2, e = indptr[r0 + rr + 1] >> 2;
        for (int64_t q = b + lane; q < e; q += kWave) {
            const float4 d = data[q];
            const uint2 ix = idx[q];
            const float x[4] = {d.x, d.y, d.z, d.w};
            int c[4];
            float2 a[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const unsigned h = k < 2 ? ix.x : ix.y;
                c[k] = (int)(((k & 1) ? (h >> 16) : (h & 0xffffu)) >> 2);
                a[k] = acc[c[k]];
            }
#pragma unroll
            for (int k = 0; k < 4; k++) acc[c[k]] = make_float2(a[k].x + x[k], fmaf(x[k], x[k], a[k].y));
        }
    }
    __syncthreads();
    const float2 *A = (const float2 *)lds;
    float *out = pr.tpart + (int64_t)bp.y * 2 * g;
    for (int i = threadIdx.x; i < g; i += 256) {
        const float2 a0 = A[i], a1 = A[gs + i], a2 = A[2 * gs + i], a3 = A[3 * gs + i];
        out[i] = (a0.x + a1.x) + (a2.x + a3.x);
        out[g + i] = (a0.y + a1.y) + (a2.y + a3.y);
    }
}

// grp[G*2g + i] = FP64 sum over the data blocks of group G, in block order.
__global__ __launch_bounds__(256) void k_reduce_groups(Tables T, int setup) {
    const int2 gp = T.groups[blockIdx.y];
    Prob &pr = T.probs[gp.x];
    if (pr.phase != (setup ? PH_SETUP : PH_LANCZOS)) return;
    const int g = pr.g;
    const int width = setup ? 2 * g : g;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= width) return;
    const int64_t b0 = (int64_t)gp.y * kGroup;
    const int64_t b1 = b0 + kGroup < pr.nb ? b0 + kGroup : pr.nb;
    const float *src = pr.tpart + i;
    double acc = 0.0;
    int64_t b = b0;
    // Sixteen loads in flight, then their sum in block order.
    for (; b + 16 <= b1; b += 16) {
        float v[16];
#pragma unroll
        for (int t = 0; t < 16; t++) v[t] = src[(b + t) * 2 * g];
#pragma unroll
        for (int t = 0; t < 16; t++) acc += (double)v[t];
    }
    for (; b < b1; b++) acc += (double)src[b * 2 * g];
    pr.grp[(int64_t)gp.y * 2 * g + i] = acc;
}

// sd = max(sqrt((sum x^2 - (sum x)^2/n)/(n-1)), 1e-8) in FP64; vs = V0/sd;
// B = 0; state reset; phase -> Lanczos.
__global__ __launch_bounds__(256) void k_setup_finish(Tables T) {
    __shared__ double tmp[4];
    Prob &pr = T.probs[blockIdx.x];
    if (pr.phase != PH_SETUP) return;
    const int g = pr.g, m = pr.m;
    const double n = (double)pr.n;
    const double *V0 = pr.V[0];
    for (int i = threadIdx.x; i < g; i += 256) {
        double s = 0.0, ss = 0.0;
        for (int64_t G = 0; G < pr.ngrp; G++) {
            s += pr.grp[G * 2 * g + i];
            ss += pr.grp[G * 2 * g + g + i];
        }
        const double var = (ss - s * s / n) / (n - 1.0);
        double sd = sqrt(var > 0.0 ? var : 0.0);
        sd = sd < 1e-8 ? 1e-8 : sd;
        pr.sd[i] = sd;
        pr.colsum[i] = s;
        pr.vs[i] = (float)(V0[i] / sd);
    }
    for (int i = threadIdx.x; i < m * (m + 1); i += 256) pr.B[i] = 0.0;
    __syncthreads();
    const double vm = vs_mean(pr, tmp);
    if (threadIdx.x == 0) {
        pr.vmean = vm;
        pr.j = 0;
        pr.jfirst = 0;
        pr.kc = 0;
        pr.iter = 0;
        pr.ubuf = 0;
        pr.vbuf = 0;
        pr.converged = 0;
        pr.nconv = 0;
        pr.svd_sweeps = 0;
        pr.smax = 2.4221817809573368e-05;  // FP32 eps^(2/3), pca.pyx:688
        pr.steps = 0;
        pr.phase = PH_LANCZOS;
    }
}

// ================================================================ u = X vs
//
// One workgroup per data block; wave w walks its 256-row sub-block.  Each
// lane loads whole aligned quads of 4 elements (16-byte values, 8-byte
// packed indices), so a 256-element window of a row costs two load
// instructions.  Canonical order of a row's sum: the row is cut into windows
// of 64 quads starting at the quad holding its first element; lane l takes
// quad l of each window and accumulates its quad's elements of the row in
// position order with FP32 fma, window after window; then the 64 lane sums
// meet in the xor butterfly 32,16,8,4,2,1.  The order is fixed by the data
// (a row's position modulo 4 included), never by the launch.

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef const f32x4 __attribute__((address_space(1))) gquad;    // 4 values
typedef float __attribute__((address_space(1))) gfloat_w;
typedef const float __attribute__((address_space(1))) gfloat_r;

// The four column indices of an aligned quad: packed to 16 bits (g <= 65535;
// element 0 in the low half of x) or 32 bits.
template <typename IX> struct IxQuad;
template <> struct IxQuad<uint16_t> {
    typedef u32x2 vec;
    typedef const u32x2 __attribute__((address_space(1))) gvec;
    __device__ static __forceinline__ int get(const vec &w, int k) {
        const unsigned h = k < 2 ? w.x : w.y;
        return (int)((k & 1) ? (h >> 16) : (h & 0xffffu));
    }
};
template <> struct IxQuad<uint32_t> {
    typedef u32x4 vec;
    typedef const u32x4 __attribute__((address_space(1))) gvec;
    __device__ static __forceinline__ int get(const vec &w, int k) {
        return (int)(k == 0 ? w.x : k == 1 ? w.y : k == 2 ? w.z : w.w);
    }
};

template <typename IX>
struct Quad {
    f32x4 d;
    typename IxQuad<IX>::vec ix;
};

template <typename IX>
__device__ __forceinline__ int quad_index(const Quad<IX> &c, int k) {
    return IxQuad<IX>::get(c.ix, k);
}

template <typename IX>
__device__ __forceinline__ float quad_value(const Quad<IX> &c, int k) {
    return k == 0 ? c.d.x : k == 1 ? c.d.y : k == 2 ? c.d.z : c.d.w;
}

// Row offsets of rows [r0, r0 + nrows], staged by the wave in its own LDS
// slice `off` as positions relative to the quad holding indptr[r0] (so
// off[0] = indptr[r0] mod 4), so that a row's bounds never wait behind the
// wave's outstanding data loads.  Returns that quad's index (nrows > 0).
__device__ inline int64_t stage_row_offsets(const int64_t *__restrict__ indptr,
                                            int64_t r0, int nrows, int lane,
                                            int *off) {
    const int64_t q0 = indptr[r0] >> 2;
    for (int q = lane; q <= nrows; q += kWave) off[q] = (int)(indptr[r0 + q] - 4 * q0);
    __builtin_amdgcn_wave_barrier();
    return q0;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ int64_t uniform64(int64_t v) { return (int64_t)uniform64((uint64_t)v); }

template <typename T>
__device__ __forceinline__ T *uniform_ptr(T *p) {
    return (T *)(uintptr_t)uniform64((uint64_t)(uintptr_t)p);
}

// A wave walks its sub-block as a stream of segments: a row is cut into
// windows of 64 quads (an empty row is one empty segment).  Segments are
// taken R at a time; the next group's loads are issued before the current
// group is reduced, so R segments of loads are always in flight per wave,
// and no load sits behind a branch (lanes past a segment's end re-load its
// last quad, a cache line the segment already touches; positions are tested
// before use).  A row's per-lane partial sum runs across its windows in
// order, so the order of summation is the one above for every R; R (2, 4
// or 8) is chosen by the autotuner.

constexpr int kWindow = 4 * kWave;  // elements per full window

// Uniform walk state: row r, its next position p and its end e.
struct Cursor {
    int r, p, e;
};

template <int R, typename IX>
struct SegGroup {
    Quad<IX> c[R];
    int b[R], e[R];  // segment bounds (positions relative to the first quad)
    int q0[R];       // first quad of the segment's window
    int row[R];      // row of the segment (nrows past the last row)
    bool fin[R];     // the segment ends row `row`
};

// Forms the next R segments from `cs` and issues their loads (non-temporal
// when NT: measured faster for X v, slightly slower for X^T u).
template <int R, bool NT, typename IX>
__device__ __forceinline__ void form_group(const int *off, int nrows,
                                           gquad *data,
                                           typename IxQuad<IX>::gvec *indices,
                                           int lane, Cursor &cs,
                                           SegGroup<R, IX> &G) {
    int W[R + 1];  // ends of rows cs.r + 1 ... cs.r + R + 1
#pragma unroll
    for (int i = 0; i <= R; i++) {
        const int r = cs.r + 2 + i;
        W[i] = uniform(off[r < nrows ? r : nrows]);
    }
#pragma unroll
    for (int q = 0; q < R; q++) {
        const int q0 = cs.p >> 2;
        const int wend = 4 * (q0 + kWave);
        const int se = wend < cs.e ? wend : cs.e;
        const bool fin = wend >= cs.e;
        const int lastq = (se > 0 ? se - 1 : 0) >> 2;
        G.b[q] = cs.p;
        G.e[q] = se;
        G.q0[q] = q0;
        G.row[q] = cs.r;
        G.fin[q] = fin && cs.r < nrows;
        const int qi = q0 + lane < lastq ? q0 + lane : lastq;
        if (NT) {
            G.c[q].d = __builtin_nontemporal_load(data + qi);
            G.c[q].ix = __builtin_nontemporal_load(indices + qi);
        } else {
            G.c[q].d = data[qi];
            G.c[q].ix = indices[qi];
        }
        if (fin) {
            cs.r = cs.r < nrows ? cs.r + 1 : nrows;
            cs.p = cs.e;
            cs.e = W[0];
#pragma unroll
            for (int i = 0; i < R; i++) W[i] = W[i + 1];
        } else {
            cs.p = wend;
        }
    }
}

// Partial sums continue across a row's segments (`carry` holds the open
// row's); the rows that end in the group are reduced together
// (store_row_totals).  Branch-free: every lane gathers all four v entries
// (a quad's indices are all below g, padding included) and elements outside
// the segment enter the chain as 0 * v.
template <int R, typename IX, bool GV>
__device__ __forceinline__ void spmv_reduce(const SegGroup<R, IX> &G,
                                            const float *lds_v, gfloat_r *gv,
                                            int lane, float &carry,
                                            gfloat_w *u, gfloat_w *sink);

// The rows of a group's slots are consecutive, so slot q's row is the first
// slot's row plus the finished slots before q.  Every lane stores -- lanes
// without a finished row into the sink slot -- so the store is not behind a
// branch and the waits for loads stay fixed counts.
//
// Halving butterfly: at offsets 32, 16, ... (log2 R steps) each lane keeps
// half of its rows and adds its partner's share of them, so after those
// steps lane l holds a partial of row l / (64 / R); the remaining offsets
// finish each row within its 64 / R lanes.  Every row is summed in the
// xor-butterfly tree 32, 16, 8, 4, 2, 1 (floating addition is commutative,
// so the bits are those of a plain per-row butterfly) with R - 1 + 6 -
// log2 R exchanges instead of 6 R.  kDpp: exchanges at offsets up to 16 by
// DPP moves and a swizzle (xor_lane) instead of LDS permutes, the same pairs.
//
template <int R, bool kDpp>
__device__ __forceinline__ void store_row_totals(float (&done)[R], const bool (&fin)[R], int row0,
                                                 int lane, gfloat_w *u, gfloat_w *sink) {
    auto xchg = [&](float v, int off) -> float {
        if (kDpp && off <= 16) return __int_as_float(xor_lane(__float_as_int(v), off));
        return __shfl_xor(v, off);
    };
    constexpr int kLanesPerRow = kWave / R;
    int off2 = 32;
#pragma unroll
    for (int half = R / 2; half >= 1; half >>= 1, off2 >>= 1) {
        const bool upper = (lane & off2) != 0;
#pragma unroll
        for (int i = 0; i < half; i++) {
            const float keep = upper ? done[i + half] : done[i];
            const float send = upper ? done[i] : done[i + half];
            done[i] = keep + xchg(send, off2);
        }
    }
#pragma unroll
    for (; off2 > 0; off2 >>= 1) done[0] += xchg(done[0], off2);
    unsigned finmask = 0;
#pragma unroll
    for (int q = 0; q < R; q++) finmask |= (unsigned)fin[q] << q;
    const int q = lane / kLanesPerRow;
    const bool wr = (lane % kLanesPerRow) == 0 && ((finmask >> q) & 1u);
    const int row = row0 + __popc(finmask & ((1u << q) - 1u));
    *(wr ? u + row : sink) = done[0];
}

template <int R, typename IX, bool GV>
__device__ __forceinline__ void spmv_reduce(const SegGroup<R, IX> &G,
                                            const float *lds_v, gfloat_r *gv,
                                            int lane, float &carry,
                                            gfloat_w *u, gfloat_w *sink) {
    float done[R];
    float acc = carry;
#pragma unroll
    for (int q = 0; q < R; q++) {
        const int a0 = 4 * (G.q0[q] + lane);
        float vv[4];
#pragma unroll
        for (int k = 0; k < 4; k++) vv[k] = GV ? gv[quad_index(G.c[q], k)] : lds_v[quad_index(G.c[q], k)];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const bool in = a0 + k >= G.b[q] && a0 + k < G.e[q];
            acc = fmaf(in ? quad_value(G.c[q], k) : 0.0f, vv[k], acc);
        }
        done[q] = acc;
        acc = G.fin[q] ? 0.0f : acc;
    }
    carry = acc;
    store_row_totals<R, false>(done, G.fin, G.row[0], lane, u, sink);
}

}  // namespace

// LDS of k_spmv: vs (g floats; none when v is read from global memory,
// g > kMaxLdsV), then 4 wave slices of kSubRows + 1 offsets.
size_t spmv_lds_bytes(int g) {
    return sizeof(float) * (size_t)(g <= kMaxLdsV ? g : 0) + sizeof(int) * 4 * (size_t)(kSubRows + 1);
}

// LDS of k_spmv_t: 4 wave accumulators of tw + 1 floats for a gene tile of
// tw = min(g, kGeneTile) genes (acc[tw] is a spare slot), 4 slices of
// kSubRows u values, 4 slices of kSubRows + 1 offsets.
size_t spmv_t_lds_bytes(int g) {
    const size_t tw = g < kGeneTile ? g : kGeneTile;
    return sizeof(float) * ((size_t)4 * (tw + 1) + 4 * kSubRows) +
           sizeof(int) * 4 * (size_t)(kSubRows + 1);
}

// Two segment groups alternate without register copies.  The loop has one
// exit, after reducing the group that finished the last row (groups formed
// past it are empty: their loads re-read the sub-block's last quad and their
// reductions write nothing).
// MODE: the problem's index width and v source (kModeLds: 16-bit indices,
// v staged in LDS; kModeGlobalV: 16-bit, v gathered from global memory, for
// g > kMaxLdsV; kModeWide: 32-bit indices, v from global memory).  Problems
// of another mode return at once.  Every mode sums in the same order.
template <int R, int MODE>
__global__ __launch_bounds__(256) void k_spmv(Tables T) {
    typedef typename std::conditional<MODE == kModeWide, uint32_t, uint16_t>::type IX;
    constexpr bool GV = MODE != kModeLds;
    extern __shared__ float lds_v[];
    const int2 bp = T.blocks[blockIdx.x];
    Prob &pr = T.probs[bp.x];
    if (pr.phase != PH_LANCZOS || pr.mode != MODE || pr.pair) return;
    const int g = pr.g;
    if (!GV)
        for (int i = threadIdx.x; i < g; i += 256) lds_v[i] = pr.vs[i];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = wave_id();
    int *off = (int *)(lds_v + (GV ? 0 : g)) + w * (kSubRows + 1);
    int64_t r0, nrows64;
    sub_block(pr.n, bp.y, w, r0, nrows64);
    const int nrows = (int)nrows64;
    const int64_t q0 = nrows > 0 ? stage_row_offsets(pr.indptr, r0, nrows, lane, off) : 0;
    __syncthreads();
    float *uj = u_of_step(pr, pr.j);
    if (nrows == 0 || off[nrows] == off[0]) {
        for (int q = lane; q < nrows; q += kWave) uj[r0 + q] = 0.0f;
        return;
    }
    gquad *data = (gquad *)pr.data + q0;
    typename IxQuad<IX>::gvec *indices = (typename IxQuad<IX>::gvec *)pr.idx + q0;
    gfloat_r *gv = (gfloat_r *)pr.vs;
    gfloat_w *u = (gfloat_w *)(uj + r0);
    gfloat_w *sink = (gfloat_w *)(uj + pr.n);
    Cursor cs{0, uniform(off[0]), uniform(off[1])};
    SegGroup<R, IX> ga, gb;
    float carry = 0.0f;
    form_group<R, true, IX>(off, nrows, data, indices, lane, cs, ga);
    for (;;) {
        form_group<R, true, IX>(off, nrows, data, indices, lane, cs, gb);
        __builtin_amdgcn_sched_barrier(0);
        spmv_reduce<R, IX, GV>(ga, lds_v, gv, lane, carry, u, sink);
        const bool b_last = cs.r >= nrows;  // gb finished the last row
        form_group<R, true, IX>(off, nrows, data, indices, lane, cs, ga);
        __builtin_amdgcn_sched_barrier(0);
        spmv_reduce<R, IX, GV>(gb, lds_v, gv, lane, carry, u, sink);
        if (b_last) break;
    }
}

// ============================================ reorthogonalisation (FP64)
//
// Block partials: lane l of a wave handles the row pairs (2l + 128t,
// 2l + 128t + 1), t = 0..7, of the data block in that order -- one 16-byte
// load per pair when the block is full and n is even (every column of U then
// starts 16-byte aligned), element loads otherwise, same order -- then the
// xor butterfly.  P[c*nb + b]: c = 0 the block sum of u, c >= 1 the dot of
// U[:, c-1] with u.

typedef double v2d __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(256) void k_reorth_partials(Tables T) {
    constexpr int R = kRowBlock / (2 * kWave);  // 8 row pairs per lane
    const int2 bp = T.blocks[blockIdx.x];
    Prob &pr = T.probs[bp.x];
    if (pr.phase != PH_LANCZOS || pr.onepass) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = wave_id();
    const int64_t n = pr.n, b = bp.y, r0 = b * kRowBlock;
    const int64_t nrows = n - r0 < kRowBlock ? n - r0 : kRowBlock;
    const bool vec = nrows == kRowBlock && (n & 1) == 0;
    const int j = pr.j;
    const float *u32 = u_of_step(pr, j);
    const double *U = pr.U[pr.ubuf];
    double ua[R], ub[R];
#pragma unroll
    for (int t = 0; t < R; t++) {
        const int r = 2 * lane + t * 2 * kWave;
        ua[t] = r < nrows ? (double)u32[r0 + r] : 0.0;
        ub[t] = r + 1 < nrows ? (double)u32[r0 + r + 1] : 0.0;
    }
    int c0 = w;
    if (w == 0) {
        double acc = 0.0;
#pragma unroll
        for (int t = 0; t < R; t++) {
            acc += ua[t];
            acc += ub[t];
        }
        acc = wave_sum_dpp(acc);
        if (lane == 0) pr.P[b] = acc;
        c0 = 4;
    }
    if (vec) {
        // Two columns (c, c + 4) per pass: 16 loads in flight per wave.
        for (int c = c0; c <= j; c += 8) {
            const bool two = c + 4 <= j;
            const double *Uc = U + (int64_t)(c - 1) * n + r0 + 2 * lane;
            const double *Ud = two ? Uc + 4 * n : Uc;
            v2d x[R], y[R];
#pragma unroll
            for (int t = 0; t < R; t++) x[t] = __builtin_nontemporal_load((const v2d *)(Uc + t * 2 * kWave));
            if (two) {
#pragma unroll
                for (int t = 0; t < R; t++) y[t] = __builtin_nontemporal_load((const v2d *)(Ud + t * 2 * kWave));
            }
            double acc = 0.0, acd = 0.0;
#pragma unroll
            for (int t = 0; t < R; t++) {
                acc = fma(x[t].x, ua[t], acc);
                acc = fma(x[t].y, ub[t], acc);
            }
            acc = wave_sum_dpp(acc);
            if (lane == 0) pr.P[(int64_t)c * pr.nb + b] = acc;
            if (two) {
#pragma unroll
                for (int t = 0; t < R; t++) {
                    acd = fma(y[t].x, ua[t], acd);
                    acd = fma(y[t].y, ub[t], acd);
                }
                acd = wave_sum_dpp(acd);
                if (lane == 0) pr.P[(int64_t)(c + 4) * pr.nb + b] = acd;
            }
        }
        return;
    }
    for (int c = c0; c <= j; c += 4) {
        const double *Uc = U + (int64_t)(c - 1) * n + r0;
        double acc = 0.0;
#pragma unroll
        for (int t = 0; t < R; t++) {
            const int r = 2 * lane + t * 2 * kWave;
            const double x0 = r < nrows ? Uc[r] : 0.0;
            const double x1 = r + 1 < nrows ? Uc[r + 1] : 0.0;
            acc = fma(x0, ua[t], acc);
            acc = fma(x1, ub[t], acc);
        }
        acc = wave_sum_dpp(acc);
        if (lane == 0) pr.P[(int64_t)c * pr.nb + b] = acc;
    }
}

// red[c] = sum over blocks of P[c*nb + b]: thread t takes blocks t, t+256,
// ... in order, then block_sum<4>.
__global__ __launch_bounds__(256) void k_reduce_columns(Tables T) {
    __shared__ double tmp[4];
    const int cols = T.max_m + 1;
    Prob &pr = T.probs[blockIdx.x / cols];
    const int c = blockIdx.x % cols;
    if (pr.phase != PH_LANCZOS || c > pr.j) return;
    const double *p = pr.P + (int64_t)c * pr.nb;
    double acc = 0.0;
#pragma unroll 8
    for (int64_t b = threadIdx.x; b < pr.nb; b += 256) acc += p[b];
    acc = block_sum<4>(acc, tmp);
    if (threadIdx.x == 0) pr.red[c] = acc;
}

// U[:, j] = u - (mean + sum_c U[:, c] red[c+1]) with the correction summed
// over c ascending; thread t handles the row pairs (r0 + 2t + 512i,
// r0 + 2t + 512i + 1), i = 0, 1 (16-byte loads and stores when the block is
// full and n is even), and the block's sum and sum of squares are
// block_sum<4> of the thread sums (rows in that order): sq[b] and sq[nb + b].
// The sum lets k_alpha remove the mean that the correction itself leaves
// (the U columns are zero-mean only to rounding, and X^T u below assumes a
// zero-mean u; uncorrected, that mean grows by about beta / alpha per step,
// which exceeds 1 when there are fewer cells than genes).
__global__ __launch_bounds__(256) void k_reorth_apply(Tables T) {
    extern __shared__ double coef[];
    __shared__ double tmp[4];
    const int2 bp = T.blocks[blockIdx.x];
    Prob &pr = T.probs[bp.x];
    if (pr.phase != PH_LANCZOS || pr.onepass) return;
    const int j = pr.j;
    const int64_t n = pr.n;
    for (int i = threadIdx.x; i <= j; i += 256) coef[i] = pr.red[i];
    __syncthreads();
    const double mean = coef[0] / (double)n;
    const int64_t r0 = (int64_t)bp.y * kRowBlock;
    double *U = pr.U[pr.ubuf];
    double *Uj = U + (int64_t)j * n;
    const float *u32 = u_of_step(pr, j);
    double acc = 0.0, sum = 0.0;
    const bool vec = r0 + kRowBlock <= n && (n & 1) == 0;
    static_assert(kRowBlock == 1024, "two 512-row halves per block");
    if (vec) {
        // Both halves advance together: two 16-byte loads per column.
        const int64_t ra = r0 + 2 * threadIdx.x, rb = ra + 512;
        double d0 = mean, d1 = mean, d2 = mean, d3 = mean;
#pragma unroll 4
        for (int c = 0; c < j; c++) {
            const v2d xa = __builtin_nontemporal_load((const v2d *)(U + (int64_t)c * n + ra));
            const v2d xb = __builtin_nontemporal_load((const v2d *)(U + (int64_t)c * n + rb));
            const double k = coef[c + 1];
            d0 = fma(xa.x, k, d0);
            d1 = fma(xa.y, k, d1);
            d2 = fma(xb.x, k, d2);
            d3 = fma(xb.y, k, d3);
        }
        const double x0 = (double)u32[ra] - d0, x1 = (double)u32[ra + 1] - d1;
        const double x2 = (double)u32[rb] - d2, x3 = (double)u32[rb + 1] - d3;
        *(v2d *)(Uj + ra) = v2d{x0, x1};
        *(v2d *)(Uj + rb) = v2d{x2, x3};
        acc = fma(x0, x0, acc);
        acc = fma(x1, x1, acc);
        acc = fma(x2, x2, acc);
        acc = fma(x3, x3, acc);
        sum = ((x0 + x1) + x2) + x3;
    } else {
#pragma unroll
        for (int t = 0; t < kRowBlock / 512; t++) {
            const int64_t r = r0 + 2 * threadIdx.x + t * 512;
#pragma unroll
            for (int e = 0; e < 2; e++) {
                if (r + e < n) {
                    double d = mean;
                    for (int c = 0; c < j; c++) d = fma(U[(int64_t)c * n + r + e], coef[c + 1], d);
                    const double x = (double)u32[r + e] - d;
                    Uj[r + e] = x;
                    acc = fma(x, x, acc);
                    sum += x;
                }
            }
        }
    }
    acc = block_sum<4>(acc, tmp);
    sum = block_sum<4>(sum, tmp);
    if (threadIdx.x == 0) {
        pr.sq[bp.y] = acc;
        pr.sq[pr.nb + bp.y] = sum;
    }
}

// One-pass reorthogonalisation (problems with Prob::onepass).  The two
// passes above read U[:, :j] twice per Lanczos step: once for the dots h =
// U^T u, once for u - U h.  Here step j reads it once and does two things
// with each element it loads:
//  * it completes the previous step's column, left pending by that step,
//      U[:, j-1] = (u_{j-1} - mean_{j-1} - U[:, :j-1] h_{j-1}) / alpha_{j-1},
//    whose coefficients are still in place (red: red[0] = the sum of u_{j-1},
//    red[c] = U[:, c-1] . u_{j-1}; inv_alpha), and
//  * it forms this step's partials: P[c] = U[:, c-1] . u_j for c = 1..j (the
//    one for the column just completed from its values in registers), the
//    block sum P[0] and the block's sum of squares sq[b] of u_j.
// k_alpha then has alpha_j = |u_j - mean - U h|  from |u - mean|^2 - |h|^2
// (U orthonormal), and X^T U[:, j] is formed in gene space (k_vside_w) from
// X^T (u_j - mean) and R = X^T U[:, :j].  A column is pending when j >
// jfirst (the first step of a sweep reads the restarted U, all of it
// written).  The last column of a sweep is never completed by a pass: the
// restart and final products take it from (u - mean) / alpha, written by
// X^T u at that step, through their coefficients (restart_coef).
// The arithmetic per element is the two-pass one, in another order; only the
// gene-space X^T U carries the previous columns' FP32 rounding forward, by a
// factor |h| / alpha per step, about sqrt(g / n) -- hence n >= 8 g.
//
// Layout: wave w takes the columns c = w (mod 4) (wave 0: 4, 8, ...; it also
// sums u), two per pass, each over the block's 1024 rows as 8 row pairs per
// lane (lane l: rows 2 l + 128 t, + 1); each lane keeps the correction
// U h of its 16 rows summed over its wave's columns in ascending order; the
// four waves' corrections meet in LDS as (w0 + w1) + (w2 + w3); thread t
// then completes rows (2 t, 2 t + 1) and (512 + 2 t, 513 + 2 t).
__global__ __launch_bounds__(256) void k_reorth_onepass(Tables T) {
    constexpr int R = kRowBlock / (2 * kWave);  // 8 row pairs per lane
    __shared__ double cbuf[4][kRowBlock];
    __shared__ double tmp[4];
    const int2 bp = T.blocks[blockIdx.x];
    Prob &pr = T.probs[bp.x];
    if (!pr.onepass) return;
    if (pr.phase != PH_LANCZOS) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = wave_id();
    const int j = pr.j;
    const bool pend = j > pr.jfirst;
    const int pc = j - 1;              // the pending column
    const int nm = pend ? pc : j;      // columns read: U[:, :nm]
    const int64_t n = pr.n, b = bp.y, r0 = b * kRowBlock, nb = pr.nb;
    const int nrows = (int)(n - r0 < kRowBlock ? n - r0 : kRowBlock);
    const bool vec = nrows == kRowBlock && (n & 1) == 0;
    const double *U = pr.U[pr.ubuf];
    const double *h = pr.red;
    const float *ucur = u_of_step(pr, j);
    double ua[R], ub[R], ca[R], cb[R];
#pragma unroll
    for (int t = 0; t < R; t++) {
        const int r = 2 * lane + t * 2 * kWave;
        ua[t] = r < nrows ? (double)ucur[r0 + r] : 0.0;
        ub[t] = r + 1 < nrows ? (double)ucur[r0 + r + 1] : 0.0;
        ca[t] = 0.0;
        cb[t] = 0.0;
    }
    int c0 = w;
    if (w == 0) {
        double s = 0.0, ss = 0.0;
#pragma unroll
        for (int t = 0; t < R; t++) {
            s += ua[t];
            s += ub[t];
            ss = fma(ua[t], ua[t], ss);
            ss = fma(ub[t], ub[t], ss);
        }
        s = wave_sum_dpp(s);
        ss = wave_sum_dpp(ss);
        if (lane == 0) {
            pr.P[b] = s;
            pr.sq[b] = ss;
        }
        c0 = 4;
    }
    if (vec) {
        for (int c = c0; c <= nm; c += 8) {
            const bool two = c + 4 <= nm;
            const double *Uc = U + (int64_t)(c - 1) * n + r0 + 2 * lane;
            const double *Ud = two ? Uc + 4 * n : Uc;
            v2d x[R], y[R];
#pragma unroll
            for (int t = 0; t < R; t++) x[t] = __builtin_nontemporal_load((const v2d *)(Uc + t * 2 * kWave));
            if (two) {
#pragma unroll
                for (int t = 0; t < R; t++) y[t] = __builtin_nontemporal_load((const v2d *)(Ud + t * 2 * kWave));
            }
            double acc = 0.0;
#pragma unroll
            for (int t = 0; t < R; t++) {
                acc = fma(x[t].x, ua[t], acc);
                acc = fma(x[t].y, ub[t], acc);
            }
            acc = wave_sum_dpp(acc);
            if (lane == 0) pr.P[(int64_t)c * nb + b] = acc;
            if (pend) {
                const double hc = h[c];
#pragma unroll
                for (int t = 0; t < R; t++) {
                    ca[t] = fma(x[t].x, hc, ca[t]);
                    cb[t] = fma(x[t].y, hc, cb[t]);
                }
            }
            if (two) {
                double acd = 0.0;
#pragma unroll
                for (int t = 0; t < R; t++) {
                    acd = fma(y[t].x, ua[t], acd);
                    acd = fma(y[t].y, ub[t], acd);
                }
                acd = wave_sum_dpp(acd);
                if (lane == 0) pr.P[(int64_t)(c + 4) * nb + b] = acd;
                if (pend) {
                    const double hd = h[c + 4];
#pragma unroll
                    for (int t = 0; t < R; t++) {
                        ca[t] = fma(y[t].x, hd, ca[t]);
                        cb[t] = fma(y[t].y, hd, cb[t]);
                    }
                }
            }
        }
    } else {
        for (int c = c0; c <= nm; c += 4) {
            const double *Uc = U + (int64_t)(c - 1) * n + r0;
            const double hc = pend ? h[c] : 0.0;
            double acc = 0.0;
#pragma unroll
            for (int t = 0; t < R; t++) {
                const int r = 2 * lane + t * 2 * kWave;
                const double x0 = r < nrows ? Uc[r] : 0.0;
                const double x1 = r + 1 < nrows ? Uc[r + 1] : 0.0;
                acc = fma(x0, ua[t], acc);
                acc = fma(x1, ub[t], acc);
                ca[t] = fma(x0, hc, ca[t]);
                cb[t] = fma(x1, hc, cb[t]);
            }
            acc = wave_sum_dpp(acc);
            if (lane == 0) pr.P[(int64_t)c * nb + b] = acc;
        }
    }
    if (!pend) return;
#pragma unroll
    for (int t = 0; t < R; t++) {
        const int r = 2 * lane + t * 2 * kWave;
        cbuf[w][r] = ca[t];
        cbuf[w][r + 1] = cb[t];
    }
    __syncthreads();
    const double mean = h[0] / (double)n, ia = pr.inv_alpha;
    const float *uprev = u_of_step(pr, pc);
    double *Up = pr.U[pr.ubuf] + (int64_t)pc * n + r0;
    double d = 0.0;
#pragma unroll
    for (int half = 0; half < 2; half++) {
        const int r = 2 * threadIdx.x + 512 * half;
        double x[2], uc[2];
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const int re = r + e;
            const double cr = (cbuf[0][re] + cbuf[1][re]) + (cbuf[2][re] + cbuf[3][re]);
            x[e] = re < nrows ? ((double)uprev[r0 + re] - mean - cr) * ia : 0.0;
            uc[e] = re < nrows ? (double)ucur[r0 + re] : 0.0;
        }
        if (vec) {
            *(v2d *)(Up + r) = v2d{x[0], x[1]};
        } else {
#pragma unroll
            for (int e = 0; e < 2; e++)
                if (r + e < nrows) Up[r + e] = x[e];
        }
        d = fma(x[0], uc[0], d);
        d = fma(x[1], uc[1], d);
    }
    d = block_sum<4>(d, tmp);
    if (threadIdx.x == 0) pr.P[(int64_t)j * nb + b] = d;
}

// mean = (sum of block sums) / n, removed from U[:, j] when k_spmv_t
// normalises it; alpha = max(sqrt(sum of squares - n mean^2), 1e-8), the
// norm of the centred column; B[j, j] = alpha.  One-pass problems: the sums
// are u's (red[0], sq[b]), and alpha^2 = |u - mean|^2 - sum_c red[c]^2, the
// squared norm of u - mean - U h for orthonormal U; mean is u's.
__global__ __launch_bounds__(256) void k_alpha(Tables T) {
    __shared__ double tmp[4];
    Prob &pr = T.probs[blockIdx.x];
    if (pr.phase != PH_LANCZOS) return;
    const bool one = pr.onepass != 0;
    double acc = 0.0, sum = 0.0;
#pragma unroll 8
    for (int64_t b = threadIdx.x; b < pr.nb; b += 256) {
        acc += pr.sq[b];
        if (!one) sum += pr.sq[pr.nb + b];
    }
    acc = block_sum<4>(acc, tmp);
    double hh = 0.0;
    if (one) {
        // |h|^2: thread t takes red[1 + t], red[257 + t], ...
        for (int c = 1 + threadIdx.x; c <= pr.j; c += 256) hh = fma(pr.red[c], pr.red[c], hh);
        hh = block_sum<4>(hh, tmp);
    } else {
        sum = block_sum<4>(sum, tmp);
    }
    if (threadIdx.x == 0) {
        if (one) sum = pr.red[0];
        const double mean = sum / (double)pr.n;
        double ss = acc - sum * mean;
        if (one) ss -= hh;
        double alpha = sqrt(ss > 0.0 ? ss : 0.0);
        if (alpha < 1e-8) alpha = 1e-8;
        pr.alpha = alpha;
        pr.inv_alpha = 1.0 / alpha;
        pr.umean = mean;
        pr.B[(int64_t)pr.j + (int64_t)pr.j * pr.m] = alpha;
    }
}

// ================================================================ X^T u

namespace {

// acc[col - c0] += x * u[row] for the segments of a loaded group, in row
// order, for the columns of the gene tile [c0, c0 + tw).
template <int R, typename IX, bool kTiled>
__device__ __forceinline__ void spmv_t_reduce(const SegGroup<R, IX> &G,
                                              const float *uval, int lane,
                                              int c0, int tw, float *acc) {
#pragma unroll
    for (int q = 0; q < R; q++) {
        // Branch-free: elements outside the segment or the tile update the
        // spare slot acc[tw] instead, so the lanes of one LDS update still
        // touch distinct live genes (a row's indices are unique).
        const float ur = uval[G.row[q] < kSubRows ? G.row[q] : 0];
        const int a0 = 4 * (G.q0[q] + lane);
        int c[4];
        float x[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int cc = quad_index(G.c[q], k) - c0;
            const bool in = a0 + k >= G.b[q] && a0 + k < G.e[q] && (!kTiled || (unsigned)cc < (unsigned)tw);
            c[k] = in ? cc : tw;
            x[k] = acc[c[k]];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) acc[c[k]] = fmaf(quad_value(G.c[q], k), ur, x[k]);
    }
}

}  // namespace

// One workgroup per data block; wave w owns a 256-row sub-block and a
// private LDS accumulator of g floats.  u = (U[:, j] - mean) / alpha is
// written back centred and normalised (one-pass problems: u = u_j - mean of
// the step's raw vector, U untouched).  Per gene the sub-block sum is the FP32 fma chain over its rows
// in order (a row's column indices are unique, so the lanes of one LDS update
// touch distinct genes); sub-blocks combine as (w0+w1)+(w2+w3).  Loads are
// the segment groups of k_spmv.
//
// With g > kGeneTile the genes go in tiles of kGeneTile, one launch per tile
// (c0 = its first gene): every launch reads the block's rows again and keeps
// the entries of its tile, so each gene's sum is the same chain whatever the
// tiling.  The launch with c0 = 0 normalises U[:, j]; later tiles read it.
template <int R, typename IX, bool kTiled>
__global__ __launch_bounds__(256) void k_spmv_t(Tables T, int c0) {
    extern __shared__ float lds[];
    const int2 bp = T.blocks[blockIdx.x];
    Prob &pr = T.probs[bp.x];
    if (pr.phase != PH_LANCZOS || pr.tsorted || pr.pair || (pr.mode == kModeWide) != (sizeof(IX) == 4)) return;
    const int g = pr.g;
    if (c0 >= g) return;
    const int tw = g - c0 < kGeneTile ? g - c0 : kGeneTile;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = wave_id();
    const int gs = tw + 1;  // accumulator stride (spare slot at [tw])
    float *acc = lds + (size_t)w * gs;
    float *uval = lds + (size_t)4 * gs + (size_t)w * kSubRows;
    int *off = (int *)(lds + (size_t)4 * gs + 4 * kSubRows) + w * (kSubRows + 1);
    for (int i = lane; i < gs; i += kWave) acc[i] = 0.0f;
    int64_t r0, nrows64;
    sub_block(pr.n, bp.y, w, r0, nrows64);
    const int nrows = (int)nrows64;
    double *Uj = pr.U[pr.ubuf] + (int64_t)pr.j * pr.n;
    const double inv_alpha = pr.inv_alpha, umean = pr.umean;
    if (pr.onepass) {
        // u - mean of this step; U[:, j] is written by the next pass, except
        // at the last step of a sweep, whose column the restart and final
        // products form from (u - mean) / alpha (restart_coef).
        const float *uj = u_of_step(pr, pr.j);
        const bool last = pr.j == pr.m - 1 && c0 == 0;
        for (int t = lane; t < nrows; t += kWave) {
            const double x = (double)uj[r0 + t] - umean;
            uval[t] = (float)x;
            if (last) Uj[r0 + t] = x * inv_alpha;
        }
    } else if (c0 == 0) {
        for (int t = lane; t < nrows; t += kWave) {
            const double x = (Uj[r0 + t] - umean) * inv_alpha;
            Uj[r0 + t] = x;
            uval[t] = (float)x;
        }
    } else {
        for (int t = lane; t < nrows; t += kWave) uval[t] = (float)Uj[r0 + t];
    }
    const int64_t q0 = nrows > 0 ? stage_row_offsets(pr.indptr, r0, nrows, lane, off) : 0;
    __syncthreads();
    if (nrows > 0 && off[nrows] > off[0]) {
        gquad *data = (gquad *)pr.data + q0;
        typename IxQuad<IX>::gvec *indices = (typename IxQuad<IX>::gvec *)pr.idx + q0;
        Cursor cs{0, uniform(off[0]), uniform(off[1])};
        SegGroup<R, IX> ga, gb;
        form_group<R, false, IX>(off, nrows, data, indices, lane, cs, ga);
        for (;;) {
            form_group<R, false, IX>(off, nrows, data, indices, lane, cs, gb);
            __builtin_amdgcn_sched_barrier(0);
            spmv_t_reduce<R, IX, kTiled>(ga, uval, lane, c0, tw, acc);
            const bool b_last = cs.r >= nrows;  // gb finished the last row
            form_group<R, false, IX>(off, nrows, data, indices, lane, cs, ga);
            __builtin_amdgcn_sched_barrier(0);
            spmv_t_reduce<R, IX, kTiled>(gb, uval, lane, c0, tw, acc);
            if (b_last) break;
        }
    }
    __syncthreads();
    float *out = pr.tpart + (int64_t)bp.y * 2 * g + c0;
    for (int i = threadIdx.x; i < tw; i += 256)
        out[i] = (lds[i] + lds[gs + i]) + (lds[2 * gs + i] + lds[3 * gs + i]);
}

// ====================================== X v and X^T u from one read of X
//
// One-pass problems with 16-bit indices and g <= kMaxPairGenes (Prob::pair)
// form both sparse products of a Lanczos step in one pass over X.  The
// X^T u of the one-pass scheme is applied to the raw step vector less its
// mean, X^T (u - mean) (k_vside_w turns it into X^T U[:, j] in gene space),
// so nothing in it waits for the reorthogonalisation: the mean of u = X vs
// is colsum . vs / n before the pass (vmean), and X^T (u - vmean) differs
// from X^T (u - mean) by (mean - vmean) colsum, which k_vside_w adds in
// FP64.  Per row the kernel forms the row total u (to u32) and then adds
// x * (u - vmean) for every element of the row to the wave's FP32
// accumulator of its gene, on the same loaded quads.
//
// A row's X^T part needs its total, known once its last segment is reduced:
// the X^T part of a group of four segments runs right after its X v part,
// and the segment table never lets a row of two segments cross a group
// boundary, so every such row is complete when its group's X^T part runs.
// Rows of three or more segments add x * 0 in the stream and are re-read
// whole from memory after their group (the loads hit the cache that just
// delivered them).

namespace {

// Segment groups of the pair kernel.  A pair problem's rows start on whole
// quads and are padded to whole quads (zeros of gene g), so a segment is a
// run of at most 64 whole quads: lane l loads quad q0 + l of its segment,
// or, past the segment's end, the zero quad after the sub-block's last row
// -- no element of a loaded quad needs a range test.  The segments come from the table built
// at admission (launch_pair_segments): four 16-bit descriptors per group,
// one scalar load, and the segment's first quad and row follow from the
// quads and row ends before it.  A row of two segments never starts in a
// group's last slot (the table puts an empty segment there), so every row of
// at most two segments lies within one group and its total is known when
// the group's X^T part runs; rows of three or more segments are flagged.
template <int R>
struct PairGroup {
    f32x4 d[R];   // values
    u32x2 ix[R];  // packed indices (gene << 2, 16 bits each)
    int row[R];   // row of the segment (nrows past the last row)
    bool fin[R];  // the segment ends row `row`
    bool lng[R];  // row `row` has three or more segments
};

struct PairCursor {
    int q;     // first quad of the next segment (relative to the sub-block)
    int row;   // row of the next segment
};

typedef const uint64_t __attribute__((address_space(4))) cu64;

template <int R>
__device__ __forceinline__ void form_group_pair(uint64_t desc, gquad *data,
                                                IxQuad<uint16_t>::gvec *indices, int zq,
                                                int lane, PairCursor &pc, PairGroup<R> &P) {
#pragma unroll
    for (int q = 0; q < R; q++) {
        const unsigned dq = (unsigned)(desc >> (16 * q)) & 0xffffu;
        const int nq = (int)(dq & 127u);
        P.row[q] = pc.row;
        P.fin[q] = (dq & 128u) != 0;
        P.lng[q] = (dq & 256u) != 0;
        const unsigned qi = (unsigned)(lane < nq ? pc.q + lane : zq);
        P.d[q] = data[qi];
        P.ix[q] = indices[qi];
        pc.q += nq;
        pc.row += P.fin[q] ? 1 : 0;
    }
}

// LDS byte offset of element k's gene slot (padding and the zero quad: g).
__device__ __forceinline__ int pair_off(const u32x2 &ix, int k) {
    const unsigned h = k < 2 ? ix.x : ix.y;
    return (int)((k & 1) ? (h >> 16) : (h & 0xffffu));
}

__device__ __forceinline__ float lds_at(const float *base, int byte_off) {
    return *(const float *)((const char *)base + byte_off);
}

__device__ __forceinline__ float f4(const f32x4 &d, int k) {
    return k == 0 ? d.x : k == 1 ? d.y : k == 2 ? d.z : d.w;
}

// X v part of a group: per lane and segment the FMA chain over the quad's
// four elements in order (padding adds 0 * v[g] = 0, v[g] being 0); a row's
// lane sum is its segments' chains added in order (one segment: the chain
// itself), then the lanes meet in store_row_totals' butterfly and the row
// totals go to u.  The segments' chains are independent of each other, so
// they overlap.  Returns, per lane, the centred value (float)(total - vmean)
// of the row whose total the lane holds (store_row_totals' lane for the
// slot ending it).
template <int R>
__device__ __forceinline__ float pair_xv(const PairGroup<R> &P, const float *lds_v, int lane,
                                         float &carry, gfloat_w *u, gfloat_w *sink,
                                         double vmean) {
    float part[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
        float vv[4];
#pragma unroll
        for (int k = 0; k < 4; k++) vv[k] = lds_at(lds_v, pair_off(P.ix[q], k));
        float a = f4(P.d[q], 0) * vv[0];
#pragma unroll
        for (int k = 1; k < 4; k++) a = fmaf(f4(P.d[q], k), vv[k], a);
        part[q] = a;
    }
    float done[R];
    float acc = carry;
#pragma unroll
    for (int q = 0; q < R; q++) {
        done[q] = acc + part[q];
        acc = P.fin[q] ? 0.0f : done[q];
    }
    carry = acc;
    store_row_totals<R, true>(done, P.fin, P.row[0], lane, u, sink);
    return (float)((double)done[0] - vmean);
}

// acc[gene] += x * ur for the four elements of one lane's quad: a row's
// genes are distinct, so one update's live lanes touch distinct genes;
// padding updates the spare slot acc[g].
__device__ __forceinline__ void pair_xt_quad(const f32x4 &d, const u32x2 &ix, float ur,
                                             float *acc) {
    float x[4];
    int o[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        o[k] = pair_off(ix, k);
        x[k] = lds_at(acc, o[k]);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) *(float *)((char *)acc + o[k]) = fmaf(f4(d, k), ur, x[k]);
}

// The X^T part of a row of three or more segments, positions [rs, re)
// (relative to the sub-block's first quad), re-read window by window.
__device__ __noinline__ void pair_xt_long_row(gquad *data, IxQuad<uint16_t>::gvec *indices,
                                              int zq, int rs, int re, float ur, int lane,
                                              float *acc) {
    for (int p = rs; p < re; p += 4 * kWave) {
        const unsigned qi = (unsigned)(p + 4 * lane < re ? (p >> 2) + lane : zq);
        const f32x4 d = data[qi];
        const u32x2 ix = indices[qi];
        pair_xt_quad(d, ix, ur, acc);
    }
}

// The X^T part of group P, after its X v part: the row ending at slot f
// has its centred total in lane f kWave / R of uc; a slot that does not
// finish its row belongs to a row ending in the next slot (two segments), to
// a row of three or more segments, or is an empty last slot.  The group's
// segments go in slot order, a row of three or more segments adding x * 0
// there; such a row is then applied whole, re-read, after the group's other
// rows (a fixed order, set by the data).
template <int R>
__device__ __forceinline__ void pair_xt(const PairGroup<R> &P, float uc, const int64_t *rowpos,
                                        int64_t pos0, gquad *data, IxQuad<uint16_t>::gvec *indices,
                                        int zq, int lane, float *acc) {
    constexpr int kLanesPerRow = kWave / R;
    const int ucb = __float_as_int(uc);
    float ur[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
        const int f = P.fin[q] || q == R - 1 ? q : q + 1;   // the slot ending the row
        ur[q] = __int_as_float(__builtin_amdgcn_readlane(ucb, f * kLanesPerRow));
        pair_xt_quad(P.d[q], P.ix[q], P.lng[q] ? 0.0f : ur[q], acc);
    }
#pragma unroll
    for (int q = 0; q < R; q++)
        if (P.lng[q] && P.fin[q])
            pair_xt_long_row(data, indices, zq, uniform((int)(rowpos[P.row[q]] - pos0)),
                             uniform((int)(rowpos[P.row[q] + 1] - pos0)), ur[q], lane, acc);
}

}  // namespace

// LDS of k_spmv_pair: vs (g floats and a zero) and 4 wave accumulators of
// g + 1 floats (a spare slot each): 40 KB at g = 2000, four workgroups per
// compute unit.
size_t spmv_pair_lds_bytes(int g) {
    return sizeof(float) * ((size_t)g + 1 + 4 * ((size_t)g + 1));
}

// One workgroup per data block, wave w on its 256-row sub-block; loads
// non-temporal (X is read once per step).  The block's partials of
// X^T (u - vmean) go to tpart as in k_spmv_t.  vs sits at LDS offset 0, so
// an element's packed index is its v slot's address.
template <int R>
__global__ __launch_bounds__(256) void k_spmv_pair(Tables T) {
    extern __shared__ float lds[];
    const int2 bp = T.blocks[blockIdx.x];
    Prob &pr = T.probs[bp.x];
    if (pr.phase != PH_LANCZOS || !pr.pair) return;
    const int g = pr.g, gs = g + 1;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = wave_id();
    float *lds_v = lds;
    float *accs = lds + gs;
    float *acc = accs + (size_t)w * gs;
    for (int i = threadIdx.x; i < gs; i += 256) lds_v[i] = i < g ? pr.vs[i] : 0.0f;
    for (int i = lane; i < gs; i += kWave) acc[i] = 0.0f;
    int64_t r0, nrows64;
    sub_block(pr.n, bp.y, w, r0, nrows64);
    const int nrows = (int)nrows64;
    // The sub-block's rows start on whole quads: positions below are
    // relative to its first (the padded row offsets are read directly; only
    // the zero quad's place and the rows of three or more segments need
    // them).
    const int64_t *rowpos = pr.indptr + r0;
    const int64_t pos0 = nrows > 0 ? uniform64(rowpos[0]) : 0;
    const int64_t q0 = pos0 >> 2;
    __syncthreads();
    float *uj = u_of_step(pr, pr.j);
    if (nrows > 0) {
        const double vmean = pr.vmean;
        // Wave-uniform bases (scalar registers), so that every load is a
        // scalar base plus a 32-bit lane offset.
        gquad *data = uniform_ptr((gquad *)pr.data + q0);
        IxQuad<uint16_t>::gvec *indices = uniform_ptr((IxQuad<uint16_t>::gvec *)pr.idx + q0);
        const int zq = (int)((uniform64(rowpos[nrows]) - pos0) >> 2) - 1;   // the sub-block's zero quad
        gfloat_w *u = (gfloat_w *)(uj + r0);
        gfloat_w *sink = (gfloat_w *)(uj + pr.n);
        const int64_t sub = (int64_t)bp.y * (kRowBlock / kSubRows) + w;
        const int64_t sb = uniform64(pr.segbase[sub]);
        const int ngroups = uniform((int)((pr.segbase[sub + 1] - sb) / R));
        // R = 4 descriptors a group, read by scalar loads (constant address
        // space, uniform address), each one group ahead of its use.
        cu64 *sd = (cu64 *)uniform_ptr((const uint64_t *)(pr.segs + sb));
        auto desc = [&](int gi) -> uint64_t {
            const uint64_t v = sd[gi < ngroups ? gi : ngroups - 1];
            return gi < ngroups ? v : 0ull;
        };
        PairCursor pc{0, 0};
        PairGroup<R> ga, gb;
        float carry = 0.0f;
        // Three groups rotate: while one is reduced, the loads of the next
        // two are in flight.  Groups past the table are empty (no quads, no
        // row ends); the loop ends after reducing the table's last group.
        PairGroup<R> gc;
        uint64_t dn = desc(2);
        form_group_pair<R>(desc(0), data, indices, zq, lane, pc, ga);
        form_group_pair<R>(desc(1), data, indices, zq, lane, pc, gb);
        for (int gi = 2;; gi += 3) {
            form_group_pair<R>(dn, data, indices, zq, lane, pc, gc);
            dn = desc(gi + 1);
            __builtin_amdgcn_sched_barrier(0);
            float uc = pair_xv<R>(ga, lds_v, lane, carry, u, sink, vmean);
            pair_xt<R>(ga, uc, rowpos, pos0, data, indices, zq, lane, acc);
            if (gi - 1 >= ngroups) break;        // ga was group gi - 2
            form_group_pair<R>(dn, data, indices, zq, lane, pc, ga);
            dn = desc(gi + 2);
            __builtin_amdgcn_sched_barrier(0);
            uc = pair_xv<R>(gb, lds_v, lane, carry, u, sink, vmean);
            pair_xt<R>(gb, uc, rowpos, pos0, data, indices, zq, lane, acc);
            if (gi >= ngroups) break;            // gb was group gi - 1
            form_group_pair<R>(dn, data, indices, zq, lane, pc, gb);
            dn = desc(gi + 3);
            __builtin_amdgcn_sched_barrier(0);
            uc = pair_xv<R>(gc, lds_v, lane, carry, u, sink, vmean);
            pair_xt<R>(gc, uc, rowpos, pos0, data, indices, zq, lane, acc);
            if (gi + 1 >= ngroups) break;        // gc was group gi
        }
    }
    __syncthreads();
    float *out = pr.tpart + (int64_t)bp.y * 2 * g;
    for (int i = threadIdx.x; i < g; i += 256)
        out[i] = (accs[i] + accs[gs + i]) + (accs[2 * gs + i] + accs[3 * gs + i]);
}

// Last step of a sweep, pair problems: U[:, m-1] = (u - mean) / alpha (what
// k_spmv_t writes for other one-pass problems; the restart and final
// coefficients add the rest of the column).
__global__ __launch_bounds__(256) void k_pair_last_column(Tables T) {
    const int2 bp = T.blocks[blockIdx.x];
    Prob &pr = T.probs[bp.x];
    if (pr.phase != PH_LANCZOS || !pr.pair || pr.j != pr.m - 1) return;
    const int64_t r0 = (int64_t)bp.y * kRowBlock;
    const int64_t nrows = pr.n - r0 < kRowBlock ? pr.n - r0 : kRowBlock;
    const float *uj = u_of_step(pr, pr.j);
    double *Uj = pr.U[pr.ubuf] + (int64_t)pr.j * pr.n;
    const double umean = pr.umean, inv_alpha = pr.inv_alpha;
    for (int64_t t = threadIdx.x; t < nrows; t += 256)
        Uj[r0 + t] = ((double)uj[r0 + t] - umean) * inv_alpha;
}

// ================================================== X^T u, gene-sorted copy
//
// For g <= kMaxSortedGenes X^T u reads a second copy of X made at admission:
// inside each 256-row sub-block the entries are ordered by (gene, row), so a
// gene's entries in the sub-block form one run, and its partial is a chain
// over that run in row order.  A wave streams its sub-block in windows of
// kTWindow entries, kTChunk consecutive entries per lane:
//   * per lane, a chain over its entries that restarts at each run start
//     (a[i] = x[i] u[row[i]] + a[i-1], or without the a[i-1] at a run start);
//   * an inclusive segmented scan over the lanes (six DPP steps, a segment
//     starting at each lane that holds a run start) of each lane's last
//     piece, which gives every lane the part of its first run that lies in
//     the lanes before it (and, through `carry`, in earlier windows);
//   * one lane per gene that ends in the window stores its total, once,
//     into the LDS slot of its run offset (no later window reads it).
// A gene's sub-block partial is therefore
//   (earlier windows' carry + the lanes' scan tree) + the lane's own chain,
// a fixed function of the entries' positions: independent of the launch.
// Sub-blocks combine as (w0 + w1) + (w2 + w3), as in the CSR path.  Each
// element costs one LDS read (u[row]) instead of the CSR path's read and
// write of a scattered accumulator, every lane of a window holds entries
// (a row-windowed CSR leaves lanes idle past each row's end), and a row
// index is one byte, so the copy streams 5 bytes an entry against 6.

namespace {

// LDS slot of sub-block row r in the u table: rows are stored as this slot,
// so that a run over consecutive rows, read kTChunk entries a lane, lands on
// distinct banks.
__host__ __device__ __forceinline__ int tsort_row_slot(int r) { return ((r & 15) << 4) | (r >> 4); }

template <int CTRL, int RM>
__device__ __forceinline__ float dpp_move(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, RM, 0xF, false));
}

// Inclusive segmented scan: lane l gets the sum of v over lanes [l - dl, l]
// (dl = distance to the first lane of its segment).  Every DPP move runs
// with all lanes active (a DPP read of an inactive lane returns 0), so each
// step adds its neighbour's value times 1 or 0 instead of selecting.
__device__ __forceinline__ float seg_scan(float v, int lane, int dl) {
    const int r = lane & 15;
    v = fmaf(dpp_move<0x111, 0xF>(v), dl >= 1 ? 1.f : 0.f, v);   // row_shr:1
    v = fmaf(dpp_move<0x112, 0xF>(v), dl >= 2 ? 1.f : 0.f, v);   // row_shr:2
    v = fmaf(dpp_move<0x114, 0xF>(v), dl >= 4 ? 1.f : 0.f, v);   // row_shr:4
    v = fmaf(dpp_move<0x118, 0xF>(v), dl >= 8 ? 1.f : 0.f, v);   // row_shr:8
    v = fmaf(dpp_move<0x142, 0xA>(v), ((lane & 16) && dl > r) ? 1.f : 0.f, v);          // row_bcast:15
    v = fmaf(dpp_move<0x143, 0xC>(v), ((lane & 32) && dl > lane - 32) ? 1.f : 0.f, v);  // row_bcast:31
    return v;
}

typedef const uint16_t __attribute__((address_space(1))) gu16_r;
typedef const u32x4 __attribute__((address_space(1))) gu32x4_r;

struct TChunk {
    f32x4 x[kTChunk / 4];
    u32x4 r;        // 16 row bytes
    unsigned h;     // 16 run-start bits
};

// Loads R lane chunks (window P + kTWindow k); chunk indices are clamped to
// the sub-block's last chunk, so no load sits behind a branch.
template <int R>
__device__ __forceinline__ void tload(gquad *V4, gu32x4_r *R16, gu16_r *H, int P, int clast, int lane,
                                      TChunk (&G)[R]) {
#pragma unroll
    for (int k = 0; k < R; k++) {
        int c = P / kTChunk + kWave * k + lane;
        c = c < clast ? c : clast;
#pragma unroll
        for (int j = 0; j < kTChunk / 4; j++) G[k].x[j] = V4[c * (kTChunk / 4) + j];
        G[k].r = R16[c];
        G[k].h = H[c];
    }
}

struct TState {
    int gcur;     // first gene whose run has not ended
    float carry;  // partial of the run open at the end of the last window
};

__device__ __forceinline__ unsigned row_word(const u32x4 &r, int k) {
    return k == 0 ? r.x : k == 1 ? r.y : k == 2 ? r.z : r.w;
}

// One window: positions [Pk, Pk + kTWindow) relative to the sub-block's
// chunk origin; valid entries are [S, E).
__device__ __forceinline__ void tsort_window(const TChunk &C, int Pk, int S, int E, int g, int lane,
                                             int *accoff, const float *uu, float *tot, float *sinv,
                                             uint8_t *fhv, TState &st) {
    const int bit0 = kTChunk * lane;
    float ui[kTChunk], xs[kTChunk];
#pragma unroll
    for (int i = 0; i < kTChunk; i++) {
        ui[i] = uu[(row_word(C.r, i >> 2) >> (8 * (i & 3))) & 255u];
        xs[i] = C.x[i >> 2][i & 3];
    }
    unsigned heads = C.h;
    if (Pk < S || Pk + kTWindow > E) {
        __asm__ volatile("" ::: "memory");  // keep this a branch: only a sub-block's first and last windows mask
#pragma unroll
        for (int i = 0; i < kTChunk; i++) {
            const int pos = bit0 + Pk + i;
            xs[i] = (pos >= S && pos < E) ? xs[i] : 0.f;
        }
        // Run starts of the neighbouring sub-blocks' entries are dropped; the
        // sub-block's first entry starts a run.
        const int lo = S - Pk - bit0, hi = E - Pk - bit0;
        unsigned vm = 0xffffffffu;
        vm = lo > 0 ? (lo >= 32 ? 0u : vm << lo) : vm;
        vm = hi < kTChunk ? (hi <= 0 ? 0u : vm & ((1u << hi) - 1u)) : vm;
        heads &= vm;
        if (lo >= 0 && lo < kTChunk) heads |= 1u << lo;
    }
    const unsigned keep = ~heads;
    float a = 0.f, as[kTChunk];
#pragma unroll
    for (int i = 0; i < kTChunk; i++) {
        const int m = ((int)(keep << (31 - i))) >> 31;  // 0 at a run start, else all ones
        a = __builtin_bit_cast(float, __builtin_bit_cast(int, a) & m);
        a = fmaf(xs[i], ui[i], a);
        as[i] = a;
    }
    heads &= (1u << kTChunk) - 1u;
    const unsigned long long hmask = __ballot(heads != 0);
    const unsigned long long below = hmask & ((2ull << lane) - 1ull);
    const int dl = below ? lane - (63 - __builtin_clzll(below)) : lane;
    const float Sinc = seg_scan(a, lane, dl);
    const float Sprev = dpp_move<0x138, 0xF>(Sinc);  // wave_shr:1
    const unsigned long long before = hmask & ((1ull << lane) - 1ull);
    const float Sin = lane == 0 ? st.carry : (before ? Sprev : st.carry + Sprev);
#pragma unroll
    for (int j = 0; j < kTChunk / 4; j++)
        *(f32x4 *)(tot + kTChunk * lane + 4 * j) = f32x4{as[4 * j], as[4 * j + 1], as[4 * j + 2], as[4 * j + 3]};
    sinv[lane] = Sin;
    fhv[lane] = heads ? (uint8_t)__builtin_ctz(heads) : (uint8_t)kTChunk;
    const float S63 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, Sinc), 63));
    st.carry = hmask ? S63 : st.carry + S63;
    __builtin_amdgcn_wave_barrier();
    // One lane per gene that ends in this window (empty genes store 0).
    for (;;) {
        const int c = st.gcur + lane;
        const int s0 = c < g ? accoff[c] : E;
        const int e0 = c < g ? accoff[c + 1] : E;
        const bool done = c < g && e0 <= Pk + kTWindow;
        float t = 0.f;
        if (done && s0 < e0) {
            const int es = e0 - 1 - Pk;
            const int lo = es / kTChunk, i = es % kTChunk;
            t = tot[es];
            if (i < fhv[lo]) t = sinv[lo] + t;
        }
        const unsigned long long dm = __ballot(done);
        __builtin_amdgcn_wave_barrier();
        if (done) accoff[c] = __builtin_bit_cast(int, t);
        __builtin_amdgcn_wave_barrier();
        st.gcur += __builtin_popcountll(dm);
        if (!((dm >> 63) & 1ull)) break;
    }
}

}  // namespace

// LDS of k_spmv_t_sorted, per wave (16-byte aligned parts): g + 1 offsets /
// totals, 256 u values, kTWindow window totals, 64 lane carries and 64
// first-run-start bytes.
__host__ __device__ inline int tsort_lds_floats(int g) { return ((g + 4) & ~3) + 256 + kTWindow + kWave + kWave / 4; }

size_t spmv_t_sorted_lds_bytes(int g) { return sizeof(float) * 4 * (size_t)tsort_lds_floats(g); }

// One workgroup per data block, wave w on sub-block w; u = (U[:, j] - mean)
// / alpha written back as in k_spmv_t (one pass: u_j - mean).  Problems without the sorted copy
// are left to k_spmv_t.
template <int R>
__global__ __launch_bounds__(256) void k_spmv_t_sorted(Tables T) {
    extern __shared__ float lds[];
    const int2 bp = T.blocks[blockIdx.x];
    Prob &pr = T.probs[bp.x];
    if (pr.phase != PH_LANCZOS || !pr.tsorted) return;
    const int g = pr.g;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = wave_id();
    const int per = tsort_lds_floats(g);
    int *accoff = (int *)(lds + (size_t)w * per);
    float *uu = lds + (size_t)w * per + ((g + 4) & ~3);
    float *tot = uu + 256;
    float *sinv = tot + kTWindow;
    uint8_t *fhv = (uint8_t *)(sinv + kWave);
    int64_t r0, nrows64;
    sub_block(pr.n, bp.y, w, r0, nrows64);
    const int nrows = (int)nrows64;
    double *Uj = pr.U[pr.ubuf] + (int64_t)pr.j * pr.n;
    const double inv_alpha = pr.inv_alpha, umean = pr.umean;
    const bool one = pr.onepass != 0;
    const bool last = pr.j == pr.m - 1;
    const float *uj = u_of_step(pr, pr.j);
    for (int t = lane; t < kSubRows; t += kWave) {
        float x = 0.f;
        if (t < nrows) {
            if (one) {
                // u - mean of this step; U[:, j] is written by the next pass
                // (at the last step of a sweep: (u - mean) / alpha, for
                // restart_coef).
                const double d = (double)uj[r0 + t] - umean;
                x = (float)d;
                if (last) Uj[r0 + t] = d * inv_alpha;
            } else {
                const double d = (Uj[r0 + t] - umean) * inv_alpha;
                Uj[r0 + t] = d;
                x = (float)d;
            }
        }
        uu[tsort_row_slot(t)] = x;
    }
    int64_t s0 = 0, e0 = 0;
    if (nrows > 0) {
        s0 = pr.indptr[r0];
        e0 = pr.indptr[r0 + nrows];
    }
    if (e0 > s0) {
        const int64_t org = s0 / kTChunk * kTChunk;
        const int S = (int)(s0 - org), E = (int)(e0 - org);
        // Run offsets: accoff[c] = S + entries of genes < c, 8 genes a lane.
        const int64_t sb = r0 / kSubRows;
        const uint4 *L8 = (const uint4 *)(pr.tlens + sb * pr.tgl);
        int base = S;
        for (int cb = 0; cb < g; cb += 8 * kWave) {
            const int c = cb + 8 * lane;
            const uint4 q = c < g ? L8[c / 8] : uint4{0, 0, 0, 0};
            const int v[8] = {(int)(q.x & 0xffff), (int)(q.x >> 16), (int)(q.y & 0xffff), (int)(q.y >> 16),
                              (int)(q.z & 0xffff), (int)(q.z >> 16), (int)(q.w & 0xffff), (int)(q.w >> 16)};
            int sum = 0;
#pragma unroll
            for (int k = 0; k < 8; k++) sum += v[k];
            int incl = sum;
#pragma unroll
            for (int d = 1; d < kWave; d <<= 1) {
                const int t = __shfl_up(incl, d);
                incl = lane >= d ? incl + t : incl;
            }
            int run = base + incl - sum;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                if (c + k < g) accoff[c + k] = run;
                run += v[k];
            }
            base += __builtin_amdgcn_readlane(incl, kWave - 1);
        }
        if (lane == 0) accoff[g] = base;
        __builtin_amdgcn_wave_barrier();
        gquad *V4 = (gquad *)(pr.tval + org);
        gu32x4_r *R16 = (gu32x4_r *)(pr.trow + org);
        gu16_r *H = (gu16_r *)(pr.thb + org / kTChunk);
        const int clast = (E - 1) / kTChunk;
        TState st{0, 0.f};
        TChunk ga[R], gb[R];
        int P = 0;
        tload<R>(V4, R16, H, P, clast, lane, ga);
        for (;;) {
            tload<R>(V4, R16, H, P + kTWindow * R, clast, lane, gb);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < R; k++)
                tsort_window(ga[k], P + kTWindow * k, S, E, g, lane, accoff, uu, tot, sinv, fhv, st);
            P += kTWindow * R;
            const bool last = P >= E;
            tload<R>(V4, R16, H, P + kTWindow * R, clast, lane, ga);
            __builtin_amdgcn_sched_barrier(0);
            if (last) break;
#pragma unroll
            for (int k = 0; k < R; k++)
                tsort_window(gb[k], P + kTWindow * k, S, E, g, lane, accoff, uu, tot, sinv, fhv, st);
            P += kTWindow * R;
            if (P >= E) break;
        }
    } else {
        for (int i = lane; i < g; i += kWave) accoff[i] = 0;
    }
    __syncthreads();
    float *out = pr.tpart + (int64_t)bp.y * 2 * g;
    for (int i = threadIdx.x; i < g; i += 256)
        out[i] = (lds[i] + lds[per + i]) + (lds[2 * per + i] + lds[3 * per + i]);
}

// ------------------------------------------------- building the sorted copy

// The gene-sorted copy of one 256-row sub-block, built in LDS by one
// workgroup of kBuildWaves waves (wave w on rows kBuildRows w and on):
//   1. each wave counts its rows' entries per gene (LDS integer atomics, an
//      order-free count), the sub-block's counts go to tlens;
//   2. start[c] = the sub-block's exclusive prefix of the counts, and each
//      wave's counter of gene c becomes start[c] + the counts of the waves
//      before it: the position of its first entry of c;
//   3. the genes are placed in passes [ga, gb) whose entries fit the LDS
//      stage: every wave walks its rows in order and puts each entry of a
//      gene in the pass at its counter (a row's genes are distinct, so the
//      lanes of one update touch distinct counters), then the whole
//      workgroup copies the stage out with consecutive stores and marks
//      the pass's run starts in thb (cleared beforehand) by atomic OR.
// Entries of one gene therefore follow in row order, as X^T u requires.
constexpr int kBuildWaves = 8;
constexpr int kBuildRows = kSubRows / kBuildWaves;
constexpr int kBuildThreads = kBuildWaves * kWave;

__global__ __launch_bounds__(kBuildThreads) void k_tsort_build(const float *__restrict__ data4,
                                                     const uint16_t *__restrict__ idx,
                                                     const int64_t *__restrict__ indptr, int64_t n,
                                                     int g, int gl, int cap,
                                                     float *__restrict__ tval, uint8_t *__restrict__ trow,
                                                     unsigned *__restrict__ thb32,
                                                     uint16_t *__restrict__ tlens) {
    extern __shared__ unsigned b_lds[];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = wave_id();
    unsigned *cntw = b_lds;                       // [kBuildWaves][gl] counts, then counters
    unsigned *start = cntw + kBuildWaves * gl;    // [gl + 1]
    unsigned *wtot = start + gl + 1;              // [kBuildWaves] wave totals of the prefix
    int *roff = (int *)(wtot + kBuildWaves);      // [kSubRows + 1]
    float *stage_v = (float *)(roff + kSubRows + 1);  // [cap]
    uint8_t *stage_r = (uint8_t *)(stage_v + cap);    // [cap]
    const int64_t sb = blockIdx.x;
    const int64_t r0 = sb * kSubRows;
    const int nr = (int)(n - r0 < kSubRows ? n - r0 : kSubRows);
    const int64_t s0 = indptr[r0];
    for (int t = threadIdx.x; t <= nr; t += kBuildThreads) roff[t] = (int)(indptr[r0 + t] - s0);
    for (int i = threadIdx.x; i < kBuildWaves * gl; i += kBuildThreads) cntw[i] = 0u;
    __syncthreads();
    const int E = roff[nr];
    const int wr0 = kBuildRows * w < nr ? kBuildRows * w : nr;
    const int wr1 = kBuildRows * (w + 1) < nr ? kBuildRows * (w + 1) : nr;
    const int wa = roff[wr0], wb = roff[wr1];  // this wave's entries
    unsigned *mycnt = cntw + w * gl;
    for (int p = wa + lane; p < wb; p += 16 * kWave) {
        unsigned c[16];
#pragma unroll
        for (int k = 0; k < 16; k++) c[k] = p + k * kWave < wb ? (unsigned)idx[s0 + p + k * kWave] : 0xffffffffu;
#pragma unroll
        for (int k = 0; k < 16; k++)
            if (c[k] < (unsigned)g) atomicAdd(&mycnt[c[k]], 1u);
    }
    __syncthreads();
    // Sub-block counts, their exclusive prefix, per-wave first positions.
    unsigned base = 0;
    for (int cb = 0; cb < gl; cb += kBuildThreads) {
        const int c = cb + threadIdx.x;
        unsigned tot = 0;
        if (c < gl) {
#pragma unroll
            for (int v = 0; v < kBuildWaves; v++) tot += cntw[v * gl + c];
            tlens[sb * gl + c] = (uint16_t)tot;
        }
        // Block-wide exclusive scan of tot (wave scans, then
        // the wave totals).
        unsigned incl = tot;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const unsigned t = __shfl_up(incl, d);
            incl = lane >= d ? incl + t : incl;
        }
        __syncthreads();
        if (lane == kWave - 1) wtot[w] = incl;
        __syncthreads();
        unsigned wbase = base, blk = 0;
        for (int v = 0; v < kBuildWaves; v++) {
            wbase += v < w ? wtot[v] : 0u;
            blk += wtot[v];
        }
        __syncthreads();
        if (c < gl) {
            const unsigned st = wbase + incl - tot;
            start[c] = st;
            unsigned run = st;
#pragma unroll
            for (int v = 0; v < kBuildWaves; v++) {
                const unsigned a = cntw[v * gl + c];
                cntw[v * gl + c] = run;
                run += a;
            }
        }
        base += blk;
        __syncthreads();
    }
    if (threadIdx.x == 0) start[gl] = base;
    __syncthreads();
    if (E == 0) return;
    // Gene passes whose entries fit the stage.
    int ga = 0;
    while (ga < g) {
        // gb: the largest gene boundary with start[gb] - start[ga] <= cap
        // (at least one gene; a gene never exceeds 256 entries <= cap).
        int lo = ga + 1, hi = g;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if ((int)(start[mid] - start[ga]) <= cap) lo = mid; else hi = mid - 1;
        }
        const int gb = lo;
        const int pa = (int)start[ga], pb = (int)start[gb];
        // Place this wave's entries of genes [ga, gb), rows in order, in
        // batches of kB x 64 entries (lane l: positions b + l + 64 k) whose
        // loads are issued two batches ahead (three rotating buffers).
        if (wb > wa) {
            constexpr int kB = 8, kStep = kB * kWave;
            struct Batch {
                unsigned c[kB];
                float x[kB];
            };
            auto load_batch = [&](int b, Batch &B) {
#pragma unroll
                for (int k = 0; k < kB; k++) {
                    int p = b + kWave * k + lane;
                    p = p < wb ? p : wb - 1;
                    B.c[k] = idx[s0 + p];
                    B.x[k] = data4[s0 + p];
                }
            };
            int r = wr0;
            auto place = [&](int b, const Batch &bt, Batch &nxt) {
                const int bn = b + 2 * kStep;
                load_batch(bn < wb ? bn : b, nxt);
#pragma unroll
                for (int k = 0; k < kB; k++) {
                    const int b0 = b + kWave * k;
                    if (b0 >= wb) break;
                    const int p = b0 + lane;
                    const unsigned c = bt.c[k];
                    const bool mine = p < wb && c >= (unsigned)ga && c < (unsigned)gb;
                    const int last = (wb - b0 < kWave ? wb - b0 : kWave) - 1 + b0;
                    while (roff[r + 1] <= b0) r++;
                    for (int rr = r;; rr++) {
                        const int ra = roff[rr], rb = roff[rr + 1];
                        if (mine && p >= ra && p < rb) {
                            const unsigned pos = mycnt[c];
                            mycnt[c] = pos + 1u;
                            stage_v[pos - pa] = bt.x[k];
                            stage_r[pos - pa] = (uint8_t)tsort_row_slot(rr);
                        }
                        __builtin_amdgcn_wave_barrier();
                        if (rb > last) break;
                    }
                }
            };
            Batch B0, B1, B2;
            load_batch(wa, B0);
            load_batch(wa + kStep < wb ? wa + kStep : wa, B1);
            for (int b = wa;;) {
                place(b, B0, B2);
                if ((b += kStep) >= wb) break;
                place(b, B1, B0);
                if ((b += kStep) >= wb) break;
                place(b, B2, B1);
                if ((b += kStep) >= wb) break;
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < pb - pa; i += kBuildThreads) {
            tval[s0 + pa + i] = stage_v[i];
            trow[s0 + pa + i] = stage_r[i];
        }
        for (int c = ga + threadIdx.x; c < gb; c += kBuildThreads) {
            if (start[c + 1] > start[c]) {
                const int64_t at = s0 + start[c];
                atomicOr(thb32 + (at >> 5), 1u << (at & 31));
            }
        }
        __syncthreads();
        ga = gb;
    }
}

// LDS of k_tsort_build for gl counter columns and a stage of cap entries.
static size_t tsort_build_lds(int gl, int cap) {
    return sizeof(unsigned) * (kBuildWaves * (size_t)gl + gl + 1 + kBuildWaves + kSubRows + 1) + 5 * (size_t)cap;
}

__global__ __launch_bounds__(256) void k_zero_words(unsigned *__restrict__ p, int64_t words) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < words; i += stride) p[i] = 0u;
}

hipError_t launch_tsort_build(const float *data4, const uint16_t *idx16, const int64_t *indptr,
                              int64_t n, int g, int gl, float *tval, uint8_t *trow,
                              uint16_t *thb, int64_t thb_words, uint16_t *tlens,
                              int64_t max_lds, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int64_t words32 = (thb_words + 1) / 2;
    k_zero_words<<<(unsigned)std::min<int64_t>(ceil_div(words32, 256), 8192), 256, 0, s>>>(
        (unsigned *)thb, words32);
    // The largest stage that fits one workgroup per compute unit.
    const size_t meta = tsort_build_lds(gl, 0);
    const int64_t budget = std::min<int64_t>(max_lds, 160 * 1024);
    int cap = (int)((budget - (int64_t)meta) / 5) & ~3;
    if (cap < kSubRows) return hipErrorInvalidValue;
    k_tsort_build<<<(unsigned)ceil_div(n, kSubRows), kBuildThreads, tsort_build_lds(gl, cap), s>>>(
        data4, idx16, indptr, n, g, gl, cap, tval, trow, (unsigned *)thb, tlens);
    return hipGetLastError();
}

// ================================================================ V side
//
// Four short launches over genes (the V side is g long, so one workgroup per
// problem would leave the step latency-bound):
//   k_vside_w    w = (sum of the group partials, in four ordered quarters)
//                    / sd - alpha V[:, j]                (FP64, per gene;
//                one pass: the sum is X^T (u - mean), corrected to X^T U[:, j]
//                first)
//   k_vside_dot  dot_c = V[:, c] . w, thread-strided over genes then
//                block_sum<4>, one workgroup per column
//   k_vside_upd  x = w - sum_c V[:, c] dot_c (c ascending); per-chunk sum
//                of squares block_sum<4> over its 256 genes
//   k_vside_fin  beta = sqrt(chunk sums in chunk order); V[:, j+1], vs and
//                B, or the remainder and rnorm at the end of a sweep.

constexpr int kVChunk = 256;  // genes per V-side workgroup

// k_vside_w: 64 genes per workgroup, lane l on gene 64 b + l; wave w sums
// the quarter [w G / 4, (w + 1) G / 4) of the G group partials in order and,
// one pass, the quarter of the correction's columns; the quarters meet as
// (q0 + q1) + (q2 + q3).
constexpr int kVwGenes = 64;

__global__ __launch_bounds__(256) void k_vside_w(Tables T) {
    __shared__ double part[2][4][kVwGenes];
    const int chunks = (T.max_g + kVwGenes - 1) / kVwGenes;
    Prob &pr = T.probs[blockIdx.x / chunks];
    if (pr.phase != PH_LANCZOS) return;
    const int g = pr.g;
    const int lane = threadIdx.x & (kWave - 1), w = wave_id();
    const int i = (blockIdx.x % chunks) * kVwGenes + lane;
    if ((blockIdx.x % chunks) * kVwGenes >= g) return;
    const int j = pr.j;
    const bool one = pr.onepass != 0;
    double s = 0.0, corr = 0.0;
    if (i < g) {
        const int64_t ng = pr.ngrp;
        const int64_t G0 = ng * w / 4, G1 = ng * (w + 1) / 4;
        const double *gp = pr.grp + i;
#pragma unroll 4
        for (int64_t G = G0; G < G1; G++) s += gp[G * 2 * g];
        if (one) {
            // X^T U[:, :j] h, h = red[1..j], c ascending within the quarter.
            const double *R = pr.R[pr.ubuf] + i;
            const int c0 = j * w / 4, c1 = j * (w + 1) / 4;
#pragma unroll 4
            for (int c = c0; c < c1; c++) corr = fma(R[(int64_t)c * g], pr.red[c + 1], corr);
        }
    }
    part[0][w][lane] = s;
    part[1][w][lane] = corr;
    __syncthreads();
    if (w != 0 || i >= g) return;
    double acc = (part[0][0][lane] + part[0][1][lane]) + (part[0][2][lane] + part[0][3][lane]);
    if (one) {
        // acc = X^T (u_j - mean) (pair problems: X^T (u_j - vmean), moved
        // here by (vmean - mean) colsum); X^T U[:, j] = (acc - X^T U[:, :j] h)
        // / alpha, kept as column j of R.
        if (pr.pair) acc = fma(pr.vmean - pr.umean, pr.colsum[i], acc);
        const double c = (part[1][0][lane] + part[1][1][lane]) + (part[1][2][lane] + part[1][3][lane]);
        acc = (acc - c) * pr.inv_alpha;
        pr.R[pr.ubuf][(int64_t)j * g + i] = acc;
    }
    const double *Vj = pr.V[pr.vbuf] + (int64_t)j * g;
    pr.vw[i] = acc / pr.sd[i] - pr.alpha * Vj[i];
}

__global__ __launch_bounds__(256) void k_vside_dot(Tables T) {
    __shared__ double tmp[4];
    const int cols = T.max_m;
    Prob &pr = T.probs[blockIdx.x / cols];
    const int c = blockIdx.x % cols;
    if (pr.phase != PH_LANCZOS) return;
    const int ncols = pr.j == pr.m - 1 ? pr.m : pr.j + 1;
    if (c >= ncols) return;
    const int g = pr.g;
    const double *Vc = pr.V[pr.vbuf] + (int64_t)c * g;
    double acc = 0.0;
    for (int i = threadIdx.x; i < g; i += 256) acc = fma(Vc[i], pr.vw[i], acc);
    acc = block_sum<4>(acc, tmp);
    if (threadIdx.x == 0) pr.vdot[c] = acc;
}

__global__ __launch_bounds__(kVChunk) void k_vside_upd(Tables T) {
    __shared__ double tmp[kVChunk / kWave];
    __shared__ double dots[kMaxSubspace];
    const int chunks = (T.max_g + kVChunk - 1) / kVChunk;
    Prob &pr = T.probs[blockIdx.x / chunks];
    if (pr.phase != PH_LANCZOS) return;
    const int g = pr.g;
    const int chunk = blockIdx.x % chunks;
    if (chunk * kVChunk >= g) return;
    const int ncols = pr.j == pr.m - 1 ? pr.m : pr.j + 1;
    for (int c = threadIdx.x; c < ncols; c += kVChunk) dots[c] = pr.vdot[c];
    __syncthreads();
    const int i = chunk * kVChunk + threadIdx.x;
    double x = 0.0;
    if (i < g) {
        const double *V = pr.V[pr.vbuf];
        double d = 0.0;
        for (int c = 0; c < ncols; c++) d = fma(V[(int64_t)c * g + i], dots[c], d);
        x = pr.vw[i] - d;
        pr.vw[i] = x;
    }
    const double sq = block_sum<kVChunk / kWave>(x * x, tmp);
    if (threadIdx.x == 0) pr.vsq[chunk] = sq;
}

__global__ __launch_bounds__(256) void k_vside_fin(Tables T) {
    __shared__ double sbeta;
    __shared__ double tmp[4];
    Prob &pr = T.probs[blockIdx.x];
    if (pr.phase != PH_LANCZOS) return;
    const int g = pr.g, m = pr.m, j = pr.j;
    if (threadIdx.x == 0) {
        const int nchunk = (g + kVChunk - 1) / kVChunk;
        double nrm = 0.0;
        for (int c = 0; c < nchunk; c++) nrm += pr.vsq[c];
        sbeta = sqrt(nrm);
    }
    __syncthreads();
    const double beta = sbeta;
    if (j != m - 1) {
        const double bclip = beta < 1e-8 ? 1e-8 : beta;
        const double inv = 1.0 / bclip;
        double *Vn = pr.V[pr.vbuf] + (int64_t)(j + 1) * g;
        for (int i = threadIdx.x; i < g; i += 256) {
            const double x = pr.vw[i] * inv;
            Vn[i] = x;
            pr.vs[i] = (float)(x / pr.sd[i]);
        }
        __syncthreads();
        const double vm = pr.pair ? vs_mean(pr, tmp) : 0.0;
        if (threadIdx.x == 0) {
            pr.vmean = vm;
            pr.B[(int64_t)j + (int64_t)(j + 1) * m] = bclip;
            pr.j = j + 1;
            pr.steps += 1;
        }
    } else {
        const double inv = beta > 0.0 ? 1.0 / beta : 0.0;
        for (int i = threadIdx.x; i < g; i += 256) pr.rem[i] = pr.vw[i] * inv;
        __syncthreads();
        if (threadIdx.x == 0) {
            pr.rnorm = beta;
            pr.steps += 1;
            pr.phase = PH_RESTART;
        }
    }
}

// ================================================================ small SVD
//
// One-sided (Hestenes) Jacobi on the columns of C = B[:, :m] (B's last column
// is zero): rotate column pairs until all are orthogonal to working
// precision (|w_p . w_q| <= sqrt(m) eps |w_p| |w_q|); then s_c = |W[:, c]|,
// Ub = W / s, right vectors = accumulated rotations.  Pairs follow the round-robin (circle) schedule, m/2 disjoint
// pairs per round, one wave per pair, so every rotation is applied in a fixed
// order.  Sweeps stop when a sweep performs no rotation.

namespace {

// pos - 1 + round < 2 (mm - 1) for pos < mm and round < mm - 1, so one
// conditional subtraction is the remainder.
__device__ inline int rr_player(int pos, int round, int mm) {
    const int v = pos - 1 + round;
    return pos == 0 ? 0 : (v >= mm - 1 ? v - (mm - 1) : v) + 1;
}

// W (m x mm) and Jv (mm x mm) column-major; m rows of W are live.
__device__ void jacobi_svd(double *W, double *Jv, int m, int mm, int *sweeps,
                           int *rot_count) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    const double tol = 2.220446049250313e-16 * sqrt((double)m);
    int sweep = 0;
    for (; sweep < 60; sweep++) {
        if (threadIdx.x == 0) *rot_count = 0;
        __syncthreads();
        for (int round = 0; round < mm - 1; round++) {
            for (int slot = wv; slot < mm / 2; slot += nw) {
                int p = rr_player(slot, round, mm);
                int q = rr_player(mm - 1 - slot, round, mm);
                if (p > q) {
                    int t = p;
                    p = q;
                    q = t;
                }
                double *wp = W + (int64_t)p * m, *wq = W + (int64_t)q * m;
                double a = 0.0, b = 0.0, c = 0.0;
                for (int i = lane; i < m; i += kWave) {
                    const double x = wp[i], y = wq[i];
                    a = fma(x, x, a);
                    b = fma(y, y, b);
                    c = fma(x, y, c);
                }
                a = wave_sum(a);
                b = wave_sum(b);
                c = wave_sum(c);
                if (c != 0.0 && fabs(c) > tol * sqrt(a * b)) {
                    const double zeta = (b - a) / (2.0 * c);
                    const double t = (zeta >= 0.0 ? 1.0 : -1.0) /
                                     (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                    const double cs = 1.0 / sqrt(1.0 + t * t);
                    const double sn = cs * t;
                    for (int i = lane; i < m; i += kWave) {
                        const double x = wp[i], y = wq[i];
                        wp[i] = cs * x - sn * y;
                        wq[i] = sn * x + cs * y;
                    }
                    double *jp = Jv + (int64_t)p * mm, *jq = Jv + (int64_t)q * mm;
                    for (int i = lane; i < mm; i += kWave) {
                        const double x = jp[i], y = jq[i];
                        jp[i] = cs * x - sn * y;
                        jq[i] = sn * x + cs * y;
                    }
                    if (lane == 0) atomicAdd(rot_count, 1);
                }
            }
            __syncthreads();
        }
        const int rc = *rot_count;
        __syncthreads();
        if (rc == 0) break;
    }
    if (threadIdx.x == 0) *sweeps = sweep + 1;
}

// Same rotations with W and Jv in LDS (mm <= 128) and one 32-lane group per
// pair, up to blockDim/32 pairs at once.  A lane holds its (at most 4) rows of
// the pair's W columns in registers: one batch of LDS loads, the partial sums
// lane-strided, then the xor butterfly 16,8,4,2,1 (a swizzle, then DPP moves
// within the 16-lane row); a rotating pair then loads its Jv columns and
// stores all four.  Rows past the end enter the sums as exact zeros.  A 32-lane
// group's 256 contiguous bytes span every LDS bank once, so the two groups of
// a wave never conflict (16-lane groups of columns 100 doubles apart did: the
// Jacobi took 1.65x as long).
__device__ void jacobi_svd_lds(double *W, double *Jv, int m, int mm,
                               int *sweeps, int *rot_count) {
    const int sub = threadIdx.x & 31;
    const int grp = threadIdx.x >> 5;
    const int ngrp = blockDim.x >> 5;
    const double tol = 2.220446049250313e-16 * sqrt((double)m);
    constexpr int kR = 4;   // rows per lane: mm <= 128 (larger mm: jacobi_svd)
    int sweep = 0;
    for (; sweep < 60; sweep++) {
        if (threadIdx.x == 0) *rot_count = 0;
        __syncthreads();
        for (int round = 0; round < mm - 1; round++) {
            for (int slot = grp; slot < mm / 2; slot += ngrp) {
                int p = rr_player(slot, round, mm);
                int q = rr_player(mm - 1 - slot, round, mm);
                if (p > q) {
                    const int t = p;
                    p = q;
                    q = t;
                }
                double *wp = W + p * m, *wq = W + q * m;
                double *jp = Jv + p * mm, *jq = Jv + q * mm;
                double a = 0.0, b = 0.0, c = 0.0;
                double x[kR], y[kR], jx[kR], jy[kR];
#pragma unroll
                for (int t = 0; t < kR; t++) {
                    const int i = sub + 32 * t;
                    x[t] = i < m ? wp[i] : 0.0;
                    y[t] = i < m ? wq[i] : 0.0;
                }
#pragma unroll
                for (int t = 0; t < kR; t++) {
                    a = fma(x[t], x[t], a);
                    b = fma(y[t], y[t], b);
                    c = fma(x[t], y[t], c);
                }
#pragma unroll
                for (int off = 16; off > 0; off >>= 1) {
                    a += xor_lane_f64(a, off);
                    b += xor_lane_f64(b, off);
                    c += xor_lane_f64(c, off);
                }
                if (c != 0.0 && fabs(c) > tol * sqrt(a * b)) {
                    const double zeta = (b - a) / (2.0 * c);
                    const double t = (zeta >= 0.0 ? 1.0 : -1.0) /
                                     (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                    const double cs = 1.0 / sqrt(1.0 + t * t);
                    const double sn = cs * t;
#pragma unroll
                    for (int u = 0; u < kR; u++) {
                        const int i = sub + 32 * u;
                        jx[u] = i < mm ? jp[i] : 0.0;
                        jy[u] = i < mm ? jq[i] : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < kR; u++) {
                        const int i = sub + 32 * u;
                        if (i < m) {
                            wp[i] = cs * x[u] - sn * y[u];
                            wq[i] = sn * x[u] + cs * y[u];
                        }
                        if (i < mm) {
                            jp[i] = cs * jx[u] - sn * jy[u];
                            jq[i] = sn * jx[u] + cs * jy[u];
                        }
                    }
                    if (sub == 0) atomicAdd(rot_count, 1);
                }
            }
            __syncthreads();
        }
        const int rc = *rot_count;
        __syncthreads();
        if (rc == 0) break;
    }
    if (threadIdx.x == 0) *sweeps = sweep + 1;
}

// Sort columns by norm (descending, ties by index) and write Sb, Ub = W/s
// and the right vectors Vb (m x m).
__device__ void svd_sort_outputs(const double *W, const double *Jv, int m,
                                 int mm, double *s, int *perm, double *Sb,
                                 double *Ub, double *Vb) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    for (int c = wv; c < m; c += nw) {
        const double *wc = W + (int64_t)c * m;
        double a = 0.0;
        for (int i = lane; i < m; i += kWave) a = fma(wc[i], wc[i], a);
        a = wave_sum(a);
        if (lane == 0) s[c] = sqrt(a);
    }
    for (int c = threadIdx.x; c < m; c += blockDim.x) perm[c] = c;
    __syncthreads();
    // Stable descending order: column c goes to position (# larger norms) +
    // (# equal norms at lower indices).
    for (int c = threadIdx.x; c < m; c += blockDim.x) {
        const double v = s[c];
        int r = 0;
        for (int j = 0; j < m; j++) {
            const double w = s[j];
            r += (w > v) || (w == v && j < c);
        }
        perm[r] = c;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < m * m; idx += blockDim.x) {
        const int t = idx / m, i = idx % m;
        const int src = perm[t];
        const double sv = s[src];
        Ub[idx] = sv > 0.0 ? W[(int64_t)src * m + i] / sv : 0.0;
        Vb[idx] = Jv[(int64_t)src * mm + i];
    }
    for (int t = threadIdx.x; t < m; t += blockDim.x) Sb[t] = s[perm[t]];
    (void)mm;
}

}  // namespace

// The sweep's projected matrix B -> Ub, Sb, right vectors (in Jv after the
// Jacobi work area), then the convergence test of pca.pyx:1181-1202 and,
// for a restart, the new B of pca.pyx:1225-1231.  kLds keeps the Jacobi
// work arrays in LDS (when the device's workgroup LDS holds them).
template <bool kLds>
__global__ __launch_bounds__(1024) void k_svd(Tables T) {
    __shared__ int sh_int[4];
    extern __shared__ double sh[];
    Prob &pr = T.probs[blockIdx.x];
    if (pr.phase != PH_SVD) return;
    const int m = pr.m, mm = m + (m & 1), k = pr.k;
    double *s = sh;                       // m
    int *perm = (int *)(sh + m);          // m
    double *W = kLds ? sh + m + (m + 1) / 2 : pr.W;
    double *Jv = kLds ? W + m * mm : pr.Jv;
    for (int idx = threadIdx.x; idx < m * mm; idx += blockDim.x) {
        const int c = idx / m;
        W[idx] = c < m ? pr.B[idx] : 0.0;  // B column-major with ld m
    }
    for (int idx = threadIdx.x; idx < mm * mm; idx += blockDim.x)
        Jv[idx] = (idx / mm == idx % mm) ? 1.0 : 0.0;
    __syncthreads();
    if (kLds)
        jacobi_svd_lds(W, Jv, m, mm, &sh_int[0], &sh_int[1]);
    else
        jacobi_svd(W, Jv, m, mm, &sh_int[0], &sh_int[1]);
    __syncthreads();
    // Right vectors go after the Jacobi work area of Jv in global memory.
    double *Vb = pr.Jv + (int64_t)mm * mm;
    svd_sort_outputs(W, Jv, m, mm, s, perm, pr.Sb, pr.Ub, Vb);
    if (threadIdx.x == 0) sh_int[2] = 0;
    __syncthreads();
    {
        // Residual test of the k leading triplets, one per thread.
        const double rnorm = pr.rnorm;
        const double smax = pr.Sb[0] > pr.smax ? pr.Sb[0] : pr.smax;
        int mine = 0;
        for (int i = threadIdx.x; i < k; i += blockDim.x) {
            const double res = rnorm * pr.Ub[(m - 1) + (int64_t)i * m];
            mine += fabs(res) < (double)pr.tol * smax;
        }
        if (mine) atomicAdd(&sh_int[2], mine);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double smax = pr.smax;
        smax = pr.Sb[0] > smax ? pr.Sb[0] : smax;
        pr.smax = smax;
        const int nconv = sh_int[2];
        pr.nconv = nconv;
        pr.svd_sweeps = sh_int[0];
        int kc = pr.kc;
        if (nconv == k) {
            pr.converged = 1;
            pr.phase = PH_FINAL;
        } else if (pr.iter == pr.maxit - 1) {
            pr.phase = PH_FINAL;
        } else {
            kc = kc > k + nconv ? kc : k + nconv;
            kc = kc < m - 1 ? kc : m - 1;
            pr.kc = kc;
            pr.iter = pr.iter + 1;
            pr.phase = PH_RESTARTING;
        }
        sh_int[2] = pr.phase;
        sh_int[3] = kc;
    }
    __syncthreads();
    if (sh_int[2] == PH_RESTARTING) {
        const int kc = sh_int[3];
        const double rnorm = pr.rnorm;
        for (int idx = threadIdx.x; idx < m * (m + 1); idx += blockDim.x) {
            const int i = idx % m, c = idx / m;
            double v = 0.0;
            if (i < kc && c == i) v = pr.Sb[i];
            if (i < kc && c == kc) v = rnorm * pr.Ub[(m - 1) + (int64_t)i * m];
            pr.B[idx] = v;
        }
    }
}

// Problems whose sweep ended in an earlier macro step take their SVD in
// this one (the SVD launch may run concurrently with the Lanczos launches,
// which never see PH_SVD).
__global__ void k_stage_svd(Tables T) {
    Prob &pr = T.probs[blockIdx.x];
    if (threadIdx.x == 0 && pr.phase == PH_RESTART) pr.phase = PH_SVD;
}

size_t svd_lds_bytes(int m, bool resident) {
    const size_t mm = m + (m & 1);
    size_t bytes = sizeof(double) * (size_t)m + sizeof(int) * (size_t)m + sizeof(double);
    if (resident) bytes += sizeof(double) * ((size_t)m * mm + mm * mm);
    return bytes;
}

// ================================================================ restart
//
// Coefficient (l, c) of the restart product U Ub[:, :kc] and of the final
// product U (Ub[:, :k] Sb) (times Sb[c] there).  For a one-pass problem the
// sweep's last column of U holds p = (u - mean) / alpha of its last step, not
// the column itself, p - U[:, :m-1] h / alpha (h = red[1..m-1] of that
// step): its second part goes into the other columns' coefficients,
// Ub[l, c] - (h[l] / alpha) Ub[m-1, c], so no pass over U completes it.
__device__ __forceinline__ double restart_coef(const Prob &pr, int l, int c) {
    const int m = pr.m;
    double v = pr.Ub[l + (int64_t)c * m];
    if (pr.onepass && l < m - 1) v = fma(-pr.red[l + 1] * pr.inv_alpha, pr.Ub[(m - 1) + (int64_t)c * m], v);
    return v;
}
//
// out[r, c] = sum_l U[r, l] coef[l, c] with the sum over l ascending; one
// thread per row, kGemmCols output columns per pass, the coefficient slab in
// LDS.  Used for U <- U Ub[:, :kc] and for PCs = U (Ub[:, :k] Sb[:k]).

template <bool kFinal>
__global__ __launch_bounds__(256) void k_tall_gemm(Tables T) {
    extern __shared__ double cs[];  // m * kGemmCols
    const int2 bp = T.blocks[blockIdx.x];
    Prob &pr = T.probs[bp.x];
    if (pr.phase != (kFinal ? PH_FINAL : PH_RESTARTING)) return;
    const int m = pr.m;
    const int kout = kFinal ? pr.k : pr.kc;
    const int64_t n = pr.n;
    const double *A = pr.U[pr.ubuf];
    double *out = pr.U[pr.ubuf ^ 1];
    const int64_t r0 = (int64_t)bp.y * kRowBlock;
    for (int c0 = 0; c0 < kout; c0 += kGemmCols) {
        const int nc = kout - c0 < kGemmCols ? kout - c0 : kGemmCols;
        __syncthreads();
        for (int i = threadIdx.x; i < m * kGemmCols; i += 256) {
            const int l = i / kGemmCols, cc = i % kGemmCols;
            double v = 0.0;
            if (cc < nc) {
                v = restart_coef(pr, l, c0 + cc);
                if (kFinal) v *= pr.Sb[c0 + cc];
            }
            cs[i] = v;
        }
        __syncthreads();
        for (int t = 0; t < kRowBlock / 256; t++) {
            const int64_t r = r0 + threadIdx.x + t * 256;
            if (r >= n) break;
            double acc[kGemmCols];
#pragma unroll
            for (int cc = 0; cc < kGemmCols; cc++) acc[cc] = 0.0;
            for (int l = 0; l < m; l++) {
                const double x = A[(int64_t)l * n + r];
#pragma unroll
                for (int cc = 0; cc < kGemmCols; cc++)
                    acc[cc] = fma(x, cs[l * kGemmCols + cc], acc[cc]);
            }
            if (kFinal) {
                float *o = pr.PCs + r * pr.k + c0;
#pragma unroll
                for (int cc = 0; cc < kGemmCols; cc++)
                    if (cc < nc) o[cc] = (float)acc[cc];
            } else {
#pragma unroll
                for (int cc = 0; cc < kGemmCols; cc++)
                    if (cc < nc) out[(int64_t)(c0 + cc) * n + r] = acc[cc];
            }
        }
    }
}

// MFMA variant (m <= 128, kout <= 128): each wave takes 16-row tiles and
// accumulates out_tile(16 x 16t) += A_tile(16 x 4) coef(4 x 16t) with
// v_mfma_f64_16x16x4f64 over k-steps of 4, k ascending.  A fragments arrive
// in chunks of kChunkSteps k-steps, the next chunk's loads issued before the
// current chunk's MFMAs (addresses clamped to the last live column, whose
// coefficient rows past m are zero, so no load needs a branch).  One
// workgroup of kGemmThreads threads per data block shares the coefficient
// slab.  Operand layout (16x16x4 f64, checked against numpy): A[i][k] in
// lane i + 16k, B[k][j] in lane j + 16k, D[i][j] in lane j + 16(i%4),
// element i/4.

typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int kMfmaMaxSteps = 32;  // m <= 128
constexpr int kMfmaMaxTiles = 8;   // kout <= 128
constexpr int kChunkSteps = 4;
constexpr int kGemmThreads = 768;  // 12 waves: 3 per SIMD, ~168 VGPRs each

// Row stride of the coefficient slab: the four k-rows read by one B operand
// load land on different LDS bank halves when the stride is an odd multiple
// of 16 doubles.
__host__ __device__ inline int slab_ld(int npad) { return (npad & 31) ? npad : npad + 16; }

// Dynamic LDS of the restart / final products for subspace m (kout <= m).
size_t gemm_lds_bytes(int m) {
    return sizeof(double) * (size_t)((m + 3) & ~3) * (size_t)slab_ld((m + 15) & ~15);
}

// Column-major FP64 outputs go through a per-wave LDS tile (16 x 16, column
// stride 17 doubles so that the lanes writing one row land in different
// banks): written in the matrix-core layout, read back so that each global
// store covers 16 consecutive rows of 4 columns, whole 128-byte lines
// instead of 32 bytes per column (measured 15% less time in the restart
// product on 10M cells; the bytes read and written are unchanged).
constexpr int kStageLd = 17;
constexpr int kStageDoubles = 16 * kStageLd;
size_t gemm_stage_lds_bytes(int m) {
    return gemm_lds_bytes(m) + sizeof(double) * (size_t)kStageDoubles * (kGemmThreads / kWave);
}

namespace {

// coef slab in LDS: cs[l * ldc + c], zero-padded to mpad x npad.  The
// output columns of a tile go in passes of up to kPassTiles 16-column tiles
// (the A fragments are read again for a second pass).  Seven tiles cover
// kout <= 112, so a restart at m = 100 reads U once; their accumulators fit
// the register budget of 12 waves per CU (16 waves with six tiles per pass
// read U twice whenever kc > 96).
constexpr int kPassTiles = 7;

template <typename Store>
__device__ inline void mfma_row_tiles(const double *__restrict__ A, int64_t n,
                                      int m, const double *cs, int ldc,
                                      int ntiles, int64_t row0, int64_t row_end,
                                      int wave, int nwaves, Store store) {
    const int lane = threadIdx.x & (kWave - 1);
    const int i = lane & 15, kq = lane >> 4;
    const int ksteps = (m + 3) >> 2;
    const int nchunks = (ksteps + kChunkSteps - 1) / kChunkSteps;
    for (int64_t r0 = row0 + (int64_t)wave * 16; r0 < row_end; r0 += (int64_t)nwaves * 16) {
        const int64_t r = r0 + i < row_end ? r0 + i : row_end - 1;
        const double *__restrict__ Ar = A + r;
        for (int t0 = 0; t0 < ntiles; t0 += kPassTiles) {
            const int nt = ntiles - t0 < kPassTiles ? ntiles - t0 : kPassTiles;
            v4d acc[kPassTiles];
#pragma unroll
            for (int t = 0; t < kPassTiles; t++) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
            // Chunk loads (clamped columns) and MFMAs; the two fragment
            // buffers alternate, the loop advancing two chunks at a time.
            auto load = [&](int c, double *ab) {
#pragma unroll
                for (int u = 0; u < kChunkSteps; u++) {
                    const int l = (c * kChunkSteps + u) * 4 + kq;
                    ab[u] = Ar[(int64_t)(l < m ? l : m - 1) * n];
                }
            };
            // A fragment that no MFMA reads is still consumed (an empty asm
            // use), so every path leaves no load pending into registers that
            // the next chunk reuses and the waits stay fixed counts.
            auto consume = [&](const double *ab) {
#pragma unroll
                for (int u = 0; u < kChunkSteps; u++) asm volatile("" ::"v"(ab[u]));
            };
            auto mma = [&](int c, const double *ab) {
#pragma unroll
                for (int u = 0; u < kChunkSteps; u++) {
                    const int ks = c * kChunkSteps + u;
                    if (ks < ksteps) {
                        const double *brow = cs + (ks * 4 + kq) * ldc + t0 * 16 + i;
#pragma unroll
                        for (int t = 0; t < kPassTiles; t++)
                            if (t < nt)
                                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ab[u], brow[t * 16], acc[t], 0, 0, 0);
                    } else {
                        asm volatile("" ::"v"(ab[u]));
                    }
                }
            };
            double ab0[kChunkSteps], ab1[kChunkSteps];
            load(0, ab0);
#pragma unroll 1
            for (int c = 0; c < nchunks; c += 2) {
                load(c + 1, ab1);
                __builtin_amdgcn_sched_barrier(0);
                mma(c, ab0);
                __builtin_amdgcn_sched_barrier(0);
                load(c + 2, ab0);
                __builtin_amdgcn_sched_barrier(0);
                if (c + 1 < nchunks)
                    mma(c + 1, ab1);
                else
                    consume(ab1);
                __builtin_amdgcn_sched_barrier(0);
            }
            consume(ab0);
#pragma unroll
            for (int t = 0; t < kPassTiles; t++)
                if (t < nt) store(r0, t0 + t, acc[t]);
        }
    }
}

// Tile stores for mfma_row_tiles: D[i][j] of a 16 x 16 tile is in lane
// j + 16 (i % 4), element i / 4.  `Elem` stores one element (row, col, v).
template <typename Elem>
struct TileByElement {
    Elem elem;
    __device__ void operator()(int64_t r0, int tile, const v4d &d) const {
        const int lane = threadIdx.x & (kWave - 1);
        const int i = lane & 15, kq = lane >> 4;
#pragma unroll
        for (int v = 0; v < 4; v++) elem(r0 + kq + 4 * v, tile * 16 + i, d[v]);
    }
};
template <typename Elem>
__device__ inline TileByElement<Elem> by_element(Elem e) { return TileByElement<Elem>{e}; }

// Column-major FP64 output (out[col * n + row], rows < row_end, cols < kout)
// through the wave's LDS tile `st` (kStageDoubles).
struct TileStaged {
    double *out;
    int64_t n, row_end;
    int kout;
    double *st;
    __device__ void operator()(int64_t r0, int tile, const v4d &d) const {
        const int lane = threadIdx.x & (kWave - 1);
        const int i = lane & 15, kq = lane >> 4;
#pragma unroll
        for (int v = 0; v < 4; v++) st[i * kStageLd + kq + 4 * v] = d[v];
        __builtin_amdgcn_wave_barrier();
        const int64_t row = r0 + i;
#pragma unroll
        for (int w = 0; w < 4; w++) {
            const int c = kq + 4 * w;
            const double y = st[c * kStageLd + i];
            const int col = tile * 16 + c;
            if (row < row_end && col < kout) out[(int64_t)col * n + row] = y;
        }
        __builtin_amdgcn_wave_barrier();
    }
};

}  // namespace

// Launched on one workgroup per compute unit (the slab and staging tiles
// take most of a CU's LDS), each workgroup taking the block list's entries
// blockIdx.x, + gridDim.x, ...: the slab is staged once per problem it meets
// rather than once per block, a launch with no problem in its phase returns
// after one look at the problem table, and the 16-row tiles of consecutive
// blocks go round the waves without restarting at wave 0 (64 tiles per
// block over 12 waves would leave four waves one tile behind per block).
template <bool kFinal, bool kStage = false>
__global__ __launch_bounds__(kGemmThreads) void k_tall_gemm_mfma(Tables T) {
    extern __shared__ double cs[];
    __shared__ int any;
    const int want = kFinal ? PH_FINAL : PH_RESTARTING;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    for (int p = threadIdx.x; p < T.num_probs; p += kGemmThreads)
        if (T.probs[p].phase == want) any = 1;
    __syncthreads();
    if (!any) return;
    const int nw = kGemmThreads / kWave;
    int cur = -1, rot = 0;
    for (int64_t item = blockIdx.x; item < T.num_blocks; item += gridDim.x) {
        const int2 bp = T.blocks[item];
        Prob &pr = T.probs[bp.x];
        if (pr.phase != want) continue;
        const int m = pr.m;
        const int kout = kFinal ? pr.k : pr.kc;
        const int mpad = (m + 3) & ~3, npad = (kout + 15) & ~15, ldc = slab_ld(npad);
        if (bp.x != cur) {
            __syncthreads();   // every wave is done with the previous slab
            for (int idx = threadIdx.x; idx < mpad * npad; idx += kGemmThreads) {
                const int l = idx / npad, c = idx % npad;
                double v = 0.0;
                if (l < m && c < kout) {
                    v = restart_coef(pr, l, c);
                    if (kFinal) v *= pr.Sb[c];
                }
                cs[l * ldc + c] = v;
            }
            __syncthreads();
            cur = bp.x;
        }
        const int64_t n = pr.n;
        const int64_t row0 = (int64_t)bp.y * kRowBlock;
        const int64_t row_end = row0 + kRowBlock < n ? row0 + kRowBlock : n;
        const int row_tiles = (int)((row_end - row0 + 15) / 16);
        const int wv = (wave_id() - rot + nw) % nw;   // the wave's first tile of this block
        rot = (rot + row_tiles) % nw;
        const double *A = pr.U[pr.ubuf];
        if (kFinal) {
            float *out = pr.PCs;
            const int k = pr.k;
            mfma_row_tiles(A, n, m, cs, ldc, npad >> 4, row0, row_end, wv, nw,
                           by_element([&](int64_t row, int col, double v) {
                               if (row < row_end && col < kout) out[row * k + col] = (float)v;
                           }));
        } else if (kStage) {
            double *st = cs + mpad * ldc + wave_id() * kStageDoubles;
            mfma_row_tiles(A, n, m, cs, ldc, npad >> 4, row0, row_end, wv, nw,
                           TileStaged{pr.U[pr.ubuf ^ 1], n, row_end, kout, st});
        } else {
            double *out = pr.U[pr.ubuf ^ 1];
            mfma_row_tiles(A, n, m, cs, ldc, npad >> 4, row0, row_end, wv, nw,
                           by_element([&](int64_t row, int col, double v) {
                               if (row < row_end && col < kout) out[(int64_t)col * n + row] = v;
                           }));
        }
    }
}

// Test entry: out (n x kout, column-major) = A (n x m) C (m x kout).
template <bool kStage>
__global__ __launch_bounds__(kGemmThreads) void k_gemm_mfma_test(const double *A,
                                                                 int64_t n, int m,
                                                                 const double *C,
                                                                 int kout,
                                                                 double *out) {
    extern __shared__ double cs[];
    const int mpad = (m + 3) & ~3, npad = (kout + 15) & ~15, ldc = slab_ld(npad);
    for (int idx = threadIdx.x; idx < mpad * npad; idx += kGemmThreads) {
        const int l = idx / npad, c = idx % npad;
        cs[l * ldc + c] = (l < m && c < kout) ? C[l + (int64_t)c * m] : 0.0;
    }
    __syncthreads();
    const int64_t row0 = (int64_t)blockIdx.x * kRowBlock;
    const int64_t row_end = row0 + kRowBlock < n ? row0 + kRowBlock : n;
    if (kStage) {
        double *st = cs + mpad * ldc + wave_id() * kStageDoubles;
        mfma_row_tiles(A, n, m, cs, ldc, npad >> 4, row0, row_end, wave_id(), kGemmThreads / kWave,
                       TileStaged{out, n, row_end, kout, st});
    } else {
        mfma_row_tiles(A, n, m, cs, ldc, npad >> 4, row0, row_end, wave_id(), kGemmThreads / kWave,
                       by_element([&](int64_t row, int col, double v) {
                           if (row < row_end && col < kout) out[(int64_t)col * n + row] = v;
                       }));
    }
}

hipError_t launch_gemm_mfma_test(const double *A, int64_t n, int m,
                                 const double *C, int kout, double *out,
                                 hipStream_t s) {
    if (m > 4 * kMfmaMaxSteps || kout > 16 * kMfmaMaxTiles || m < 1 || kout < 1) return hipErrorInvalidValue;
    const size_t lds = sizeof(double) * (size_t)((m + 3) & ~3) * slab_ld((kout + 15) & ~15);
    const size_t slds = lds + sizeof(double) * (size_t)kStageDoubles * (kGemmThreads / kWave);
    int dev = 0, max_lds = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return e;
    const unsigned grid = (unsigned)ceil_div(n, kRowBlock);
    if ((int64_t)slds <= (int64_t)max_lds)
        k_gemm_mfma_test<true><<<grid, kGemmThreads, slds, s>>>(A, n, m, C, kout, out);
    else
        k_gemm_mfma_test<false><<<grid, kGemmThreads, lds, s>>>(A, n, m, C, kout, out);
    return hipGetLastError();
}

// Calibration kernels: each wave runs `iters` rounds of 8 independent FP64
// operations on registers (MFMA 16x16x4: 2048 flops each; VALU fma: 128
// per wave), and stores one sum per lane so that nothing is dropped.
__global__ __launch_bounds__(256) void k_fp64_rate_mfma(double *out, int iters) {
    v4d acc[8];
#pragma unroll
    for (int t = 0; t < 8; t++) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
    const double a = threadIdx.x * 1e-3, b = blockIdx.x * 1e-3;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int t = 0; t < 8; t++) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
    }
    double s = 0.0;
#pragma unroll
    for (int t = 0; t < 8; t++) s += (acc[t][0] + acc[t][1]) + (acc[t][2] + acc[t][3]);
    out[(int64_t)blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_fp64_rate_valu(double *out, int iters) {
    double a[8];
#pragma unroll
    for (int t = 0; t < 8; t++) a[t] = threadIdx.x * 1e-3 + t;
    const double x = 1.0000001, y = 1e-9;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int t = 0; t < 8; t++) a[t] = fma(a[t], x, y);
    }
    double s = 0.0;
#pragma unroll
    for (int t = 0; t < 8; t++) s += a[t];
    out[(int64_t)blockIdx.x * 256 + threadIdx.x] = s;
}

hipError_t launch_fp64_rate(int which, int blocks, int iters, double *out,
                            double *flops, hipStream_t s) {
    if (blocks < 1 || iters < 1 || (which != 0 && which != 1)) return hipErrorInvalidValue;
    const double waves = 4.0 * blocks;
    if (which == 0) {
        k_fp64_rate_mfma<<<blocks, 256, 0, s>>>(out, iters);
        *flops = waves * iters * 8 * 2048.0;
    } else {
        k_fp64_rate_valu<<<blocks, 256, 0, s>>>(out, iters);
        *flops = waves * iters * 8 * 128.0;
    }
    return hipGetLastError();
}

// V <- V Vb[:, :kc], V[:, kc] = remainder, vs = remainder / sd (and, one
// pass, R <- R Ub[:, :kc]).  Grid:
// (problem, chunk of 64 genes, group of 4 columns); thread t computes gene
// t % 64 of column 4 * group + t / 64, summing over l ascending.
__global__ __launch_bounds__(256) void k_restart_v(Tables T) {
    const int chunks = (T.max_g + 63) / 64, cgroups = (T.max_m + 3) / 4;
    const int per_prob = chunks * cgroups;
    Prob &pr = T.probs[blockIdx.x / per_prob];
    if (pr.phase != PH_RESTARTING) return;
    const int g = pr.g, m = pr.m, kc = pr.kc, mm = m + (m & 1);
    const int chunk = (blockIdx.x % per_prob) % chunks, cgrp = (blockIdx.x % per_prob) / chunks;
    const int i = chunk * 64 + (threadIdx.x & 63);
    if (i >= g) return;
    const double *V = pr.V[pr.vbuf];
    double *V2 = pr.V[pr.vbuf ^ 1];
    const double *Vb = pr.Jv + (int64_t)mm * mm;
    const int c = cgrp * 4 + (threadIdx.x >> 6);
    if (c < kc) {
        double acc = 0.0;
        for (int l = 0; l < m; l++) acc = fma(V[(int64_t)l * g + i], Vb[l + (int64_t)c * m], acc);
        V2[(int64_t)c * g + i] = acc;
        if (pr.onepass) {
            // X^T U follows U: R <- R Ub[:, :kc].
            const double *R = pr.R[pr.ubuf];
            double ar = 0.0;
            for (int l = 0; l < m; l++) ar = fma(R[(int64_t)l * g + i], pr.Ub[l + (int64_t)c * m], ar);
            pr.R[pr.ubuf ^ 1][(int64_t)c * g + i] = ar;
        }
    }
    if (cgrp == 0 && threadIdx.x < 64) {
        V2[(int64_t)kc * g + i] = pr.rem[i];
        pr.vs[i] = (float)(pr.rem[i] / pr.sd[i]);
    }
}

// Flip the U and V buffers and continue the Lanczos sweep at j = kc.
__global__ __launch_bounds__(256) void k_restart_commit(Tables T) {
    __shared__ double tmp[4];
    Prob &pr = T.probs[blockIdx.x];
    if (pr.phase != PH_RESTARTING) return;
    const double vm = pr.pair ? vs_mean(pr, tmp) : 0.0;  // vs from k_restart_v
    if (threadIdx.x != 0) return;
    pr.vmean = vm;
    pr.ubuf ^= 1;
    pr.vbuf ^= 1;
    pr.j = pr.kc;
    pr.jfirst = pr.kc;
    pr.phase = PH_LANCZOS;
}

__global__ void k_finish(Tables T, int *done_counter) {
    Prob &pr = T.probs[blockIdx.x];
    if (threadIdx.x != 0 || pr.phase != PH_FINAL) return;
    pr.phase = PH_DONE;
    atomicAdd(done_counter, 1);
}

// ================================================================ launchers

// Dynamic LDS of a sparse-product launch: what the kernel needs, raised when
// `wg` > 0 so that wg + 1 workgroups no longer fit on a CU of `lds_per_cu`
// bytes (and wg still do), unless that exceeds what a workgroup may request.
// Fewer resident waves contend less for LDS banks in X^T u's
// read-modify-write accumulation; the autotuner picks the cap.
static size_t capped_lds(size_t need, int wg, const Tables &T) {
    if (wg <= 0 || T.lds_per_cu <= 0) return need;
    const size_t cap = (size_t)T.lds_per_cu / (size_t)(wg + 1) + 256;
    if (cap > (size_t)T.max_lds) return need;
    return need > cap ? need : cap;
}

size_t lds_bytes_needed(int g, int m) {
    size_t need = spmv_t_lds_bytes(g);
    if (g <= kMaxSortedGenes) need = std::max(need, spmv_t_sorted_lds_bytes(g));
    if (g <= kMaxPairGenes) need = std::max(need, spmv_pair_lds_bytes(g));
    if (g <= kMaxPairGenes) need = std::max(need, sizeof(float) * 8 * ((size_t)g + 1));  // pair colstats
    need = std::max(need, spmv_lds_bytes(g));
    need = std::max(need, sizeof(float) * 4 * ((size_t)std::min(g, kGeneTile) + 1));  // colstats
    need = std::max(need, sizeof(double) * (size_t)m * kGemmCols);   // gemm
    need = std::max(need, sizeof(double) * (size_t)m + sizeof(int) * (size_t)m);
    return need;
}

namespace {

template <int R>
void launch_spmv_modes(const Tables &T, size_t lds_ldsv, size_t lds_global, hipStream_t s) {
    const unsigned nblk = (unsigned)T.num_blocks;
    if (T.vmodes & (1 << kModeLds)) k_spmv<R, kModeLds><<<nblk, 256, lds_ldsv, s>>>(T);
    if (T.vmodes & (1 << kModeGlobalV)) k_spmv<R, kModeGlobalV><<<nblk, 256, lds_global, s>>>(T);
    if (T.vmodes & (1 << kModeWide)) k_spmv<R, kModeWide><<<nblk, 256, lds_global, s>>>(T);
}

template <int R>
void launch_spmv_t_tiles(const Tables &T, size_t lds, hipStream_t s) {
    const unsigned nblk = (unsigned)T.num_blocks;
    const bool narrow = (T.tmodes & ((1 << kModeLds) | (1 << kModeGlobalV))) != 0;
    const bool wide = (T.tmodes & (1 << kModeWide)) != 0;
    if (T.max_g <= kGeneTile) {   // one tile: no range test per element
        if (narrow) k_spmv_t<R, uint16_t, false><<<nblk, 256, lds, s>>>(T, 0);
        return;
    }
    for (int c0 = 0; c0 < T.max_g; c0 += kGeneTile) {
        if (narrow) k_spmv_t<R, uint16_t, true><<<nblk, 256, lds, s>>>(T, c0);
        if (wide) k_spmv_t<R, uint32_t, true><<<nblk, 256, lds, s>>>(T, c0);
    }
}

}  // namespace

// X v for every problem of the table (one launch per index mode present),
// `group` segments per load group, at most `wg` workgroups per CU.
// Problems with the one-pass pair (kPairBit) get both products here, in
// k_spmv_pair (groups of kPairGroup segments, the table's).
static void launch_spmv_all(const Tables &T, int group, int wg, hipStream_t s) {
    if (T.tmodes & kPairBit) {
        const size_t lp = capped_lds(spmv_pair_lds_bytes(std::min(T.max_g, kMaxPairGenes)), wg, T);
        k_spmv_pair<kPairGroup><<<(unsigned)T.num_blocks, 256, lp, s>>>(T);
    }
    if (!T.vmodes) return;
    const size_t lv = capped_lds(spmv_lds_bytes(std::min(T.max_g, kMaxLdsV)), wg, T);
    const size_t lg = capped_lds(spmv_lds_bytes(kMaxLdsV + 1), wg, T);
    switch (group) {
        case 2: launch_spmv_modes<2>(T, lv, lg, s); break;
        case 8: launch_spmv_modes<8>(T, lv, lg, s); break;
        default: launch_spmv_modes<4>(T, lv, lg, s); break;
    }
}

// Block partials of X^T u for every problem (one launch per gene tile and
// index width present).
static void launch_spmv_t_all(const Tables &T, int group, int wg, hipStream_t s) {
    if (T.tmodes & kSortedBit)
        k_spmv_t_sorted<2><<<(unsigned)T.num_blocks, 256,
                             spmv_t_sorted_lds_bytes(std::min(T.max_g, kMaxSortedGenes)), s>>>(T);
    if (!(T.tmodes & ~(kSortedBit | kPairBit))) return;
    const size_t lds = capped_lds(spmv_t_lds_bytes(T.max_g), wg, T);
    switch (group) {
        case 2: launch_spmv_t_tiles<2>(T, lds, s); break;
        case 8: launch_spmv_t_tiles<8>(T, lds, s); break;
        default: launch_spmv_t_tiles<4>(T, lds, s); break;
    }
}

hipError_t launch_spmv_variant(const Tables &T, int which, int group, int wg,
                               hipStream_t s) {
    if (which == 0)
        launch_spmv_all(T, group, wg, s);
    else
        launch_spmv_t_all(T, group, wg, s);
    return hipGetLastError();
}

hipError_t launch_setup(const Tables &T, hipStream_t s) {
    if (T.num_probs == 0) return hipSuccess;
    const size_t lds4g = sizeof(float) * 4 * ((size_t)std::min(T.max_g, kGeneTile) + 1);
    const unsigned nblk = (unsigned)T.num_blocks;
    const bool narrow = (T.modes & ((1 << kModeLds) | (1 << kModeGlobalV))) != 0;
    const bool wide = (T.modes & (1 << kModeWide)) != 0;
    if (T.tmodes & kPairBit)
        k_colstats_pair<<<nblk, 256, sizeof(float) * 8 * ((size_t)std::min(T.max_g, kMaxPairGenes) + 1), s>>>(T);
    if (T.max_g <= kGeneTile) {   // one tile: no range test per element
        k_colstats<false, uint16_t><<<nblk, 256, lds4g, s>>>(T, 0);
        k_colstats<true, uint16_t><<<nblk, 256, lds4g, s>>>(T, 0);
    } else {
        for (int c0 = 0; c0 < T.max_g; c0 += kGeneTile) {
            if (narrow) {
                k_colstats<false, uint16_t><<<nblk, 256, lds4g, s>>>(T, c0);
                k_colstats<true, uint16_t><<<nblk, 256, lds4g, s>>>(T, c0);
            }
            if (wide) {
                k_colstats<false, uint32_t><<<nblk, 256, lds4g, s>>>(T, c0);
                k_colstats<true, uint32_t><<<nblk, 256, lds4g, s>>>(T, c0);
            }
        }
    }
    k_reduce_groups<<<dim3((unsigned)ceil_div(2 * T.max_g, 256), (unsigned)T.num_groups), 256, 0, s>>>(T, 1);
    k_setup_finish<<<T.num_probs, 256, 0, s>>>(T);
    return hipGetLastError();
}

hipError_t launch_macro_step(const Tables &T, int *done_counter,
                             hipStream_t s, const LaunchHook *hook,
                             hipStream_t side, hipEvent_t fork,
                             hipEvent_t join) {
    if (T.num_probs == 0) return hipSuccess;
    const unsigned nblk = (unsigned)T.num_blocks;
    const int g = T.max_g, m = T.max_m;
    auto mark = [&](int kind, int end) {
        if (hook) hook->mark(hook->user, kind, end);
    };
    // Kernel classes follow enum bpca_kernel_class in bpca.h.
    auto svd = [&](hipStream_t st) {
        if ((int64_t)svd_lds_bytes(m, true) <= T.max_lds && m <= 128)
            k_svd<true><<<T.num_probs, 1024, svd_lds_bytes(m, true), st>>>(T);
        else
            k_svd<false><<<T.num_probs, 1024, svd_lds_bytes(m, false), st>>>(T);
    };
    k_stage_svd<<<T.num_probs, 64, 0, s>>>(T);
    const bool forked = side && fork && join;
    if (forked) {
        hipError_t e = hipEventRecord(fork, s);
        if (e == hipSuccess) e = hipStreamWaitEvent(side, fork, 0);
        if (e != hipSuccess) return e;
        svd(side);
        e = hipEventRecord(join, side);
        if (e != hipSuccess) return e;
    } else {
        mark(4, 0);
        svd(s);
        mark(4, 1);
    }
    mark(0, 0);
    launch_spmv_all(T, T.spmv_group, T.spmv_wg, s);
    mark(0, 1);
    mark(2, 0);
    const bool two_pass = (T.rmodes & kReorthTwoBit) != 0, one_pass = (T.rmodes & kReorthOneBit) != 0;
    if (two_pass) k_reorth_partials<<<nblk, 256, 0, s>>>(T);
    if (one_pass) k_reorth_onepass<<<nblk, 256, 0, s>>>(T);
    k_reduce_columns<<<T.num_probs * (m + 1), 256, 0, s>>>(T);
    if (two_pass) k_reorth_apply<<<nblk, 256, sizeof(double) * (m + 1), s>>>(T);
    k_alpha<<<T.num_probs, 256, 0, s>>>(T);
    if (T.tmodes & kPairBit) k_pair_last_column<<<nblk, 256, 0, s>>>(T);
    mark(2, 1);
    mark(1, 0);
    launch_spmv_t_all(T, T.spmvt_group, T.spmvt_wg, s);
    mark(1, 1);
    mark(3, 0);
    k_reduce_groups<<<dim3((unsigned)ceil_div(g, 256), (unsigned)T.num_groups), 256, 0, s>>>(T, 0);
    const unsigned vchunks = (unsigned)ceil_div(g, kVChunk);
    k_vside_w<<<T.num_probs * (unsigned)ceil_div(g, kVwGenes), 256, 0, s>>>(T);
    k_vside_dot<<<T.num_probs * (unsigned)m, 256, 0, s>>>(T);
    k_vside_upd<<<T.num_probs * vchunks, kVChunk, 0, s>>>(T);
    k_vside_fin<<<T.num_probs, 256, 0, s>>>(T);
    mark(3, 1);
    if (forked) {
        const hipError_t e = hipStreamWaitEvent(s, join, 0);
        if (e != hipSuccess) return e;
    } else {
        // Nothing to overlap with: sweeps that ended in this step restart in
        // this step (same arithmetic, one macro step sooner).
        mark(4, 0);
        k_stage_svd<<<T.num_probs, 64, 0, s>>>(T);
        svd(s);
        mark(4, 1);
    }
    mark(4, 0);
    const size_t mfma_lds = gemm_lds_bytes(m), stage_lds = gemm_stage_lds_bytes(m);
    const bool mfma = m <= 4 * kMfmaMaxSteps && (int64_t)mfma_lds <= T.max_lds;
    // One workgroup per compute unit (the LDS request allows no second).
    const unsigned ngemm = T.num_cus > 0 && (int64_t)T.num_cus < T.num_blocks ? (unsigned)T.num_cus : nblk;
    if (mfma && (int64_t)stage_lds <= T.max_lds)
        k_tall_gemm_mfma<false, true><<<ngemm, kGemmThreads, stage_lds, s>>>(T);
    else if (mfma)
        k_tall_gemm_mfma<false><<<ngemm, kGemmThreads, mfma_lds, s>>>(T);
    else
        k_tall_gemm<false><<<nblk, 256, sizeof(double) * m * kGemmCols, s>>>(T);
    k_restart_v<<<T.num_probs * (unsigned)(ceil_div(g, 64) * ceil_div(m, 4)), 256, 0, s>>>(T);
    k_restart_commit<<<T.num_probs, 256, 0, s>>>(T);
    mark(4, 1);
    mark(5, 0);
    if (mfma)
        k_tall_gemm_mfma<true><<<ngemm, kGemmThreads, mfma_lds, s>>>(T);
    else
        k_tall_gemm<true><<<nblk, 256, sizeof(double) * m * kGemmCols, s>>>(T);
    k_finish<<<T.num_probs, 64, 0, s>>>(T, done_counter);
    mark(5, 1);
    return hipGetLastError();
}

// ------------------------------------------------------- validation etc.

template <typename IP, typename IX, bool kWrite = true>
__global__ __launch_bounds__(256) void k_validate_pack(
    const float *__restrict__ data, const int32_t *__restrict__ indices,
    const IP *__restrict__ indptr, int64_t n, int64_t nnz, int g,
    float *__restrict__ data4, IX *__restrict__ idx, int *flag) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t row = (int64_t)blockIdx.x * 4 + wave_id();
    if (row >= n) return;
    const int64_t b = indptr[row], e = indptr[row + 1];
    int bad = 0;
    if (row == 0 && b != 0) bad = 1;
    if (e < b || b < 0 || e > nnz) bad = 1;
    if (!bad) {
        for (int64_t p = b + lane; p < e; p += kWave) {
            const int32_t c = indices[p];
            const int32_t prev = p > b ? indices[p - 1] : -1;
            if (c < 0 || c >= g || c <= prev) bad = 1;
            if (kWrite) {
                idx[p] = (IX)c;
                data4[p] = data[p];
            }
        }
    }
    if (bad) atomicOr(flag, 1);
}

hipError_t launch_validate_pack(const float *data, const int32_t *indices,
                                const void *indptr, int indptr_width, int64_t n,
                                int64_t nnz, int g, float *data4, void *idx,
                                int *flag, hipStream_t s, bool check_only) {
    dim3 grid((unsigned)ceil_div(n, 4));
    const bool wide = g > kMaxPacked16;
    if (check_only) {
        if (indptr_width == 8)
            k_validate_pack<int64_t, uint16_t, false><<<grid, 256, 0, s>>>(data, indices, (const int64_t *)indptr, n, nnz, g, nullptr, nullptr, flag);
        else
            k_validate_pack<int32_t, uint16_t, false><<<grid, 256, 0, s>>>(data, indices, (const int32_t *)indptr, n, nnz, g, nullptr, nullptr, flag);
        return hipGetLastError();
    }
    if (indptr_width == 8 && wide)
        k_validate_pack<int64_t, uint32_t><<<grid, 256, 0, s>>>(data, indices, (const int64_t *)indptr, n, nnz, g, data4, (uint32_t *)idx, flag);
    else if (indptr_width == 8)
        k_validate_pack<int64_t, uint16_t><<<grid, 256, 0, s>>>(data, indices, (const int64_t *)indptr, n, nnz, g, data4, (uint16_t *)idx, flag);
    else if (wide)
        k_validate_pack<int32_t, uint32_t><<<grid, 256, 0, s>>>(data, indices, (const int32_t *)indptr, n, nnz, g, data4, (uint32_t *)idx, flag);
    else
        k_validate_pack<int32_t, uint16_t><<<grid, 256, 0, s>>>(data, indices, (const int32_t *)indptr, n, nnz, g, data4, (uint16_t *)idx, flag);
    return hipGetLastError();
}

// ------------------------------------------ packed copy padded to quads

namespace {

// Exclusive prefix sum over a 256-thread block of int64 values, in thread
// order; tmp holds 4 int64.  Returns the block total in *total.
__device__ inline int64_t block_exclusive_scan(int64_t v, int64_t *tmp, int64_t *total) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    int64_t x = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int64_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == kWave - 1) tmp[w] = x;
    __syncthreads();
    int64_t before = 0;
    for (int i = 0; i < w; i++) before += tmp[i];
    *total = tmp[0] + tmp[1] + tmp[2] + tmp[3];
    __syncthreads();
    return before + x - v;
}

// Row r's length rounded up to whole quads, plus the sub-block's zero quad
// after the last row of each 256-row sub-block.
template <typename IP>
__device__ __forceinline__ int64_t padded_len(const IP *indptr, int64_t r, int64_t n) {
    const int64_t q = (((int64_t)indptr[r + 1] - (int64_t)indptr[r]) + 3) & ~(int64_t)3;
    return q + ((r % kSubRows == kSubRows - 1 || r == n - 1) ? 4 : 0);
}

}  // namespace

// Sum of the padded lengths of the 1024 rows of block b.
template <typename IP>
__global__ __launch_bounds__(256) void k_pad_block_sums(const IP *__restrict__ indptr, int64_t n,
                                                        int64_t *__restrict__ bsum) {
    __shared__ int64_t tmp[4];
    int64_t v = 0;
    for (int t = 0; t < 4; t++) {
        const int64_t r = (int64_t)blockIdx.x * kRowBlock + 4 * threadIdx.x + t;
        if (r < n) v += padded_len(indptr, r, n);
    }
    int64_t total;
    (void)block_exclusive_scan(v, tmp, &total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// bsum[0..nb) -> exclusive prefix sums starting at `start`, in place;
// bsum[nb] = the total plus start.
__global__ __launch_bounds__(256) void k_scan_in_place(int64_t *__restrict__ bsum, int64_t nb,
                                                       int64_t start) {
    __shared__ int64_t tmp[4];
    int64_t carry = start;
    for (int64_t b0 = 0; b0 < nb; b0 += 256) {
        const int64_t b = b0 + threadIdx.x;
        const int64_t v = b < nb ? bsum[b] : 0;
        int64_t total;
        const int64_t x = block_exclusive_scan(v, tmp, &total);
        if (b < nb) bsum[b] = carry + x;
        carry += total;
    }
    if (threadIdx.x == 0) bsum[nb] = carry;
}

// poff[r] for the rows of block b (4 consecutive rows a thread), and
// poff[n].
template <typename IP>
__global__ __launch_bounds__(256) void k_pad_offsets(const IP *__restrict__ indptr, int64_t n,
                                                     const int64_t *__restrict__ boff,
                                                     int64_t *__restrict__ poff) {
    __shared__ int64_t tmp[4];
    int64_t len[4], v = 0;
    for (int t = 0; t < 4; t++) {
        const int64_t r = (int64_t)blockIdx.x * kRowBlock + 4 * threadIdx.x + t;
        len[t] = r < n ? padded_len(indptr, r, n) : 0;
        v += len[t];
    }
    int64_t total;
    int64_t x = boff[blockIdx.x] + block_exclusive_scan(v, tmp, &total);
    for (int t = 0; t < 4; t++) {
        const int64_t r = (int64_t)blockIdx.x * kRowBlock + 4 * threadIdx.x + t;
        if (r < n) poff[r] = x;
        x += len[t];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) poff[n] = boff[gridDim.x];
}

// One wave per row, one output quad a lane: the row's elements at poff[r]
// (a multiple of 4), indices as gene << 2, then zeros of gene g to
// poff[r + 1] (the next multiple of 4, and the sub-block's zero quad after
// its last row); each lane writes its quad with one 16-byte and one 8-byte
// store.
template <typename IP>
__global__ __launch_bounds__(256) void k_pack_padded(const float *__restrict__ data,
                                                     const int32_t *__restrict__ indices,
                                                     const IP *__restrict__ indptr, int64_t n, int g,
                                                     const int64_t *__restrict__ poff,
                                                     float *__restrict__ data4,
                                                     uint16_t *__restrict__ idx) {
    const int lane = threadIdx.x & (kWave - 1);
    const unsigned pad = (unsigned)(g << 2);
    const int64_t row = (int64_t)blockIdx.x * 4 + wave_id();
    if (row >= n) return;
    const int64_t b = indptr[row], len = (int64_t)indptr[row + 1] - b;
    const int64_t o = poff[row], plen = poff[row + 1] - o;
    f32x4 *dq = (f32x4 *)(data4 + o);
    u32x2 *iq = (u32x2 *)(idx + o);
    for (int64_t t = 4 * lane; t < plen; t += 4 * kWave) {
        float x[4];
        unsigned c[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const bool in = t + k < len;
            x[k] = in ? data[b + t + k] : 0.0f;
            c[k] = in ? (unsigned)indices[b + t + k] << 2 : pad;
        }
        dq[t >> 2] = f32x4{x[0], x[1], x[2], x[3]};
        iq[t >> 2] = u32x2{c[0] | (c[1] << 16), c[2] | (c[3] << 16)};
    }
}

int64_t pad_packed_bound(int64_t nnz, int64_t n) {
    return (nnz + 3 * n + 4 * ceil_div(n, kSubRows) + 3) / 4 * 4;
}

hipError_t launch_pack_padded(const float *data, const int32_t *indices,
                              const void *indptr, int indptr_width, int64_t n,
                              int g, int64_t *poff, int64_t *scratch,
                              float *data4, uint16_t *idx, hipStream_t s) {
    if (n <= 0 || (g << 2) > 0xffff) return hipErrorInvalidValue;
    const int64_t nb = ceil_div(n, kRowBlock);
    if (indptr_width == 8) {
        const int64_t *ip = (const int64_t *)indptr;
        k_pad_block_sums<<<(unsigned)nb, 256, 0, s>>>(ip, n, scratch);
        k_scan_in_place<<<1, 256, 0, s>>>(scratch, nb, 0);
        k_pad_offsets<<<(unsigned)nb, 256, 0, s>>>(ip, n, scratch, poff);
        k_pack_padded<<<(unsigned)ceil_div(n, 4), 256, 0, s>>>(data, indices, ip, n, g, poff, data4, idx);
    } else {
        const int32_t *ip = (const int32_t *)indptr;
        k_pad_block_sums<<<(unsigned)nb, 256, 0, s>>>(ip, n, scratch);
        k_scan_in_place<<<1, 256, 0, s>>>(scratch, nb, 0);
        k_pad_offsets<<<(unsigned)nb, 256, 0, s>>>(ip, n, scratch, poff);
        k_pack_padded<<<(unsigned)ceil_div(n, 4), 256, 0, s>>>(data, indices, ip, n, g, poff, data4, idx);
    }
    return hipGetLastError();
}

// One thread per 256-row sub-block walks its rows (the pair kernel's rule:
// a row of two segments is not started in a group's last slot) and counts
// its segments (kWrite = false: into segbase[sub]) or writes them from
// segbase[sub].
template <bool kWrite>
__global__ __launch_bounds__(256) void k_pair_segments(const int64_t *__restrict__ poff, int64_t n,
                                                       int64_t *__restrict__ segbase,
                                                       uint16_t *__restrict__ segs) {
    const int64_t sub = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t nsub = (n + kSubRows - 1) / kSubRows;
    if (sub >= nsub) return;
    const int64_t r0 = sub * kSubRows;
    const int nrows = (int)(n - r0 < kSubRows ? n - r0 : kSubRows);
    const int64_t base = kWrite ? segbase[sub] : 0;
    int64_t cnt = 0;
    int slot = 0;
    auto emit = [&](unsigned d) {
        if (kWrite) segs[base + cnt] = (uint16_t)d;
        cnt++;
        slot = (slot + 1) & (kPairGroup - 1);
    };
    for (int r = 0; r < nrows; r++) {
        // Quads of the row (the sub-block's zero quad follows its last row).
        const int64_t nq = ((poff[r0 + r + 1] - poff[r0 + r]) >> 2) - (r == nrows - 1 ? 1 : 0);
        const bool lng = nq > 2 * kWave;                              // three or more segments
        if (nq > kWave && !lng && slot == kPairGroup - 1) emit(0);     // empty last slot
        int64_t left = nq;
        do {
            const int64_t take = left < kWave ? left : kWave;
            left -= take;
            emit((unsigned)take | (left == 0 ? 128u : 0u) | (lng ? 256u : 0u));
        } while (left > 0);
    }
    while (slot != 0) emit(0);
    if (!kWrite) segbase[sub] = cnt;
}

int64_t pair_segments_bound(int64_t n, int64_t padded_total) {
    // A row of q quads has ceil(q / 64) segments (one if empty), at most
    // 1 + q / 64; empty slots before two-segment rows number at most q / 64
    // over the rows; each sub-block adds at most kPairGroup - 1.
    return n + 2 * (padded_total / 4 / kWave) + (kPairGroup - 1) * ceil_div(n, kSubRows) + kPairGroup;
}

hipError_t launch_pair_segments(const int64_t *poff, int64_t n, int64_t *segbase,
                                uint16_t *segs, hipStream_t s) {
    if (n <= 0) return hipErrorInvalidValue;
    const int64_t nsub = ceil_div(n, kSubRows);
    const unsigned grid = (unsigned)ceil_div(nsub, 256);
    k_pair_segments<false><<<grid, 256, 0, s>>>(poff, n, segbase, segs);
    k_scan_in_place<<<1, 256, 0, s>>>(segbase, nsub, 0);
    k_pair_segments<true><<<grid, 256, 0, s>>>(poff, n, segbase, segs);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_widen(const int32_t *__restrict__ in,
                                               int64_t count,
                                               int64_t *__restrict__ out) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += stride)
        out[i] = (int64_t)in[i];
}

hipError_t launch_widen_indptr(const int32_t *in, int64_t count, int64_t *out,
                               hipStream_t s) {
    if (count <= 0) return hipSuccess;
    const int grid = (int)std::min<int64_t>(ceil_div(count, 256), 65536);
    k_widen<<<grid, 256, 0, s>>>(in, count, out);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_narrow(const int64_t *__restrict__ in,
                                                int64_t nnz, int g,
                                                int32_t *__restrict__ out,
                                                int *flag) {
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
                                 int32_t *out, int *flag, hipStream_t s) {
    if (nnz <= 0) return hipSuccess;
    const int grid = (int)std::min<int64_t>(ceil_div(nnz, 256), 65536);
    k_narrow<<<grid, 256, 0, s>>>(in, nnz, g, out, flag);
    return hipGetLastError();
}

// Byte fill with zeros by a kernel (callers that must not record memset
// nodes or allocate use it for counters, flags and padding).
__global__ __launch_bounds__(256) void k_zero(unsigned char *__restrict__ p,
                                              size_t bytes) {
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < bytes; i += stride) p[i] = 0;
}

hipError_t launch_zero(void *p, size_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    const unsigned grid = (unsigned)std::min<size_t>((bytes + 255) / 256, 4096);
    k_zero<<<grid, 256, 0, s>>>((unsigned char *)p, bytes);
    return hipGetLastError();
}

// Single-matrix SVD through the same device code (test entry point).
template <bool kLds>
__global__ __launch_bounds__(1024) void k_svd_single(const double *B, int m,
                                                     double *Wg, double *Jg,
                                                     double *Ub, double *Sb,
                                                     double *Vb, int *sweeps) {
    __shared__ int sh_int[2];
    extern __shared__ double sh[];
    const int mm = m + (m & 1);
    double *W = kLds ? sh + m + (m + 1) / 2 : Wg;
    double *Jv = kLds ? W + m * mm : Jg;
    for (int idx = threadIdx.x; idx < m * mm; idx += blockDim.x)
        W[idx] = idx / m < m ? B[idx] : 0.0;
    for (int idx = threadIdx.x; idx < mm * mm; idx += blockDim.x)
        Jv[idx] = (idx / mm == idx % mm) ? 1.0 : 0.0;
    __syncthreads();
    if (kLds)
        jacobi_svd_lds(W, Jv, m, mm, &sh_int[0], &sh_int[1]);
    else
        jacobi_svd(W, Jv, m, mm, &sh_int[0], &sh_int[1]);
    __syncthreads();
    svd_sort_outputs(W, Jv, m, mm, sh, (int *)(sh + m), Sb, Ub, Vb);
    if (threadIdx.x == 0) *sweeps = sh_int[0];
}

hipError_t launch_svd_single(const double *B, int m, double *W, double *Jv,
                             double *Ub, double *Sb, int *sweeps,
                             int64_t max_lds, hipStream_t s) {
    const int mm = m + (m & 1);
    double *Vb = Jv + (int64_t)mm * mm;
    if ((int64_t)svd_lds_bytes(m, true) <= max_lds && m <= 128)
        k_svd_single<true><<<1, 1024, svd_lds_bytes(m, true), s>>>(B, m, W, Jv, Ub, Sb, Vb, sweeps);
    else
        k_svd_single<false><<<1, 1024, svd_lds_bytes(m, false), s>>>(B, m, W, Jv, Ub, Sb, Vb, sweeps);
    return hipGetLastError();
}

}  // namespace bpca
