// Device side of the IRLBA engine (kernels.hip, synth.hip).
//
// Every kernel is batch-native: it reads a device-resident table of Prob
// descriptors and works on the problems whose phase it serves, so one solve,
// a batch of uneven problems and a serving mix all run the same arithmetic.
//
// Data layout in HBM for one problem (n cells, g genes, m subspace, k PCs):
//   X        CSR: data float32[nnz4], idx uint16[nnz4] (uint32 when g > 65535) -- copies made at
//            admission, 16-byte aligned and zero-padded to nnz4 = nnz rounded
//            up to a multiple of 4, so every aligned quad of 4 elements can be
//            loaded whole -- and indptr int64[n+1]
//   U[0/1]   float64 column-major n x m (ld n): left Lanczos vectors
//   V[0/1]   float64 column-major g x (m+1) (ld g): right Lanczos vectors
//   u32      float32[2 (n + 64)]: raw X v' of the current and the previous step
//   R[0/1]   float64 column-major g x (m+1) (ld g): X^T U (one-pass problems)
//   vs       float32[g]: v_j / sd, the X v' input
//   B        float64 column-major m x (m+1): projected matrix
//   W, Jv    float64 m x m: Jacobi work / right singular vectors
//   Ub, Sb   float64 m x m / m: left singular vectors and values of B
//   P        float64 (m+1) x nb: reorthogonalisation partials per data block
//   tpart    float32 nb x 2g: X^T u (or column-statistics) block partials
//   grp      float64 ngrp x 2g: first-level group sums of tpart
//   vw, vdot, vsq  float64 g / m / ceil(g/256): V-side work vectors
// Reductions over cells are keyed to fixed data blocks of kRowBlock rows
// (and fixed sub-blocks / groups inside them), never to the launch shape.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>
#include <cstdint>

namespace bpca {

constexpr int kWave = 64;

// The value of lane ^ stride (stride a power of two below 64, known once
// the loops around a call are unrolled): strides 1, 2, 4 and 8 by DPP moves
// within a row of 16 lanes (quad permutations, the half-row mirror then a
// quad reversal, a rotation by 8), 16 by a swizzle (xor mask within 32
// lanes), 32 by a permute.
__device__ __forceinline__ int xor_lane(int v, int stride) {
    switch (stride) {
        case 1: return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);   // quad_perm 1,0,3,2
        case 2: return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);   // quad_perm 2,3,0,1
        case 4: {
            const int t = __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);  // row_half_mirror
            return __builtin_amdgcn_update_dpp(0, t, 0x1B, 0xF, 0xF, false);          // quad_perm 3,2,1,0
        }
        case 8: return __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, false);  // row_ror 8
        case 16: return __builtin_amdgcn_ds_swizzle(v, 0x401F);                  // xor 16
        default: return __shfl_xor(v, stride);
    }
}

// xor_lane of a double, one 32-bit half at a time.
__device__ __forceinline__ double xor_lane_f64(double v, int stride) {
    const long long b = __double_as_longlong(v);
    const int lo = xor_lane((int)(b & 0xffffffffLL), stride);
    const int hi = xor_lane((int)(b >> 32), stride);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
constexpr int kRowBlock = 1024;   // rows per data block
constexpr int kSubRows = 256;     // rows per wave sub-block of X^T u
constexpr int kGroup = 64;        // data blocks per first-level group
constexpr int kGeneTile = 8192;   // genes per X^T u / column-statistics launch (LDS accumulators)
constexpr int kMaxLdsV = 32768;   // X v stages v in LDS up to this many genes
constexpr int kMaxPacked16 = 65535;  // column indices packed to 16 bits up to this g
constexpr int kMaxSortedGenes = 4096;  // X^T u from the gene-sorted copy up to this g
constexpr int kTChunk = 16;         // entries per lane in the gene-sorted X^T u
constexpr int kTWindow = kWave * kTChunk;  // entries per wave step of that kernel
constexpr int kSortedBit = 1 << 3;  // Tables::tmodes: some problem uses the gene-sorted X^T u
constexpr int kPairBit = 1 << 4;    // Tables::tmodes: some problem forms X v and X^T u in one pass
constexpr int kMaxPairGenes = 4096; // X v and X^T u in one pass up to this g (LDS)
constexpr int kPairGroup = 4;       // segments per load group of the pair kernel
constexpr int kReorthTwoBit = 1;    // Tables::rmodes: some problem reorthogonalises in two passes
constexpr int kReorthOneBit = 2;    //                 some problem in one pass
constexpr int64_t kOnePassMinCells = 65536;  // automatic one-pass choice: n >= max(8 g, this)

// Index modes of a problem (Prob::mode): how its sparse products load.
enum IndexMode : int32_t {
    kModeLds = 0,      // 16-bit column indices, v staged in LDS (g <= kMaxLdsV)
    kModeGlobalV = 1,  // 16-bit column indices, v read from global memory
    kModeWide = 2      // 32-bit column indices (g > kMaxPacked16), v from global
};
constexpr int kGemmCols = 32;     // output columns per restart-GEMM pass
constexpr int kMaxSubspace = 1024;  // largest subspace_size (m) accepted

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

enum Phase : int32_t {
    PH_IDLE = 0,        // slot unused or finished and collected
    PH_SETUP = 1,       // column statistics pending
    PH_LANCZOS = 2,     // Lanczos step j pending
    PH_RESTART = 3,     // sweep done, waiting for the next macro step's SVD
    PH_RESTARTING = 4,  // U, V restart products pending
    PH_FINAL = 5,       // PCs product pending
    PH_DONE = 6,        // results written
    PH_SVD = 7          // small SVD + convergence test run in this macro step
};

struct Prob {
    // Input CSR and output.
    const float *data;       // packed copy of the values (see above)
    const void *idx;         // packed column indices: uint16 (mode < kModeWide) or uint32
    const int64_t *indptr;
    // Gene-sorted copy of X for X^T u (tsorted: g <= kMaxSortedGenes), made
    // at admission.  Within each 256-row sub-block the entries are ordered by
    // (gene, row), at the sub-block's own CSR positions:
    const float *tval;       // values
    const uint8_t *trow;     // row within the sub-block, stored as tsort_row_slot(row)
    const uint16_t *thb;     // run starts: bit j of word q marks position 16 q + j
    const uint16_t *tlens;   // entries per (sub-block, gene), row stride tgl
    float *PCs;  // n x k float32 row-major
    int64_t n, nb, ngrp;
    int32_t g, k, m, maxit;
    float tol;
    int32_t id;
    // Workspace.
    double *U[2];
    double *V[2];
    float *u32;
    double *P, *red, *sq;   // sq: block sums of squares [nb], then block sums [nb]
    float *tpart;
    double *grp;
    double *B, *sd, *rem;
    float *vs;
    double *W, *Jv, *Ub, *Sb;
    double *vw, *vdot, *vsq;  // V-side work: w (g), dots (m), chunk sums
    // State, advanced by the kernels.
    int32_t mode;            // IndexMode
    int32_t phase, j, kc, iter;
    int32_t ubuf, vbuf, converged, nconv;
    int32_t svd_sweeps, pad0;
    int32_t tsorted, tgl;    // gene-sorted X^T u in use; tlens row stride (g rounded up to 8)
    // One-pass reorthogonalisation (onepass != 0, n >= 8 g): column j of U is
    // written by step j + 1's pass over U (the last column of a sweep holds
    // (u - mean) / alpha instead, which the restart and final coefficients
    // correct), and X^T U[:, j] is formed in gene space from X^T (u - mean)
    // and R.
    int32_t onepass;
    int32_t jfirst;          // first step of the current sweep (0, then kc)
    double *R[2];            // float64 g x (m+1) (ld g): X^T U, flipped with U
    // X v and X^T u from one read of X (pair != 0: one-pass problems with
    // 16-bit indices and g <= kMaxPairGenes).  X^T is applied to
    // u - vmean, vmean = colsum . vs / n being the mean of X vs known before
    // the pass; k_vside_w moves the result to X^T (u - mean of u).
    int32_t pair;
    // Pair problems: X is packed with every row padded to whole quads
    // (padding: value 0, gene g) and a zero quad after the last row of each
    // 256-row sub-block, and each packed index holds gene << idx_shift (2:
    // the byte offset of the gene's FP32 slot in LDS).  0 for other
    // problems.
    int32_t idx_shift;
    // Pair problems: the segment table of each 256-row sub-block (built at
    // admission, launch_pair_segments): sub-block s's segments are
    // segs[segbase[s] .. segbase[s + 1]), a multiple of kPairGroup.
    const uint16_t *segs;
    const int64_t *segbase;
    double *colsum;          // float64 g: column sums of X
    double vmean;            // colsum . vs / n for the current vs
    double alpha, inv_alpha, rnorm, smax;
    double umean;         // two passes: mean left in U[:, j]; one pass: mean of u
    int64_t steps;
};

// The raw X v' of Lanczos step j: two buffers of n + kWave floats (a sink
// slot for masked lanes at [n]), alternating with j, so that a one-pass
// step still reads the previous step's vector.
__device__ __forceinline__ float *u_of_step(const Prob &pr, int j) {
    return pr.u32 + (int64_t)(j & 1) * (pr.n + kWave);
}

struct Tables {
    Prob *probs;
    int num_probs;
    const int2 *blocks;   // (problem, data block) pairs, problem-major
    int64_t num_blocks;
    const int2 *groups;   // (problem, group) pairs
    int64_t num_groups;
    int max_g, max_m;
    int modes;            // bit set of the IndexModes of the problems
    int vmodes;           // X v: IndexModes of the problems without the one-pass pair
    int tmodes;           // X^T u: IndexModes of the CSR-path problems, kSortedBit, kPairBit
    int rmodes;           // reorthogonalisation: kReorthTwoBit, kReorthOneBit
    int64_t max_lds;      // dynamic LDS bytes a workgroup may use
    int spmv_group;       // segments per load group in X v   (2, 4 or 8)
    int spmvt_group;      // segments per load group in X^T u (2, 4 or 8)
    int spmv_wg;          // workgroups per CU at most for X v   (0 = no cap)
    int spmvt_wg;         // workgroups per CU at most for X^T u (0 = no cap)
    int64_t lds_per_cu;   // LDS bytes of one compute unit
    int num_cus;          // compute units (0 = unknown)
};

// Optional timing hook around each class of launches (null when captured
// into a graph): mark(user, kind, 0) before and mark(user, kind, 1) after.
struct LaunchHook {
    void (*mark)(void *user, int kind, int end);
    void *user;
};

// Fixed launch sequences (queued on one stream; capturable in a graph).
hipError_t launch_setup(const Tables &T, hipStream_t s);
// With a side stream (and its fork/join events), the small SVDs of the
// problems whose sweep ended in an earlier step run on it concurrently with
// the other problems' Lanczos launches; without one (side = nullptr) every
// launch goes to s and a sweep that ends in a step restarts in that step.
// Each problem sees the same sequence of operations either way, so its
// results are the same.
hipError_t launch_macro_step(const Tables &T, int *done_counter,
                             hipStream_t s, const LaunchHook *hook,
                             hipStream_t side = nullptr,
                             hipEvent_t fork = nullptr,
                             hipEvent_t join = nullptr);
// One launch of X v or X^T u alone (autotuner timing; problems must be in the
// Lanczos phase).  which = 0: X v, 1: X^T u; group = 2, 4 or 8; wg = at most
// that many workgroups per CU (0 = no cap).
hipError_t launch_spmv_variant(const Tables &T, int which, int group, int wg,
                               hipStream_t s);
// Largest dynamic LDS request of any kernel for g genes and subspace m.
size_t lds_bytes_needed(int g, int m);
// Canonical-CSR check of one problem's operand fused with packing it: column
// indices to 16 bits (32 bits when g > kMaxPacked16; idx) and the values into the aligned copy (data4);
// both copies must hold nnz rounded up to 4 elements, the tail zeroed by the
// caller.  Sets *flag when indptr does not start at 0, decreases or exceeds
// nnz, or a row's indices are not strictly increasing in [0, g).
// check_only: validate without writing data4 / idx.
hipError_t launch_validate_pack(const float *data, const int32_t *indices,
                                const void *indptr, int indptr_width, int64_t n,
                                int64_t nnz, int g, float *data4, void *idx,
                                int *flag, hipStream_t s, bool check_only = false);
// The packed copy of a pair problem's CSR (X v and X^T u in one pass; the
// CSR must have passed launch_validate_pack with check_only): row r
// occupies positions [poff[r], poff[r + 1]) (poff[0] = 0), its elements then
// zeros of gene g up to a multiple of 4 and, after the last row of each
// 256-row sub-block, one more quad of them (the sub-block's zero quad); idx
// holds gene << 2 (g <= 16383).  poff: n + 1 int64;
// scratch: ceil(n / 1024) + 1 int64.  data4 and idx must hold
// pad_packed_bound(nnz, n) elements.
int64_t pad_packed_bound(int64_t nnz, int64_t n);
// The pair kernel's segment table for a padded packed copy (row offsets
// poff, n rows): per 256-row sub-block, its rows cut into segments of at
// most 64 quads, one uint16 each (bits 0-6: quads, bit 7: the segment ends
// its row, bit 8: its row has three or more segments), with an empty
// segment before any two-segment row that would start a group's last slot,
// and empty segments to a multiple of kPairGroup.  segbase: nsub + 1 int64
// (nsub = ceil(n / 256)); segs: pair_segments_bound(n, poff[n]) entries.
int64_t pair_segments_bound(int64_t n, int64_t padded_total);
hipError_t launch_pair_segments(const int64_t *poff, int64_t n, int64_t *segbase,
                                uint16_t *segs, hipStream_t s);
hipError_t launch_pack_padded(const float *data, const int32_t *indices,
                              const void *indptr, int indptr_width, int64_t n,
                              int g, int64_t *poff, int64_t *scratch,
                              float *data4, uint16_t *idx, hipStream_t s);
// The gene-sorted copy of one problem's packed CSR (g <= kMaxSortedGenes,
// 16-bit indices, int64 row offsets): per 256-row sub-block the entry counts
// of each gene (tlens, row stride gl), then its entries ordered by (gene,
// row) at the sub-block's CSR positions (tval, trow) and the run-start bits
// (thb, cleared here first; thb_words 16-bit words).  tval and trow hold at
// least nnz + 16 entries.  max_lds: the workgroup LDS limit of the device.
hipError_t launch_tsort_build(const float *data4, const uint16_t *idx16, const int64_t *indptr,
                              int64_t n, int g, int gl, float *tval, uint8_t *trow,
                              uint16_t *thb, int64_t thb_words, uint16_t *tlens,
                              int64_t max_lds, hipStream_t s);
// indptr (int32) -> int64 and int64 indices -> int32 (flagging range errors).
hipError_t launch_widen_indptr(const int32_t *in, int64_t count, int64_t *out,
                               hipStream_t s);
hipError_t launch_narrow_indices(const int64_t *in, int64_t nnz, int g,
                                 int32_t *out, int *flag, hipStream_t s);
// Zero `bytes` bytes at p with a kernel (no memset node, no allocation).
hipError_t launch_zero(void *p, size_t bytes, hipStream_t s);
// One-sided Jacobi SVD of a single m x (m+1) matrix (test entry point).
hipError_t launch_svd_single(const double *B, int m, double *W, double *Jv,
                             double *Ub, double *Sb, int *sweeps,
                             int64_t max_lds, hipStream_t s);

// out (n x kout, column-major) = A (n x m) C (m x kout) through the MFMA
// restart kernel's tile code (test entry point; m <= 128, kout <= 128).
hipError_t launch_gemm_mfma_test(const double *A, int64_t n, int m,
                                 const double *C, int kout, double *out,
                                 hipStream_t s);
// FP64 rate with register operands only (calibration): which = 0 MFMA
// 16x16x4, 1 VALU fma; `blocks` workgroups of 256 threads, `iters` rounds
// of 8 independent operations per wave; out: blocks x 256 doubles.
// Returns the floating-point operations one launch performs in *flops.
hipError_t launch_fp64_rate(int which, int blocks, int iters, double *out,
                            double *flops, hipStream_t s);

// QC x HVG subset of a CSR matrix (subset.hip).  Output row r is input row
// rows[r] (rows = nullptr: all rows); input column c becomes col_map[c] or is
// dropped when col_map[c] < 0 (col_map = nullptr: identity).  The count pass
// validates the rows, row offsets and column indices it reads (*bad) and
// writes the kept entries per output row; the fill writes them in ascending
// new column order, by compaction when col_map is increasing over the kept
// columns (ordered != 0), else by a per-row sort in LDS (rows of at most
// subset_sort_max() kept entries).
hipError_t launch_subset_count(const void *indptr, int indptr_width,
                               const void *indices, int index_width,
                               int64_t num_rows, int64_t nnz, int64_t num_cols,
                               const int64_t *rows, int64_t num_out_rows,
                               const int32_t *col_map, int64_t *row_nnz,
                               int *bad, hipStream_t s);
hipError_t launch_subset_fill(const float *data, const void *indptr,
                              int indptr_width, const void *indices,
                              int index_width, const int64_t *rows,
                              int64_t num_out_rows, const int32_t *col_map,
                              int ordered, const int64_t *out_indptr,
                              float *out_data, int32_t *out_indices,
                              hipStream_t s);
int subset_sort_max();

// Exact k nearest neighbours among the rows of X (n x d, row-major FP32;
// knn.hip): squared Euclidean distances ||q||^2 + (||c||^2 - 2 q.c), self
// excluded, each row ascending (ties by index).  k <= 64; norms from
// launch_row_norms.
hipError_t launch_row_norms(const float *X, int64_t n, int d, float *norms,
                            hipStream_t s);
hipError_t launch_knn_exact(const float *X, const float *norms, int64_t n,
                            int d, int k, uint32_t *out_idx, float *out_dist,
                            hipStream_t s);
size_t knn_lds_bytes(int d, int kmax);
int knn_kmax(int k);
// Matrix-core path for num_dims <= kKnnMfmaMaxDims: Xp is scratch of
// (n + kKnnPadRows) * knn_mfma_stride(d) floats (the padded copy of X);
// norms (n floats) from launch_row_norms.
constexpr int kKnnMfmaMaxDims = 256;
constexpr int kKnnPadRows = 64;
constexpr int64_t kKnnMaxRows = (1ll << 31) - 1 - kKnnPadRows;  // 32-bit row positions
int knn_mfma_ksteps(int d);   // compiled k-steps covering ceil((d + 2) / 4): x, 1, ||x||^2
int knn_mfma_stride(int d);   // 4 ksteps + 4 floats
size_t knn_mfma_lds_bytes(int d);
hipError_t launch_knn_mfma(const float *X, const float *norms, int64_t n,
                           int d, int k, float *Xp, uint32_t *out_idx,
                           float *out_dist, hipStream_t s);

// The tile engine behind launch_knn_mfma, for the inverted-file search and
// the centroid rankings of k-means (knn.hip).  Q and C are padded copies
// (launch_knn_pack: row [x, 1, ||x||^2, 0 ...], stride knn_mfma_stride(d),
// kKnnPadRows zero rows after the last).  Query q's k best candidates go
// to row q (inverted file: ids[q]) of out_idx / out_dist, ascending by
// (distance, id); a candidate's id is c (inverted file: ids[c]).  ids is
// read by inverted-file launches only.
struct KnnTiles {
    const float *Q;        // padded queries
    const float *C;        // padded candidates
    int qnorm;             // 1: ||q||^2 + ||c||^2 - 2 q.c; 0: partial ||c||^2 - 2 q.c
    int d;                 // dimensions
    int64_t nq, nc;
    int st, ksteps, k;     // knn_mfma_stride(d), knn_mfma_ksteps(d), neighbours (<= 64)
    int self;              // queries are the candidates: skip q == c
    // Inverted file (ivf launches): workgroup b (queries 64b..64b+63) scans
    // entries [wg_off[b], wg_off[b + 1]); entry e is four words ent[4e ..]:
    // candidate rows [first, end) and a 64-bit mask (low, high word): the
    // queries whose bit (q - 64b) is set search those rows.
    const int64_t *wg_off;
    const uint32_t *ent;
    const uint32_t *ids;
    // Exhaustive launches: optional, per query; the scan of query block b
    // starts at the 64-row tile holding row start_hint[64 b] and wraps.
    const uint32_t *start_hint;
    // With start_hint, optional: 64 padded rows per start_hint value (the
    // first boot_rows of them candidates) scanned first; their k-th distance
    // per query caps its threshold for the scan proper.  boot_rows >= k.
    const float *boot;
    int boot_rows;
    // Optional, per output row (row q, or ids[q] for inverted-file
    // launches): a floor (floor_d[row * floor_ld], floor_i[row * floor_ld]);
    // only candidates after it in (distance, id) order are kept.  Searches
    // for more than 64 neighbours run in passes of 64, each pass's floor the
    // last entry of the pass before (launch_knn_tiles_k).
    const float *floor_d;
    const uint32_t *floor_i;
    int floor_ld;
    uint32_t *out_idx;
    float *out_dist;
};
// Row c of Xp = [x_src, 1, norms[src], 0 ...], src = order ? order[c] : c.
hipError_t launch_knn_pack(const float *X, const uint32_t *order, int64_t n,
                           int d, const float *norms, float *Xp, hipStream_t s);
hipError_t launch_knn_tiles(const KnnTiles &T, bool ivf, hipStream_t s);
// The same search for any k (T.k ignored; k >= 1): k <= 64 is one launch;
// more runs ceil(k / 64) launches of at most 64 neighbours each, every one
// after the first with the previous launch's last neighbour of each row as
// its floor, into pass buffers copied to columns 64 p .. of out_idx /
// out_dist (row stride k).  Synchronous when k > 64.
hipError_t launch_knn_tiles_k(const KnnTiles &T, bool ivf, int k, hipStream_t s);

// SingleCell.neighbors() with its inverted-file search (ivf.hip).
struct IvfParams {
    int64_t n;          // cells
    int d, k;           // dimensions (<= kKnnMfmaMaxDims), neighbours (<= 64)
    int num_clusters, searched, iterations;
    float tolerance;    // k-means relative inertia change
    int64_t chunk;      // k-means chunk size (the M-step's summation order)
    uint64_t seed;
    // When kmeans_labels is set: stop after k-means and copy out the labels
    // (n; from the final E-step against the final centroids), centroids
    // (num_clusters x d) and counts (num_clusters, of those labels), as
    // kmeans() leaves them.
    uint32_t *kmeans_labels = nullptr;
    float *kmeans_centroids = nullptr;
    uint32_t *kmeans_counts = nullptr;
    // k-means|| initialisation (kmeans_barbar=True) instead of random cells.
    int barbar = 0;
    int init_iterations = 5;
    float oversampling = 1.0f;
};
struct IvfInfo {
    int clusters;       // non-empty clusters
    int searched;       // clusters searched per cell
    int iterations;     // k-means iterations run
    int exhaustive;     // every cluster searched: the exhaustive search ran
    int64_t entries;    // (workgroup, cluster) pairs scanned
    int64_t candidates; // k-means||: candidates selected (0 otherwise)
};
// 0: done; 1: invalid input (*msg); 2: HIP error (*herr).  Synchronous.
int run_knn_ivf(const float *X, const IvfParams &P, uint32_t *out_idx,
                float *out_dist, IvfInfo *info, std::string *msg,
                hipError_t *herr, hipStream_t s);
std::vector<uint32_t> kmeans_random_init_cells(int64_t n, int num_clusters,
                                               uint64_t seed);
// k-means|| initialisation (brisc/kmeans.pyx:41-360) of P.num_clusters
// cells from P's n x d cells X (device), with P.init_iterations,
// P.oversampling, P.chunk and P.seed: the chosen cells go to *cells (host),
// the number of candidates selected to *candidates.  norms and Xp: X's
// squared norms (launch_row_norms) and padded copy (launch_knn_pack), or
// null to have them made here.  0: done; 1: the reference's input errors
// (*msg); 2: HIP error (*herr).  Synchronous.
int kmeans_barbar_init_cells(const float *X, const float *norms, const float *Xp,
                             const IvfParams &P, std::vector<uint32_t> *cells,
                             int64_t *candidates, std::string *msg,
                             hipError_t *herr, hipStream_t s);

// Transpose of a sparse matrix stored by rows (transpose.hip): input
// num_rows x num_cols with int32 or int64 offsets / indices (widths 4 or 8),
// output num_cols x num_rows with int64 offsets (num_cols + 1) and int32
// indices, each output row in ascending input-row order.  0: done; 1: an
// offset or index out of range; 2: HIP error (*herr).  Synchronous.
int run_csr_transpose(const float *data, const void *indices, int index_width,
                      const void *indptr, int indptr_width, int64_t num_rows,
                      int64_t num_cols, int64_t nnz, float *out_data,
                      int32_t *out_indices, int64_t *out_indptr,
                      hipError_t *herr, hipStream_t s);

// SingleCell.normalize() on the device (normalize.hip).
struct NormalizeArgs {
    const void *data;
    int data_type;        // 0 int32, 1 uint32, 2 int64, 3 uint64, 4 float32, 5 float64
    const void *indices;  // CSR: gene of each entry; CSC: cell of each entry
    int index_width;      // 4 or 8
    const void *indptr;   // CSR: num_cells + 1 offsets; CSC: num_genes + 1
    int indptr_width;     // 4 or 8
    int is_csr;
    int64_t num_cells, num_genes, nnz;
    const uint8_t *qc;    // cells passing QC (nonzero), or null for all
    int method;           // 0 logCP10k, 1 log1pPF, 2 PFlog1pPF
    float *out;           // nnz normalised values (may alias float32 data)
    uint64_t *row_sums;   // num_cells row sums of the counts
};
// 0: done; 1: a CSC row index out of range; 2: HIP error (*herr).
// Synchronous.
int run_normalize(const NormalizeArgs &A, hipError_t *herr, hipStream_t s);
// y = the host library's log1pf of x, elementwise (test of log1pf.h).
hipError_t launch_log1pf_eval(const float *x, int64_t n, float *y, hipStream_t s);

// Synthetic workload generator and log1pPF (synth.hip).
hipError_t launch_synth_row_nnz(uint64_t seed, uint64_t row0, int64_t rows,
                                int g, const double *tgm, const int32_t *type,
                                const double *factor, int64_t *row_nnz,
                                hipStream_t s);
hipError_t launch_synth_fill(uint64_t seed, uint64_t row0, int64_t rows, int g,
                             const double *tgm, const int32_t *type,
                             const double *factor, const int64_t *indptr,
                             int32_t *indices, int32_t *counts, hipStream_t s);
hipError_t launch_row_sums(const int32_t *counts, const int64_t *indptr,
                           int64_t n, uint64_t *row_sums, hipStream_t s);
hipError_t launch_sum_u64(const uint64_t *x, int64_t n, uint64_t *out,
                          hipStream_t s);
hipError_t launch_log1ppf(const int32_t *counts, const int64_t *indptr,
                          int64_t n, const uint64_t *row_sums,
                          const uint64_t *total, float *out, hipStream_t s);

}  // namespace bpca
