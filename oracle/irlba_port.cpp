// ORACLE (test infrastructure and CPU baseline only; never linked into the
// product library).
//
// C++/OpenMP restatement of the reference's *parallel* IRLBA path
// (brisc/pca.pyx:973-1254, selected when num_threads > 1 or
// match_parallel=True), in FP32 with the reference's order of operations:
//
//   * clipped_stddev_csc           pca.pyx:51-103  (sample std, ddof=1, FP32)
//   * start vector                 pca.pyx:801-806, cyutils.pyx:1153-1191,
//                                  1228-1232 (PCG-XSH-RR 64/32 + Box-Muller)
//   * matvec_csr_parallel_raw      pca.pyx:440-475
//   * fused pass 1 / pass 2 with fixed chunk_size row blocks and serial
//     block-order reductions      pca.pyx:1006-1090
//   * rmatvec_csc_parallel_raw     pca.pyx:478-511 on a row-sorted CSC copy
//                                  (cysparse.pyx:1179-1295 ordering)
//   * V update + single-pass CGS   pca.pyx:1103-1171
//   * sgesdd of B, convergence, k_current, restart   pca.pyx:1178-1231
//   * final PCs = U (Ub[:, :k] * Sb[:k])              pca.pyx:1233-1254
//
// LAPACK sgesdd and BLAS sgemm are the same routines the reference calls
// (SciPy's cython_lapack / cython_blas); the caller passes their addresses.
// The non-converged exit reproduces the reference as written: when
// max_iterations runs out, U has already been restarted, and the final
// product applies Ub again (pca.pyx:1210-1215 followed by 1239-1254), so only
// converged solves are meaningful for comparison.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

typedef void (*sgesdd_fn)(char *, int *, int *, float *, int *, float *,
                          float *, int *, float *, int *, float *, int *,
                          int *, int *);
typedef void (*sgemm_fn)(char *, char *, int *, int *, int *, float *,
                         float *, int *, float *, int *, float *, float *,
                         int *);

// PCG-XSH-RR 64/32 (O'Neill 2014) with the multiplier and increment used by
// cyutils.pyx:1153-1158 and the seeding of cyutils.pyx:1228-1232.
struct Pcg32 {
    uint64_t state;
    explicit Pcg32(uint64_t seed) : state(seed + 1442695040888963407ULL) {
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + 1442695040888963407ULL;
        uint32_t xorshifted = (uint32_t)((old ^ (old >> 18)) >> 27);
        uint32_t rot = (uint32_t)(old >> 59);
        return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
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

float norm_f32(const float *x, size_t len) {
    float s = 0.0f;
    for (size_t i = 0; i < len; i++) s += x[i] * x[i];
    return std::sqrt(s);
}

}  // namespace

extern "C" int irlba_port(const float *data, const int32_t *indices,
                          const int64_t *indptr, uint64_t num_cells,
                          uint32_t num_genes, uint32_t k,
                          uint32_t subspace_size, float tolerance,
                          uint32_t max_iterations, uint64_t seed,
                          uint32_t num_threads, uint64_t chunk_size,
                          float *PCs, void *sgesdd_ptr, void *sgemm_ptr,
                          int64_t *stats, float *singular_values) {
    const sgesdd_fn sgesdd = (sgesdd_fn)sgesdd_ptr;
    const sgemm_fn sgemm = (sgemm_fn)sgemm_ptr;
    const uint64_t n = num_cells;
    const uint32_t g = num_genes, m = subspace_size;
    const float clip = 1e-8f;
    omp_set_num_threads((int)std::max<uint32_t>(1, num_threads));
    const uint64_t nnz = (uint64_t)indptr[n];

    // Row-sorted CSC copy (column entries in ascending cell order).  The
    // reference receives it from pca() (single_cell.py:12050); its build
    // time is reported apart (stats[3], nanoseconds).
    const double t_csc0 = omp_get_wtime();
    std::vector<int64_t> cptr(g + 1, 0);
    std::vector<int32_t> crow(nnz);
    std::vector<float> cval(nnz);
    for (uint64_t e = 0; e < nnz; e++) cptr[indices[e] + 1]++;
    for (uint32_t j = 0; j < g; j++) cptr[j + 1] += cptr[j];
    {
        std::vector<int64_t> fill(cptr.begin(), cptr.end() - 1);
        for (uint64_t i = 0; i < n; i++)
            for (int64_t e = indptr[i]; e < indptr[i + 1]; e++) {
                int64_t pos = fill[indices[e]]++;
                crow[pos] = (int32_t)i;
                cval[pos] = data[e];
            }
    }

    const double t_csc = omp_get_wtime() - t_csc0;

    // Per-gene clipped sample standard deviation, FP32.
    std::vector<float> sd(g);
    const float inv_pairs = (float)(1.0 / (double)(n * (n - 1)));
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t j = 0; j < (int64_t)g; j++) {
        float s = 0.0f, ss = 0.0f;
        for (int64_t e = cptr[j]; e < cptr[j + 1]; e++) {
            float v = cval[e];
            s = s + v;
            ss = ss + v * v;
        }
        float v = std::sqrt(inv_pairs * ((float)n * ss - s * s));
        sd[j] = v < clip ? clip : v;
    }

    std::vector<float> Ubuf(n * m, 0.0f), Unew(n * m, 0.0f);
    std::vector<float> Vbuf((size_t)g * m, 0.0f), Vnew((size_t)g * m, 0.0f);
    float *U = Ubuf.data(), *U2 = Unew.data();
    float *V = Vbuf.data(), *V2 = Vnew.data();
    std::vector<float> B((size_t)m * (m + 1), 0.0f), Ub((size_t)m * m),
        Sb(m), Vbt((size_t)m * (m + 1)), rem(g), vs(g);
    const uint64_t nb = (n + chunk_size - 1) / chunk_size;
    std::vector<float> block_sums(nb), partial((size_t)nb * m);

    // sgesdd workspace query, then the working arrays.
    char jobz = 'S';
    int bm = (int)m, bn = (int)m + 1, lwork = -1, info = 0;
    std::vector<int> iwork(8 * m);
    float wsz = 0.0f;
    sgesdd(&jobz, &bm, &bn, B.data(), &bm, Sb.data(), Ub.data(), &bm,
           Vbt.data(), &bm, &wsz, &lwork, iwork.data(), &info);
    lwork = (int)wsz;
    std::vector<float> work(std::max<int>(lwork, 4 * m * m + 7 * m));

    Pcg32 rng(seed);
    for (uint32_t i = 0; i < g; i++) V[i] = rng.normal();
    {
        float inv = 1.0f / norm_f32(V, g);
        for (uint32_t i = 0; i < g; i++) V[i] *= inv;
    }

    float Smax = 2.4221817809573368e-05f;
    uint32_t kc = 0, iteration = 0;
    int64_t matvecs = 0;
    bool converged = false;
    for (iteration = 0; iteration < max_iterations; iteration++) {
        uint32_t j = kc;
        while (true) {
            // u = X (v_j / sd), raw (centering deferred to the fused pass).
            float *u = U + (size_t)j * n;
            const float *vj = V + (size_t)j * g;
            for (uint32_t i = 0; i < g; i++) vs[i] = vj[i] / sd[i];
            #pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < (int64_t)n; i++) {
                float acc = 0.0f;
                for (int64_t e = indptr[i]; e < indptr[i + 1]; e++)
                    acc = acc + data[e] * vs[indices[e]];
                u[i] = acc;
            }
            matvecs++;
            float mean = 0.0f;
            if (j > 0) {
                #pragma omp parallel for schedule(static)
                for (int64_t b = 0; b < (int64_t)nb; b++) {
                    uint64_t r0 = b * chunk_size,
                             r1 = std::min<uint64_t>(r0 + chunk_size, n);
                    float p = 0.0f;
                    for (uint64_t r = r0; r < r1; r++) p = p + u[r];
                    block_sums[b] = p;
                    for (uint32_t c = 0; c < j; c++) {
                        const float *Uc = U + (size_t)c * n;
                        float q = 0.0f;
                        for (uint64_t r = r0; r < r1; r++) q = q + Uc[r] * u[r];
                        partial[(size_t)b * m + c] = q;
                    }
                }
                for (uint64_t b = 0; b < nb; b++) mean += block_sums[b];
                mean /= (float)n;
                for (uint32_t c = 0; c < j; c++) {
                    float d = 0.0f;
                    for (uint64_t b = 0; b < nb; b++)
                        d = d + partial[(size_t)b * m + c];
                    work[c] = d;
                }
                #pragma omp parallel for schedule(static)
                for (int64_t b = 0; b < (int64_t)nb; b++) {
                    uint64_t r0 = b * chunk_size,
                             r1 = std::min<uint64_t>(r0 + chunk_size, n);
                    float p = 0.0f;
                    for (uint64_t r = r0; r < r1; r++) {
                        float d = mean;
                        for (uint32_t c = 0; c < j; c++)
                            d = d + U[(size_t)c * n + r] * work[c];
                        u[r] = u[r] - d;
                        p = p + u[r] * u[r];
                    }
                    block_sums[b] = p;
                }
            } else {
                #pragma omp parallel for schedule(static)
                for (int64_t b = 0; b < (int64_t)nb; b++) {
                    uint64_t r0 = b * chunk_size,
                             r1 = std::min<uint64_t>(r0 + chunk_size, n);
                    float p = 0.0f;
                    for (uint64_t r = r0; r < r1; r++) p = p + u[r];
                    block_sums[b] = p;
                }
                for (uint64_t b = 0; b < nb; b++) mean += block_sums[b];
                mean /= (float)n;
                #pragma omp parallel for schedule(static)
                for (int64_t b = 0; b < (int64_t)nb; b++) {
                    uint64_t r0 = b * chunk_size,
                             r1 = std::min<uint64_t>(r0 + chunk_size, n);
                    float p = 0.0f;
                    for (uint64_t r = r0; r < r1; r++) {
                        u[r] = u[r] - mean;
                        p = p + u[r] * u[r];
                    }
                    block_sums[b] = p;
                }
            }
            float sq = 0.0f;
            for (uint64_t b = 0; b < nb; b++) sq += block_sums[b];
            float alpha = std::sqrt(sq);
            if (alpha < clip) alpha = clip;
            float inv_alpha = 1.0f / alpha;
            #pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < (int64_t)n; i++) u[i] *= inv_alpha;
            B[(size_t)j + (size_t)j * m] = alpha;

            // w = (X^T u_j) / sd on the CSC copy (u_j is zero-mean).
            const bool last = (j == m - 1);
            float *w = last ? rem.data() : V + (size_t)(j + 1) * g;
            #pragma omp parallel for schedule(dynamic, 16)
            for (int64_t c = 0; c < (int64_t)g; c++) {
                float acc = 0.0f;
                for (int64_t e = cptr[c]; e < cptr[c + 1]; e++)
                    acc = acc + cval[e] * u[crow[e]];
                w[c] = acc;
            }
            for (uint32_t i = 0; i < g; i++) w[i] /= sd[i];
            matvecs++;
            for (uint32_t i = 0; i < g; i++) w[i] -= alpha * V[(size_t)j * g + i];
            const uint32_t ncols = last ? m : j + 1;
            for (uint32_t c = 0; c < ncols; c++) {
                float d = 0.0f;
                for (uint32_t r = 0; r < g; r++)
                    d = d + V[(size_t)c * g + r] * w[r];
                work[c] = d;
            }
            for (uint32_t r = 0; r < g; r++) {
                float d = 0.0f;
                for (uint32_t c = 0; c < ncols; c++)
                    d = d + V[(size_t)c * g + r] * work[c];
                w[r] -= d;
            }
            if (last) break;
            float beta = norm_f32(w, g);
            if (beta < clip) beta = clip;
            float inv_beta = 1.0f / beta;
            for (uint32_t i = 0; i < g; i++) w[i] *= inv_beta;
            B[(size_t)j + (size_t)(j + 1) * m] = beta;
            j++;
        }

        sgesdd(&jobz, &bm, &bn, B.data(), &bm, Sb.data(), Ub.data(), &bm,
               Vbt.data(), &bm, work.data(), &lwork, iwork.data(), &info);
        float rnorm = norm_f32(rem.data(), g);
        for (uint32_t i = 0; i < g; i++) rem[i] /= rnorm;
        Smax = std::max(Sb[0], Smax);
        uint32_t nconv = 0;
        for (uint32_t i = 0; i < k; i++) {
            float res = rnorm * Ub[(m - 1) + (size_t)i * m];
            nconv += std::fabs(res) < tolerance * Smax;
        }
        if (nconv == k) {
            converged = true;
            break;
        }
        kc = std::min(std::max(kc, k + nconv), m - 1);

        // U[:, :kc] = U Ub[:, :kc]
        #pragma omp parallel for schedule(static)
        for (int64_t r = 0; r < (int64_t)n; r++)
            for (uint32_t c = 0; c < kc; c++) {
                float d = 0.0f;
                for (uint32_t l = 0; l < m; l++)
                    d = d + U[(size_t)l * n + r] * Ub[l + (size_t)c * m];
                U2[(size_t)c * n + r] = d;
            }
        std::swap(U, U2);
        {
            char tn = 'N', tt = 'T';
            int gi = (int)g, kci = (int)kc, mi = (int)m;
            float one = 1.0f, zero = 0.0f;
            sgemm(&tn, &tt, &gi, &kci, &mi, &one, V, &gi, Vbt.data(), &mi,
                  &zero, V2, &gi);
        }
        std::swap(V, V2);
        for (uint32_t i = 0; i < g; i++) V[(size_t)kc * g + i] = rem[i];
        std::fill(B.begin(), B.end(), 0.0f);
        for (uint32_t i = 0; i < kc; i++) {
            B[i + (size_t)i * m] = Sb[i];
            B[i + (size_t)kc * m] = rnorm * Ub[(m - 1) + (size_t)i * m];
        }
    }

    for (uint32_t c = 0; c < k; c++) {
        float s = Sb[c];
        if (singular_values) singular_values[c] = s;
        for (uint32_t l = 0; l < m; l++) Ub[l + (size_t)c * m] *= s;
    }
    #pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < (int64_t)n; r++)
        for (uint32_t c = 0; c < k; c++) {
            float d = 0.0f;
            for (uint32_t l = 0; l < m; l++)
                d = d + U[(size_t)l * n + r] * Ub[l + (size_t)c * m];
            PCs[(size_t)r * k + c] = d;
        }
    if (stats) {
        stats[0] = (int64_t)std::min(iteration + 1, max_iterations);
        stats[1] = matvecs;
        stats[2] = converged;
        stats[3] = (int64_t)(t_csc * 1e9);
    }
    return converged ? 1 : 0;
}
