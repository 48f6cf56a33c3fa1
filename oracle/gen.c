/*
 * ORACLE (test infrastructure only; never linked into the product library).
 *
 * CPU generator for the synthetic negative-binomial count workloads and the
 * log1pPF normalisation that turns them into the CSR float32 input of the
 * sparse-PCA path.  The device generator in brisc_amd/csrc/synth.hip must
 * produce the same integer counts bit for bit; the tests compare the two.
 *
 * Per entry (cell i, gene j):
 *   m = type_gene_mean[type[i], j] * cell_factor[i]              (double)
 *   h = SplitMix64 finaliser(seed_key + (i*g + j) * golden-ratio constant)
 *   u = (h >> 11) * 2^-53                                          in [0,1)
 *   count = smallest k with u < CDF_NB(k; mean m, size 5), by inversion with
 *           P(0) = (5/(5+m))^5,  P(k+1) = P(k) * p * (k+5) / (k+1), p = m/(5+m)
 * Compile with -ffp-contract=off so that no multiply-add is fused: the
 * device code is compiled the same way and evaluates the same expressions.
 *
 * log1pPF follows brisc/normalize.pyx:9-128 (method 'log1pPF', no QC column):
 *   row_sums (uint64), factor = (float)total_sum / num_cells,
 *   inverse_size_factor = factor / row_sum (float), x' = log1pf(x * isf).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define NB_SIZE 5.0
#define NB_MAX_K 1000000

static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t gen_seed_key(uint64_t seed) { return mix64(seed + 0x632BE59BD9B4E019ULL); }

static inline int nb_draw(uint64_t key, uint64_t cell, uint32_t gene,
                          uint32_t num_genes, double m) {
    uint64_t h = mix64(key + (cell * (uint64_t)num_genes + gene) *
                                 0x9E3779B97F4A7C15ULL);
    double u = (double)(h >> 11) * 0x1.0p-53;
    double q = NB_SIZE / (NB_SIZE + m);
    double q2 = q * q;
    double p0 = q2 * q2 * q;
    if (u < p0) return 0;
    double p = m / (NB_SIZE + m);
    double pk = p0, cdf = p0;
    int k = 0;
    while (k < NB_MAX_K) {
        pk = pk * p * ((double)k + NB_SIZE) / ((double)k + 1.0);
        k = k + 1;
        cdf = cdf + pk;
        if (u < cdf) break;
    }
    return k;
}

/* Number of nonzeros of rows [row0, row0 + num_rows). */
void gen_row_nnz(uint64_t seed, uint64_t row0, uint64_t num_rows,
                 uint32_t num_genes, const double *type_gene_mean,
                 const int32_t *cell_type, const double *cell_factor,
                 int64_t *row_nnz) {
    uint64_t key = gen_seed_key(seed);
    #pragma omp parallel for schedule(dynamic, 256)
    for (int64_t r = 0; r < (int64_t)num_rows; r++) {
        uint64_t cell = row0 + (uint64_t)r;
        const double *base = type_gene_mean + (size_t)cell_type[r] * num_genes;
        double s = cell_factor[r];
        int64_t count = 0;
        for (uint32_t j = 0; j < num_genes; j++) {
            double m = base[j] * s;
            count += nb_draw(key, cell, j, num_genes, m) != 0;
        }
        row_nnz[r] = count;
    }
}

/* Fill column indices and integer counts; indptr[r] is the offset of row r. */
void gen_fill(uint64_t seed, uint64_t row0, uint64_t num_rows,
              uint32_t num_genes, const double *type_gene_mean,
              const int32_t *cell_type, const double *cell_factor,
              const int64_t *indptr, int32_t *indices, int32_t *counts) {
    uint64_t key = gen_seed_key(seed);
    #pragma omp parallel for schedule(dynamic, 256)
    for (int64_t r = 0; r < (int64_t)num_rows; r++) {
        uint64_t cell = row0 + (uint64_t)r;
        const double *base = type_gene_mean + (size_t)cell_type[r] * num_genes;
        double s = cell_factor[r];
        int64_t pos = indptr[r];
        for (uint32_t j = 0; j < num_genes; j++) {
            double m = base[j] * s;
            int x = nb_draw(key, cell, j, num_genes, m);
            if (x != 0) {
                indices[pos] = (int32_t)j;
                counts[pos] = x;
                pos++;
            }
        }
    }
}

/* log1pPF of integer counts into float32 (normalize.pyx:9-128 semantics). */
void normalize_log1ppf(const int32_t *counts, const int64_t *indptr,
                       uint64_t num_cells, float *out, uint64_t *row_sums) {
    unsigned long long total = 0;
    for (uint64_t i = 0; i < num_cells; i++) {
        unsigned long long s = 0;
        for (int64_t e = indptr[i]; e < indptr[i + 1]; e++)
            s += (unsigned long long)counts[e];
        row_sums[i] = s;
        total += s;
    }
    float factor = (float)total / (float)num_cells;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)num_cells; i++) {
        volatile float isf = factor / (float)row_sums[i];
        float f = isf;
        for (int64_t e = indptr[i]; e < indptr[i + 1]; e++)
            out[e] = log1pf((float)counts[e] * f);
    }
}

/* The host C library's log1pf, elementwise (the function the reference's
   normalize calls, normalize.pyx:70). */
void log1pf_array(const float *x, int64_t n, float *y) {
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) y[i] = log1pf(x[i]);
}
