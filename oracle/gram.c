/*
 * ORACLE (test infrastructure only; never linked into the product library).
 *
 * Exact FP64 Gram matrix of a CSR float32 matrix: G = X^T X (g x g, full
 * symmetric, row-major) and the column sums s = X^T 1, both accumulated in
 * double.  The FP64 oracle of the PCA path builds the implicitly centred and
 * scaled Gram  Z^T Z = D^-1 (G - n mu mu^T) D^-1  from these (mu = s / n,
 * D = diag of the per-gene sample standard deviation) and diagonalises it,
 * which fixes the singular values and right singular vectors of
 * Z = (X - 1 mu^T) D^-1 that IRLBA approximates.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

int gram_f64(const float *data, const int32_t *indices, const int64_t *indptr,
             int64_t num_rows, int32_t num_genes, double *gram, double *colsum) {
    int nt = omp_get_max_threads();
    size_t gg = (size_t)num_genes * num_genes;
    double *acc = (double *)calloc((size_t)nt * gg, sizeof(double));
    double *cs = (double *)calloc((size_t)nt * num_genes, sizeof(double));
    if (!acc || !cs) { free(acc); free(cs); return -1; }
    #pragma omp parallel
    {
        int t = omp_get_thread_num();
        double *G = acc + (size_t)t * gg;
        double *c = cs + (size_t)t * num_genes;
        #pragma omp for schedule(dynamic, 512)
        for (int64_t r = 0; r < num_rows; r++) {
            int64_t b = indptr[r], e = indptr[r + 1];
            for (int64_t p = b; p < e; p++) {
                double xp = data[p];
                int32_t a = indices[p];
                c[a] += xp;
                double *row = G + (size_t)a * num_genes;
                for (int64_t q = p; q < e; q++)
                    row[indices[q]] += xp * (double)data[q];
            }
        }
    }
    memset(gram, 0, gg * sizeof(double));
    memset(colsum, 0, (size_t)num_genes * sizeof(double));
    for (int t = 0; t < nt; t++) {
        const double *G = acc + (size_t)t * gg;
        for (size_t i = 0; i < gg; i++) gram[i] += G[i];
        for (int32_t j = 0; j < num_genes; j++) colsum[j] += cs[(size_t)t * num_genes + j];
    }
    /* Rows hold sorted column indices, so only entries (a, b) with a <= b
       were accumulated; mirror them.  Unsorted rows put some entries below
       the diagonal; folding both triangles handles either case. */
    for (int32_t a = 0; a < num_genes; a++)
        for (int32_t b = a + 1; b < num_genes; b++) {
            double v = gram[(size_t)a * num_genes + b] + gram[(size_t)b * num_genes + a];
            gram[(size_t)a * num_genes + b] = v;
            gram[(size_t)b * num_genes + a] = v;
        }
    free(acc);
    free(cs);
    return 0;
}
