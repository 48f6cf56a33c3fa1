/* ORACLE -- test infrastructure only; the product library never links it.
 *
 * k-means|| initialisation (kmeans_barbar_init, brisc/kmeans.pyx:41-360;
 * the parallel version kmeans_barbar_init_parallel, 362-720, selects the
 * same cells in the same order and sums the same integer weights), in the
 * operation order the device path defines:
 *
 *   - distances by explicit differences, dimensions in order, the first
 *     square a product and each later one fused (fmaf), as the reference's
 *     `distance += difference * difference` contracts;
 *   - squared norms the same way;
 *   - the partial distance ||c||^2 - 2 x.c of the candidate updates
 *     (kmeans.pyx:166-224) as the fused sum of (-2 x_j) c_j, dimensions in
 *     order, plus ||c||^2 (the device forms it on the matrix cores, whose
 *     internal order differs: bit-equal on inputs whose products and sums
 *     are exact, such as small integers);
 *   - every cost of the selection rounds (the reference's first one is one
 *     serial sum, later ones per chunk) as chunk sums, cells in order, then
 *     the chunk sums in order;
 *   - the k-means++ cost as per-256-candidate halving trees, then the tree
 *     sums in order (the reference sums serially);
 *   - the k-means++ draw exactly as the reference: probability -=
 *     weight_distance * inverse_cost, fused, candidates in order, the last
 *     candidate if it never goes negative.
 *
 * Random draws: PCG-XSH-RR with brisc's srand (cyutils.pyx:1153-1158,
 * 1228-1232), Lemire's bounded draw (1161-1172) and random_uniform
 * (1175-1177).  Compiled with -ffp-contract=off: every fused operation is an
 * explicit fmaf. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static uint32_t pcg_next(uint64_t *state) {
    const uint64_t x = *state;
    *state = x * 6364136223846793005ULL + 1442695040888963407ULL;
    const uint32_t s = (uint32_t)((x ^ (x >> 18)) >> 27);
    const uint32_t rot = (uint32_t)(x >> 59);
    return (s >> rot) | (s << ((-rot) & 31u));
}

static uint64_t pcg_srand(uint64_t seed) {
    uint64_t state = seed + 1442695040888963407ULL;
    pcg_next(&state);
    return state;
}

static uint32_t pcg_bounded(uint32_t bound, uint64_t *state) {
    uint64_t m = (uint64_t)pcg_next(state) * bound;
    uint32_t l = (uint32_t)m;
    if (l < bound) {
        const uint32_t threshold = (0u - bound) % bound;
        while (l < threshold) {
            m = (uint64_t)pcg_next(state) * bound;
            l = (uint32_t)m;
        }
    }
    return (uint32_t)(m >> 32);
}

static float pcg_uniform(uint64_t *state) {
    return (float)pcg_next(state) / (float)UINT32_MAX;
}

static float sq_dist(const float *a, const float *b, int d) {
    float diff = a[0] - b[0];
    float s = diff * diff;
    for (int j = 1; j < d; j++) {
        diff = a[j] - b[j];
        s = fmaf(diff, diff, s);
    }
    return s;
}

static float sq_norm(const float *a, int d) {
    float s = a[0] * a[0];
    for (int j = 1; j < d; j++) s = fmaf(a[j], a[j], s);
    return s;
}

static float chunk_cost(const float *md, int64_t n, int64_t chunk) {
    float cost = 0.0f;
    for (int64_t c0 = 0; c0 < n; c0 += chunk) {
        const int64_t c1 = c0 + chunk < n ? c0 + chunk : n;
        float s = 0.0f;
        for (int64_t i = c0; i < c1; i++) s += md[i];
        cost += s;
    }
    return cost;
}

static float tree_cost(const float *m, int64_t S) {
    float a[256];
    float cost = 0.0f;
    for (int64_t b0 = 0; b0 < S; b0 += 256) {
        for (int t = 0; t < 256; t++) a[t] = b0 + t < S ? m[b0 + t] : 0.0f;
        for (int h = 128; h > 0; h >>= 1)
            for (int t = 0; t < h; t++) a[t] += a[t + h];
        cost += a[0];
    }
    return cost;
}

/* Returns 0 (cells[0..C) written, *num_candidates = candidates selected),
 * 1 (every cell at the same point), 2 (fewer distinct cells than clusters),
 * 3 (out of memory). */
int kmeans_barbar_init(const float *X, int64_t n, int32_t d, int32_t C,
                       int32_t num_init_iterations, float oversampling,
                       int64_t chunk, uint64_t seed, uint32_t *cells,
                       int64_t *num_candidates) {
    const float l = oversampling * (float)(uint32_t)C;
    float *md = malloc(sizeof(float) * n), *xn = malloc(sizeof(float) * n);
    uint32_t *best = calloc(n, sizeof(uint32_t)), *sel = malloc(sizeof(uint32_t) * n);
    float *cn = malloc(sizeof(float) * n);
    if (!md || !xn || !best || !sel || !cn) return 3;
    int64_t S = 0;
    uint64_t state = pcg_srand(seed - 1);
    const uint32_t rc = pcg_bounded((uint32_t)n, &state);
    sel[S++] = rc;
    for (int64_t i = 0; i < n; i++) {
        md[i] = sq_dist(X + i * d, X + (int64_t)rc * d, d);
        xn[i] = sq_norm(X + i * d, d);
    }
    float cost = chunk_cost(md, n, chunk);
    int rcode = 0;
    if (cost == 0.0f) {
        rcode = 1;
        goto done;
    }
    float loc = l / cost;
    for (int64_t i = 0; i < n; i++) {
        uint64_t st = pcg_srand(seed + (uint64_t)(uint32_t)i);
        if (loc * md[i] >= pcg_uniform(&st)) {
            sel[S++] = (uint32_t)i;
            md[i] = 0.0f;
        }
    }
    int64_t prev = 1, nnew = S - 1;
    for (int64_t p = prev; p < S; p++) best[sel[p]] = (uint32_t)p;
    uint32_t iteration = 1;
    for (;;) {
        if (nnew > 0) {
            for (int64_t j = 0; j < nnew; j++) cn[j] = xn[sel[prev + j]];
            for (int64_t i = 0; i < n; i++) {
                const float *x = X + i * d;
                float bd = INFINITY;
                int64_t bj = -1;
                for (int64_t j = 0; j < nnew; j++) {
                    const float *c = X + (int64_t)sel[prev + j] * d;
                    float acc = (-2.0f * x[0]) * c[0];
                    for (int t = 1; t < d; t++) acc = fmaf(-2.0f * x[t], c[t], acc);
                    const float pd = acc + cn[j];
                    if (pd < bd) {
                        bd = pd;
                        bj = j;
                    }
                }
                float m = md[i] - xn[i];
                uint32_t b = best[i];
                if (bj >= 0 && bd < m) {
                    m = bd;
                    b = (uint32_t)(prev + bj);
                }
                md[i] = m + xn[i];
                best[i] = b;
            }
            cost = chunk_cost(md, n, chunk);
        }
        loc = l / cost;
        const int64_t before = S;
        for (int64_t i = 0; i < n; i++) {
            uint64_t st = pcg_srand(seed + (uint64_t)((uint32_t)i * iteration));
            if (loc * md[i] >= pcg_uniform(&st)) {
                sel[S++] = (uint32_t)i;
                md[i] = 0.0f;
            }
        }
        prev = before;
        nnew = S - prev;
        for (int64_t p = prev; p < S; p++) best[sel[p]] = (uint32_t)p;
        if (nnew == 0 && cost == 0.0f) {
            rcode = 2;
            goto done;
        }
        iteration++;
        if ((iteration == (uint32_t)num_init_iterations && S >= C) || S == n) break;
    }
    {
        uint32_t *w = calloc(S, sizeof(uint32_t));
        float *m2 = malloc(sizeof(float) * S);
        if (!w || !m2) {
            free(w);
            free(m2);
            rcode = 3;
            goto done;
        }
        for (int64_t i = 0; i < n; i++) w[best[i]]++;
        state = pcg_srand(seed - 2);
        int64_t pick = pcg_bounded((uint32_t)S, &state);
        cells[0] = sel[pick];
        if (C > 1) {
            for (int64_t i = 0; i < S; i++)
                m2[i] = sq_dist(X + (int64_t)sel[i] * d, X + (int64_t)sel[pick] * d, d) * (float)w[i];
            cost = tree_cost(m2, S);
            for (int32_t c = 1;; c++) {
                const float inv = 1.0f / cost;
                float p = pcg_uniform(&state);
                pick = S - 1;
                for (int64_t i = 0; i < S; i++) {
                    p = fmaf(-m2[i], inv, p);
                    if (p < 0.0f) {
                        pick = i;
                        break;
                    }
                }
                m2[pick] = 0.0f;
                cells[c] = sel[pick];
                if (c == C - 1) break;
                const float *cen = X + (int64_t)sel[pick] * d;
                for (int64_t i = 0; i < S; i++) {
                    const float dd = sq_dist(X + (int64_t)sel[i] * d, cen, d) * (float)w[i];
                    if (dd < m2[i]) m2[i] = dd;
                }
                cost = tree_cost(m2, S);
            }
        }
        free(w);
        free(m2);
    }
done:
    *num_candidates = S;
    free(md);
    free(xn);
    free(best);
    free(sel);
    free(cn);
    return rcode;
}
