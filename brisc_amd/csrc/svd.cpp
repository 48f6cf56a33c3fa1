// Small dense FP64 SVD for the IRLBA restart (the role sgesdd plays at
// brisc/pca.pyx:514-545, 1178, here in double precision).
//
// Golub-Reinsch algorithm (Golub & Reinsch, Numer. Math. 14 (1970) 403-420):
// Householder reduction to upper bidiagonal form, accumulation of the left
// and right transformations, then implicitly shifted QR sweeps on the
// bidiagonal with the shift taken from the trailing 2x2 block, splitting
// where a superdiagonal or diagonal entry becomes negligible.
#include "svd.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace bpca {

namespace {

inline double sign_of(double magnitude, double s) {
    return s >= 0.0 ? std::fabs(magnitude) : -std::fabs(magnitude);
}

}  // namespace

// A: rows x cols column-major, rows >= cols.  On return A holds U (rows x
// cols), S the singular values in descending order, V (cols x cols,
// column-major) the right singular vectors: A_in = U diag(S) V^T.
int golub_reinsch_svd(int rows, int cols, double *A, double *S, double *V) {
    const int M = rows, N = cols;
    if (M < N || N <= 0) return -2;
    auto a = [&](int i, int j) -> double & { return A[i + (size_t)j * M]; };
    auto v = [&](int i, int j) -> double & { return V[i + (size_t)j * N]; };
    std::vector<double> e(N, 0.0);
    double g = 0.0, scale = 0.0, anorm = 0.0;
    int l = 0;

    // 1. Householder bidiagonalisation: S holds the diagonal, e the
    //    superdiagonal (e[i] couples columns i-1 and i).
    for (int i = 0; i < N; i++) {
        l = i + 1;
        e[i] = scale * g;
        g = scale = 0.0;
        double s = 0.0;
        for (int r = i; r < M; r++) scale += std::fabs(a(r, i));
        if (scale != 0.0) {
            for (int r = i; r < M; r++) {
                a(r, i) /= scale;
                s += a(r, i) * a(r, i);
            }
            double f = a(i, i);
            g = -sign_of(std::sqrt(s), f);
            double h = f * g - s;
            a(i, i) = f - g;
            for (int j = l; j < N; j++) {
                double dot = 0.0;
                for (int r = i; r < M; r++) dot += a(r, i) * a(r, j);
                double coef = dot / h;
                for (int r = i; r < M; r++) a(r, j) += coef * a(r, i);
            }
            for (int r = i; r < M; r++) a(r, i) *= scale;
        }
        S[i] = scale * g;
        g = s = scale = 0.0;
        if (i != N - 1) {
            for (int c = l; c < N; c++) scale += std::fabs(a(i, c));
            if (scale != 0.0) {
                for (int c = l; c < N; c++) {
                    a(i, c) /= scale;
                    s += a(i, c) * a(i, c);
                }
                double f = a(i, l);
                g = -sign_of(std::sqrt(s), f);
                double h = f * g - s;
                a(i, l) = f - g;
                for (int c = l; c < N; c++) e[c] = a(i, c) / h;
                for (int r = l; r < M; r++) {
                    double dot = 0.0;
                    for (int c = l; c < N; c++) dot += a(r, c) * a(i, c);
                    for (int c = l; c < N; c++) a(r, c) += dot * e[c];
                }
                for (int c = l; c < N; c++) a(i, c) *= scale;
            }
        }
        anorm = std::max(anorm, std::fabs(S[i]) + std::fabs(e[i]));
    }

    // 2. Right transformations into V.
    for (int i = N - 1; i >= 0; i--) {
        if (i < N - 1) {
            if (g != 0.0) {
                for (int j = l; j < N; j++) v(j, i) = (a(i, j) / a(i, l)) / g;
                for (int j = l; j < N; j++) {
                    double dot = 0.0;
                    for (int c = l; c < N; c++) dot += a(i, c) * v(c, j);
                    for (int c = l; c < N; c++) v(c, j) += dot * v(c, i);
                }
            }
            for (int j = l; j < N; j++) v(i, j) = v(j, i) = 0.0;
        }
        v(i, i) = 1.0;
        g = e[i];
        l = i;
    }

    // 3. Left transformations into A (becomes U).
    for (int i = N - 1; i >= 0; i--) {
        l = i + 1;
        g = S[i];
        for (int j = l; j < N; j++) a(i, j) = 0.0;
        if (g != 0.0) {
            g = 1.0 / g;
            for (int j = l; j < N; j++) {
                double dot = 0.0;
                for (int r = l; r < M; r++) dot += a(r, i) * a(r, j);
                double f = (dot / a(i, i)) * g;
                for (int r = i; r < M; r++) a(r, j) += f * a(r, i);
            }
            for (int r = i; r < M; r++) a(r, i) *= g;
        } else {
            for (int r = i; r < M; r++) a(r, i) = 0.0;
        }
        a(i, i) += 1.0;
    }

    // 4. Implicit-shift QR on the bidiagonal, from the bottom up.
    const double tiny = 4.0 * 2.220446049250313e-16 * anorm;
    for (int k = N - 1; k >= 0; k--) {
        for (int its = 0;; its++) {
            bool cancel = true;
            int nm = 0;
            for (l = k; l >= 0; l--) {
                nm = l - 1;
                if (l == 0 || std::fabs(e[l]) <= tiny) {
                    cancel = false;
                    break;
                }
                if (std::fabs(S[nm]) <= tiny) break;
            }
            if (cancel) {
                // S[nm] is negligible: chase e[l] out with rotations.
                double c = 0.0, s = 1.0;
                for (int i = l; i <= k; i++) {
                    double f = s * e[i];
                    e[i] = c * e[i];
                    if (std::fabs(f) <= tiny) break;
                    g = S[i];
                    double h = std::hypot(f, g);
                    S[i] = h;
                    h = 1.0 / h;
                    c = g * h;
                    s = -f * h;
                    for (int r = 0; r < M; r++) {
                        double y = a(r, nm), z = a(r, i);
                        a(r, nm) = y * c + z * s;
                        a(r, i) = z * c - y * s;
                    }
                }
            }
            double z = S[k];
            if (l == k) {
                if (z < 0.0) {
                    S[k] = -z;
                    for (int r = 0; r < N; r++) v(r, k) = -v(r, k);
                }
                break;
            }
            if (its >= 90) return -1;
            double x = S[l];
            nm = k - 1;
            double y = S[nm];
            g = e[nm];
            double h = e[k];
            double f = ((y - z) * (y + z) + (g - h) * (g + h)) / (2.0 * h * y);
            g = std::hypot(f, 1.0);
            f = ((x - z) * (x + z) + h * ((y / (f + sign_of(g, f))) - h)) / x;
            double c = 1.0, s = 1.0;
            for (int j = l; j <= nm; j++) {
                int i = j + 1;
                g = e[i];
                y = S[i];
                h = s * g;
                g = c * g;
                z = std::hypot(f, h);
                e[j] = z;
                c = f / z;
                s = h / z;
                f = x * c + g * s;
                g = g * c - x * s;
                h = y * s;
                y *= c;
                for (int r = 0; r < N; r++) {
                    double xv = v(r, j), zv = v(r, i);
                    v(r, j) = xv * c + zv * s;
                    v(r, i) = zv * c - xv * s;
                }
                z = std::hypot(f, h);
                S[j] = z;
                if (z != 0.0) {
                    z = 1.0 / z;
                    c = f * z;
                    s = h * z;
                }
                f = c * g + s * y;
                x = c * y - s * g;
                for (int r = 0; r < M; r++) {
                    double ya = a(r, j), za = a(r, i);
                    a(r, j) = ya * c + za * s;
                    a(r, i) = za * c - ya * s;
                }
            }
            e[l] = 0.0;
            e[k] = f;
            S[k] = x;
        }
    }

    // 5. Descending order (stable, so equal values keep their order).
    std::vector<int> order(N);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int p, int q) { return S[p] > S[q]; });
    std::vector<double> S2(N), U2((size_t)M * N), V2((size_t)N * N);
    for (int t = 0; t < N; t++) {
        int src = order[t];
        S2[t] = S[src];
        std::copy(A + (size_t)src * M, A + (size_t)(src + 1) * M,
                  U2.begin() + (size_t)t * M);
        std::copy(V + (size_t)src * N, V + (size_t)(src + 1) * N,
                  V2.begin() + (size_t)t * N);
    }
    std::copy(S2.begin(), S2.end(), S);
    std::copy(U2.begin(), U2.end(), A);
    std::copy(V2.begin(), V2.end(), V);
    return 0;
}

// SVD of the m x (m+1) IRLBA matrix B (column-major, ld m):
// B = Ub diag(Sb) Vbt with Ub m x m, Vbt m x (m+1) (both column-major).
int bidiag_svd(int m, const double *B, double *Ub, double *Sb, double *Vbt) {
    const int M = m + 1, N = m;
    std::vector<double> A((size_t)M * N), V((size_t)N * N);
    // A = B^T ((m+1) x m).
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m + 1; j++) A[j + (size_t)i * M] = B[i + (size_t)j * m];
    int rc = golub_reinsch_svd(M, N, A.data(), Sb, V.data());
    if (rc != 0) return rc;
    // B = A^T = V diag(S) U_A^T  ->  Ub = V, Vbt = U_A^T.
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) Ub[i + (size_t)j * m] = V[i + (size_t)j * N];
    for (int s = 0; s < m; s++)
        for (int j = 0; j < m + 1; j++) Vbt[s + (size_t)j * m] = A[j + (size_t)s * M];
    return 0;
}

}  // namespace bpca

extern "C" int bpca_test_bidiag_svd(int m, const double *B, double *Ub,
                                    double *Sb, double *Vbt) {
    return bpca::bidiag_svd(m, B, Ub, Sb, Vbt);
}
