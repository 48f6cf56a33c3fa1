"""ORACLE -- test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker (or the timed
CPU baseline).  The product library never links or calls it.

Contents:

* ``log1pf`` -- the host C library's log1pf, elementwise.
* ``generate_csr`` -- CPU generator of the synthetic NB count workloads
  (``gen.c``), bit-identical in its integer counts to the device generator,
  followed by the log1pPF restatement of ``brisc/normalize.pyx:9-128``.
* ``irlba_port`` -- FP32 restatement of the reference's parallel IRLBA path
  (``brisc/pca.pyx:973-1254``) calling the same SciPy LAPACK ``sgesdd`` and
  BLAS ``sgemm`` the reference calls.
* ``fp64_pca`` -- the FP64 oracle: exact Gram ``Z^T Z`` of the implicitly
  centred and scaled matrix, diagonalised with ``numpy.linalg.eigh``; gives
  singular values, right singular vectors and PCs (``Z V_k``).
* ``fp64_pca_cells`` -- the same through the cell Gram ``Z Z^T`` (for many
  genes and few cells).
* ``clipped_stddev`` -- closed form of ``clipped_stddev_csc``
  (``brisc/pca.pyx:51-103``) in FP64.
* ``kmeans_fp64`` -- Lloyd's k-means as ``kmeans()`` runs it
  (``brisc/kmeans.pyx:889-1432``: given initial cells, E-step by partial
  distance with ties to the lower index, M-step means, empty clusters moved
  to the farthest cells, a final E-step), in FP64.
* ``kmeans_barbar_init`` -- k-means|| initialisation
  (``brisc/kmeans.pyx:41-360``, ``kmeans_barbar.c``) in the operation order
  the device path defines: the cells taken as initial centroids.
* ``knn_fp64`` -- brute-force k nearest neighbours in FP64 (squared
  Euclidean, self excluded, ties by index): what ``knn_self``
  (``brisc/knn.pyx:13-516``) returns when every cluster is searched, up to the
  FP32 rounding of its distances.

Pinning (what fixes each function independently of its own formula):
golden vectors written by the reference itself (``tests/golden``), a dense
FP64 SVD on tiny inputs, SVD identities, SciPy sparse products in FP64, and
exact integer checks of the generator (row sums, SHA-256 of the arrays).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.run(['make', '-s', '-C', _HERE], check=True)


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, 'liboracle.so')
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        p = ctypes.c_void_p
        u64, u32, i64 = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int64
        L.gen_row_nnz.argtypes = [u64, u64, u64, u32, p, p, p, p]
        L.gen_fill.argtypes = [u64, u64, u64, u32, p, p, p, p, p, p]
        L.normalize_log1ppf.argtypes = [p, p, u64, p, p]
        L.log1pf_array.argtypes = [p, i64, p]
        L.gram_f64.argtypes = [p, p, p, i64, ctypes.c_int32, p, p]
        L.gram_f64.restype = ctypes.c_int
        L.irlba_port.argtypes = [p, p, p, u64, u32, u32, u32, ctypes.c_float,
                                 u32, u64, u32, u64, p, p, p, p, p]
        L.irlba_port.restype = ctypes.c_int
        L.kmeans_barbar_init.argtypes = [p, i64, ctypes.c_int32, ctypes.c_int32,
                                         ctypes.c_int32, ctypes.c_float, i64,
                                         u64, p, p]
        L.kmeans_barbar_init.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def generate_counts(w, row0: int = 0, num_rows: int | None = None):
    """Integer counts of rows [row0, row0+num_rows) of workload ``w``.

    Returns ``(indptr int64, indices int32, counts int32)``.
    """
    L = lib()
    if num_rows is None:
        num_rows = w.num_cells - row0
    ct = np.ascontiguousarray(w.cell_type[row0:row0 + num_rows])
    cf = np.ascontiguousarray(w.cell_factor[row0:row0 + num_rows])
    tgm = np.ascontiguousarray(w.type_gene_mean)
    row_nnz = np.empty(num_rows, dtype=np.int64)
    L.gen_row_nnz(w.seed, row0, num_rows, w.num_genes, _ptr(tgm), _ptr(ct),
                  _ptr(cf), _ptr(row_nnz))
    indptr = np.zeros(num_rows + 1, dtype=np.int64)
    np.cumsum(row_nnz, out=indptr[1:])
    nnz = int(indptr[-1])
    indices = np.empty(nnz, dtype=np.int32)
    counts = np.empty(nnz, dtype=np.int32)
    L.gen_fill(w.seed, row0, num_rows, w.num_genes, _ptr(tgm), _ptr(ct),
               _ptr(cf), _ptr(indptr), _ptr(indices), _ptr(counts))
    return indptr, indices, counts


def log1ppf(counts: np.ndarray, indptr: np.ndarray) -> np.ndarray:
    """log1pPF of integer CSR counts (brisc/normalize.pyx:9-128)."""
    n = len(indptr) - 1
    out = np.empty(len(counts), dtype=np.float32)
    row_sums = np.empty(n, dtype=np.uint64)
    # Converted copies are bound to names so they outlive the C call.
    counts = np.ascontiguousarray(counts, np.int32)
    indptr = np.ascontiguousarray(indptr, np.int64)
    lib().normalize_log1ppf(_ptr(counts), _ptr(indptr), n,
                            _ptr(out), _ptr(row_sums))
    return out


def log1pf(x: np.ndarray) -> np.ndarray:
    """The host C library's log1pf of float32 ``x`` (what the reference's
    normalize calls)."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    lib().log1pf_array(_ptr(x), x.size, _ptr(y))
    return y


def generate_csr(w):
    """Normalised CSR (data float32, indices int32, indptr int64) of ``w``."""
    indptr, indices, counts = generate_counts(w)
    return log1ppf(counts, indptr), indices, indptr


def _capsule_pointer(capsule) -> int:
    get_name = ctypes.pythonapi.PyCapsule_GetName
    get_name.restype = ctypes.c_char_p
    get_name.argtypes = [ctypes.py_object]
    get_ptr = ctypes.pythonapi.PyCapsule_GetPointer
    get_ptr.restype = ctypes.c_void_p
    get_ptr.argtypes = [ctypes.py_object, ctypes.c_char_p]
    return get_ptr(capsule, get_name(capsule))


def scipy_blas_lapack():
    """Addresses of SciPy's Fortran sgesdd and sgemm (as the reference uses)."""
    from scipy.linalg import cython_blas, cython_lapack
    return (_capsule_pointer(cython_lapack.__pyx_capi__['sgesdd']),
            _capsule_pointer(cython_blas.__pyx_capi__['sgemm']))


def irlba_port(data, indices, indptr, num_cells, num_genes, k,
               subspace_size=100, tolerance=1e-6, max_iterations=100, seed=0,
               num_threads=None, chunk_size=1024):
    """Run the FP32 restatement of the reference parallel IRLBA.

    Returns ``(PCs, converged, stats)`` with ``stats`` = dict(restarts,
    matvecs, singular_values, csc_seconds): the last is the time spent
    building the CSC copy that the reference receives ready-made from
    ``pca()``.
    """
    if num_threads is None:
        num_threads = os.cpu_count() or 1
    subspace_size = min(subspace_size, num_cells, num_genes)
    data = np.ascontiguousarray(data, np.float32)
    indices = np.ascontiguousarray(indices, np.int32)
    indptr = np.ascontiguousarray(indptr, np.int64)
    PCs = np.empty((num_cells, k), dtype=np.float32)
    stats = np.zeros(4, dtype=np.int64)
    sv = np.zeros(k, dtype=np.float32)
    sgesdd, sgemm = scipy_blas_lapack()
    conv = lib().irlba_port(_ptr(data), _ptr(indices), _ptr(indptr),
                            num_cells, num_genes, k, subspace_size,
                            tolerance, max_iterations, seed, num_threads,
                            chunk_size, _ptr(PCs), sgesdd, sgemm, _ptr(stats),
                            _ptr(sv))
    return PCs, bool(conv), dict(restarts=int(stats[0]),
                                 matvecs=int(stats[1]), singular_values=sv,
                                 csc_seconds=stats[3] * 1e-9)


def clipped_stddev(data, indices, indptr, num_cells, num_genes,
                   clip=1e-8):
    """FP64 per-gene sample std (ddof=1) clipped below at ``clip``."""
    s = np.bincount(indices, weights=data.astype(np.float64),
                    minlength=num_genes)
    ss = np.bincount(indices, weights=data.astype(np.float64) ** 2,
                     minlength=num_genes)
    var = (ss - s * s / num_cells) / (num_cells - 1)
    return np.maximum(np.sqrt(np.maximum(var, 0.0)), clip)


def gram(data, indices, indptr, num_genes):
    """FP64 (X^T X, X^T 1) of a CSR float32 matrix."""
    n = len(indptr) - 1
    G = np.empty((num_genes, num_genes), dtype=np.float64)
    cs = np.empty(num_genes, dtype=np.float64)
    # Converted copies are bound to names so they outlive the C call.
    data = np.ascontiguousarray(data, np.float32)
    indices = np.ascontiguousarray(indices, np.int32)
    indptr = np.ascontiguousarray(indptr, np.int64)
    rc = lib().gram_f64(_ptr(data), _ptr(indices), _ptr(indptr), n,
                        num_genes, _ptr(G), _ptr(cs))
    if rc != 0:
        raise MemoryError('gram_f64 allocation failed')
    return G, cs


def fp64_pca(data, indices, indptr, num_cells, num_genes, k,
             want_pcs=True):
    """FP64 oracle of the top-k PCA of Z = (X - 1 mu^T) diag(1/sd).

    Returns dict(singular_values (k,), right (g,k), PCs (n,k) or None, sd).
    Singular values are sqrt of the top eigenvalues of the exact FP64 Gram
    Z^T Z = D^-1 (X^T X - n mu mu^T) D^-1.  PCs = Z V_k (= U_k S_k).
    """
    G, cs = gram(data, indices, indptr, num_genes)
    sd = clipped_stddev(data, indices, indptr, num_cells, num_genes)
    mu = cs / num_cells
    C = (G - num_cells * np.outer(mu, mu)) / np.outer(sd, sd)
    evals, evecs = np.linalg.eigh(C)
    order = np.argsort(evals)[::-1][:k]
    lam = np.maximum(evals[order], 0.0)
    Vk = evecs[:, order]
    out = dict(singular_values=np.sqrt(lam), right=Vk, sd=sd, PCs=None)
    if want_pcs:
        import scipy.sparse as sp
        X = sp.csr_array((data.astype(np.float64), indices, indptr),
                         shape=(num_cells, num_genes))
        W = Vk / sd[:, None]
        out['PCs'] = X @ W - (mu @ W)[None, :]
    return out


def fp64_pca_cells(data, indices, indptr, num_cells, num_genes, k):
    """FP64 oracle through the cell-by-cell Gram, for many genes and few cells
    (the gene Gram of fp64_pca would be num_genes^2): Z Z^T (n x n) of
    Z = (X - 1 mu^T) D^-1 from the sparse product Y Y^T, Y = X D^-1, with the
    centring applied in closed form, diagonalised with numpy.linalg.eigh.
    Returns dict(singular_values (k,), PCs (n, k) = U_k S_k)."""
    import scipy.sparse as sp
    sd = clipped_stddev(data, indices, indptr, num_cells, num_genes)
    X = sp.csr_array((np.asarray(data, np.float64), np.asarray(indices, np.int64),
                      np.asarray(indptr, np.int64)), shape=(num_cells, num_genes))
    Y = X @ sp.diags_array(1.0 / sd)
    mu = np.asarray(X.sum(axis=0)).ravel() / num_cells
    w = mu / sd                                   # D^-1 mu
    G = (Y @ Y.T).toarray()
    a = Y @ w                                     # Y D^-1 mu
    G = G - a[:, None] - a[None, :] + float(w @ w)
    evals, evecs = np.linalg.eigh(G)
    order = np.argsort(evals)[::-1][:k]
    lam = np.maximum(evals[order], 0.0)
    return dict(singular_values=np.sqrt(lam), PCs=evecs[:, order] * np.sqrt(lam))


def dense_z(data, indices, indptr, num_cells, num_genes):
    """Dense FP64 Z = (X - 1 mu^T) diag(1/sd) for tiny brute-force checks."""
    X = np.zeros((num_cells, num_genes), dtype=np.float64)
    for i in range(num_cells):
        for e in range(indptr[i], indptr[i + 1]):
            X[i, indices[e]] += data[e]
    sd = clipped_stddev(data, indices, indptr, num_cells, num_genes)
    return (X - X.mean(0)) / sd


def compare_up_to_sign(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """Per-column relative error of A vs B after aligning each column's sign."""
    A = np.asarray(A, np.float64)
    B = np.asarray(B, np.float64)
    sign = np.sign(np.sum(A * B, axis=0))
    sign[sign == 0] = 1
    diff = np.linalg.norm(A * sign - B, axis=0)
    return diff / np.maximum(np.linalg.norm(B, axis=0), 1e-300)


def knn_fp64(X: np.ndarray, k: int, block: int = 2048, rows=None):
    """Exact k nearest neighbours of every row of ``X`` (or of the rows
    ``rows``) among the rows of ``X``, in FP64: ``(indices int64 x k,
    squared distances float64)``,
    each row ascending by distance, ties by index, a row never its own
    neighbour (``brisc/knn.pyx:240-252`` skips the query cell itself).

    Distances are ``sum_j (x_qj - x_cj)^2`` accumulated directly (no norm
    identity), so they carry no cancellation error; a stable sort on
    distance keeps equal distances in index order."""
    X = np.asarray(X, np.float64)
    n = X.shape[0]
    if not 1 <= k < n:
        raise ValueError('k must be in [1, n)')
    norms = np.einsum('ij,ij->i', X, X)
    qrows = np.arange(n) if rows is None else np.asarray(rows, np.int64)
    nq = len(qrows)
    idx = np.empty((nq, k), np.int64)
    dist = np.empty((nq, k), np.float64)
    for q0 in range(0, nq, block):
        q1 = min(nq, q0 + block)
        qr = qrows[q0:q1]
        Q = X[qr]
        # Candidate ranking from the norm identity, widened to 2k + 8 so
        # that the exact re-evaluation below decides the order.
        D = norms[qr, None] + norms[None, :] - 2.0 * (Q @ X.T)
        D[np.arange(q1 - q0), qr] = np.inf
        w = min(n - 1, 2 * k + 8)
        cand = np.argpartition(D, w - 1, axis=1)[:, :w]
        exact = ((Q[:, None, :] - X[cand]) ** 2).sum(axis=2)
        exact[cand == qr[:, None]] = np.inf
        order = np.lexsort((cand, exact), axis=1)[:, :k]
        r = np.arange(q1 - q0)[:, None]
        idx[q0:q1] = cand[r, order]
        dist[q0:q1] = exact[r, order]
    return idx, dist


def kmeans_fp64(X: np.ndarray, init_cells, iterations: int):
    """Labels, centroids and counts of ``kmeans()`` with the given initial
    cells (``brisc/kmeans.pyx:889-1432``), in FP64: ``iterations`` E/M
    steps, then a final E-step against the final centroids."""
    X = np.asarray(X, np.float64)
    cent = X[np.asarray(init_cells, np.int64)].copy()
    C = len(cent)

    def e_step(cent):
        cn = (cent * cent).sum(1)
        labels = np.empty(len(X), np.int64)
        for b in range(0, len(X), 8192):
            pd = cn[None, :] - 2.0 * (X[b:b + 8192] @ cent.T)
            labels[b:b + 8192] = pd.argmin(1)   # first minimum: lower index
        return labels

    for _ in range(iterations):
        labels = e_step(cent)
        counts = np.bincount(labels, minlength=C)
        sums = np.zeros_like(cent)
        np.add.at(sums, labels, X)
        empty = np.flatnonzero(counts == 0)
        if empty.size:
            far = ((X - cent[labels]) ** 2).sum(1)
            order = np.lexsort((np.arange(len(X)), -far))[:empty.size]
            for j, cell in zip(empty, order):
                sums[labels[cell]] -= X[cell]
                sums[j] = X[cell]
                counts[labels[cell]] -= 1
                counts[j] += 1
        nz = counts > 0
        sums[nz] /= counts[nz, None]
        cent = sums
    labels = e_step(cent)
    return labels, cent, np.bincount(labels, minlength=C)


def kmeans_barbar_init(X: np.ndarray, num_clusters: int,
                       num_init_iterations: int = 5,
                       oversampling_factor: float = 1.0,
                       chunk_size: int | None = None, seed: int = 0):
    """The cells k-means|| initialisation takes as initial centroids
    (``brisc/kmeans.pyx:41-360``), and the number of candidates it selected
    before the weighted k-means++ pass.  Raises ValueError where the
    reference does: every cell at one point, or fewer distinct cells than
    clusters."""
    X = np.ascontiguousarray(X, np.float32)
    n, d = X.shape
    chunk = min(4096, n) if chunk_size is None else chunk_size
    cells = np.empty(num_clusters, np.uint32)
    S = ctypes.c_int64(0)
    rc = lib().kmeans_barbar_init(_ptr(X), n, d, num_clusters,
                                  num_init_iterations,
                                  float(oversampling_factor), chunk,
                                  seed & (2 ** 64 - 1), _ptr(cells),
                                  ctypes.addressof(S))
    if rc == 1:
        raise ValueError('all cells have the same principal component '
                         'loadings')
    if rc == 2:
        raise ValueError(f'num_clusters ({num_clusters:,}) is greater than '
                         f'the number of cells with distinct principal '
                         f'component loadings; decrease num_clusters')
    if rc != 0:
        raise MemoryError('kmeans_barbar_init: out of memory')
    return cells, int(S.value)
