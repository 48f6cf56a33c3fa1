"""k nearest neighbours of every cell among all cells, on device-resident PCs.

``SingleCell.neighbors()`` (``brisc/single_cell.py:12139-12665``) finds each
QC'd cell's ``num_neighbors`` nearest cells in PC space with ``knn_self``
(``brisc/knn.pyx:13-516``).  :func:`knn_exact` is the exhaustive form of that
search (every cluster searched), through ``bpca_knn_exact``
(``include/bpca.h``): squared Euclidean distances formed in FP32 as the
reference forms them, ``||q||^2 + (||c||^2 - 2 q.c)``
(``brisc/cyutils.pyx:1121-1150``, ``brisc/knn.pyx:222-252``), a cell never
its own neighbour, each row ascending by distance (ties by index).
"""
from __future__ import annotations

from . import _lib

# Neighbours per cell found in one pass of the tile engine; more run in
# passes of this many (matrix-core path, num_dims <= IVF_MAX_DIMS), each
# keeping only candidates after the previous pass's last.
MAX_NEIGHBORS_PER_PASS = 64
MAX_DIMS = 256


def knn_exact(X, num_neighbors: int, context=None):
    """Exact ``num_neighbors`` nearest neighbours of each row of ``X``.

    ``X``: float32 device tensor, ``num_cells x num_dims`` (``num_dims <=
    256``).  Returns ``(neighbors, distances)`` device tensors of shape
    ``num_cells x num_neighbors``: int32 row indices (the uint32 values of the
    reference; ``num_cells < 2**31`` here) and float32 squared distances.
    """
    import torch

    from .irlba import default_context
    if not isinstance(X, torch.Tensor) or X.device.type != 'cuda':
        raise TypeError('X must be a device tensor')
    if X.dtype != torch.float32 or X.dim() != 2:
        raise ValueError('X must be a 2-D float32 tensor')
    n, d = X.shape
    k = int(num_neighbors)
    if n < 2:
        raise ValueError('there is only one cell, so nearest neighbors cannot '
                         'be calculated')
    if not 1 <= k <= n - 1:
        raise ValueError(f'num_neighbors is {k}, but must be between 1 and '
                         f'the number of cells - 1')
    if not 1 <= d <= MAX_DIMS:
        raise ValueError(f'X has {d} columns; at most {MAX_DIMS} are supported')
    if d > IVF_MAX_DIMS and k > MAX_NEIGHBORS_PER_PASS:
        raise ValueError(f'more than {MAX_NEIGHBORS_PER_PASS} neighbours need '
                         f'at most {IVF_MAX_DIMS} columns')
    if n >= 2 ** 31:
        raise ValueError('at most 2**31 - 1 cells are supported')
    X = X.contiguous()
    ctx = context or default_context(X.device.index or 0)
    neighbors = torch.empty((n, k), dtype=torch.int32, device=X.device)
    distances = torch.empty((n, k), dtype=torch.float32, device=X.device)
    _lib.check(_lib.lib().bpca_knn_exact(ctx.handle, X.data_ptr(), n, d, k,
                                         neighbors.data_ptr(),
                                         distances.data_ptr()))
    return neighbors, distances


IVF_MAX_DIMS = 256


def default_num_clusters(num_cells: int) -> int:
    """``ceil(min(4 sqrt(n), n / 100))``, at least 1
    (``brisc/single_cell.py:12461-12465``)."""
    import math
    return max(1, int(math.ceil(min(4 * math.sqrt(num_cells),
                                    num_cells / 100))))


def _barbar_args(kmeans_barbar, num_init_iterations, oversampling_factor):
    if not isinstance(kmeans_barbar, bool):
        raise TypeError('kmeans_barbar must be Boolean')
    if int(num_init_iterations) < 1:
        raise ValueError('num_init_iterations must be a positive integer')
    if not float(oversampling_factor) > 0:
        raise ValueError('oversampling_factor must be a positive number')
    return int(kmeans_barbar), int(num_init_iterations), \
        float(oversampling_factor)


def knn_ivf(X, num_neighbors: int, num_clusters: int | None = None,
            num_clusters_searched: int | None = None,
            num_kmeans_iterations: int = 2, kmeans_tolerance: float = 1e-2,
            chunk_size_kmeans: int | None = None, seed: int = 0,
            context=None, kmeans_barbar: bool = False,
            num_init_iterations: int = 5, oversampling_factor: float = 1.0):
    """``SingleCell.neighbors()``'s default search on the device
    (``bpca_knn_ivf_ex``): k-means (random initial cells, or k-means||
    with ``kmeans_barbar=True``), then each cell searches the cells of its
    ``num_clusters_searched`` nearest clusters.

    ``X``: float32 device tensor, ``num_cells x num_dims`` (``num_dims <=
    256``).  Defaults as the reference's: ``num_clusters`` =
    ``ceil(min(4 sqrt(n), n / 100))``, ``num_clusters_searched`` =
    ``min(64, num_clusters)``, ``chunk_size_kmeans`` = ``min(4096, n)``.
    Returns ``(neighbors int32, distances float32, info)`` with ``info`` a
    dict: non-empty clusters, clusters searched, k-means iterations run,
    whether the exhaustive search ran (every cluster searched), and the
    number of (workgroup, cluster) pairs scanned, and the k-means||
    candidates selected (0 with random initial cells).
    """
    import ctypes

    import numpy as np
    import torch

    from .irlba import default_context
    barbar = _barbar_args(kmeans_barbar, num_init_iterations,
                          oversampling_factor)
    if not isinstance(X, torch.Tensor) or X.device.type != 'cuda':
        raise TypeError('X must be a device tensor')
    if X.dtype != torch.float32 or X.dim() != 2:
        raise ValueError('X must be a 2-D float32 tensor')
    n, d = X.shape
    k = int(num_neighbors)
    if n < 2:
        raise ValueError('there is only one cell, so nearest neighbors cannot '
                         'be calculated')
    if not 1 <= k <= n - 1:
        raise ValueError(f'num_neighbors is {k}, but must be between 1 and '
                         f'the number of cells - 1')
    if not 1 <= d <= IVF_MAX_DIMS:
        raise ValueError(f'X has {d} columns; at most {IVF_MAX_DIMS} are '
                         f'supported')
    C = default_num_clusters(n) if num_clusters is None else int(num_clusters)
    searched = min(64, C) if num_clusters_searched is None \
        else int(num_clusters_searched)
    chunk = min(4096, n) if chunk_size_kmeans is None \
        else int(chunk_size_kmeans)
    X = X.contiguous()
    ctx = context or default_context(X.device.index or 0)
    neighbors = torch.empty((n, k), dtype=torch.int32, device=X.device)
    distances = torch.empty((n, k), dtype=torch.float32, device=X.device)
    info = np.zeros(6, np.int64)
    _lib.check(_lib.lib().bpca_knn_ivf_ex(
        ctx.handle, X.data_ptr(), n, d, k, C, searched,
        int(num_kmeans_iterations), ctypes.c_float(kmeans_tolerance), chunk,
        ctypes.c_uint64(int(seed) & 0xffffffffffffffff), barbar[0],
        barbar[1], ctypes.c_float(barbar[2]),
        neighbors.data_ptr(), distances.data_ptr(), info.ctypes.data))
    return neighbors, distances, dict(
        num_clusters=int(info[0]), num_clusters_searched=int(info[1]),
        kmeans_iterations=int(info[2]), exhaustive=bool(info[3]),
        entries=int(info[4]), kmeans_candidates=int(info[5]))


def kmeans_init_cells(num_cells: int, num_clusters: int, seed: int = 0):
    """The cells ``kmeans()`` takes as its initial centroids with random
    initialisation (``brisc/kmeans.pyx:13-39``), drawn on the host with the
    reference's generator (``bpca_kmeans_init_cells``)."""
    import ctypes

    import numpy as np
    out = np.empty(int(num_clusters), np.uint32)
    _lib.check(_lib.lib().bpca_kmeans_init_cells(
        int(num_cells), int(num_clusters),
        ctypes.c_uint64(int(seed) & 0xffffffffffffffff), out.ctypes.data))
    return out


def kmeans_barbar_init(X, num_clusters: int, num_init_iterations: int = 5,
                       oversampling_factor: float = 1.0,
                       chunk_size: int | None = None, seed: int = 0,
                       context=None):
    """The cells k-means|| initialisation takes as centroids
    (``brisc/kmeans.pyx:41-360``) from a float32 device tensor ``X``
    (``num_dims <= 256``), through ``bpca_kmeans_barbar_init``.  Returns
    ``(cells uint32 (num_clusters,), candidates selected)``."""
    import ctypes

    import numpy as np
    import torch

    from .irlba import default_context
    _barbar_args(True, num_init_iterations, oversampling_factor)
    if not isinstance(X, torch.Tensor) or X.device.type != 'cuda' or \
            X.dtype != torch.float32 or X.dim() != 2:
        raise TypeError('X must be a 2-D float32 device tensor')
    n, d = X.shape
    X = X.contiguous()
    ctx = context or default_context(X.device.index or 0)
    cells = np.empty(int(num_clusters), np.uint32)
    S = ctypes.c_int64(0)
    _lib.check(_lib.lib().bpca_kmeans_barbar_init(
        ctx.handle, X.data_ptr(), n, d, int(num_clusters),
        int(num_init_iterations), ctypes.c_float(oversampling_factor),
        min(4096, n) if chunk_size is None else int(chunk_size),
        ctypes.c_uint64(int(seed) & 0xffffffffffffffff), cells.ctypes.data,
        ctypes.byref(S)))
    return cells, int(S.value)


def kmeans(X, num_clusters: int, num_kmeans_iterations: int = 2,
           kmeans_tolerance: float = 1e-2, chunk_size: int | None = None,
           seed: int = 0, context=None, kmeans_barbar: bool = False,
           num_init_iterations: int = 5, oversampling_factor: float = 1.0):
    """``kmeans()`` (``brisc/kmeans.pyx:889-1100``; random initial cells, or
    k-means|| with ``kmeans_barbar=True``) on a float32 device tensor ``X``
    (``num_dims <= 256``), through ``bpca_kmeans_ex``.  Returns ``(labels
    int32 (n,), centroids float32
    (num_clusters, d), counts int32 (num_clusters,), iterations run)`` --
    the labels and counts of a final E-step against the final centroids,
    as the reference leaves them."""
    import ctypes

    import torch

    from .irlba import default_context
    barbar = _barbar_args(kmeans_barbar, num_init_iterations,
                          oversampling_factor)
    if not isinstance(X, torch.Tensor) or X.device.type != 'cuda' or \
            X.dtype != torch.float32 or X.dim() != 2:
        raise TypeError('X must be a 2-D float32 device tensor')
    n, d = X.shape
    X = X.contiguous()
    ctx = context or default_context(X.device.index or 0)
    labels = torch.empty(n, dtype=torch.int32, device=X.device)
    cent = torch.empty((int(num_clusters), d), dtype=torch.float32,
                       device=X.device)
    counts = torch.empty(int(num_clusters), dtype=torch.int32, device=X.device)
    run = ctypes.c_int32(0)
    _lib.check(_lib.lib().bpca_kmeans_ex(
        ctx.handle, X.data_ptr(), n, d, int(num_clusters),
        int(num_kmeans_iterations), ctypes.c_float(kmeans_tolerance),
        min(4096, n) if chunk_size is None else int(chunk_size),
        ctypes.c_uint64(int(seed) & 0xffffffffffffffff), barbar[0],
        barbar[1], ctypes.c_float(barbar[2]), labels.data_ptr(),
        cent.data_ptr(), counts.data_ptr(), ctypes.byref(run)))
    return labels, cent, counts, int(run.value)
