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

MAX_NEIGHBORS = 64
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
    if not 1 <= k <= min(MAX_NEIGHBORS, n - 1):
        raise ValueError(f'num_neighbors is {k}, but must be between 1 and '
                         f'min({MAX_NEIGHBORS}, number of cells - 1)')
    if not 1 <= d <= MAX_DIMS:
        raise ValueError(f'X has {d} columns; at most {MAX_DIMS} are supported')
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
