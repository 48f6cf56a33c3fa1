"""QC x HVG subset of a device-resident CSR matrix (operand preparation).

``SingleCell.pca()`` computes PCA on ``X[np.ix_(QC rows, HVG columns)]``
(``brisc/single_cell.py:11853-11908``); :func:`csr_subset` forms that matrix
on the device through ``bpca_csr_subset_count`` / ``bpca_csr_subset_fill``
(``include/bpca.h``): a validated count of the kept entries per output row,
an exclusive prefix sum into int64 row offsets, then the kept values and
remapped int32 column indices in ascending column order.  Integer work only:
the result equals SciPy's ``X[rows][:, genes]`` with sorted indices, bit for
bit.
"""
from __future__ import annotations

import numpy as np

from . import _lib


def column_map(num_cols: int, genes, device='cuda'):
    """int32 map for :func:`csr_subset`: ``genes`` (positions of the kept
    input columns, in output order; a boolean mask keeps its True columns in
    order) -> ``map[genes[i]] = i``, other columns -1.  Returns the map as a
    device tensor and whether it increases over the kept columns."""
    import torch
    genes = np.asarray(genes)
    if genes.dtype == bool:
        if genes.shape != (num_cols,):
            raise ValueError('a boolean gene mask must have one entry per column')
        genes = np.flatnonzero(genes)
    genes = genes.astype(np.int64, copy=False)
    if genes.size and (genes.min() < 0 or genes.max() >= num_cols):
        raise ValueError('gene positions must lie in [0, num_cols)')
    if np.unique(genes).size != genes.size:
        raise ValueError('gene positions must be unique')
    cmap = np.full(num_cols, -1, np.int32)
    cmap[genes] = np.arange(genes.size, dtype=np.int32)
    increasing = bool(np.all(np.diff(genes) > 0))
    return torch.from_numpy(cmap).to(device), increasing


def csr_subset(data, indices, indptr, num_cols: int, rows=None, genes=None,
               context=None):
    """Rows ``rows`` (int64 positions or a boolean mask, as NumPy arrays or
    device tensors; None = all) and
    columns ``genes`` (see :func:`column_map`; None = all) of a CSR matrix
    held in device tensors (float32 data; int32/int64 indices and indptr).

    Returns ``(data float32, indices int32, indptr int64)`` device tensors.
    Input rows must have strictly increasing column indices.
    """
    import torch

    from .irlba import default_context
    ctx = context or default_context(data.device.index or 0)
    dev = data.device
    L = _lib.lib()
    num_rows = indptr.numel() - 1
    nnz = indices.numel()
    if data.dtype != torch.float32 or data.numel() != nnz:
        raise ValueError('data must be float32 with one value per index')
    if indices.dtype not in (torch.int32, torch.int64) or indptr.dtype not in (torch.int32, torch.int64):
        raise ValueError('indices and indptr must be int32 or int64')
    for t in (data, indices, indptr):
        if not t.is_contiguous() or t.device != dev:
            raise ValueError('data, indices and indptr must be contiguous tensors on one device')
    if rows is None:
        d_rows, n_out = None, num_rows
    elif isinstance(rows, torch.Tensor):
        # Device-resident row selection: a boolean mask or int64 positions.
        r = rows.to(dev)
        if r.dtype == torch.bool:
            if r.shape != (num_rows,):
                raise ValueError('a boolean row mask must have one entry per row')
            r = torch.nonzero(r).flatten()
        d_rows = r.to(torch.int64).contiguous()
        n_out = int(d_rows.numel())
    else:
        r = np.asarray(rows)
        if r.dtype == bool:
            if r.shape != (num_rows,):
                raise ValueError('a boolean row mask must have one entry per row')
            r = np.flatnonzero(r)
        d_rows = torch.from_numpy(r.astype(np.int64)).to(dev)
        n_out = int(r.size)
    if genes is None:
        d_map, increasing = None, True
    else:
        d_map, increasing = column_map(num_cols, genes, dev)

    row_nnz = torch.empty(max(n_out, 1), dtype=torch.int64, device=dev)
    _lib.check(L.bpca_csr_subset_count(
        ctx.handle, indptr.data_ptr(), indptr.element_size(), indices.data_ptr(),
        indices.element_size(), num_rows, nnz, num_cols,
        d_rows.data_ptr() if d_rows is not None else None, n_out,
        d_map.data_ptr() if d_map is not None else None, row_nnz.data_ptr()))
    out_indptr = torch.zeros(n_out + 1, dtype=torch.int64, device=dev)
    if n_out:
        torch.cumsum(row_nnz[:n_out], 0, out=out_indptr[1:])
    out_nnz = int(out_indptr[-1].item())
    max_row = int(row_nnz[:n_out].max().item()) if n_out else 0
    out_data = torch.empty(max(out_nnz, 1), dtype=torch.float32, device=dev)
    out_indices = torch.empty(max(out_nnz, 1), dtype=torch.int32, device=dev)
    _lib.check(L.bpca_csr_subset_fill(
        ctx.handle, data.data_ptr(), indptr.data_ptr(), indptr.element_size(),
        indices.data_ptr(), indices.element_size(),
        d_rows.data_ptr() if d_rows is not None else None, n_out,
        d_map.data_ptr() if d_map is not None else None, int(increasing),
        out_indptr.data_ptr(), max_row, out_data.data_ptr(), out_indices.data_ptr()))
    return out_data[:out_nnz], out_indices[:out_nnz], out_indptr


def csr_transpose(data, indices, indptr, num_cols: int, context=None):
    """Transpose of a sparse matrix stored by rows, on the device: the CSR of
    X from its CSC (or the CSC from the CSR), as brisc's csr_tocsc / scipy's
    tocsr do (``brisc/cysparse.pyx:1179-1295``).  ``data`` float32,
    ``indices`` and ``indptr`` int32/int64 device tensors; ``num_cols`` is the
    input's minor dimension.  Returns ``(data float32, indices int32, indptr
    int64)`` device tensors, each output row in ascending input-row order.
    """
    import torch

    from .irlba import default_context
    ctx = context or default_context(data.device.index or 0)
    dev = data.device
    num_rows = indptr.numel() - 1
    nnz = indices.numel()
    if data.dtype != torch.float32 or data.numel() != nnz:
        raise ValueError('data must be float32 with one value per index')
    if indices.dtype not in (torch.int32, torch.int64) or indptr.dtype not in (torch.int32, torch.int64):
        raise ValueError('indices and indptr must be int32 or int64')
    for t in (data, indices, indptr):
        if not t.is_contiguous() or t.device != dev:
            raise ValueError('data, indices and indptr must be contiguous tensors on one device')
    out_data = torch.empty(nnz, dtype=torch.float32, device=dev)
    out_indices = torch.empty(nnz, dtype=torch.int32, device=dev)
    out_indptr = torch.empty(int(num_cols) + 1, dtype=torch.int64, device=dev)
    _lib.check(_lib.lib().bpca_csr_transpose(
        ctx.handle, data.data_ptr(), indices.data_ptr(), indices.element_size(),
        indptr.data_ptr(), indptr.element_size(), num_rows, int(num_cols),
        out_data.data_ptr(), out_indices.data_ptr(), out_indptr.data_ptr()))
    return out_data, out_indices, out_indptr
