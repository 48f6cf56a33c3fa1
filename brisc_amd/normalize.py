"""``SingleCell.normalize()`` on the device.

``normalize_device`` is the device counterpart of the reference's
``normalize_csr`` / ``normalize_csc`` (``brisc/normalize.pyx:9-460``): counts
held in device tensors (any of the reference's count types), an optional QC
mask, one of the methods ``'logCP10k'``, ``'log1pPF'`` (the default) or
``'PFlog1pPF'``; it returns the float32 normalised values, entry for entry
in the input's layout, and each cell's library size.  The values are the
reference's, bit for bit: the same single-precision steps in the same order
(PFlog1pPF's per-cell sums and their total included), and log1p
bit-identical to the host C library's log1pf.
"""
from __future__ import annotations

from . import _lib

METHODS = {'logCP10k': 0, 'log1pPF': 1, 'PFlog1pPF': 2}


def _dtype_code(t) -> int:
    import torch
    codes = {torch.int32: 0, torch.int64: 2, torch.float32: 4, torch.float64: 5}
    for name, code in (('uint32', 1), ('uint64', 3)):
        dt = getattr(torch, name, None)
        if dt is not None:
            codes[dt] = code
    if t not in codes:
        raise TypeError(f'counts must be int32, uint32, int64, uint64, float32 or '
                        f'float64, not {t}')
    return codes[t]


def normalize_device(data, indices, indptr, num_cells: int, num_genes: int,
                     is_csr: bool = True, QC=None, method: str = 'log1pPF',
                     out=None, context=None):
    """Normalise the counts of a CSR (``is_csr``) or CSC matrix held in
    device tensors.  ``QC`` is an optional boolean device tensor (or NumPy
    array) of cells passing QC.  ``out`` may be ``data`` itself for float32
    counts (in-place).  Returns ``(normalized float32, row_sums int64)``
    device tensors."""
    import numpy as np
    import torch

    from .irlba import default_context
    if method not in METHODS:
        raise ValueError("method must be one of 'PFlog1pPF', 'log1pPF', or "
                         "'logCP10k'")
    dev = data.device
    ctx = context or default_context(dev.index or 0)
    code = _dtype_code(data.dtype)
    for t, name in ((data, 'data'), (indices, 'indices'), (indptr, 'indptr')):
        if not t.is_cuda or t.dim() != 1 or not t.is_contiguous():
            raise ValueError(f'{name} must be a contiguous 1-D device tensor')
    if indices.dtype not in (torch.int32, torch.int64) or \
            indptr.dtype not in (torch.int32, torch.int64):
        raise TypeError('indices and indptr must be int32 or int64')
    outer = num_cells if is_csr else num_genes
    if indptr.numel() != outer + 1:
        raise ValueError(f'indptr must have {"num_cells" if is_csr else "num_genes"} + 1 entries')
    if out is None:
        out = torch.empty(data.numel(), dtype=torch.float32, device=dev)
    elif out.dtype != torch.float32 or out.numel() != data.numel():
        raise ValueError('out must be float32 with one value per entry')
    qc = None
    if QC is not None:
        qc = QC if isinstance(QC, torch.Tensor) else torch.from_numpy(np.asarray(QC))
        qc = qc.to(device=dev, dtype=torch.uint8).contiguous()
        if qc.numel() != num_cells:
            raise ValueError('QC must have one entry per cell')
    row_sums = torch.empty(num_cells, dtype=torch.int64, device=dev)
    _lib.check(_lib.lib().bpca_normalize(
        ctx.handle, data.data_ptr(), code, indices.data_ptr(),
        indices.element_size(), indptr.data_ptr(), indptr.element_size(),
        int(bool(is_csr)), int(num_cells), int(num_genes),
        qc.data_ptr() if qc is not None else None, METHODS[method],
        out.data_ptr(), row_sums.data_ptr()))
    return out, row_sums


def log1pf_device(x):
    """The device's log1pf (the restated host-library function normalize
    uses) of a float32 device tensor."""
    import torch

    from .irlba import default_context
    y = torch.empty_like(x)
    _lib.check(_lib.lib().bpca_test_log1pf(
        default_context(x.device.index or 0).handle, x.data_ptr(), x.numel(),
        y.data_ptr()))
    return y
