"""MI355X-native sparse PCA (IRLBA) for single-cell count matrices.

Drop-in for the IRLBA hot path of brisc (``brisc/pca.pyx:661-1257`` behind
``SingleCell.pca()``): hand-written gfx950 HIP kernels behind a C ABI
(``include/bpca.h``, ``libbrisc_pca.so``), reached from Python through
ctypes.

* :func:`irlba` -- same keyword signature as ``brisc.pca.irlba``.
* :func:`irlba_device` -- device-resident CSR (torch tensors) in HBM.
* :class:`SingleCell` -- ``pca()`` with the reference's signature, checks and
  ``obsm`` output layout.
* :mod:`brisc_amd.synth` -- synthetic NB count workloads (host parameters and
  the device generator).
"""
from .irlba import (Context, Engine, default_context, irlba, irlba_batched,
                    irlba_device, small_svd)
from . import serve, synth
from .single_cell import SingleCell

__all__ = ['Context', 'Engine', 'SingleCell', 'default_context', 'irlba',
           'irlba_batched', 'irlba_device', 'small_svd']
