"""``SingleCell.pca()`` on the GPU: the caller side of the IRLBA path.

``SingleCell`` here is a minimal container (X, obs, var, obsm, uns) whose
``pca()`` reproduces ``brisc/single_cell.py:11625-12137``: the same signature
and defaults, the same argument validation and error messages, the QC x HVG
subset, joint PCA over several datasets, the ``subspace_size`` clip, and the
output contract -- ``obsm[PC_key]`` is a float32 C-contiguous
``(len(obs), num_PCs)`` array of scores ``U_k S_k`` with NaN rows for cells
that fail QC.  ``pca()`` writes nothing to ``varm``, as the reference.

``obs`` and ``var`` are pandas DataFrames or dicts of 1-D arrays (the
reference uses polars, which this package does not depend on).  QC and HVG
columns may be a column name, a 1-D boolean array, or a function of the
dataset returning one, as in the reference.
"""
from __future__ import annotations

import os
from typing import Any

import numpy as np
import scipy.sparse as sp


def check_type(variable: Any, variable_name: str, expected_types,
               expected_type_name: str) -> None:
    """Type check with the reference's message (``brisc/utils.py:326-358``)."""
    if isinstance(variable, expected_types):
        return
    if not isinstance(expected_types, tuple):
        expected_types = expected_types,
    for t in expected_types:
        if t is int and isinstance(variable, np.integer):
            return
        if t is float and isinstance(variable, np.floating):
            return
        if t is bool and isinstance(variable, np.bool_):
            return
    raise TypeError(f'{variable_name} must be {expected_type_name}, but has '
                    f'type {type(variable).__name__!r}')


def check_bounds(variable, variable_name: str, lower_bound=None,
                 upper_bound=None, *, left_open: bool = False,
                 right_open: bool = False) -> None:
    """Bound check with the reference's message (``brisc/utils.py:197-234``)."""
    if lower_bound is not None and (variable <= lower_bound if left_open
                                    else variable < lower_bound) or \
            upper_bound is not None and (variable >= upper_bound if right_open
                                         else variable > upper_bound):
        msg = f'{variable_name} is {variable:,}, but must be'
        if lower_bound is not None:
            msg += f' {">" if left_open else "≥"} {lower_bound:,}'
            if upper_bound is not None:
                msg += ' and'
        if upper_bound is not None:
            msg += f' {"<" if right_open else "≤"} {upper_bound:,}'
        raise ValueError(msg)


def _columns(frame) -> dict:
    if frame is None:
        return {}
    if hasattr(frame, 'columns') and hasattr(frame, '__getitem__'):
        return {c: np.asarray(frame[c]) for c in frame.columns}
    return {k: np.asarray(v) for k, v in dict(frame).items()}


class SingleCell:
    """Cells x genes counts with obs/var annotations and obsm results."""

    def __init__(self, X=None, obs=None, var=None, obsm=None, varm=None,
                 uns=None, num_threads: int = 1):
        if isinstance(X, (str, os.PathLike)):
            # An .h5ad file (single_cell.py:83-1158): X, obs, var, obsm.
            from .h5ad import read_h5ad
            loaded = read_h5ad(os.fspath(X))
            X, obs, var = loaded._X, obs if obs is not None else loaded._obs, \
                var if var is not None else loaded._var
            obsm = obsm if obsm is not None else loaded._obsm
        if X is not None and not isinstance(X, (sp.csr_array, sp.csc_array,
                                                sp.csr_matrix, sp.csc_matrix)):
            raise TypeError('X must be a CSR or CSC sparse array')
        self._X = X
        if X is not None:
            n, g = X.shape
        else:  # annotations and obsm only (e.g. PCs for neighbors())
            n = len(next(iter(_columns(obs).values()), ())) if obs is not None \
                else 0
            g = 0
        self._obs = obs if obs is not None else {'cell': np.arange(n)}
        self._var = var if var is not None else {
            'gene': np.array([f'gene{i}' for i in range(g)])}
        self._obsm = dict(obsm or {})
        self._varm = dict(varm or {})
        self._uns = dict(uns or {})
        self._uns.setdefault('normalized', False)
        self._num_threads = os.cpu_count() if num_threads == -1 else num_threads
        for key, value in self._obsm.items():
            if len(value) != n:
                raise ValueError(f'obsm[{key!r}] has length {len(value)}, '
                                 f'but there are {n:,} cells')

    X = property(lambda self: self._X)
    obs = property(lambda self: self._obs)
    var = property(lambda self: self._var)
    obsm = property(lambda self: self._obsm)
    varm = property(lambda self: self._varm)
    uns = property(lambda self: self._uns)
    num_threads = property(lambda self: self._num_threads)

    def __len__(self) -> int:
        return self._X.shape[0] if self._X is not None else \
            len(next(iter(_columns(self._obs).values()), ()))

    @property
    def var_names(self) -> np.ndarray:
        cols = _columns(self._var)
        return np.asarray(next(iter(cols.values())))

    def _process_num_threads(self, num_threads) -> int:
        """``single_cell.py:1787-1817``."""
        if num_threads is None:
            return self._num_threads
        check_type(num_threads, 'num_threads', int,
                   'a positive integer, -1, or None')
        if num_threads == -1:
            return os.cpu_count()
        num_threads = int(num_threads)
        if num_threads <= 0:
            raise ValueError(f'num_threads is {num_threads:,}, but must be a '
                             f'positive integer, -1, or None')
        return num_threads

    def save(self, path, overwrite: bool = False) -> None:
        """Write X, obs, var and obsm to an .h5ad file
        (``single_cell.py:3954``)."""
        from .h5ad import write_h5ad
        write_h5ad(self, os.fspath(path), overwrite=overwrite)

    def skip_qc(self) -> 'SingleCell':
        """The dataset with ``uns['QCed']`` set to True
        (``single_cell.py:9888-9896``)."""
        uns = dict(self._uns)
        uns['QCed'] = True
        return SingleCell(X=self._X, obs=self._obs, var=self._var,
                          obsm=self._obsm, varm=self._varm, uns=uns,
                          num_threads=self._num_threads)

    def normalize(self, *, QC_column='passed_QC', method='log1pPF',
                  inplace: bool = False, num_threads=None) -> 'SingleCell':
        """Normalise the counts on the GPU (``single_cell.py:11450-11624``):
        the same signature, checks and messages; ``method`` one of
        ``'PFlog1pPF'``, ``'log1pPF'`` (default) or ``'logCP10k'``; cells
        failing QC are normalised but do not count towards the mean library
        size.  X keeps its layout (CSR or CSC) and its indices; the values
        become float32.  Returns a new dataset with ``uns['normalized']``
        True, or this one normalised in place (float32 counts only)."""
        import torch

        from .normalize import normalize_device
        if self._X is None:
            raise ValueError('X is None, so normalizing is not possible')
        if not self._uns.get('QCed', False):
            raise ValueError(
                "uns['QCed'] is False; did you forget to run qc() (and "
                "possibly hvg()) before normalize()? Set uns['QCed'] = True "
                "or run skip_qc() to bypass this check.")
        if self._uns['normalized']:
            raise ValueError(
                "uns['normalized'] is True; did you already run normalize()?")
        qc = self._get_column('obs', QC_column, 'QC_column',
                              QC_column == 'passed_QC', None, len(self)) \
            if QC_column is not None else None
        if qc is not None and not qc.any():
            raise ValueError('no cells pass QC: QC_column is all-False')
        if method not in ('PFlog1pPF', 'log1pPF', 'logCP10k'):
            raise ValueError("method must be one of 'PFlog1pPF', 'log1pPF', or "
                             "'logCP10k'")
        check_type(inplace, 'inplace', bool, 'Boolean')
        if inplace and self._X.dtype != np.float32:
            raise TypeError(f'inplace normalization requires floating-point '
                            f'count matrices, but X has data type '
                            f'{str(self._X.dtype)!r}')
        self._process_num_threads(num_threads)
        X = self._X
        is_csr = isinstance(X, (sp.csr_array, sp.csr_matrix))
        dev = torch.device('cuda', 0)
        d, i, p = (torch.from_numpy(np.ascontiguousarray(a)).to(dev)
                   for a in (X.data, X.indices, X.indptr))
        out, _ = normalize_device(d, i, p, X.shape[0], X.shape[1],
                                  is_csr=is_csr, QC=qc, method=method,
                                  out=d if inplace else None)
        values = out.cpu().numpy()
        if inplace:
            X.data[:] = values
            self._uns['normalized'] = True
            return self
        cls = type(X) if isinstance(X, (sp.csr_array, sp.csc_array)) else \
            (sp.csr_array if is_csr else sp.csc_array)
        Y = cls((values, X.indices, X.indptr), shape=X.shape)
        uns = dict(self._uns)
        uns['normalized'] = True
        return SingleCell(X=Y, obs=self._obs, var=self._var, obsm=self._obsm,
                          varm=self._varm, uns=uns,
                          num_threads=self._num_threads)

    def _get_column(self, where: str, column, name: str, allow_missing: bool,
                    custom_error: str | None, length: int):
        if column is None:
            return None
        if callable(column):
            column = column(self)
        if isinstance(column, str):
            cols = _columns(self._obs if where == 'obs' else self._var)
            if column not in cols:
                if allow_missing:
                    return None
                if custom_error is not None:
                    raise ValueError(custom_error.format(repr(column)))
                raise ValueError(f'{name} {column!r} is not a column of '
                                 f'{where}')
            column = cols[column]
        column = np.asarray(column)
        if column.dtype != np.bool_:
            raise TypeError(f'{name} must be Boolean, but has data type '
                            f'{str(column.dtype)!r}')
        if column.ndim != 1 or len(column) != length:
            raise ValueError(f'{name} has length {len(column):,}, but must '
                             f'have length {length:,}')
        return column

    def pca(self, *others: 'SingleCell', QC_column='passed_QC',
            hvg_column='highly_variable', PC_key: str = 'pca',
            num_PCs: int = 50, subspace_size: int = 100,
            tolerance: float = 1e-6, max_iterations: int = 100,
            chunk_size: int = 1024, seed: int = 0,
            match_parallel: bool = False, overwrite: bool = False,
            verbose: bool = False, num_threads: int | None = None):
        """Top ``num_PCs`` principal components across cells, on the GPU.

        Same contract as ``brisc/single_cell.py:11625-12137``; see the module
        docstring.  Returns a new ``SingleCell`` (or a tuple, one per dataset,
        when ``others`` is given) with ``obsm[PC_key]`` added.
        """
        if others:
            for o in others:
                if not isinstance(o, SingleCell):
                    raise TypeError('all elements of others must be SingleCell '
                                    'datasets')
        datasets = [self] + list(others)
        check_type(PC_key, 'PC_key', str, 'a string')
        check_type(overwrite, 'overwrite', bool, 'Boolean')
        suffix = ' for at least one dataset' if others else ''
        for dataset in datasets:
            if not overwrite and PC_key in dataset._obsm:
                raise ValueError(
                    f'PC_key {PC_key!r} is already a key of obsm{suffix}; did '
                    f'you already run PCA()? Set overwrite=True to overwrite.')
        if any(dataset._X is None for dataset in datasets):
            raise ValueError(f'X is None{suffix}, so finding principal '
                             f'components is not possible')
        n_sets = len(datasets)

        def per_dataset(column):
            if isinstance(column, (list, tuple)) and not isinstance(column, str):
                if len(column) != n_sets:
                    raise ValueError(f'column sequence must have length '
                                     f'{n_sets}')
                return list(column)
            return [column] * n_sets

        QC_columns = [d._get_column('obs', c, 'QC_column', True, None, len(d))
                      for d, c in zip(datasets, per_dataset(QC_column))]
        if any(c is not None and not c.any() for c in QC_columns):
            raise ValueError(f'no cells pass QC{suffix}: QC_column is '
                             f'all-False')
        hvg_error = (f'hvg_column {{}} is not a column of var{suffix}; did '
                     f'you forget to run hvg() (and possibly normalize()) '
                     f'before PCA()?')
        hvg_columns = [d._get_column('var', c, 'hvg_column', False, hvg_error,
                                     d._X.shape[1])
                       for d, c in zip(datasets, per_dataset(hvg_column))]
        if not all(dataset._uns['normalized'] for dataset in datasets):
            raise ValueError(
                f"PCA() requires normalized counts but uns['normalized'] is "
                f"False{suffix}; did you forget to run normalize() before "
                f"PCA()?")
        for dataset in datasets:
            if dataset._X.dtype != np.float32:
                raise TypeError(
                    f'PCA() requires normalized counts with data type '
                    f'float32, but X has data type '
                    f'{str(dataset._X.dtype)!r}{suffix}; did you forget to '
                    f'run normalize() before PCA()?')
        check_type(num_PCs, 'num_PCs', int, 'a positive integer')
        check_bounds(num_PCs, 'num_PCs', 1)
        check_type(subspace_size, 'subspace_size', int, 'a positive integer')
        if subspace_size < num_PCs:
            raise ValueError(f'subspace_size is {subspace_size:,}, but must be '
                             f'≥ num_PCs ({num_PCs:,})')
        check_type(tolerance, 'tolerance', (int, float), 'a positive number')
        check_bounds(tolerance, 'tolerance', 0, left_open=True)
        check_type(max_iterations, 'max_iterations', int,
                   'a positive integer')
        check_bounds(max_iterations, 'max_iterations', 1)
        check_type(chunk_size, 'chunk_size', int, 'a positive integer')
        check_bounds(chunk_size, 'chunk_size', 1)
        check_type(seed, 'seed', int, 'an integer')
        check_type(verbose, 'verbose', bool, 'Boolean')
        num_threads = self._process_num_threads(num_threads)
        check_type(match_parallel, 'match_parallel', bool, 'Boolean')
        if match_parallel and num_threads != 1:
            raise ValueError(
                'match_parallel must be False unless num_threads is 1')

        # QC x HVG subset; several datasets are aligned on shared gene names.
        if others:
            names = [d.var_names[h] if h is not None else d.var_names
                     for d, h in zip(datasets, hvg_columns)]
            common = sorted(set(names[0]).intersection(*names[1:]))
            gene_indices = []
            for d, h in zip(datasets, hvg_columns):
                pos = {name: i for i, name in enumerate(d.var_names)}
                idx = np.array([pos[name] for name in common], dtype=np.int64)
                gene_indices.append(idx)
        else:
            gene_indices = [np.flatnonzero(h) if h is not None else None
                            for h in hvg_columns]
        # Operand preparation on the device: each dataset's CSR is uploaded
        # once and the QC x HVG subset is formed there (device subset of
        # brisc_amd.subset; the reference takes X[np.ix_(QC, HVG)] on the
        # host, single_cell.py:11853-11908), then the datasets' rows are
        # stacked in order.  A gene order that is not increasing (joint PCA)
        # is sorted per row on the device when rows keep at most 4096 genes,
        # else that dataset is subset on the host.
        import torch

        from .irlba import irlba_device
        from .subset import csr_subset, csr_transpose
        dev = torch.device('cuda', 0)
        parts = []
        for d, q, genes in zip(datasets, QC_columns, gene_indices):
            X = d._X
            if isinstance(X, (sp.csc_array, sp.csc_matrix)) and X.has_canonical_format and \
                    not (genes is not None and len(genes) > 4096 and not np.all(np.diff(genes) > 0)):
                # CSC: uploaded as it is and turned into the CSR on the
                # device (the reference's X.tocsr(), single_cell.py:12066).
                upload = [torch.from_numpy(np.ascontiguousarray(a)).to(dev)
                          for a in (X.data.astype(np.float32, copy=False),
                                    X.indices, X.indptr)]
                csr = csr_transpose(*upload, X.shape[0])
                del upload
                parts.append(csr_subset(*csr, X.shape[1], rows=q, genes=genes))
                del csr
                continue
            if not isinstance(X, (sp.csr_array, sp.csr_matrix)):
                X = sp.csr_array(X)
            if not X.has_canonical_format:
                X = sp.csr_array(X, copy=True)
                X.sum_duplicates()
                X.sort_indices()
            if genes is not None and len(genes) > 4096 and \
                    not np.all(np.diff(genes) > 0):
                X = sp.csr_array(X[np.flatnonzero(q)] if q is not None else X)[:, genes]
                X = sp.csr_array(X)
                X.sort_indices()
                q = genes = None
            upload = [torch.from_numpy(np.ascontiguousarray(a)).to(dev)
                      for a in (X.data.astype(np.float32, copy=False),
                                X.indices, X.indptr)]
            parts.append(csr_subset(*upload, X.shape[1], rows=q, genes=genes))
            del upload
        num_cells_per_dataset = np.array([p[2].numel() - 1 for p in parts])
        num_cells = int(num_cells_per_dataset.sum())
        num_genes = (len(gene_indices[0]) if gene_indices[0] is not None
                     else datasets[0]._X.shape[1])
        if len(parts) == 1:
            data, indices, indptr = parts[0]
        else:
            offsets = np.cumsum([0] + [int(p[2][-1].item()) for p in parts])
            data = torch.cat([p[0] for p in parts])
            indices = torch.cat([p[1] for p in parts])
            indptr = torch.cat([parts[0][2][:1]] +
                               [p[2][1:] + int(o) for p, o in zip(parts, offsets[:-1])])
        del parts
        if num_cells < num_PCs:
            raise ValueError(f'num_PCs is {num_PCs:,}, but must be ≤ the '
                             f'number of cells ({num_cells:,})')
        if num_genes < num_PCs:
            raise ValueError(f'num_PCs is {num_PCs:,}, but must be ≤ the '
                             f'number of genes ({num_genes:,})')
        if num_cells == 1:
            raise ValueError('there is only one cell, so principal components '
                             'cannot be calculated')
        if num_genes == 1:
            raise ValueError('there is only one gene, so principal components '
                             'cannot be calculated')
        subspace_size = min(subspace_size, num_cells, num_genes)
        P, _, stats = irlba_device(data, indices, indptr, num_cells, num_genes,
                                   num_PCs, subspace_size=subspace_size,
                                   tolerance=tolerance,
                                   max_iterations=max_iterations, seed=seed)
        del data, indices, indptr
        PCs = P.cpu().numpy()
        converged = bool(stats['converged'])
        if not converged and verbose:
            print(f'PCA did not converge to a tolerance of {tolerance:.2g} '
                  f'after {max_iterations:,} iterations; consider increasing '
                  f'max_iterations or tolerance')

        out = []
        start = 0
        for d, q, n in zip(datasets, QC_columns, num_cells_per_dataset):
            part = PCs[start:start + n]
            start += n
            if q is not None:
                full = np.full((len(d), num_PCs), np.nan, dtype=np.float32)
                full[q] = part
                part = full
            elif len(datasets) > 1:
                part = np.ascontiguousarray(part)
            out.append(SingleCell(X=d._X, obs=d._obs, var=d._var,
                                  obsm=d._obsm | {PC_key: part},
                                  varm=d._varm, uns=d._uns,
                                  num_threads=d._num_threads))
        return out[0] if not others else tuple(out)

    def neighbors(self, *, QC_column='passed_QC', PC_key: str = 'pca',
                  neighbors_key: str = 'neighbors',
                  distances_key: str = 'distances', num_neighbors: int = 20,
                  num_clusters: int | None = None,
                  num_clusters_searched: int | None = None,
                  num_kmeans_iterations: int = 2,
                  kmeans_tolerance: float = 1e-2, kmeans_barbar: bool = False,
                  num_init_iterations: int = 5,
                  oversampling_factor: float = 1,
                  chunk_size_kmeans: int | None = None,
                  chunk_size_search: int | None = None, seed: int = 0,
                  overwrite: bool = False, verbose: bool = False,
                  num_threads: int | None = None) -> 'SingleCell':
        """The ``num_neighbors`` nearest neighbours of each cell in PC space,
        on the GPU.

        Same signature, checks and outputs as
        ``brisc/single_cell.py:12139-12665``: ``obsm[neighbors_key]`` holds
        uint32 indices (among QC'd cells) of each cell's nearest cells, not
        counting the cell itself, ascending by distance;
        ``obsm[distances_key]`` holds the FP32 squared Euclidean distances;
        cells failing QC get ``UINT32_MAX`` and NaN.

        The search is the reference's: k-means (random initialisation drawn
        with the reference's generator, or k-means|| with
        ``kmeans_barbar=True``; ``num_kmeans_iterations`` Lloyd iterations)
        into ``num_clusters`` clusters, then each cell searches
        the cells of its ``num_clusters_searched`` nearest clusters
        (``brisc_amd.knn.knn_ivf``).  When every cluster is searched, or
        more than 64 would be, the exhaustive search runs instead, which
        returns what searching every cluster returns.  More than 64
        neighbours are found in passes of 64 (each keeping only candidates
        after the previous pass's last distance and index).  Limits on the
        GPU: the PCs have at most 256 columns.
        """
        import torch

        from .knn import (IVF_MAX_DIMS, MAX_DIMS, MAX_NEIGHBORS_PER_PASS,
                          knn_exact, knn_ivf)
        check_type(neighbors_key, 'neighbors_key', str, 'a string')
        check_type(distances_key, 'distances_key', str, 'a string')
        check_type(overwrite, 'overwrite', bool, 'Boolean')
        if not overwrite and neighbors_key in self._obsm:
            raise ValueError(
                f'neighbors_key {neighbors_key!r} is already a key of obsm; '
                f'did you already run neighbors()? Set overwrite=True to '
                f'overwrite.')
        if QC_column is not None:
            QC_column = self._get_column('obs', QC_column, 'QC_column',
                                         QC_column == 'passed_QC', None,
                                         len(self))
            if QC_column is not None and not QC_column.any():
                raise ValueError('no cells pass QC: QC_column is all-False')
        check_type(PC_key, 'PC_key', str, 'a string')
        if PC_key not in self._obsm:
            message = f'PC_key {PC_key!r} is not a key of obsm'
            if PC_key == 'pca':
                message += '; did you forget to run PCA() before neighbors()?'
            raise ValueError(message)
        PCs = self._obsm[PC_key]
        if PCs.dtype != np.float32:
            raise TypeError(f'obsm[{PC_key!r}].dtype is {PCs.dtype!r}, but '
                            f'must be float32')
        if not PCs.flags['C_CONTIGUOUS']:
            raise ValueError(
                f'obsm[{PC_key!r}] is not C-contiguous; make it C-contiguous '
                f'with pipe_obsm_key({PC_key!r}, np.ascontiguousarray)')
        check_type(num_kmeans_iterations, 'num_kmeans_iterations', int,
                   'a positive integer')
        check_bounds(num_kmeans_iterations, 'num_kmeans_iterations', 1)
        check_type(kmeans_tolerance, 'kmeans_tolerance', (int, float),
                   'a positive number')
        check_bounds(kmeans_tolerance, 'kmeans_tolerance', 0, left_open=True)
        check_type(kmeans_barbar, 'kmeans_barbar', bool, 'Boolean')
        check_type(num_init_iterations, 'num_init_iterations', int,
                   'a positive integer')
        check_bounds(num_init_iterations, 'num_init_iterations', 1)
        check_type(oversampling_factor, 'oversampling_factor', (int, float),
                   'a positive number')
        check_bounds(oversampling_factor, 'oversampling_factor', 0,
                     left_open=True)
        if not kmeans_barbar:
            if num_init_iterations != 5:
                raise ValueError('num_init_iterations can only be specified '
                                 'when kmeans_barbar=True')
            if oversampling_factor != 1:
                raise ValueError('oversampling_factor can only be specified '
                                 'when kmeans_barbar=True')
        check_type(seed, 'seed', int, 'an integer')
        check_type(verbose, 'verbose', bool, 'Boolean')
        self._process_num_threads(num_threads)
        if QC_column is not None:
            PCs = PCs[QC_column]
        num_cells = len(PCs)
        if num_cells == 1:
            raise ValueError('there is only one cell, so nearest neighbors '
                             'cannot be calculated')
        check_type(num_neighbors, 'num_neighbors', int, 'a positive integer')
        if not 1 <= num_neighbors < num_cells:
            raise ValueError(
                f'num_neighbors is {num_neighbors:,}, but must be ≥ 1 and '
                f'less than the number of cells ({num_cells:,})')
        if num_clusters is not None:
            check_type(num_clusters, 'num_clusters', int, 'a positive integer')
            if not 1 <= num_clusters <= num_cells:
                raise ValueError(
                    f'num_clusters is {num_clusters:,}, but must be ≥ 1 and '
                    f'≤ the number of cells ({num_cells:,})')
        else:
            num_clusters = max(1, int(np.ceil(min(4 * np.sqrt(num_cells),
                                                  num_cells / 100))))
        if num_clusters_searched is not None:
            check_type(num_clusters_searched, 'num_clusters_searched', int,
                       'a positive integer')
            if not 1 <= num_clusters_searched <= num_clusters:
                raise ValueError(
                    f'num_clusters_searched is {num_clusters_searched:,}, but '
                    f'must be ≥ 1 and ≤ num_clusters ({num_clusters:,})')
        for name, value in (('chunk_size_kmeans', chunk_size_kmeans),
                            ('chunk_size_search', chunk_size_search)):
            if value is not None:
                check_type(value, name, int, 'a positive integer')
                if not 1 <= value <= num_cells:
                    raise ValueError(
                        f'{name} is {value:,}, but must be ≥ 1 and ≤ the '
                        f'number of cells ({num_cells:,})')
        if num_clusters_searched is None:
            num_clusters_searched = min(64, num_clusters)
        if chunk_size_kmeans is None:
            chunk_size_kmeans = min(4096, num_cells)
        exhaustive = num_clusters_searched == num_clusters
        max_dims = MAX_DIMS if exhaustive and \
            num_neighbors <= MAX_NEIGHBORS_PER_PASS else IVF_MAX_DIMS
        if PCs.shape[1] > max_dims:
            raise ValueError(f'obsm[{PC_key!r}] has {PCs.shape[1]:,} columns; '
                             f'at most {max_dims} are supported on the GPU')
        X = torch.from_numpy(np.ascontiguousarray(PCs)).to('cuda')
        if exhaustive:
            nb, dist = knn_exact(X, num_neighbors)
        else:
            nb, dist, info = knn_ivf(
                X, num_neighbors, num_clusters=num_clusters,
                num_clusters_searched=num_clusters_searched,
                num_kmeans_iterations=num_kmeans_iterations,
                kmeans_tolerance=kmeans_tolerance,
                chunk_size_kmeans=chunk_size_kmeans, seed=seed,
                kmeans_barbar=kmeans_barbar,
                num_init_iterations=num_init_iterations,
                oversampling_factor=oversampling_factor)
            if verbose:
                print(f'neighbors: {info["num_clusters"]:,} clusters, '
                      f'{info["num_clusters_searched"]} searched per cell, '
                      f'{info["kmeans_iterations"]} k-means iterations')
        neighbors = nb.cpu().numpy().view(np.uint32)
        distances = dist.cpu().numpy()
        del X, nb, dist
        if QC_column is not None:
            full_nb = np.full((len(self), num_neighbors), 4_294_967_295,
                              dtype=np.uint32)
            full_nb[QC_column] = neighbors
            full_d = np.full((len(self), num_neighbors), np.nan,
                             dtype=np.float32)
            full_d[QC_column] = distances
            neighbors, distances = full_nb, full_d
        return SingleCell(X=self._X, obs=self._obs, var=self._var,
                          obsm=self._obsm | {neighbors_key: neighbors,
                                             distances_key: distances},
                          varm=self._varm, uns=self._uns,
                          num_threads=self._num_threads)
