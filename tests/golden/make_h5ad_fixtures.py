"""Write the .h5ad fixtures of the h5ad tests with h5py, in the on-disk
layout anndata defines and brisc's SingleCell reads
(brisc/single_cell.py:600-760, 2217-2300): X as a csr_matrix / csc_matrix
group (data, indices, indptr; shape attribute) or a dense dataset, obs and
var as dataframe groups (_index, column-order; boolean, integer, float,
string and categorical columns), obsm arrays.

Run with the image's h5py (the test interpreter has none):
    /opt/conda/bin/python3.9 tests/golden/make_h5ad_fixtures.py

The matrix is a small seeded sparse count matrix drawn here with NumPy; the
tests regenerate it from the same seed with h5ad_fixture_counts().
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def h5ad_fixture_counts(seed=7, n=300, g=40, density=0.2):
    """(indptr int64, indices int32, counts int32) of the fixture matrix, rows
    with sorted column indices; rows 10-14 empty."""
    rng = np.random.Generator(np.random.PCG64(seed))
    mask = rng.random((n, g)) < density
    mask[10:15] = False
    vals = rng.integers(1, 30, size=(n, g))
    indptr = np.zeros(n + 1, np.int64)
    indptr[1:] = np.cumsum(mask.sum(1))
    rows, cols = np.nonzero(mask)
    return indptr, cols.astype(np.int32), vals[rows, cols].astype(np.int32)


def _str_attr(obj, name, value):
    import h5py
    obj.attrs.create(name, value, dtype=h5py.string_dtype('utf-8'))


def _strings(group, name, values):
    import h5py
    ds = group.create_dataset(name, data=np.array(values, dtype=object),
                              dtype=h5py.string_dtype('utf-8'))
    _str_attr(ds, 'encoding-type', 'string-array')
    _str_attr(ds, 'encoding-version', '0.2.0')
    return ds


def _array(group, name, values):
    ds = group.create_dataset(name, data=values)
    _str_attr(ds, 'encoding-type', 'array')
    _str_attr(ds, 'encoding-version', '0.2.0')
    return ds


def _dataframe(f, name, index, columns):
    import h5py
    grp = f.create_group(name)
    _str_attr(grp, 'encoding-type', 'dataframe')
    _str_attr(grp, 'encoding-version', '0.2.0')
    _str_attr(grp, '_index', '_index')
    grp.attrs.create('column-order', np.array(list(columns), dtype=object),
                     dtype=h5py.string_dtype('utf-8'))
    _strings(grp, '_index', index)
    for key, value in columns.items():
        if isinstance(value, tuple):          # categorical: (codes, categories)
            codes, cats = value
            cg = grp.create_group(key)
            _str_attr(cg, 'encoding-type', 'categorical')
            _str_attr(cg, 'encoding-version', '0.2.0')
            cg.attrs['ordered'] = False
            _array(cg, 'codes', codes)
            _strings(cg, 'categories', cats)
        elif value.dtype == object:
            _strings(grp, key, value)
        else:
            _array(grp, key, value)


def write(path, layout):
    import h5py
    n, g = 300, 40
    indptr, indices, counts = h5ad_fixture_counts()
    rng = np.random.Generator(np.random.PCG64(8))
    with h5py.File(path, 'w') as f:
        _str_attr(f, 'encoding-type', 'anndata')
        _str_attr(f, 'encoding-version', '0.1.0')
        import scipy.sparse as sp
        X = sp.csr_matrix((counts, indices, indptr), shape=(n, g))
        if layout == 'dense':
            _array(f, 'X', X.toarray().astype(np.float32))
        else:
            M = X if layout == 'csr' else X.tocsc()
            M.sort_indices()
            grp = f.create_group('X')
            _str_attr(grp, 'encoding-type', f'{layout}_matrix')
            _str_attr(grp, 'encoding-version', '0.1.0')
            grp.attrs['shape'] = np.array([n, g], dtype=np.int64)
            grp.create_dataset('data', data=M.data.astype(np.int32))
            grp.create_dataset('indices', data=M.indices.astype(np.int32))
            grp.create_dataset('indptr', data=M.indptr.astype(np.int64))
        qc = rng.random(n) < 0.85
        types = rng.integers(0, 3, n).astype(np.int8)
        _dataframe(f, 'obs', [f'cell{i}' for i in range(n)],
                   {'passed_QC': qc, 'n_counts': np.asarray(X.sum(1)).ravel().astype(np.int64),
                    'cell_type': (types, ['B', 'T', 'NK']),
                    'score': rng.random(n).astype(np.float64)})
        hvg = np.zeros(g, bool)
        hvg[rng.choice(g, 25, replace=False)] = True
        _dataframe(f, 'var', [f'gene{j}' for j in range(g)],
                   {'highly_variable': hvg,
                    'symbol': np.array([f'G{j:03d}' for j in range(g)], dtype=object)})
        obsm = f.create_group('obsm')
        _str_attr(obsm, 'encoding-type', 'dict')
        _str_attr(obsm, 'encoding-version', '0.1.0')
        _array(obsm, 'X_umap', rng.random((n, 2)).astype(np.float32))
        for name in ('varm', 'obsp', 'varp', 'layers', 'uns'):
            grp = f.create_group(name)
            _str_attr(grp, 'encoding-type', 'dict')
            _str_attr(grp, 'encoding-version', '0.1.0')


if __name__ == '__main__':
    for layout in ('csr', 'csc', 'dense'):
        write(os.path.join(HERE, f'fixture_{layout}.h5ad'), layout)
    print('written', file=sys.stderr)
