""".h5ad input and output (libbrisc_h5ad.so): fixtures written by the
image's h5py in the anndata layout brisc reads (tests/golden/
make_h5ad_fixtures.py) read back exactly; files written here read back and
read by h5py; obsm written back into an existing file."""
import os
import re
import shutil
import subprocess
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from brisc_amd import SingleCell, h5ad

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, 'tests', 'golden')
sys.path.insert(0, GOLDEN)
from make_h5ad_fixtures import h5ad_fixture_counts  # noqa: E402

H5PY = '/opt/conda/bin/python3.9'


def _expected():
    indptr, indices, counts = h5ad_fixture_counts()
    return sp.csr_array((counts, indices, indptr), shape=(300, 40))


def test_library_exports_every_declared_symbol():
    text = open(os.path.join(ROOT, 'include', 'bpca_h5ad.h')).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    names = sorted(set(re.findall(r'\b(bh5_[a-z0-9_]+)\s*\(', text)))
    L = h5ad.lib()
    assert 'bh5_read_rows' in names
    assert [n for n in names if not hasattr(L, n)] == []


@pytest.mark.parametrize('layout', ['csr', 'csc', 'dense'])
def test_read_fixture(layout):
    want = _expected()
    sc = SingleCell(os.path.join(GOLDEN, f'fixture_{layout}.h5ad'))
    X = sc.X
    assert isinstance(X, sp.csc_array if layout == 'csc' else sp.csr_array)
    assert X.shape == (300, 40)
    Y = sp.csr_array(X)
    Y.sort_indices()
    assert np.array_equal(Y.indptr, want.indptr) and np.array_equal(Y.indices, want.indices)
    assert np.array_equal(Y.data.astype(np.int64), want.data.astype(np.int64))
    obs, var = sc.obs, sc.var
    assert list(obs['_index'][:3]) == ['cell0', 'cell1', 'cell2']
    assert obs['passed_QC'].dtype == np.bool_ and 200 < obs['passed_QC'].sum() < 300
    assert np.array_equal(obs['n_counts'], want.sum(axis=1))
    assert set(obs['cell_type']) == {'B', 'T', 'NK'}
    assert obs['score'].dtype == np.float64 and len(obs['score']) == 300
    assert list(sc.var_names[:2]) == ['gene0', 'gene1']
    assert var['highly_variable'].sum() == 25 and var['symbol'][39] == 'G039'
    assert sc.obsm['X_umap'].shape == (300, 2) and sc.obsm['X_umap'].dtype == np.float32


def test_read_slices_and_errors():
    path = os.path.join(GOLDEN, 'fixture_csr.h5ad')
    want = _expected()
    with h5ad.H5File(path) as f:
        assert h5ad.x_layout(f, 'X') == ('csr', (300, 40))
        part = f.read('X/indices', 100, 50)
        assert np.array_equal(part, want.indices[100:150])
        with pytest.raises(OSError, match='out of range'):
            f.read('X/data', len(want.data) - 1, 5)
        with pytest.raises(ValueError, match='not present'):
            h5ad.read_x(f, 'layers/none')
    with pytest.raises(OSError, match='cannot open'):
        h5ad.H5File(os.path.join(GOLDEN, 'no_such.h5ad'))


def test_write_roundtrip_and_obsm_write_back(tmp_path):
    src = os.path.join(GOLDEN, 'fixture_csr.h5ad')
    sc = SingleCell(src)
    sc.obsm['pca'] = np.arange(600, dtype=np.float32).reshape(300, 2)
    out = str(tmp_path / 'out.h5ad')
    sc.save(out)
    with pytest.raises(FileExistsError):
        sc.save(out)
    back = SingleCell(out)
    assert np.array_equal(back.X.indptr, sc.X.indptr)
    assert np.array_equal(back.X.data, sc.X.data)
    assert np.array_equal(back.obs['passed_QC'], sc.obs['passed_QC'])
    assert list(back.var_names) == list(sc.var_names)
    assert np.array_equal(back.obsm['pca'], sc.obsm['pca'])
    # write back into a copy of the h5py-written file
    dst = str(tmp_path / 'copy.h5ad')
    shutil.copy(src, dst)
    pcs = np.random.default_rng(0).standard_normal((300, 5)).astype(np.float32)
    h5ad.write_obsm(dst, 'X_pca', pcs)
    h5ad.write_obsm(dst, 'X_pca', pcs * 2)          # replaced, not duplicated
    again = SingleCell(dst)
    assert np.array_equal(again.obsm['X_pca'], pcs * 2)
    assert np.array_equal(again.obsm['X_umap'], sc.obsm['X_umap'])
    if os.path.exists(H5PY):
        # the written files open in h5py with the anndata attributes
        code = ('import h5py, sys, numpy as np\n'
                'f = h5py.File(sys.argv[1], "r")\n'
                'assert f["X"].attrs["encoding-type"] == "csr_matrix"\n'
                'assert tuple(f["X"].attrs["shape"]) == (300, 40)\n'
                'assert f["obs"].attrs["encoding-type"] == "dataframe"\n'
                'assert f["obs/_index"][0] in (b"cell0", "cell0")\n'
                'print(f["obsm"][sys.argv[2]].shape, f["obsm"][sys.argv[2]].dtype)\n')
        for path, key in ((out, 'pca'), (dst, 'X_pca')):
            r = subprocess.run([H5PY, '-c', code, path, key], capture_output=True, text=True,
                               env={'PATH': '/usr/bin:/bin'})
            assert r.returncode == 0, r.stderr
            assert '(300,' in r.stdout and 'float32' in r.stdout
