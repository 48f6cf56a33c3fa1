"""An .h5ad file through the device path: X streamed to the device through
pinned buffers equals X read on the host; normalize() -> pca() on the loaded
dataset, the PCs written back into the file as obsm and read again."""
import os
import shutil

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from brisc_amd import SingleCell, h5ad

import helpers

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


@pytest.mark.parametrize('layout', ['csr', 'csc'])
def test_stream_x_to_device_matches_host_read(layout):
    import torch
    assert torch.cuda.is_available()
    path = os.path.join(GOLDEN, f'fixture_{layout}.h5ad')
    print(f'case stream {layout}', flush=True)
    with h5ad.H5File(path) as f:
        X = h5ad.read_x(f)
    # small slices: many rounds through the two pinned buffers
    got_layout, shape, d, i, p = h5ad.stream_x_to_device(path, chunk_bytes=1000)
    assert got_layout == layout and shape == X.shape
    assert np.array_equal(d.cpu().numpy(), X.data)
    assert np.array_equal(i.cpu().numpy(), X.indices)
    assert np.array_equal(p.cpu().numpy(), X.indptr)


def test_file_to_pcs_and_back(tmp_path):
    dst = str(tmp_path / 'pcs.h5ad')
    shutil.copy(os.path.join(GOLDEN, 'fixture_csr.h5ad'), dst)
    print('case h5ad -> normalize -> pca -> obsm', flush=True)
    sc = SingleCell(dst).skip_qc()
    norm = sc.normalize()                           # QC column passed_QC
    out = norm.pca(num_PCs=10, subspace_size=30)    # QC x HVG (highly_variable)
    P = out.obsm['pca']
    qc = sc.obs['passed_QC']
    assert np.isnan(P[~qc]).all() and np.isfinite(P[qc]).all()
    h5ad.write_obsm(dst, 'pca', P)
    back = SingleCell(dst)
    assert np.array_equal(np.nan_to_num(back.obsm['pca']).view(np.uint32),
                          np.nan_to_num(P).view(np.uint32))
    # the PCs against the FP64 oracle on the same QC x HVG rows
    hvg = sc.var['highly_variable']
    Y = sp.csr_array(norm.X[np.flatnonzero(qc)][:, np.flatnonzero(hvg)])
    Y.sort_indices()
    ref = oracle.fp64_pca(Y.data, Y.indices, Y.indptr.astype(np.int64), Y.shape[0],
                          Y.shape[1], 10, want_pcs=False)['singular_values']
    err = helpers.relative_sv_error(P[qc], ref).max()
    print(f'  sv rel err vs FP64 oracle {err:.2e}', flush=True)
    assert err < 1e-6
