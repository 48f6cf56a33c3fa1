"""The drop-in's input domain beyond the 2000-gene workloads, on the GPU.

* More genes than one X^T u launch holds in LDS (g > 8192: gene tiles), more
  than X v stages in LDS (g > 32768: v read from global memory) and more
  than 16-bit packed indices hold (g > 65535: 32-bit indices), against the
  FP64 oracle through the cell Gram (``oracle.fp64_pca_cells``), and the
  host entry point with int64 indices giving the same bits.
* The reference's serial CSC path (``is_csr=False, num_threads=1,
  match_parallel=False``: the matvec operand is the CSC of X,
  ``brisc/single_cell.py:12054-12069``) through the C ABI: the same bits as
  the CSR call.
* The device transpose (``bpca_csr_transpose``) against scipy's
  ``tocsc``/``tocsr``, bit for bit, int32 and int64 inputs, empty rows and
  columns, and its refusal of out-of-range indices.

Tolerances: singular values within 1e-6 of the FP64 oracle, PCs with a > 1%
gap within 1e-4 up to sign; transposes and path equivalences bit-exact.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from brisc_amd import irlba, irlba_device, synth

import helpers

pytestmark = pytest.mark.gpu

SV_RTOL = 1e-6
COL_RTOL = 1e-4


def _torch():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    return torch


def _dev(*arrays):
    torch = _torch()
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrays]


@pytest.mark.parametrize('n,g,k', [(6000, 8193, 20), (4000, 20000, 20),
                                   (2500, 40000, 15), (2000, 70000, 15)])
def test_many_genes_against_fp64_oracle(n, g, k):
    w = synth.make_workload(900 + g % 97, n, num_types=6, num_genes=g, quick=True)
    data, indices, indptr = oracle.generate_csr(w)
    print(f'case genes n={n} g={g} k={k} nnz={len(data)}', flush=True)
    d, i, p = _dev(data, indices, indptr)
    P, sv, st = irlba_device(d, i, p, n, g, k)
    print('stats', st, flush=True)
    assert st['converged'] == 1
    out = oracle.fp64_pca_cells(data, indices, indptr, n, g, k)
    ref = out['singular_values']
    PCs = P.cpu().numpy()
    err = np.max(np.abs(sv - ref) / ref)
    err_col = helpers.relative_sv_error(PCs, ref).max()
    ws = helpers.well_separated(ref, 1e-2)
    col = helpers.column_errors_up_to_sign(PCs[:, ws], out['PCs'][:, ws]).max() if len(ws) else 0.0
    print(f'  sv rel err vs FP64 oracle {err:.2e} (PC norms {err_col:.2e}); '
          f'separated PCs ({len(ws)}) {col:.2e}', flush=True)
    assert err < SV_RTOL and err_col < SV_RTOL and col < COL_RTOL
    # host entry point, int64 indices: the same bits
    PCh = np.empty((n, k), np.float32)
    conv = irlba(data, indices.astype(np.int64), indptr.astype(np.int64), None, None, None,
                 n, g, k, min(100, n, g), 1e-6, 100, 0, False, True, 8, 1024, PCh)
    assert conv and np.array_equal(PCh.view(np.uint32), PCs.view(np.uint32))


@pytest.mark.parametrize('width', [4, 8])
def test_serial_csc_path_matches_csr_call(width):
    n, g, k = 20000, 2000, 20
    w = synth.make_workload(81, n, num_types=8)
    data, indices, indptr = oracle.generate_csr(w)
    X = sp.csr_array((data, indices, indptr), shape=(n, g))
    C = X.tocsc()
    C.sort_indices()
    t = np.int32 if width == 4 else np.int64
    print(f'case serial CSC path width={width}', flush=True)
    P1 = np.empty((n, k), np.float32)
    c1 = irlba(data, indices.astype(t), indptr.astype(t), None, None, None, n, g, k, 100,
               1e-6, 100, 0, False, True, 8, 1024, P1)
    P2 = np.empty((n, k), np.float32)
    c2 = irlba(C.data, C.indices.astype(t), C.indptr.astype(t), C.data, C.indices.astype(t),
               C.indptr.astype(t), n, g, k, 100, 1e-6, 100, 0, False, False, 1, 1024, P2)
    assert c1 and c2
    assert np.array_equal(P1.view(np.uint32), P2.view(np.uint32))
    out = oracle.fp64_pca(data, indices, indptr, n, g, k, want_pcs=False)
    assert helpers.relative_sv_error(P2, out['singular_values']).max() < SV_RTOL


@pytest.mark.parametrize('wide', [False, True])
def test_device_transpose_matches_scipy(wide):
    from brisc_amd.subset import csr_transpose
    rng = np.random.default_rng(5)
    X = sp.random(3001, 517, density=0.05, format='csr', random_state=rng, dtype=np.float32)
    X = X.tolil()
    X[100:140, :] = 0                 # empty rows
    X[:, 7] = 0                       # an empty column
    X = sp.csr_array(X)
    X.sort_indices()
    t = np.int64 if wide else np.int32
    print(f'case transpose wide={wide}', flush=True)
    d, i, p = _dev(X.data, X.indices.astype(t), X.indptr.astype(t))
    cd, ci, cp = csr_transpose(d, i, p, X.shape[1])
    C = X.tocsc()
    C.sort_indices()
    assert np.array_equal(cp.cpu().numpy(), C.indptr.astype(np.int64))
    assert np.array_equal(ci.cpu().numpy(), C.indices.astype(np.int32))
    assert np.array_equal(cd.cpu().numpy().view(np.uint32), C.data.view(np.uint32))
    # and back: the CSR again
    rd, ri, rp = csr_transpose(cd, ci, cp, X.shape[0])
    assert np.array_equal(rp.cpu().numpy(), X.indptr.astype(np.int64))
    assert np.array_equal(ri.cpu().numpy(), X.indices.astype(np.int32))
    assert np.array_equal(rd.cpu().numpy().view(np.uint32), X.data.view(np.uint32))
    bad = X.indices.astype(t).copy()
    bad[len(bad) // 2] = X.shape[1]
    with pytest.raises(ValueError, match='indices must lie'):
        csr_transpose(d, *_dev(bad), p, X.shape[1])
