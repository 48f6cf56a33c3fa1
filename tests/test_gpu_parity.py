"""Parity of the HIP path with the reference's outputs and the FP64 oracle.

Tolerances (written here, applied throughout):
  * singular values (column norms of the PCs): relative error < 1e-4 against
    the reference's own irlba output and against the FP64 oracle;
  * individual PCs whose singular value is separated by > 1% from its
    neighbours: relative error < 1e-2 up to sign;
  * every PC inside a near-degenerate cluster (window 1e-3): projection
    deficit onto the oracle's cluster < 1e-3;
  * integer work (generator counts, CSR structure): bit-exact;
  * run-to-run: bit-identical.
"""
import numpy as np
import pytest

import oracle
from brisc_amd import SingleCell, irlba, irlba_device, synth

import helpers

pytestmark = pytest.mark.gpu

SV_RTOL = 1e-4
COL_RTOL = 1e-2
CLUSTER_TOL = 1e-3
TINY = ['tiny_a', 'tiny_b', 'tiny_c', 'tiny_d']


def _torch():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    return torch


def to_device(data, indices, indptr, wide=False):
    torch = _torch()
    return (torch.from_numpy(np.ascontiguousarray(data, np.float32)).cuda(),
            torch.from_numpy(np.ascontiguousarray(indices, np.int32)).cuda(),
            torch.from_numpy(np.ascontiguousarray(
                indptr, np.int64 if wide else np.int32)).cuda())


def run_host(data, indices, indptr, n, g, k, m=100, width=4, **kw):
    idx_t = np.int32 if width == 4 else np.int64
    PCs = np.empty((n, k), np.float32)
    conv = irlba(data_matvec=data, indices_matvec=indices.astype(idx_t),
                 indptr_matvec=indptr.astype(idx_t), data_rmatvec=None,
                 indices_rmatvec=None, indptr_rmatvec=None, num_cells=n,
                 num_genes=g, k=k, subspace_size=min(m, n, g),
                 tolerance=kw.get('tolerance', 1e-6),
                 max_iterations=kw.get('max_iterations', 100),
                 seed=kw.get('seed', 0), match_parallel=False, is_csr=True,
                 num_threads=8, chunk_size=kw.get('chunk_size', 1024), PCs=PCs)
    return PCs, conv


def check_against(PCs, ref_sv, ref_pcs=None, oracle_out=None):
    assert np.isfinite(PCs).all()
    assert helpers.relative_sv_error(PCs, ref_sv).max() < SV_RTOL
    if ref_pcs is not None:
        ws = helpers.well_separated(ref_sv, 1e-2)
        if len(ws):
            assert helpers.column_errors_up_to_sign(PCs[:, ws], ref_pcs[:, ws]).max() < COL_RTOL
    if oracle_out is not None:
        sv = oracle_out['singular_values']
        assert helpers.relative_sv_error(PCs, sv).max() < SV_RTOL
        assert helpers.cluster_projection_deficit(PCs, oracle_out['PCs'], sv, 1e-3).max() < CLUSTER_TOL


@pytest.mark.parametrize('name', TINY)
def test_tiny_matches_reference_and_oracle(golden, name):
    meta, tiny, _ = golden
    case = next(c for c in meta['tiny'] if c['name'] == name)
    data, indices, indptr, ref = (tiny[f'{name}_{s}'] for s in ('data', 'indices', 'indptr', 'pcs'))
    n, g, k, m = case['num_cells'], case['num_genes'], case['k'], case['subspace_size']
    print(f'case {name} n={n} g={g} k={k} m={m}', flush=True)
    PCs, conv = run_host(data, indices, indptr, n, g, k, m,
                         chunk_size=case['chunk_size'])
    assert conv
    out = oracle.fp64_pca(data, indices, indptr, n, g, k)
    check_against(PCs, helpers.singular_values(ref), ref, out)


def test_cells50k_matches_reference_and_oracle(golden, cells50k):
    meta, _, large = golden
    w, data, indices, indptr = cells50k
    k = meta['large']['k']
    print(f'case cells50k nnz={len(data)}', flush=True)
    P, sv, st = irlba_device(*to_device(data, indices, indptr), w.num_cells,
                             w.num_genes, k)
    PCs = P.cpu().numpy()
    print('stats', st, flush=True)
    assert st['converged'] == 1
    ref_sv = large['singular_values']
    assert np.max(np.abs(sv - ref_sv) / ref_sv) < SV_RTOL
    check_against(PCs, ref_sv)
    rows = large['rows']
    ws = helpers.well_separated(ref_sv, 1e-2)
    assert helpers.column_errors_up_to_sign(PCs[rows][:, ws], large['pcs_rows'][:, ws]).max() < COL_RTOL
    out = oracle.fp64_pca(data, indices, indptr, w.num_cells, w.num_genes, k)
    check_against(PCs, out['singular_values'], oracle_out=out)


def test_bitwise_reproducible_and_index_width_invariant(golden):
    meta, tiny, _ = golden
    w = synth.make_workload(77, 30_000, num_types=12)
    data, indices, indptr = oracle.generate_csr(w)
    print('case repro n=30000', flush=True)
    a = irlba_device(*to_device(data, indices, indptr), w.num_cells, w.num_genes, 30)
    b = irlba_device(*to_device(data, indices, indptr), w.num_cells, w.num_genes, 30)
    c = irlba_device(*to_device(data, indices, indptr, wide=True), w.num_cells, w.num_genes, 30)
    pa, pb, pc = (x[0].cpu().numpy() for x in (a, b, c))
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32))
    assert np.array_equal(pa.view(np.uint32), pc.view(np.uint32))
    assert a[2]['restarts'] == b[2]['restarts'] == c[2]['restarts']
    # host entry point with int64 indices gives the same bits
    ph, conv = run_host(data, indices, indptr, w.num_cells, w.num_genes, 30, width=8)
    assert conv and np.array_equal(pa.view(np.uint32), ph.view(np.uint32))


def test_device_generator_matches_cpu_generator():
    torch = _torch()
    w = synth.make_workload(31, 40_000, num_types=9)
    print('case generator n=40000', flush=True)
    d, i, p, cnt = synth.generate_device(w, keep_counts=True, row_chunk=15_000)
    cp, ci, cc = oracle.generate_counts(w)
    assert np.array_equal(p.cpu().numpy(), cp)
    assert np.array_equal(i.cpu().numpy(), ci)
    assert np.array_equal(cnt.cpu().numpy(), cc)
    ref = oracle.log1ppf(cc, cp)
    dev = d.cpu().numpy()
    ulps = np.abs(dev.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64))
    assert ulps.max() <= 2, ulps.max()
    torch.cuda.synchronize()


def test_zero_variance_gene_and_empty_rows(golden):
    meta, tiny, _ = golden
    name = 'tiny_b'
    data, indices, indptr = (tiny[f'{name}_{s}'] for s in ('data', 'indices', 'indptr'))
    import scipy.sparse as sp
    X = sp.csr_array((data, indices, indptr), shape=(1000, 120)).tolil()
    X[:, 3] = 0                 # a gene never expressed: sd clipped to 1e-8
    X[10:20, :] = 0             # empty rows
    X = sp.csr_array(X)
    X.sort_indices()
    print('case zero-variance', flush=True)
    PCs, conv = run_host(X.data, X.indices, X.indptr, 1000, 120, 10, 25)
    out = oracle.fp64_pca(X.data, X.indices, X.indptr.astype(np.int64), 1000, 120, 10)
    assert conv
    check_against(PCs, out['singular_values'], oracle_out=out)


def test_single_pc_and_clipped_subspace(golden):
    meta, tiny, _ = golden
    data, indices, indptr = (tiny[f'tiny_a_{s}'] for s in ('data', 'indices', 'indptr'))
    print('case k=1 / clipped subspace', flush=True)
    PCs, conv = run_host(data, indices, indptr, 300, 40, 1, 100)   # m -> 40
    out = oracle.fp64_pca(data, indices, indptr, 300, 40, 1)
    assert conv
    check_against(PCs, out['singular_values'], oracle_out=out)
    # fewer cells than the subspace size: m is clipped to the cell count
    sub = slice(0, 31)
    ip = indptr[:32] - indptr[0]
    d, ix = data[indptr[0]:indptr[31]], indices[indptr[0]:indptr[31]]
    PCs, conv = run_host(d, ix, ip, 31, 40, 5, 100)
    out = oracle.fp64_pca(d, ix, ip.astype(np.int64), 31, 40, 5)
    assert conv
    check_against(PCs, out['singular_values'], oracle_out=out)


def test_nonconverged_exit_returns_ritz_vectors(cells50k):
    """max_iterations=1: converged is False and the PCs are the Ritz
    approximation (the reference's exit applies Ub twice; this path does
    not), so each Ritz value is a lower bound of the true singular value
    and the Rayleigh quotient of each PC reproduces it."""
    w, data, indices, indptr = cells50k
    print('case max_iterations=1', flush=True)
    PCs, conv = run_host(data, indices, indptr, w.num_cells, w.num_genes, 10,
                         max_iterations=1)
    assert not conv
    out = oracle.fp64_pca(data, indices, indptr, w.num_cells, w.num_genes, 10,
                          want_pcs=False)
    sv = helpers.singular_values(PCs)
    assert np.all(sv <= out['singular_values'] * (1 + 1e-6))
    assert np.all(sv >= 0.5 * out['singular_values'])
    # U is orthonormal: columns of PCs are mutually orthogonal
    G = (PCs.astype(np.float64).T @ PCs.astype(np.float64)) / np.outer(sv, sv)
    assert np.abs(G - np.eye(10)).max() < 1e-4


def test_noncanonical_rows_are_rejected(golden):
    meta, tiny, _ = golden
    data, indices, indptr = (tiny[f'tiny_a_{s}'].copy() for s in ('data', 'indices', 'indptr'))
    s, e = indptr[5], indptr[6]
    indices[s:e] = indices[s:e][::-1].copy()     # unsorted row
    print('case noncanonical', flush=True)
    with pytest.raises(ValueError, match='not canonical'):
        run_host(data, indices, indptr, 300, 40, 5, 12)


def test_single_cell_qc_backfill_matches_subset(golden):
    meta, tiny, _ = golden
    import scipy.sparse as sp
    data, indices, indptr = (tiny[f'tiny_d_{s}'] for s in ('data', 'indices', 'indptr'))
    X = sp.csr_array((data, indices, indptr), shape=(2500, 200))
    rng = np.random.default_rng(0)
    qc = rng.random(2500) < 0.8
    hvg = np.ones(200, bool)
    hvg[::7] = False
    print('case SingleCell QC back-fill', flush=True)
    sc = SingleCell(X, obs={'passed_QC': qc}, var={'gene': np.arange(200).astype(str), 'highly_variable': hvg},
                    uns={'normalized': True})
    full = sc.pca(num_PCs=12, subspace_size=30)
    P = full.obsm['pca']
    assert P.dtype == np.float32 and P.flags.c_contiguous and P.shape == (2500, 12)
    assert np.isnan(P[~qc]).all() and np.isfinite(P[qc]).all()
    sub = SingleCell(sp.csr_array(X[np.flatnonzero(qc)]), var={'gene': np.arange(200).astype(str), 'highly_variable': hvg},
                     uns={'normalized': True})
    Ps = sub.pca(QC_column=None, num_PCs=12, subspace_size=30).obsm['pca']
    assert np.array_equal(P[qc].view(np.uint32), Ps.view(np.uint32))
    # CSC input gives the same bits (the CSR of X is what the kernels read)
    sc_csc = SingleCell(sp.csc_array(X), obs={'passed_QC': qc},
                        var={'gene': np.arange(200).astype(str), 'highly_variable': hvg},
                        uns={'normalized': True})
    Pc = sc_csc.pca(num_PCs=12, subspace_size=30).obsm['pca']
    assert np.array_equal(np.nan_to_num(P).view(np.uint32), np.nan_to_num(Pc).view(np.uint32))
    # joint PCA over two datasets splits the rows back per dataset
    a, b = sc.pca(sub, num_PCs=5, subspace_size=20, QC_column=[qc, None])
    assert a.obsm['pca'].shape == (2500, 5) and b.obsm['pca'].shape == (int(qc.sum()), 5)


@pytest.mark.slow
def test_one_million_cells_against_fp64_oracle():
    """1M cells x 2000 genes (~140M nnz), generated on the device: singular
    values within 1e-4 of the FP64 Gram oracle and run-to-run identical."""
    w = synth.make_workload(**synth.WORKLOADS['cells1m'])
    print('case cells1m', flush=True)
    d, i, p = synth.generate_device(w)
    P1, sv1, st1 = irlba_device(d, i, p, w.num_cells, w.num_genes, 50)
    print('stats', st1, flush=True)
    assert st1['converged'] == 1
    P2, sv2, st2 = irlba_device(d, i, p, w.num_cells, w.num_genes, 50)
    assert np.array_equal(P1.cpu().numpy().view(np.uint32), P2.cpu().numpy().view(np.uint32))
    data, indices, indptr = d.cpu().numpy(), i.cpu().numpy(), p.cpu().numpy()
    out = oracle.fp64_pca(data, indices, indptr, w.num_cells, w.num_genes, 50,
                          want_pcs=False)
    assert np.max(np.abs(sv1 - out['singular_values']) / out['singular_values']) < SV_RTOL
    PCs = P1.cpu().numpy()
    assert helpers.relative_sv_error(PCs, out['singular_values']).max() < SV_RTOL
