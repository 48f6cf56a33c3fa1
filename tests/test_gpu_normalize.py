"""SingleCell.normalize() on the device against the reference's own
normalize_csr / normalize_csc (brisc/normalize.pyx:9-460).

Fixtures (tests/golden/ref_normalize.npz and ref_vectors.json['normalize'],
written by make_ref_vectors.py --normalize from the reference compiled from
its sources): SHA-256 of the normalised values and of the row sums for every
method (logCP10k, log1pPF, PFlog1pPF), count type (int32, int64, uint32,
float32, fractional float32 and float64), QC mask or none, CSR and CSC; the
full PFlog1pPF outputs.  Bar: bit-exact.  The device log1pf is checked bit
for bit against the host C library's on a dense sweep of floats.
"""
import hashlib
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from brisc_amd import SingleCell, synth
from brisc_amd.normalize import log1pf_device, normalize_device

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def _torch():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    return torch


def _inputs():
    """The fixture's inputs, regenerated from the stored seeds (as
    make_ref_vectors.normalize_inputs)."""
    with open(os.path.join(HERE, 'ref_vectors.json')) as f:
        meta = json.load(f)['normalize']
    w = synth.make_workload(meta['seed'], meta['num_cells'], num_types=meta['num_types'],
                            num_genes=meta['num_genes'], target_density=meta['density'])
    indptr, indices, counts = oracle.generate_counts(w)
    qc = np.random.Generator(np.random.PCG64(meta['qc_seed'])).random(meta['num_cells']) \
        < meta['qc_fraction']
    data = {'int32': counts.astype(np.int32), 'int64': counts.astype(np.int64),
            'uint32': counts.astype(np.uint32), 'float32': counts.astype(np.float32),
            'float32_frac': (counts * 0.75).astype(np.float32),
            'float64_frac': counts.astype(np.float64) * 0.75}
    return meta, indptr.astype(np.int64), indices.astype(np.int32), data, qc


def test_device_log1pf_matches_host_library():
    torch = _torch()
    # Every 61st float in (0, +inf) and in [-1, 0): 52M arguments.
    pos = np.arange(1, 0x7f800000, 61, dtype=np.uint32).view(np.float32)
    neg = (np.arange(0x80000000, 0xbf800000, 61, dtype=np.uint64).astype(np.uint32)).view(np.float32)
    x = np.concatenate([pos, neg, np.array([0.0, -0.0, 1e-30, 3.4e38], np.float32)])
    print(f'case log1pf sweep {x.size}', flush=True)
    y = log1pf_device(torch.from_numpy(x).cuda()).cpu().numpy()
    ref = oracle.log1pf(x)
    same = (y.view(np.uint32) == ref.view(np.uint32)) | (np.isnan(y) & np.isnan(ref))
    assert same.all(), (x[~same][:5], y[~same][:5], ref[~same][:5])


@pytest.mark.parametrize('layout', ['csr', 'csc'])
def test_normalize_matches_reference_bit_for_bit(layout):
    torch = _torch()
    meta, indptr, indices, data, qc = _inputs()
    n, g = meta['num_cells'], meta['num_genes']
    full = np.load(os.path.join(HERE, 'ref_normalize.npz'))
    checked = 0
    for case in meta['cases']:
        X = sp.csr_array((data[case['dtype']], indices, indptr), shape=(n, g))
        M = X if layout == 'csr' else X.tocsc()
        if layout == 'csc':
            M.sort_indices()
        d, i, p = (torch.from_numpy(np.ascontiguousarray(a)).cuda()
                   for a in (M.data, M.indices, M.indptr))
        out, rs = normalize_device(d, i, p, n, g, is_csr=layout == 'csr',
                                   QC=qc if case['qc'] else None, method=case['method'])
        vals = out.cpu().numpy()
        sums = rs.cpu().numpy().view(np.uint64)
        want = case[layout]
        tag = f"{case['dtype']} {case['method']} qc={case['qc']} {layout}"
        assert hashlib.sha256(sums.tobytes()).hexdigest() == want['row_sums_sha256'], tag
        ok = hashlib.sha256(vals.tobytes()).hexdigest() == want['values_sha256']
        key = f"pf_{case['dtype']}_qc{int(case['qc'])}"
        if not ok and case['method'] == 'PFlog1pPF' and layout == 'csr':
            d_ulp = np.abs(vals.view(np.int32).astype(np.int64) - full[key].view(np.int32).astype(np.int64))
            print(tag, 'max ulp', d_ulp.max(), 'entries differing', int((d_ulp > 0).sum()), flush=True)
        assert ok, tag
        checked += 1
    print(f'{checked} cases bit-exact ({layout})', flush=True)


def test_normalize_in_place_and_single_cell():
    """SingleCell.normalize(): the reference's checks, CSR and CSC kept,
    inplace on float32 counts, and the device values equal to the oracle's
    log1pPF (the generator's normalisation) on the 50k-cell workload."""
    w = synth.make_workload(**synth.WORKLOADS['cells50k'])
    indptr, indices, counts = oracle.generate_counts(w)
    X = sp.csr_array((counts, indices, indptr), shape=(w.num_cells, w.num_genes))
    sc = SingleCell(X)
    print('case SingleCell.normalize', flush=True)
    with pytest.raises(ValueError, match="uns\\['QCed'\\] is False"):
        sc.normalize()
    sc = sc.skip_qc()
    with pytest.raises(ValueError, match='method must be one of'):
        sc.normalize(method='CP10k')
    with pytest.raises(TypeError, match='inplace normalization requires'):
        sc.normalize(inplace=True)
    norm = sc.normalize(QC_column=None)
    assert norm.uns['normalized'] and isinstance(norm.X, sp.csr_array)
    ref = oracle.log1ppf(counts, indptr)
    assert np.array_equal(norm.X.data.view(np.uint32), ref.view(np.uint32))
    with pytest.raises(ValueError, match='already run normalize'):
        norm.normalize()
    # CSC stays CSC with the same values per entry
    C = X.tocsc()
    C.sort_indices()
    nc = SingleCell(C).skip_qc().normalize(QC_column=None)
    assert isinstance(nc.X, sp.csc_array)
    assert np.array_equal(sp.csr_array(nc.X).data.view(np.uint32), ref.view(np.uint32))
    # in place on float32 counts
    F = sp.csr_array((counts.astype(np.float32), indices, indptr), shape=X.shape)
    sf = SingleCell(F).skip_qc()
    same = sf.normalize(QC_column=None, inplace=True)
    assert same is sf and sf.uns['normalized']
    assert np.array_equal(F.data.view(np.uint32), ref.view(np.uint32))
    # QC column: all-False is refused
    with pytest.raises(ValueError, match='no cells pass QC'):
        SingleCell(X, obs={'passed_QC': np.zeros(w.num_cells, bool)}).skip_qc().normalize()
