"""k-means|| initialisation (kmeans_barbar=True) on the device against the
oracle and the reference.

Tolerances (written here):
  * integer-lattice inputs (every distance and cost exact in FP32): the
    initial cells and the candidate count equal the oracle's exactly, and
    the centroids after one Lloyd iteration equal the reference's bit for
    bit (fixtures written by the reference; the oracle is pinned against
    them on the CPU in test_oracle.py);
  * k-means labels after that iteration (a last E-step against non-integer
    means): at least 99.9% equal to the reference's;
  * real-valued PCs: the initial cells equal the oracle's (the candidate
    updates' partial distances are formed on the matrix cores, the
    oracle's in sequence; on these inputs no draw lands within their
    rounding of its threshold);
  * the inverted-file search seeded by k-means||: recall against the exact
    search at least the random initialisation's minus 0.005;
  * the reference's two input errors, with its messages.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import helpers
import oracle

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(__file__), 'golden')


def _torch():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    return torch


def _dev(X):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(X, np.float32)).cuda()


def _fixtures():
    with open(os.path.join(HERE, 'ref_vectors.json')) as f:
        meta = json.load(f)
    return meta, np.load(os.path.join(HERE, 'ref_kmeans_barbar.npz'))


def _points(m):
    X = helpers.lattice_points(m['seed'], m['num_cells'], m['num_dims'],
                               m['centres'], *m['lattice'])
    assert hashlib.sha256(X.tobytes()).hexdigest() == m['X_sha256']
    return X


def test_lattice_init_matches_oracle_and_reference():
    from brisc_amd.knn import kmeans, kmeans_barbar_init
    meta, ref = _fixtures()
    for m in meta['kmeans_barbar']:
        name, C = m['name'], m['num_clusters']
        X = _points(m)
        Xd = _dev(X)
        for chunk in m['chunk_sizes']:
            print(f'case k-means|| {name} n={len(X)} d={X.shape[1]} C={C} '
                  f'chunk={chunk}', flush=True)
            args = (C, m['num_init_iterations'], m['oversampling_factor'],
                    chunk, m['kmeans_seed'])
            cells, S = kmeans_barbar_init(Xd, *args)
            want, S_want = oracle.kmeans_barbar_init(X, *args)
            assert S == S_want and np.array_equal(cells, want), name
            labels, cent, counts, it = kmeans(
                Xd, C, 1, chunk_size=chunk, seed=m['kmeans_seed'],
                kmeans_barbar=True,
                num_init_iterations=m['num_init_iterations'],
                oversampling_factor=m['oversampling_factor'])
            cent = cent.cpu().numpy()
            assert it == 1
            assert np.array_equal(cent.view(np.uint32),
                                  ref[f'{name}_centroids'].view(np.uint32)), name
            same = (labels.cpu().numpy() == ref[f'{name}_labels']).mean()
            print(f'  {S} candidates; labels equal to the reference: '
                  f'{same:.6f}', flush=True)
            assert same >= 0.999
            if same == 1.0:
                assert np.array_equal(counts.cpu().numpy(), ref[f'{name}_counts'])


def test_duplicate_points_and_errors_as_reference():
    from brisc_amd.knn import kmeans, kmeans_barbar_init
    meta, ref = _fixtures()
    X = ref['kb_dup_X'].astype(np.float32)
    Xd = _dev(X)
    for C, outcome in meta['kmeans_barbar_duplicates'].items():
        C = int(C)
        print(f'case k-means|| five distinct points C={C}', flush=True)
        if outcome == 'centroids':
            _, cent, _, _ = kmeans(Xd, C, 1, kmeans_barbar=True)
            assert np.array_equal(cent.cpu().numpy().view(np.uint32),
                                  ref[f'kb_dup_{C}_centroids'].view(np.uint32))
        else:
            with pytest.raises(ValueError) as e:
                kmeans_barbar_init(Xd, C)
            assert outcome in str(e.value)
    print('case k-means|| every cell at one point', flush=True)
    with pytest.raises(ValueError, match=meta['kmeans_barbar_same_point_error']):
        kmeans_barbar_init(_dev(np.ones((50, 4), np.float32)), 5)


@pytest.mark.parametrize('seed,n,d,centres', [(51, 20000, 30, 40),
                                              (52, 6000, 50, 25),
                                              (54, 5000, 200, 20)])
def test_real_valued_init_matches_oracle(seed, n, d, centres):
    from brisc_amd.knn import default_num_clusters, kmeans_barbar_init
    X = helpers.knn_points(seed, n, d, centres)
    C = default_num_clusters(n)
    print(f'case k-means|| real-valued n={n} d={d} C={C}', flush=True)
    cells, S = kmeans_barbar_init(_dev(X), C, seed=seed)
    want, S_want = oracle.kmeans_barbar_init(X, C, seed=seed)
    print(f'  {S} candidates (oracle {S_want}); first difference at '
          f'{int(np.argmax(cells != want)) if not np.array_equal(cells, want) else None}',
          flush=True)
    assert S == S_want and np.array_equal(cells, want)


def test_ivf_search_with_kmeans_barbar():
    from brisc_amd.knn import knn_exact, knn_ivf
    with open(os.path.join(HERE, 'ref_vectors.json')) as f:
        m = {m['name']: m for m in json.load(f)['knn']}['knn_c']
    X = helpers.knn_points(m['seed'], m['num_cells'], m['num_dims'],
                           m['centers'])
    k = m['k']
    Xd = _dev(X)
    print(f'case inverted-file search with k-means|| n={len(X)} k={k}',
          flush=True)
    ex, _ = knn_exact(Xd, k)
    ex = ex.cpu().numpy()
    nb_r, _, info_r = knn_ivf(Xd, k)
    nb_b, d_b, info_b = knn_ivf(Xd, k, kmeans_barbar=True)
    nb_b2, d_b2, _ = knn_ivf(Xd, k, kmeans_barbar=True)
    torch = _torch()
    assert torch.equal(nb_b, nb_b2) and torch.equal(d_b.view(torch.int32),
                                                    d_b2.view(torch.int32))

    def recall(nb):
        nb = nb.cpu().numpy()
        return float(np.mean([len(set(a) & set(b)) / k for a, b in
                              zip(nb.tolist(), ex.tolist())]))
    r_rand, r_bar = recall(nb_r), recall(nb_b)
    print(f'  recall: random init {r_rand:.5f}, k-means|| {r_bar:.5f}; '
          f'{info_b["kmeans_candidates"]} candidates', flush=True)
    assert info_r['kmeans_candidates'] == 0
    assert info_b['kmeans_candidates'] >= info_b['num_clusters']
    assert r_bar >= r_rand - 0.005


def test_single_cell_neighbors_with_kmeans_barbar():
    from brisc_amd import SingleCell
    from brisc_amd.knn import knn_ivf
    X = helpers.knn_points(53, 8000, 20, 16)
    print('case SingleCell.neighbors(kmeans_barbar=True) n=8000', flush=True)
    sc = SingleCell(obs={'cell': np.arange(len(X)),
                         'passed_QC': np.ones(len(X), bool)},
                    obsm={'pca': X})
    out = sc.neighbors(kmeans_barbar=True, num_init_iterations=3,
                       oversampling_factor=2.0, num_neighbors=10)
    nb, _, _ = knn_ivf(_dev(X), 10, kmeans_barbar=True, num_init_iterations=3,
                       oversampling_factor=2.0)
    assert np.array_equal(out.obsm['neighbors'],
                          nb.cpu().numpy().view(np.uint32))
