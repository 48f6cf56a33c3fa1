"""Exact k nearest neighbours on the device against the FP64 oracle.

The device forms squared distances in FP32 as the reference does
(``||q||^2 + (||c||^2 - 2 q.c)``, ``brisc/knn.pyx:222-252``), so they carry
rounding error; the oracle's are exact.  Tolerances (written here):
  * each returned distance: |d_gpu - d_exact| <= DIST_RTOL * (||q||^2 + ||c||^2)
    + 1e-30, the error bound of the norm identity in FP32 for these sizes;
  * neighbour sets: every oracle neighbour whose exact distance is below the
    device's k-th exact distance by more than the per-row slack (2 x the
    distance bound) is returned, and every returned neighbour lies within the
    slack of the oracle's k-th distance -- equal sets except near-ties;
  * order: the exact distances along each returned row never fall by more
    than the slack;
  * structure (bit-exact): no self, no repeats, rows ascending by distance
    with ties in ascending index, indices in range;
  * run-to-run: bit-identical.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

DIST_RTOL = 2e-5


def _torch():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    return torch


def _knn(X, k):
    torch = _torch()
    from brisc_amd.knn import knn_exact
    Xd = torch.from_numpy(np.ascontiguousarray(X, np.float32)).cuda()
    nb, dist = knn_exact(Xd, k)
    torch.cuda.synchronize()
    return nb.cpu().numpy().astype(np.int64), dist.cpu().numpy()


def _clustered(seed, n, d, centers=12, spread=0.3):
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((centers, d)) * 3.0
    lab = rng.integers(0, centers, n)
    return (C[lab] + spread * rng.standard_normal((n, d))).astype(np.float32)


def check_knn(X, nb, dist, k):
    n = X.shape[0]
    X64 = X.astype(np.float64)
    norms = (X64 * X64).sum(1)
    assert nb.shape == (n, k) and dist.shape == (n, k)
    assert nb.min() >= 0 and nb.max() < n
    assert not (nb == np.arange(n)[:, None]).any(), 'a cell is its own neighbour'
    for row in nb[: min(n, 4096)]:
        assert np.unique(row).size == k, 'repeated neighbour'
    # Ascending by distance, ties by index (bit-exact structure).
    dd, ii = dist[:, 1:] - dist[:, :-1], nb[:, 1:] - nb[:, :-1]
    assert ((dd > 0) | ((dd == 0) & (ii > 0))).all()
    # Returned distances against the exact FP64 distance of the same pair.
    exact_pair = ((X64[:, None, :] - X64[nb]) ** 2).sum(2)
    bound = DIST_RTOL * (norms[:, None] + norms[nb]) + 1e-30
    err = np.abs(dist - exact_pair)
    assert (err <= bound).all(), float((err / bound).max())
    # Neighbour sets against the oracle, up to near-ties.
    oi, od = oracle.knn_fp64(X, k)
    slack = 2.0 * DIST_RTOL * (norms + norms.max())
    # The device order follows the exact order except within the slack.
    assert (exact_pair[:, 1:] >= exact_pair[:, :-1] - slack[:, None]).all()
    gpu_kth = exact_pair.max(1)
    assert (exact_pair <= od[:, -1:] + slack[:, None]).all()
    missing = od < (gpu_kth - slack)[:, None]
    for q in np.flatnonzero(missing.any(1)):
        need = set(oi[q][missing[q]].tolist())
        assert need <= set(nb[q].tolist()), q
    same_order = float((nb == oi).all(1).mean())
    same_set = float((np.sort(nb, 1) == np.sort(oi, 1)).all(1).mean())
    return same_order, same_set


@pytest.mark.parametrize('n,d,k', [
    (2, 1, 1), (65, 3, 1), (257, 50, 20), (1000, 50, 20), (3000, 50, 64),
    (4099, 256, 16), (5000, 10, 32), (700, 7, 33), (1000, 60, 5),
    (1500, 61, 64), (2000, 128, 20), (999, 129, 7), (20000, 50, 20),
    (3000, 47, 20), (2500, 49, 64), (3000, 200, 20), (1200, 161, 64)])
def test_exact_knn_matches_fp64_oracle(n, d, k):
    X = _clustered(n * 7 + d, n, d)
    nb, dist = _knn(X, k)
    same_order, same_set = check_knn(X, nb, dist, k)
    # Rows that differ from the oracle do so only inside the near-tie slack
    # (checked above); the fractions are reported.
    print(f'n={n} d={d} k={k}: rows identical to the oracle {same_order:.4f}, '
          f'same neighbour set {same_set:.4f}')


@pytest.mark.parametrize('n,d,k', [(3000, 30, 100), (2500, 50, 150), (1000, 7, 200), (700, 3, 65),
                                   (2000, 160, 100), (1500, 256, 65)])
def test_exact_knn_more_than_64_neighbours(n, d, k):
    # Passes of 64: the first 64 columns are the 64-neighbour search's,
    # bit for bit; the whole row against the FP64 oracle as above (with
    # lattice-like duplicates in the 3-dimensional case).
    X = _clustered(n * 3 + d, n, d)
    if d == 3:
        X = np.round(X * 2.0) / 2.0
    nb, dist = _knn(X, k)
    nb64, dist64 = _knn(X, 64)
    assert np.array_equal(nb[:, :64], nb64)
    assert np.array_equal(dist[:, :64].view(np.uint32), dist64.view(np.uint32))
    same_order, same_set = check_knn(X, nb, dist, k)
    print(f'n={n} d={d} k={k}: rows identical to the oracle {same_order:.4f}, '
          f'same neighbour set {same_set:.4f}')


def test_exact_knn_duplicates_and_ties():
    # Exact duplicates: distance 0 up to rounding of the norm identity; each
    # duplicate group member lists the other members first.
    rng = np.random.default_rng(5)
    base = rng.standard_normal((200, 16)).astype(np.float32)
    X = np.concatenate([base, base[:50], base[:50]])
    nb, dist = _knn(X, 5)
    for q in range(50):
        group = {q, 200 + q, 250 + q} - {q}
        assert group <= set(nb[q, :2].tolist())
        assert np.abs(dist[q, :2]).max() <= DIST_RTOL * 2 * (base[q] ** 2).sum()
    # Integer-valued points on a lattice: many exact ties, broken by index.
    L = np.stack(np.meshgrid(np.arange(8), np.arange(8), indexing='ij'),
                 -1).reshape(-1, 2).astype(np.float32)
    nb, dist = _knn(L, 4)
    assert (dist == 1.0).sum() > 0
    oi, od = oracle.knn_fp64(L, 4)
    assert np.array_equal(dist, od.astype(np.float32))
    assert np.array_equal(nb, oi)


def test_exact_knn_is_deterministic():
    X = _clustered(3, 6000, 50)
    a = _knn(X, 20)
    b = _knn(X, 20)
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))


def test_exact_knn_rejects_bad_arguments():
    torch = _torch()
    from brisc_amd.knn import knn_exact
    X = torch.zeros((10, 4), device='cuda')
    with pytest.raises(ValueError):
        knn_exact(X, 10)
    with pytest.raises(ValueError):
        knn_exact(X, 65)
    with pytest.raises(ValueError):
        knn_exact(torch.zeros((10, 257), device='cuda'), 3)
    with pytest.raises(ValueError):
        knn_exact(X[:1], 1)


def _ref_fixture(name):
    import hashlib
    import json
    import os
    import helpers
    here = os.path.join(os.path.dirname(__file__), 'golden')
    with open(os.path.join(here, 'ref_vectors.json')) as f:
        meta = {m['name']: m for m in json.load(f)['knn']}[name]
    X = helpers.knn_points(meta['seed'], meta['num_cells'], meta['num_dims'],
                           meta['centers'])
    assert hashlib.sha256(X.tobytes()).hexdigest() == meta['x_sha256']
    ref = np.load(os.path.join(here, 'ref_knn.npz'))
    rows = ref[f'{name}_rows'] if f'{name}_rows' in ref.files else \
        np.arange(len(X))
    return (meta, X, rows, ref[f'{name}_neighbors'].astype(np.int64),
            ref[f'{name}_distances'],
            ref[f'{name}_default_neighbors'].astype(np.int64))


@pytest.mark.parametrize('name', ['knn_a', 'knn_b', 'knn_c'])
def test_exact_knn_matches_reference_exhaustive_search(name):
    # The reference's knn_self with every cluster searched, run on the same
    # points (fixtures written by the reference itself).
    import helpers
    meta, X, rows, ref_nb, ref_dist, ref_default = _ref_fixture(name)
    k = meta['k']
    nb, dist = _knn(X, k)
    nb, dist = nb[rows], dist[rows]
    ok, same_rows, same_sets = helpers.knn_sets_agree(X, nb, ref_nb,
                                                      DIST_RTOL, rows)
    assert ok
    X64 = X.astype(np.float64)
    norms = (X64 * X64).sum(1)
    same_pair = nb == ref_nb
    bound = 2 * DIST_RTOL * (norms[rows][:, None] + norms[nb])
    assert (np.abs(dist - ref_dist)[same_pair] <= bound[same_pair]).all()
    # The reference's default search (64 nearest of its k-means clusters)
    # against the exact search: its recall.
    recall = np.mean([len(set(a) & set(b)) / k for a, b in
                      zip(ref_default.tolist(), nb.tolist())])
    print(f'{name}: rows identical to the reference {same_rows:.4f}, same '
          f'sets {same_sets:.4f}; reference default-search recall against '
          f'the device exact search {recall:.5f} (reference, against FP64: '
          f'{meta["recall_default_vs_fp64"]:.5f})')
    assert same_sets > 0.99


def test_single_cell_neighbors_with_qc():
    from brisc_amd import SingleCell
    meta, X, rows, ref_nb, _, _ = _ref_fixture('knn_a')
    n, k = len(X), meta['k']
    qc = np.random.default_rng(4).random(n) < 0.9
    sc = SingleCell(obs={'cell': np.arange(n), 'passed_QC': qc},
                    obsm={'pca': X})
    out = sc.neighbors(num_neighbors=k)
    nb, dist = out.obsm['neighbors'], out.obsm['distances']
    assert nb.dtype == np.uint32 and dist.dtype == np.float32
    assert nb.shape == (n, k) and dist.shape == (n, k)
    assert (nb[~qc] == np.uint32(4_294_967_295)).all()
    assert np.isnan(dist[~qc]).all()
    sub_nb, sub_dist = _knn(X[qc], k)
    assert np.array_equal(nb[qc].astype(np.int64), sub_nb)
    assert np.array_equal(dist[qc], sub_dist)
    # Without QC every cell is a query and a candidate: the reference's
    # exhaustive output up to near-ties.
    out = sc.neighbors(QC_column=None, num_neighbors=k)
    import helpers
    ok, _, _ = helpers.knn_sets_agree(X, out.obsm['neighbors'].astype(np.int64),
                                      ref_nb, DIST_RTOL)
    assert ok
    with pytest.raises(ValueError, match='already a key'):
        out.neighbors()
    with pytest.raises(ValueError, match='PC_key'):
        sc.neighbors(PC_key='missing')
