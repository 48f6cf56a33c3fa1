"""The oracle, pinned against the reference's own outputs and the mathematics.

Every function of ``oracle/`` is checked by something other than its own
formula: the reference's golden vectors (written by the compiled reference
``irlba`` / ``normalize_csr``, ``tests/golden``), a dense FP64 SVD on tiny
inputs, closed forms and exact integer identities.
"""
import os

import numpy as np
import pytest

import oracle
from brisc_amd import synth

import helpers

SV_RTOL = 1e-4          # singular values, relative
COL_RTOL = 1e-2         # individual PCs with a >=1% singular-value gap
CLUSTER_TOL = 1e-3      # projection deficit onto near-degenerate clusters


def _tiny_case(golden, name):
    meta, tiny, _ = golden
    case = next(c for c in meta['tiny'] if c['name'] == name)
    return case, (tiny[f'{name}_data'], tiny[f'{name}_indices'],
                  tiny[f'{name}_indptr'], tiny[f'{name}_pcs'])


TINY = ['tiny_a', 'tiny_b', 'tiny_c', 'tiny_d']


# ----------------------------------------------------------- generator

def test_generator_matches_reference_normalisation_bytes(golden, cells50k):
    """CPU generator + log1pPF restatement reproduce, byte for byte, the
    arrays the reference's normalize_csr produced for the 50k workload."""
    meta, _, _ = golden
    w, data, indices, indptr = cells50k
    assert w.scale == pytest.approx(meta['large']['scale'], rel=0, abs=0)
    assert synth.csr_sha256(data, indices, indptr) == meta['large']['csr_sha256']
    assert len(data) == meta['large']['nnz']


@pytest.mark.parametrize('name', TINY)
def test_tiny_inputs_regenerate_exactly(golden, name):
    case, (data, indices, indptr, _) = _tiny_case(golden, name)
    w = synth.make_workload(case['seed'], case['num_cells'],
                            num_types=case['num_types'],
                            num_genes=case['num_genes'],
                            target_density=case['density'])
    d2, i2, p2 = oracle.generate_csr(w)
    assert synth.csr_sha256(d2, i2, p2) == case['csr_sha256']
    assert np.array_equal(d2.view(np.uint32), data.view(np.uint32))


def test_generator_row_ranges_are_independent():
    """Counts of a row range do not depend on where generation starts."""
    w = synth.make_workload(5, 3000, num_types=7)
    full = oracle.generate_counts(w)
    part = oracle.generate_counts(w, row0=1234, num_rows=777)
    s, e = full[0][1234], full[0][1234 + 777]
    assert np.array_equal(full[1][s:e], part[1])
    assert np.array_equal(full[2][s:e], part[2])
    assert np.array_equal(full[0][1234:1234 + 778] - s, part[0])


def test_generator_statistics():
    """Density near the calibrated 7%, sorted unique columns, positive counts,
    and the NB mean/variance relation on a high-mean gene."""
    w = synth.make_workload(21, 20000)
    indptr, indices, counts = oracle.generate_counts(w)
    density = len(counts) / (w.num_cells * w.num_genes)
    assert abs(density - w.expected_density) < 0.003
    assert (counts > 0).all()
    rows = np.repeat(np.arange(w.num_cells), np.diff(indptr))
    key = rows.astype(np.int64) * w.num_genes + indices
    assert (np.diff(key) > 0).all()
    # mean of one gene within one type vs its NB mean, over cells
    gene = int(np.argmax(w.type_gene_mean[0]))
    sel = (w.cell_type == 0)
    dense = np.zeros(w.num_cells)
    hit = indices == gene
    dense[rows[hit]] = counts[hit]
    m = w.type_gene_mean[0, gene] * w.cell_factor[sel]
    assert abs(dense[sel].mean() - m.mean()) < 4 * np.sqrt((m + 0.2 * m * m).mean() / sel.sum())


def test_log1ppf_closed_form():
    """log1pPF equals log1p(x * mean_row_sum / row_sum) to FP32 rounding."""
    w = synth.make_workload(8, 500, num_types=3, num_genes=200, target_density=0.2)
    indptr, indices, counts = oracle.generate_counts(w)
    out = oracle.log1ppf(counts, indptr)
    rs = np.add.reduceat(counts.astype(np.float64), indptr[:-1])
    rs[np.diff(indptr) == 0] = 0
    mean = rs.sum() / w.num_cells
    ref = np.log1p(counts * np.repeat(mean / np.where(rs > 0, rs, 1), np.diff(indptr)))
    assert np.allclose(out, ref, rtol=3e-7, atol=0)


# ----------------------------------------------------------- FP64 oracle

@pytest.mark.parametrize('name', TINY)
def test_fp64_oracle_matches_dense_svd(golden, name):
    case, (data, indices, indptr, _) = _tiny_case(golden, name)
    n, g, k = case['num_cells'], case['num_genes'], case['k']
    out = oracle.fp64_pca(data, indices, indptr, n, g, k)
    Z = oracle.dense_z(data, indices, indptr, n, g)
    U, S, Vt = np.linalg.svd(Z, full_matrices=False)
    assert np.allclose(out['singular_values'], S[:k], rtol=1e-10, atol=0)
    ws = helpers.well_separated(S[:k], 1e-3)
    errs = helpers.column_errors_up_to_sign(out['PCs'][:, ws], (U * S)[:, ws])
    assert errs.max() < 1e-8


@pytest.mark.parametrize('name', TINY)
def test_cell_gram_oracle_matches_dense_svd(golden, name):
    """fp64_pca_cells (Z Z^T with the centring in closed form) against the
    dense SVD of Z on the tiny inputs."""
    case, (data, indices, indptr, _) = _tiny_case(golden, name)
    n, g, k = case['num_cells'], case['num_genes'], case['k']
    out = oracle.fp64_pca_cells(data, indices, indptr, n, g, k)
    Z = oracle.dense_z(data, indices, indptr, n, g)
    U, S, Vt = np.linalg.svd(Z, full_matrices=False)
    assert np.allclose(out['singular_values'], S[:k], rtol=1e-10, atol=0)
    ws = helpers.well_separated(S[:k], 1e-3)
    errs = helpers.column_errors_up_to_sign(out['PCs'][:, ws], (U * S)[:, ws])
    assert errs.max() < 1e-8


def test_cell_gram_oracle_planted_centring_mistake_is_caught():
    """Both oracles agree on a matrix with more genes than cells, and the
    cell Gram without its centring terms (a plausible mistake) does not."""
    from brisc_amd import synth
    import scipy.sparse as sp
    w = synth.make_workload(71, 600, num_types=4, num_genes=1500)
    data, indices, indptr = oracle.generate_csr(w)
    a = oracle.fp64_pca_cells(data, indices, indptr, 600, 1500, 20)['singular_values']
    b = oracle.fp64_pca(data, indices, indptr, 600, 1500, 20, want_pcs=False)['singular_values']
    assert np.allclose(a, b, rtol=1e-9, atol=0)
    sd = oracle.clipped_stddev(data, indices, indptr, 600, 1500)
    X = sp.csr_array((data.astype(np.float64), indices, indptr), shape=(600, 1500))
    Y = X @ sp.diags_array(1.0 / sd)
    wrong = np.sqrt(np.sort(np.linalg.eigvalsh((Y @ Y.T).toarray()))[::-1][:20])
    assert np.max(np.abs(wrong - a) / a) > 100 * SV_RTOL


def test_stddev_closed_form():
    rng = np.random.default_rng(0)
    X = (rng.random((60, 9)) < 0.3) * rng.random((60, 9)).astype(np.float32)
    X[:, 4] = 0.0                      # zero-variance gene -> clipped
    X[:, 7] = 2.5                      # constant gene -> clipped
    import scipy.sparse as sp
    A = sp.csr_array(X.astype(np.float32))
    sd = oracle.clipped_stddev(A.data, A.indices, A.indptr, 60, 9)
    ref = np.maximum(X.astype(np.float64).std(axis=0, ddof=1), 1e-8)
    assert np.allclose(sd, ref, rtol=1e-12, atol=1e-15)
    assert sd[4] == 1e-8 and sd[7] == 1e-8


def test_planted_centering_mistake_is_caught(golden):
    """Dropping the centring term from the oracle's Gram (a plausible
    mistake) moves the singular values far outside the tolerance."""
    case, (data, indices, indptr, _) = _tiny_case(golden, 'tiny_b')
    n, g, k = case['num_cells'], case['num_genes'], case['k']
    G, cs = oracle.gram(data, indices, indptr, g)
    sd = oracle.clipped_stddev(data, indices, indptr, n, g)
    wrong = np.sqrt(np.sort(np.linalg.eigvalsh(G / np.outer(sd, sd)))[::-1][:k])
    right = oracle.fp64_pca(data, indices, indptr, n, g, k)['singular_values']
    assert np.max(np.abs(wrong - right) / right) > 100 * SV_RTOL


# ----------------------------------------------------------- FP32 port

@pytest.mark.parametrize('name', TINY)
def test_port_and_reference_agree_with_fp64_oracle(golden, name):
    case, (data, indices, indptr, ref_pcs) = _tiny_case(golden, name)
    n, g, k = case['num_cells'], case['num_genes'], case['k']
    out = oracle.fp64_pca(data, indices, indptr, n, g, k)
    P, conv, _ = oracle.irlba_port(data, indices, indptr, n, g, k,
                                   subspace_size=case['subspace_size'],
                                   chunk_size=case['chunk_size'],
                                   num_threads=2)
    assert conv and case['converged']
    sv = out['singular_values']
    assert helpers.relative_sv_error(ref_pcs, sv).max() < SV_RTOL
    assert helpers.relative_sv_error(P, sv).max() < SV_RTOL
    ws = helpers.well_separated(sv, 1e-2)
    if len(ws):
        assert helpers.column_errors_up_to_sign(P[:, ws], ref_pcs[:, ws]).max() < COL_RTOL
        assert helpers.column_errors_up_to_sign(ref_pcs[:, ws], out['PCs'][:, ws]).max() < COL_RTOL
    assert helpers.cluster_projection_deficit(P, out['PCs'], sv, 1e-3).max() < CLUSTER_TOL


@pytest.mark.slow
def test_reference_port_oracle_agree_at_50k(golden, cells50k):
    meta, _, large = golden
    w, data, indices, indptr = cells50k
    k = meta['large']['k']
    out = oracle.fp64_pca(data, indices, indptr, w.num_cells, w.num_genes, k)
    ref_sv = large['singular_values']
    assert np.max(np.abs(ref_sv - out['singular_values']) / out['singular_values']) < SV_RTOL
    P, conv, st = oracle.irlba_port(data, indices, indptr, w.num_cells,
                                    w.num_genes, k, num_threads=8)
    assert conv
    assert helpers.relative_sv_error(P, ref_sv).max() < SV_RTOL
    rows = large['rows']
    ws = helpers.well_separated(ref_sv, 1e-2)
    errs = helpers.column_errors_up_to_sign(P[rows][:, ws], large['pcs_rows'][:, ws])
    assert errs.max() < COL_RTOL


def test_principal_angles_closed_form():
    """The subspace-angle helper on bases rotated by known angles."""
    rng = np.random.default_rng(4)
    Q, _ = np.linalg.qr(rng.standard_normal((60, 6)))
    t = np.array([0.0, 1e-6, 1e-3, 0.3, 0.7, 1.2])
    A = Q[:, :3]
    B = np.cos(t[3:]) * Q[:, :3] + np.sin(t[3:]) * Q[:, 3:]
    ang = helpers.principal_angles(A @ rng.standard_normal((3, 3)), B)
    assert np.allclose(ang, np.sort(t[3:]), atol=1e-12)
    assert helpers.principal_angles(A, A @ np.diag([2.0, -1.0, 3.0])).max() < 1e-7


@pytest.mark.slow
def test_fp64_oracle_matches_reference_at_one_million_cells():
    """The FP64 oracle against the reference's own irlba singular values on
    the 1M-cell workload (same bytes: SHA-256 of the generated CSR)."""
    import json
    import os
    here = os.path.join(os.path.dirname(__file__), 'golden')
    meta = json.load(open(os.path.join(here, 'ref_vectors.json')))['cells1m']
    ref = np.load(os.path.join(here, 'ref_irlba_cells1m.npz'))
    w = synth.make_workload(**synth.WORKLOADS['cells1m'])
    indptr, indices, counts = oracle.generate_counts(w)
    data = oracle.log1ppf(counts, indptr)
    assert synth.csr_sha256(data, indices, indptr) == meta['csr_sha256']
    out = oracle.fp64_pca(data, indices, indptr, w.num_cells, w.num_genes, 50,
                          want_pcs=False)
    err = np.abs(out['singular_values'] - ref['singular_values']) / ref['singular_values']
    assert err.max() < SV_RTOL


@pytest.mark.parametrize('n,d,k', [(2, 1, 1), (40, 3, 5), (300, 6, 39)])
def test_knn_oracle_matches_brute_force(n, d, k):
    # Brute force over every pair in pure Python (exact squared distances,
    # self excluded, ties by index) pins the oracle's candidate widening and
    # tie order; a small integer lattice gives many exact ties.
    rng = np.random.default_rng(n + d)
    X = rng.integers(-3, 4, size=(n, d)).astype(np.float32)
    idx, dist = oracle.knn_fp64(X, k, block=17)
    Xl = X.astype(np.float64).tolist()
    for q in range(n):
        pairs = sorted((sum((a - b) ** 2 for a, b in zip(Xl[q], Xl[c])), c)
                       for c in range(n) if c != q)[:k]
        assert [c for _, c in pairs] == idx[q].tolist()
        assert [s for s, _ in pairs] == dist[q].tolist()


def test_knn_oracle_planted_self_mistake_is_caught():
    # A restatement that forgot to exclude the query itself would return it
    # first at distance 0; the brute-force check above rejects that.
    X = np.random.default_rng(1).standard_normal((50, 4)).astype(np.float32)
    idx, dist = oracle.knn_fp64(X, 3)
    assert not (idx == np.arange(50)[:, None]).any()
    assert (dist > 0).all()


def _knn_fixture(golden, name):
    import hashlib
    meta = {m['name']: m for m in golden[0]['knn']}[name]
    X = helpers.knn_points(meta['seed'], meta['num_cells'], meta['num_dims'],
                           meta['centers'])
    assert hashlib.sha256(X.tobytes()).hexdigest() == meta['x_sha256']
    ref = np.load(os.path.join(os.path.dirname(__file__), 'golden',
                               'ref_knn.npz'))
    rows = ref[f'{name}_rows'] if f'{name}_rows' in ref.files else \
        np.arange(len(X))
    return (meta, X, rows, ref[f'{name}_neighbors'].astype(np.int64),
            ref[f'{name}_distances'])


@pytest.mark.parametrize('name', ['knn_a', 'knn_b', 'knn_c'])
def test_knn_oracle_matches_reference_exhaustive_search(golden, name):
    # The reference's knn_self with every cluster searched (fixtures written
    # by the reference itself) against the FP64 brute-force oracle: the same
    # neighbour sets up to near-ties, and the reference's FP32 distances
    # within the FP32 error of the norm identity of the exact ones.
    meta, X, rows, nb, dist = _knn_fixture(golden, name)
    k = meta['k']
    oi, od = oracle.knn_fp64(X, k, rows=rows)
    ok, same_rows, same_sets = helpers.knn_sets_agree(X, nb, oi, 2e-5, rows)
    assert ok
    assert same_sets > 0.99
    X64 = X.astype(np.float64)
    norms = (X64 * X64).sum(1)
    exact = ((X64[rows][:, None, :] - X64[nb]) ** 2).sum(2)
    assert (np.abs(dist - exact) <=
            2e-5 * (norms[rows][:, None] + norms[nb])).all()


def test_kmeans_init_draw_matches_reference(golden):
    # The library's draw of the initial cells (the reference's PCG32 and
    # Lemire generator, host code of the library) fed to the FP64 k-means
    # restatement reproduces the reference's own k-means labels after one
    # and two iterations (fixtures written by the reference), and its
    # centroids; another seed's draw does not.
    from brisc_amd.knn import kmeans_init_cells
    meta, X, rows, _, _ = _knn_fixture(golden, 'knn_c')
    ref = np.load(os.path.join(os.path.dirname(__file__), 'golden',
                               'ref_knn.npz'))
    C = meta['num_clusters']
    cells = kmeans_init_cells(len(X), C, 0).astype(np.int64)
    assert len(np.unique(cells)) == C
    l1, _, _ = oracle.kmeans_fp64(X, cells, 1)
    l2, c2, n2 = oracle.kmeans_fp64(X, cells, 2)
    same1 = (l1 == ref['knn_c_kmeans1_labels']).mean()
    same2 = (l2 == ref['knn_c_kmeans2_labels']).mean()
    assert same1 >= 0.999 and same2 >= 0.999, (same1, same2)
    # Centroids: equal to FP32 rounding except for clusters that a cell
    # near a tie joined or left.
    rc2 = ref['knn_c_kmeans2_centroids'].astype(np.float64)
    close = (np.abs(c2 - rc2) <= 1e-5 * np.abs(rc2).max()).all(1)
    assert close.mean() >= 0.95, close.mean()
    other = kmeans_init_cells(len(X), C, 1).astype(np.int64)
    lo, _, _ = oracle.kmeans_fp64(X, other, 1)
    assert (lo == ref['knn_c_kmeans1_labels']).mean() < 0.5


# ----------------------------------------------------------- k-means||

def _kb_fixtures():
    import json
    here = os.path.join(os.path.dirname(__file__), 'golden')
    meta = json.load(open(os.path.join(here, 'ref_vectors.json')))
    return meta['kmeans_barbar'], meta['kmeans_barbar_same_point_error'], \
        np.load(os.path.join(here, 'ref_kmeans_barbar.npz'))


def _kb_points(m):
    import hashlib
    X = helpers.lattice_points(m['seed'], m['num_cells'], m['num_dims'],
                               m['centres'], *m['lattice'])
    assert hashlib.sha256(X.tobytes()).hexdigest() == m['X_sha256']
    return X


def one_lloyd_step_fp32(X, cells):
    """Centroids after one Lloyd iteration from the given initial cells, as
    ``kmeans()`` computes them on integer-lattice points (where every
    distance and sum is exact): nearest initial centroid (lower index on
    ties), exact sums, times the reciprocal count rounded to FP32."""
    X64 = X.astype(np.float64)
    cent = X64[np.asarray(cells, np.int64)]
    lab = ((cent * cent).sum(1)[None, :] - 2.0 * X64 @ cent.T).argmin(1)
    C = len(cells)
    cnt = np.bincount(lab, minlength=C)
    sums = np.zeros((C, X.shape[1]))
    np.add.at(sums, lab, X64)
    assert (cnt > 0).all()
    return sums.astype(np.float32) * \
        (1.0 / cnt).astype(np.float32)[:, None]


def test_kmeans_barbar_oracle_matches_reference():
    # On integer lattices the reference's k-means|| is exact except for its
    # fused k-means++ probability chain: the centroids after one Lloyd
    # iteration from the oracle's cells must equal the reference's bit for
    # bit, for every chunk size it was run with (serial and parallel
    # versions, which the generator checked agree).  kb_f is large enough
    # that the chain's rounding decides some draws.
    meta, _, ref = _kb_fixtures()
    for m in meta:
        name = m['name']
        X = _kb_points(m)
        for chunk in m['chunk_sizes']:
            cells, S = oracle.kmeans_barbar_init(
                X, m['num_clusters'], m['num_init_iterations'],
                m['oversampling_factor'], chunk, m['kmeans_seed'])
            assert len(np.unique(cells)) == m['num_clusters']
            assert S >= m['num_clusters']
            c = one_lloyd_step_fp32(X, cells)
            assert np.array_equal(c.view(np.uint32),
                                  ref[f'{name}_centroids'].view(np.uint32)), \
                (name, chunk)


def test_kmeans_barbar_oracle_other_seed_differs():
    meta, _, ref = _kb_fixtures()
    m = next(m for m in meta if m['name'] == 'kb_a')
    X = _kb_points(m)
    cells, _ = oracle.kmeans_barbar_init(
        X, m['num_clusters'], m['num_init_iterations'],
        m['oversampling_factor'], None, m['kmeans_seed'] + 1)
    c = one_lloyd_step_fp32(X, cells)
    same = (c.view(np.uint32) == ref['kb_a_centroids'].view(np.uint32)).all(1)
    assert same.mean() < 0.5


def test_kmeans_barbar_oracle_special_cases():
    # Every cell at one point, and five distinct points (40 copies each)
    # with 1, 4, 5 and 8 clusters: the reference's error or its centroids
    # after one iteration (recorded by the fixture generator).  With five
    # clusters the reference stalls before its fifth round and raises, as
    # the oracle must.
    _, same_point_error, ref = _kb_fixtures()
    with pytest.raises(ValueError, match='same principal component'):
        oracle.kmeans_barbar_init(np.ones((50, 4), np.float32), 5)
    assert 'same principal component' in same_point_error
    import json
    meta = json.load(open(os.path.join(os.path.dirname(__file__), 'golden',
                                       'ref_vectors.json')))
    X = ref['kb_dup_X'].astype(np.float32)
    for C, outcome in meta['kmeans_barbar_duplicates'].items():
        C = int(C)
        if outcome == 'centroids':
            cells, _ = oracle.kmeans_barbar_init(X, C)
            c = one_lloyd_step_fp32(X, cells)
            assert np.array_equal(c.view(np.uint32),
                                  ref[f'kb_dup_{C}_centroids'].view(np.uint32)), C
        else:
            with pytest.raises(ValueError) as e:
                oracle.kmeans_barbar_init(X, C)
            assert str(e.value) == outcome, C


def test_kmeans_barbar_first_rounds_closed_form():
    # Two rounds by hand on small integer points (every distance exact): the
    # first candidate is randint(n) from srand(seed - 1); round one takes
    # each cell whose uniform from srand(seed + i) lies at or below
    # l d_i / cost; round two (seeds seed + i * 1, the same uniforms) does
    # the same with each distance lowered to the nearest new candidate and
    # taken cells at 0; with num_init_iterations = 2 the selection then
    # stops.  With one cluster the output is candidate randint(S) from
    # srand(seed - 2).
    def pcg(state):
        x = state
        state = (x * 6364136223846793005 + 1442695040888963407) % 2 ** 64
        s = ((x ^ (x >> 18)) >> 27) & 0xffffffff
        rot = x >> 59
        return ((s >> rot) | (s << ((-rot) & 31))) & 0xffffffff, state

    def srand(seed):
        _, st = pcg((seed + 1442695040888963407) % 2 ** 64)
        return st

    def bounded(bound, st):
        r, st = pcg(st)
        m = r * bound
        lo = m & 0xffffffff
        if lo < bound:
            thr = (2 ** 32 - bound) % bound
            while lo < thr:
                r, st = pcg(st)
                m = r * bound
                lo = m & 0xffffffff
        return m >> 32, st

    rng = np.random.default_rng(3)
    X = rng.integers(-3, 4, (40, 3)).astype(np.float32)
    seed = 17
    l = np.float32(2.0)
    u = np.array([np.float32(pcg(srand(seed + i))[0]) / np.float32(2 ** 32)
                  for i in range(40)], np.float32)
    first, _ = bounded(40, srand((seed - 1) % 2 ** 64))
    d = ((X - X[first]) ** 2).sum(1).astype(np.float32)
    cand = [first]
    for rnd in range(2):
        if rnd:
            for c in new:
                d = np.minimum(d, ((X - X[c]) ** 2).sum(1).astype(np.float32))
        cost = np.float32(d.sum())
        new = [i for i in range(40) if np.float32(l / cost) * d[i] >= u[i]]
        d[new] = 0
        cand += new
    cells, S = oracle.kmeans_barbar_init(X, 1, num_init_iterations=2,
                                         oversampling_factor=2.0, seed=seed)
    assert S == len(cand)
    pick, _ = bounded(S, srand((seed - 2) % 2 ** 64))
    assert cells[0] == cand[pick]
