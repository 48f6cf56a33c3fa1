"""Write the reference's own IRLBA / log1pPF outputs as small fixtures.

Run in the development container only (it reads the reference sources):

    python tests/golden/make_ref_vectors.py [--reference /root/reference]
    python tests/golden/make_ref_vectors.py --cells1m   (adds the 1M-cell fixture)
    python tests/golden/make_ref_vectors.py --knn       (adds the neighbour fixtures)
    python tests/golden/make_ref_vectors.py --knn-large (adds the 1M-cell neighbour fixture)
    python tests/golden/make_ref_vectors.py --pipeline1m (adds the 1M-cell pipeline fixture)
    python tests/golden/make_ref_vectors.py --knn-approx (adds few-clusters-searched fixtures)
    python tests/golden/make_ref_vectors.py --rng       (adds the start-vector draws)
    python tests/golden/make_ref_vectors.py --calibrate (times the reference against the port)
    python tests/golden/make_ref_vectors.py --normalize (adds the normalisation fixtures)
    python tests/golden/make_ref_vectors.py --kmeans-barbar (adds the k-means|| fixtures)
    python tests/golden/make_ref_vectors.py --knn-large-barbar (times the 1M-cell k-means|| search)

It cythonizes ``brisc/cyutils.pyx``, ``brisc/pca.pyx``,
``brisc/normalize.pyx``, ``brisc/kmeans.pyx`` and ``brisc/knn.pyx`` from the
reference tree into the git-ignored scratch
directory ``tests/golden/_refbuild`` (never shipped anywhere), compiles them
with the reference's Linux flags (``setup.py:93-98``) and compiler directives
(``setup.py:148-151``), imports them under the package name ``brisc``, and
records what only the reference itself can pin:

* the exact FP32 output of ``irlba`` (parallel path, ``pca.pyx:973-1254``) --
  full PCs for tiny inputs, singular values (column norms of the PCs) and 64
  PC rows for the 50k-cell workload, plus the converged flag;
* the exact bytes of ``normalize_csr`` (log1pPF, ``normalize.pyx:9-128``) as a
  SHA-256 of the normalised data.

The inputs are regenerated deterministically by the tests from the seeds
stored beside the outputs, so no reference code or large array is committed.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import subprocess
import sys
import sysconfig
import time
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'tests'))

from brisc_amd import synth  # noqa: E402
import oracle  # noqa: E402
import helpers  # noqa: E402

BUILD = os.path.join(HERE, '_refbuild')
MODULES = ('cyutils', 'pca', 'normalize', 'kmeans', 'knn')

# Tiny inputs: (name, seed, cells, genes, types, density, k, subspace, chunk)
TINY = [
    ('tiny_a', 11, 300, 40, 4, 0.30, 5, 12, 1024),
    ('tiny_b', 12, 1000, 120, 6, 0.15, 10, 25, 1024),
    ('tiny_c', 13, 257, 64, 5, 0.20, 8, 20, 7),
    ('tiny_d', 14, 2500, 200, 8, 0.10, 20, 45, 100),
]
LARGE = ('cells50k', 50)
ROW_SAMPLE = 64


def build_reference(reference: str) -> None:
    os.makedirs(os.path.join(BUILD, 'brisc'), exist_ok=True)
    suffix = sysconfig.get_config_var('EXT_SUFFIX')
    np_inc = np.get_include()
    py_inc = sysconfig.get_paths()['include']
    procs = []
    for m in MODULES:
        src = os.path.join(reference, 'brisc', f'{m}.pyx')
        cpp = os.path.join(BUILD, f'{m}.cpp')
        so = os.path.join(BUILD, 'brisc', f'{m}{suffix}')
        if os.path.exists(so) and os.path.getmtime(so) > os.path.getmtime(src):
            continue
        subprocess.run([sys.executable, '-m', 'cython', '-3', '--cplus',
                        '-X', 'boundscheck=False', '-X', 'wraparound=False',
                        '-X', 'cdivision=True', '-X', 'initializedcheck=False',
                        src, '-o', cpp], check=True)
        procs.append(subprocess.Popen(
            ['g++', '-Ofast', '-funroll-loops', '-std=c++17', '-g0',
             '-fopenmp', '-fPIC', '-shared', '-march=native',
             f'-I{py_inc}', f'-I{np_inc}',
             '-DNPY_NO_DEPRECATED_API=NPY_1_17_API_VERSION', cpp, '-o', so]))
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError('compiling the reference module failed')


def import_reference():
    pkg = types.ModuleType('brisc')
    pkg.__path__ = [os.path.join(BUILD, 'brisc')]
    sys.modules['brisc'] = pkg
    return (importlib.import_module('brisc.pca'),
            importlib.import_module('brisc.normalize'))


def ref_normalize(norm_mod, counts, indices, indptr, num_threads=8):
    n = len(indptr) - 1
    out = np.empty(len(counts), np.float32)
    row_sums = np.empty(n, np.uint64)
    norm_mod.normalize_csr(counts.astype(np.int32), indices.astype(np.int32),
                           indptr.astype(np.int32), np.empty(0, np.int8), out,
                           row_sums, n, 1, num_threads)
    return out


def ref_irlba(pca_mod, data, indices, indptr, n, g, k, m, chunk,
              num_threads=8, seed=0, tolerance=1e-6, max_iterations=100):
    import scipy.sparse as sp
    X = sp.csr_array((data, indices, indptr), shape=(n, g))
    Xc = X.tocsc()
    Xc.sort_indices()
    PCs = np.empty((n, k), np.float32)
    m = min(m, n, g)
    t0 = time.perf_counter()
    conv = pca_mod.irlba(
        data_matvec=data, indices_matvec=indices.astype(np.int32),
        indptr_matvec=indptr.astype(np.int32), data_rmatvec=Xc.data,
        indices_rmatvec=Xc.indices.astype(np.int32),
        indptr_rmatvec=Xc.indptr.astype(np.int32), num_cells=n, num_genes=g,
        k=k, subspace_size=m, tolerance=tolerance,
        max_iterations=max_iterations, seed=seed,
        match_parallel=num_threads == 1, is_csr=True,
        num_threads=num_threads, chunk_size=chunk, PCs=PCs)
    return PCs, bool(conv), time.perf_counter() - t0


def cells1m(pca_mod, norm_mod):
    """The 1M-cell configuration: the reference's singular values, 64 PC
    rows and converged flag, written beside the other fixtures."""
    w = synth.make_workload(**synth.WORKLOADS['cells1m'])
    indptr, indices, counts = oracle.generate_counts(w)
    data = ref_normalize(norm_mod, counts, indices, indptr)
    PCs, conv, t8 = ref_irlba(pca_mod, data, indices, indptr, w.num_cells,
                              w.num_genes, 50, 100, 1024, num_threads=8)
    rows = np.random.Generator(np.random.PCG64(1)).choice(
        w.num_cells, ROW_SAMPLE, replace=False)
    rows.sort()
    np.savez_compressed(
        os.path.join(HERE, 'ref_irlba_cells1m.npz'),
        singular_values=np.linalg.norm(PCs.astype(np.float64), axis=0),
        rows=rows, pcs_rows=PCs[rows])
    path = os.path.join(HERE, 'ref_vectors.json')
    with open(path) as f:
        meta = json.load(f)
    meta['cells1m'] = dict(
        workload='cells1m', k=50, subspace_size=100, chunk_size=1024,
        tolerance=1e-6, max_iterations=100, seed=0, converged=conv,
        nnz=int(len(data)), scale=w.scale,
        csr_sha256=synth.csr_sha256(data, indices, indptr),
        seconds_8_threads=round(t8, 3), host_cpus=os.cpu_count())
    with open(path, 'w') as f:
        json.dump(meta, f, indent=1)
    print(json.dumps(meta['cells1m'], indent=1))


# Neighbour-search inputs: (name, seed, cells, dims, centres, k, full rows)
KNN = [
    ('knn_a', 31, 2000, 50, 30, 20, True),
    ('knn_b', 32, 500, 10, 8, 64, True),
    ('knn_c', 33, 50000, 50, 60, 20, False),
]
KNN_ROW_SAMPLE = 256


def import_knn():
    return (importlib.import_module('brisc.kmeans'),
            importlib.import_module('brisc.knn'))


def ref_neighbors(kmeans_mod, knn_mod, X, k, num_clusters_searched=None,
                  num_threads=8, seed=0, num_clusters=None,
                  kmeans_barbar=False):
    """The steps of SingleCell.neighbors() (single_cell.py:12452-12650) with
    its defaults, on PCs ``X``; ``num_clusters_searched='all'`` searches every
    cluster (the exhaustive search).  Returns neighbours (uint32), squared
    distances (float32) and the number of clusters."""
    from threadpoolctl import threadpool_limits
    n, d = X.shape
    C = max(1, int(np.ceil(min(4 * np.sqrt(n), n / 100)))) \
        if num_clusters is None else num_clusters
    labels = np.zeros(n, np.uint32)
    min_dist = np.zeros(n, np.float32)
    norms = np.zeros(n, np.float32)
    cent = np.empty((C, d), np.float32)
    cent_new = np.empty((C, d), np.float32)
    counts = np.empty(C, np.uint32)
    with threadpool_limits(1, user_api='blas'):
        it = kmeans_mod.kmeans(
            X=X, cluster_labels=labels, centroids=cent, centroids_new=cent_new,
            num_cells_per_cluster=counts, min_distances=min_dist,
            cell_norms=norms, kmeans_barbar=kmeans_barbar,
            num_init_iterations=5,
            num_kmeans_iterations=2, tolerance=1e-2, oversampling_factor=1,
            seed=seed, chunk_size=min(4096, n), is_mac=False,
            num_threads=num_threads)
    if it & 1:
        cent = cent_new
    keep = counts > 0
    if not keep.all():
        cent, counts = cent[keep], counts[keep]
        remap = np.zeros(len(keep), np.uint32)
        remap[keep] = np.arange(keep.sum(), dtype=np.uint32)
        labels = remap[labels]
    C = len(cent)
    searched = C if num_clusters_searched == 'all' else \
        min(64, C) if num_clusters_searched is None else num_clusters_searched
    smallest = np.sort(counts).cumsum(dtype=np.uint64)
    searched = min(max(searched, int(np.searchsorted(smallest, k + 1)) + 1), C)
    corder = np.argsort(cent[:, 0]).astype(np.uint32)
    cent = np.ascontiguousarray(cent[corder])
    counts = counts[corder]
    inv = np.empty_like(corder)
    inv[corder] = np.arange(C, dtype=np.uint32)
    labels = inv[labels]
    order = np.argsort(labels).astype(np.uint32)
    Xs = np.ascontiguousarray(X[order])
    norms_s = np.ascontiguousarray(norms[order])
    nb = np.zeros((n, k), np.uint32)
    dist = np.zeros((n, k), np.float32)
    with threadpool_limits(1, user_api='blas'):
        knn_mod.knn_self(
            X=Xs, sorted_order=order, centroids=cent,
            num_cells_per_cluster=counts, neighbors=nb, distances=dist,
            cell_norms=norms_s,
            centroid_distances=np.zeros((n, searched), np.float32),
            nearest_clusters=np.zeros((n, searched), np.uint32),
            num_neighbors=k, num_clusters_searched=searched,
            chunk_size_kmeans=min(4096, n), chunk_size_search=min(256, n),
            is_mac=False, num_threads=num_threads)
    out_nb = np.empty_like(nb)
    out_nb[order] = nb
    out_dist = np.empty_like(dist)
    out_dist[order] = dist
    return out_nb, out_dist, C


def ref_kmeans(kmeans_mod, X, num_clusters, iterations, seed=0,
               num_threads=8, kmeans_barbar=False, num_init_iterations=5,
               oversampling_factor=1.0, chunk_size=None):
    """kmeans() as neighbors() calls it (random initialisation unless
    ``kmeans_barbar``): labels (last E-step), centroids, counts and
    iterations run."""
    from threadpoolctl import threadpool_limits
    n, d = X.shape
    labels = np.zeros(n, np.uint32)
    cent = np.empty((num_clusters, d), np.float32)
    cent_new = np.empty((num_clusters, d), np.float32)
    counts = np.empty(num_clusters, np.uint32)
    with threadpool_limits(1, user_api='blas'):
        it = kmeans_mod.kmeans(
            X=X, cluster_labels=labels, centroids=cent, centroids_new=cent_new,
            num_cells_per_cluster=counts, min_distances=np.zeros(n, np.float32),
            cell_norms=np.zeros(n, np.float32), kmeans_barbar=kmeans_barbar,
            num_init_iterations=num_init_iterations,
            num_kmeans_iterations=iterations, tolerance=1e-2,
            oversampling_factor=oversampling_factor, seed=seed,
            chunk_size=min(4096, n) if chunk_size is None else chunk_size,
            is_mac=False, num_threads=num_threads)
    return labels, (cent_new if it & 1 else cent), counts, it


# k-means|| inputs on integer lattices (every distance and cost exact in FP32):
# (name, seed, cells, dims, centres, clusters, init iterations, oversampling,
#  k-means seed, chunk sizes)
KMEANS_BARBAR = [
    ('kb_a', 41, 3000, 8, 12, 40, 5, 1.0, 0, (3000, 512)),
    ('kb_b', 42, 20000, 10, 30, 120, 3, 2.0, 5, (4096, 1000)),
    ('kb_c', 43, 60, 3, 4, 10, 5, 1.0, 7, (60, 7)),
    ('kb_d', 44, 5000, 6, 8, 1, 5, 0.5, 3, (4096,)),
    ('kb_e', 45, 8000, 16, 40, 300, 2, 1.5, 11, (4096, 333)),
    ('kb_f', 46, 100000, 6, 50, 1000, 5, 2.0, 13, (4096,)),
]


def kmeans_barbar_fixtures():
    """kmeans() with k-means|| initialisation (kmeans.pyx:41-720) and one
    Lloyd iteration on integer-lattice points, serial (one thread) and
    parallel (several chunks, eight threads): the centroids after the
    iteration (means of the cells nearest each initial centroid, which pins
    the initial centroids), the last E-step's labels and counts.  The inputs
    are regenerated by helpers.lattice_points (SHA-256 recorded)."""
    kmeans_mod, _ = import_knn()
    arrays, meta = {}, []
    for (name, seed, n, d, centres, C, iters, over, kseed, chunks) in KMEANS_BARBAR:
        lattice = (2, 1.0) if n > 50000 else (4, 1.5)
        X = helpers.lattice_points(seed, n, d, centres, *lattice)
        assert np.abs(X).max() < 128
        runs = []
        for chunk in chunks:
            for threads in (1, 8):
                t0 = time.perf_counter()
                out = ref_kmeans(kmeans_mod, X, C, 1, seed=kseed,
                                 num_threads=threads, kmeans_barbar=True,
                                 num_init_iterations=iters,
                                 oversampling_factor=over, chunk_size=chunk)
                runs.append((chunk, threads, time.perf_counter() - t0, out))
        l0, c0, n0, _ = runs[0][3]
        for chunk, threads, _, (l, c, cnt, it) in runs:
            assert it == 1
            assert np.array_equal(c.view(np.uint32), c0.view(np.uint32)), \
                (name, chunk, threads, 'centroids differ')
            assert np.array_equal(l, l0) and np.array_equal(cnt, n0)
        import hashlib
        arrays[f'{name}_centroids'] = c0
        arrays[f'{name}_labels'] = l0.astype(np.uint16)
        arrays[f'{name}_counts'] = n0
        meta.append(dict(name=name, seed=seed, num_cells=n, num_dims=d,
                         centres=centres, lattice=list(lattice),
                         X_sha256=hashlib.sha256(X.tobytes()).hexdigest(),
                         num_clusters=C,
                         num_init_iterations=iters, oversampling_factor=over,
                         kmeans_seed=kseed, chunk_sizes=list(chunks),
                         seconds=[round(r[2], 4) for r in runs]))
        print(name, [round(r[2], 3) for r in runs], flush=True)
    # Every cell at one point: the reference's error.
    Xs = np.ones((50, 4), np.float32)
    try:
        ref_kmeans(kmeans_mod, Xs, 5, 1, kmeans_barbar=True)
        same_error = None
    except ValueError as e:
        same_error = str(e)
    # Five distinct points, 40 copies each: the reference's outcome (error
    # message or centroids after one iteration) for a few cluster counts.
    Xd = np.ascontiguousarray(np.repeat(
        np.array([[0, 0], [3, 1], [-2, 4], [5, -3], [1, 1]], np.float32),
        40, axis=0))
    arrays['kb_dup_X'] = Xd.astype(np.int8)
    dup = {}
    for C in (1, 4, 5, 8):
        try:
            _, c, _, _ = ref_kmeans(kmeans_mod, Xd, C, 1, kmeans_barbar=True)
            arrays[f'kb_dup_{C}_centroids'] = c
            dup[str(C)] = 'centroids'
        except ValueError as e:
            dup[str(C)] = str(e)
    np.savez_compressed(os.path.join(HERE, 'ref_kmeans_barbar.npz'), **arrays)
    path = os.path.join(HERE, 'ref_vectors.json')
    allmeta = json.load(open(path))
    allmeta['kmeans_barbar'] = meta
    allmeta['kmeans_barbar_same_point_error'] = same_error
    allmeta['kmeans_barbar_duplicates'] = dup
    print(dup)
    json.dump(allmeta, open(path, 'w'), indent=1)
    print(json.dumps(meta, indent=1), same_error)


def knn_fixtures():
    """Exhaustive and default (approximate) neighbour searches of the
    reference on seeded PC-space mixtures: full outputs for the small sets,
    a row sample, SHA-256 digests and recall figures for the large one."""
    import hashlib
    kmeans_mod, knn_mod = import_knn()
    arrays, meta = {}, []
    for (name, seed, n, d, centers, k, full) in KNN:
        X = helpers.knn_points(seed, n, d, centers)
        t0 = time.perf_counter()
        nb_all, dist_all, C = ref_neighbors(kmeans_mod, knn_mod, X, k, 'all')
        t_all = time.perf_counter() - t0
        t0 = time.perf_counter()
        nb_def, dist_def, _ = ref_neighbors(kmeans_mod, knn_mod, X, k)
        t_def = time.perf_counter() - t0
        nb1, dist1, _ = ref_neighbors(kmeans_mod, knn_mod, X, k, 'all',
                                      num_threads=1)
        exact_i, _ = oracle.knn_fp64(X, k)
        recall_def = float(np.mean([len(set(a) & set(b)) / k for a, b in
                                    zip(nb_def.tolist(), exact_i.tolist())]))
        recall_all = float(np.mean([len(set(a) & set(b)) / k for a, b in
                                    zip(nb_all.tolist(), exact_i.tolist())]))
        entry = dict(name=name, seed=seed, num_cells=n, num_dims=d,
                     centers=centers, k=k, num_clusters=C,
                     x_sha256=hashlib.sha256(X.tobytes()).hexdigest(),
                     exhaustive_sha256=hashlib.sha256(nb_all.tobytes()).hexdigest(),
                     thread_invariant=bool(np.array_equal(nb_all, nb1) and
                                           np.array_equal(dist_all, dist1)),
                     recall_default_vs_fp64=recall_def,
                     recall_exhaustive_vs_fp64=recall_all,
                     seconds_exhaustive_8_threads=round(t_all, 3),
                     seconds_default_8_threads=round(t_def, 3))
        if full:
            arrays[f'{name}_neighbors'] = nb_all.astype(np.uint16)
            arrays[f'{name}_distances'] = dist_all
            arrays[f'{name}_default_neighbors'] = nb_def.astype(np.uint16)
        else:
            # The k-means stage on its own: labels after one and two
            # iterations, and the centroids / counts after two.
            l1, _, _, _ = ref_kmeans(kmeans_mod, X, C, 1)
            l2, c2, n2, it2 = ref_kmeans(kmeans_mod, X, C, 2)
            arrays[f'{name}_kmeans1_labels'] = l1.astype(np.uint16)
            arrays[f'{name}_kmeans2_labels'] = l2.astype(np.uint16)
            arrays[f'{name}_kmeans2_centroids'] = c2
            arrays[f'{name}_kmeans2_counts'] = n2
            entry['kmeans_iterations_run'] = int(it2)
            rows = np.random.Generator(np.random.PCG64(2)).choice(
                n, KNN_ROW_SAMPLE, replace=False)
            rows.sort()
            arrays[f'{name}_rows'] = rows
            arrays[f'{name}_neighbors'] = nb_all[rows]
            arrays[f'{name}_distances'] = dist_all[rows]
            arrays[f'{name}_default_neighbors'] = nb_def[rows]
        meta.append(entry)
        print(json.dumps(entry), flush=True)
    np.savez_compressed(os.path.join(HERE, 'ref_knn.npz'), **arrays)
    path = os.path.join(HERE, 'ref_vectors.json')
    with open(path) as f:
        allmeta = json.load(f)
    allmeta['knn'] = meta
    with open(path, 'w') as f:
        json.dump(allmeta, f, indent=1)


KNN_LARGE = ('pcs1m', 34, 1_000_000, 50, 100, 20, 1.0)

# Approximate searches with non-default settings on the 50k-cell set:
# (name, num_clusters, num_clusters_searched, k, seed)
KNN_APPROX = [
    ('approx_c100_s8', 100, 8, 20, 0),
    ('approx_c300_s3', 300, 3, 10, 7),
]


def knn_approx():
    """The reference's search with few clusters searched (recall well below
    1): a 4096-row sample of its neighbours and the SHA-256 of all of them,
    for parity of the approximate path."""
    kmeans_mod, knn_mod = import_knn()
    path = os.path.join(HERE, 'ref_vectors.json')
    with open(path) as f:
        meta = json.load(f)
    base = {m['name']: m for m in meta['knn']}['knn_c']
    X = helpers.knn_points(base['seed'], base['num_cells'], base['num_dims'],
                           base['centers'])
    arrays, out = {}, []
    for (name, C, searched, k, seed) in KNN_APPROX:
        nb, dist, Ck = ref_neighbors(kmeans_mod, knn_mod, X, k, searched,
                                     seed=seed, num_clusters=C)
        exact_i, _ = oracle.knn_fp64(X, k)
        recall = float(np.mean([len(set(a) & set(b)) / k for a, b in
                                zip(nb.tolist(), exact_i.tolist())]))
        import hashlib
        rows = np.random.Generator(np.random.PCG64(4)).choice(
            len(X), 4096, replace=False)
        rows.sort()
        arrays[f'{name}_rows'] = rows.astype(np.uint32)
        arrays[f'{name}_neighbors'] = nb[rows].astype(np.uint16)
        out.append(dict(name=name, base='knn_c', num_clusters=C,
                        num_clusters_searched=searched, k=k, seed=seed,
                        clusters_kept=Ck, recall_vs_fp64=recall,
                        neighbors_sha256=hashlib.sha256(nb.tobytes()).hexdigest()))
        print(json.dumps(out[-1]), flush=True)
    np.savez_compressed(os.path.join(HERE, 'ref_knn_approx.npz'), **arrays)
    meta['knn_approx'] = out
    with open(path, 'w') as f:
        json.dump(meta, f, indent=1)


def knn_large():
    """The reference's default neighbour search on the 1M-cell PC mixture
    that bench.py --mode knn times: seconds on this host's 8 threads, the
    SHA-256 of its neighbour indices, a row sample of its output."""
    import hashlib
    kmeans_mod, knn_mod = import_knn()
    name, seed, n, d, centers, k, spread = KNN_LARGE
    X = synth.pc_mixture(seed, n, d, centers, spread)
    t0 = time.perf_counter()
    nb, dist, C = ref_neighbors(kmeans_mod, knn_mod, X, k)
    secs = time.perf_counter() - t0
    rows = np.random.Generator(np.random.PCG64(3)).choice(
        n, KNN_ROW_SAMPLE, replace=False)
    rows.sort()
    np.savez_compressed(os.path.join(HERE, f'ref_knn_{name}.npz'), rows=rows,
                        neighbors=nb[rows], distances=dist[rows])
    path = os.path.join(HERE, 'ref_vectors.json')
    with open(path) as f:
        meta = json.load(f)
    meta['knn_large'] = dict(
        name=name, seed=seed, num_cells=n, num_dims=d, centers=centers,
        spread=spread, k=k, num_clusters=C,
        x_sha256=hashlib.sha256(X.tobytes()).hexdigest(),
        default_sha256=hashlib.sha256(nb.tobytes()).hexdigest(),
        seconds_default_8_threads=round(secs, 3), host_cpus=os.cpu_count())
    with open(path, 'w') as f:
        json.dump(meta, f, indent=1)
    print(json.dumps(meta['knn_large'], indent=1))


def knn_large_barbar():
    """The reference's neighbour search with k-means|| initialisation
    (kmeans_barbar=True) on the same 1M-cell PC mixture: seconds on this
    host's 8 threads and the recall of a row sample against the exact
    neighbours (FP64)."""
    kmeans_mod, knn_mod = import_knn()
    name, seed, n, d, centers, k, spread = KNN_LARGE
    X = synth.pc_mixture(seed, n, d, centers, spread)
    t0 = time.perf_counter()
    nb, dist, C = ref_neighbors(kmeans_mod, knn_mod, X, k, kmeans_barbar=True)
    secs = time.perf_counter() - t0
    rows = np.random.Generator(np.random.PCG64(3)).choice(
        n, KNN_ROW_SAMPLE, replace=False)
    rows.sort()
    exact_i, _ = oracle.knn_fp64(X, k, rows=rows)
    recall = float(np.mean([len(set(a) & set(b)) / k for a, b in
                            zip(nb[rows].tolist(), exact_i.tolist())]))
    path = os.path.join(HERE, 'ref_vectors.json')
    with open(path) as f:
        meta = json.load(f)
    meta['knn_large']['seconds_kmeans_barbar_8_threads'] = round(secs, 3)
    meta['knn_large']['recall_kmeans_barbar_rows'] = round(recall, 5)
    with open(path, 'w') as f:
        json.dump(meta, f, indent=1)
    print(json.dumps(meta['knn_large'], indent=1))


def pipeline1m(pca_mod, norm_mod):
    """The reference's normalize -> irlba -> neighbors chain on the 1M-cell
    workload (SingleCell.pca() + SingleCell.neighbors() defaults, all genes,
    no QC): stage times on this host's 8 threads and a row sample of the
    neighbours found on the reference's own PCs."""
    import hashlib
    kmeans_mod, knn_mod = import_knn()
    w = synth.make_workload(**synth.WORKLOADS['cells1m'])
    indptr, indices, counts = oracle.generate_counts(w)
    t0 = time.perf_counter()
    data = ref_normalize(norm_mod, counts, indices, indptr)
    t_norm = time.perf_counter() - t0
    PCs, conv, t_pca = ref_irlba(pca_mod, data, indices, indptr, w.num_cells,
                                 w.num_genes, 50, 100, 1024, num_threads=8)
    t0 = time.perf_counter()
    nb, dist, C = ref_neighbors(kmeans_mod, knn_mod, np.ascontiguousarray(PCs), 20)
    t_knn = time.perf_counter() - t0
    rows = np.random.Generator(np.random.PCG64(5)).choice(
        w.num_cells, KNN_ROW_SAMPLE, replace=False)
    rows.sort()
    np.savez_compressed(os.path.join(HERE, 'ref_pipeline_cells1m.npz'),
                        rows=rows, neighbors=nb[rows], distances=dist[rows])
    path = os.path.join(HERE, 'ref_vectors.json')
    with open(path) as f:
        meta = json.load(f)
    meta['pipeline_cells1m'] = dict(
        workload='cells1m', k=50, num_neighbors=20, num_clusters=C,
        converged=conv, seconds_normalize_8_threads=round(t_norm, 3),
        seconds_pca_8_threads=round(t_pca, 3),
        seconds_neighbors_8_threads=round(t_knn, 3),
        pcs_sha256=hashlib.sha256(PCs.tobytes()).hexdigest(),
        neighbors_sha256=hashlib.sha256(nb.tobytes()).hexdigest(),
        host_cpus=os.cpu_count())
    with open(path, 'w') as f:
        json.dump(meta, f, indent=1)
    print(json.dumps(meta['pipeline_cells1m'], indent=1))


# A few lines of this script's own Cython that call the reference's
# srand / random_normal (declared in brisc/cyutils.pxd, compiled into the
# reference's cyutils module above) to record the IRLBA start-vector draws.
RNG_SHIM = """
# distutils: language = c++
from brisc.cyutils cimport srand, random_normal
import numpy as np

def normal_draws(unsigned long long seed, unsigned count):
    cdef unsigned long long state = srand(seed)
    out = np.empty(count, np.float32)
    cdef float[::1] o = out
    cdef unsigned i
    for i in range(count):
        o[i] = random_normal(&state)
    return out
"""

RNG_SEEDS = [(0, 2000), (1, 300), (42, 300), (2 ** 63 + 12345, 300),
             (2 ** 64 - 1, 300), (1_000_003, 20_000)]


def rng_fixture(reference: str):
    """The reference's start-vector draws (pca.pyx:801-806: srand(seed),
    then random_normal per gene) for a few seeds, as float32 arrays."""
    suffix = sysconfig.get_config_var('EXT_SUFFIX')
    src = os.path.join(BUILD, 'rngshim.pyx')
    cpp = os.path.join(BUILD, 'rngshim.cpp')
    so = os.path.join(BUILD, f'rngshim{suffix}')
    with open(src, 'w') as f:
        f.write(RNG_SHIM)
    subprocess.run([sys.executable, '-m', 'cython', '-3', '--cplus',
                    '-I', reference, src, '-o', cpp], check=True)
    subprocess.run(['g++', '-Ofast', '-std=c++17', '-fPIC', '-shared',
                    '-march=native', f"-I{sysconfig.get_paths()['include']}",
                    f'-I{np.get_include()}', cpp, '-o', so], check=True)
    sys.path.insert(0, BUILD)
    shim = importlib.import_module('rngshim')
    out = {}
    for seed, count in RNG_SEEDS:
        out[f'seed_{seed}'] = shim.normal_draws(seed, count)
    np.savez_compressed(os.path.join(HERE, 'ref_start_vector.npz'), **out)
    print({k: v[:3] for k, v in out.items()})


NORM_DTYPES = ('int32', 'int64', 'uint32', 'float32', 'float32_frac', 'float64_frac')
NORM_METHODS = ('logCP10k', 'log1pPF', 'PFlog1pPF')


def normalize_inputs():
    """Counts of the normalisation fixtures: a 500-cell x 200-gene workload
    (5 types, 10% nonzeros) from the seeded generator, every count type the
    reference's normalize accepts (fractional float counts are the integer
    counts times 0.75, so the library sizes truncate), and a QC mask."""
    w = synth.make_workload(17, 500, num_types=5, num_genes=200,
                            target_density=0.10)
    indptr, indices, counts = oracle.generate_counts(w)
    qc = np.random.Generator(np.random.PCG64(17)).random(500) < 0.8
    data = {'int32': counts.astype(np.int32), 'int64': counts.astype(np.int64),
            'uint32': counts.astype(np.uint32), 'float32': counts.astype(np.float32),
            'float32_frac': (counts * 0.75).astype(np.float32),
            'float64_frac': counts.astype(np.float64) * 0.75}
    return w, indptr.astype(np.int64), indices.astype(np.int32), data, qc


def normalize_fixture(norm_mod):
    """The reference's normalize_csr and normalize_csc outputs for every
    method, count type and QC setting (SHA-256 of the normalised values and
    of the row sums), with 8 threads and 1, and the full PFlog1pPF outputs
    (CSR order)."""
    import hashlib
    import scipy.sparse as sp
    w, indptr, indices, data, qc = normalize_inputs()
    n, g = 500, 200
    cases, full = [], {}
    for dname in NORM_DTYPES:
        X = sp.csr_array((data[dname], indices, indptr), shape=(n, g))
        C = X.tocsc()
        C.sort_indices()
        for mi, method in enumerate(NORM_METHODS):
            for use_qc in (False, True):
                q = qc.astype(np.int8) if use_qc else np.empty(0, np.int8)
                res = {}
                for layout, M, fn in (('csr', X, norm_mod.normalize_csr),
                                      ('csc', C, norm_mod.normalize_csc)):
                    outs = []
                    for threads in (8, 1):
                        out = np.empty(len(M.data), np.float32)
                        # zeroed, as pca() passes it (the parallel CSC
                        # path adds into it)
                        rs = np.zeros(n, np.uint64)
                        fn(M.data, M.indices.astype(np.int32), M.indptr.astype(np.int32),
                           q, out, rs, n, mi, threads)
                        outs.append((out, rs))
                    same_threads = bool(np.array_equal(outs[0][0].view(np.uint32),
                                                       outs[1][0].view(np.uint32)))
                    out, rs = outs[0]
                    res[layout] = dict(
                        values_sha256=hashlib.sha256(out.tobytes()).hexdigest(),
                        row_sums_sha256=hashlib.sha256(rs.tobytes()).hexdigest(),
                        thread_invariant=same_threads)
                    if method == 'PFlog1pPF' and layout == 'csr':
                        full[f'pf_{dname}_qc{int(use_qc)}'] = out
                cases.append(dict(dtype=dname, method=method, qc=use_qc, **res))
                print(dname, method, use_qc, res['csr']['thread_invariant'],
                      res['csc']['thread_invariant'], flush=True)
    np.savez_compressed(os.path.join(HERE, 'ref_normalize.npz'), **full)
    path = os.path.join(HERE, 'ref_vectors.json')
    with open(path) as f:
        meta = json.load(f)
    meta['normalize'] = dict(seed=17, num_cells=n, num_genes=g, num_types=5,
                             density=0.10, qc_seed=17, qc_fraction=0.8,
                             threads=[8, 1], cases=cases)
    with open(path, 'w') as f:
        json.dump(meta, f, indent=1)


def calibrate_port(pca_mod, norm_mod, repeats=3, threads=8):
    """Time the reference's irlba and the oracle's C++ restatement of it
    (oracle.irlba_port) on the same 50k-cell input with the same thread
    count, best of `repeats` each; the CSC copy is built outside the timed
    call for the reference (as pca() does) and subtracted for the port.  The
    ratio port / reference calibrates the CPU baseline that bench.py times
    on the GPU host, where the reference does not run."""
    w = synth.make_workload(**synth.WORKLOADS['cells50k'])
    indptr, indices, counts = oracle.generate_counts(w)
    data = ref_normalize(norm_mod, counts, indices, indptr)
    ref_s, port_s = [], []
    for _ in range(repeats):
        _, _, t = ref_irlba(pca_mod, data, indices, indptr, w.num_cells,
                            w.num_genes, 50, 100, 1024, num_threads=threads)
        ref_s.append(t)
        t0 = time.perf_counter()
        _, _, st = oracle.irlba_port(data, indices, indptr, w.num_cells,
                                     w.num_genes, 50, num_threads=threads)
        port_s.append(time.perf_counter() - t0 - st['csc_seconds'])
    path = os.path.join(HERE, 'ref_vectors.json')
    with open(path) as f:
        meta = json.load(f)
    meta['port_calibration'] = dict(
        workload='cells50k', k=50, threads=threads, host_cpus=os.cpu_count(),
        reference_seconds=round(min(ref_s), 3), port_seconds=round(min(port_s), 3),
        port_over_reference=round(min(port_s) / min(ref_s), 3))
    with open(path, 'w') as f:
        json.dump(meta, f, indent=1)
    print(json.dumps(meta['port_calibration'], indent=1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--reference', default='/root/reference')
    ap.add_argument('--cells1m', action='store_true',
                    help='only add the 1M-cell fixture (ref_irlba_cells1m.npz)')
    ap.add_argument('--knn', action='store_true',
                    help='only add the neighbour-search fixtures (ref_knn.npz)')
    ap.add_argument('--knn-large', action='store_true',
                    help='only add the 1M-cell neighbour fixture and timing')
    ap.add_argument('--pipeline1m', action='store_true',
                    help='only add the 1M-cell normalize -> pca -> neighbors fixture')
    ap.add_argument('--knn-approx', action='store_true',
                    help='only add the few-clusters-searched neighbour fixtures')
    ap.add_argument('--rng', action='store_true',
                    help='only add the start-vector draws (ref_start_vector.npz)')
    ap.add_argument('--calibrate', action='store_true',
                    help='only time the reference against the oracle port (50k cells)')
    ap.add_argument('--normalize', action='store_true',
                    help='only add the normalisation fixtures (ref_normalize.npz)')
    ap.add_argument('--knn-large-barbar', action='store_true',
                    help='only time the 1M-cell search with k-means|| initialisation')
    ap.add_argument('--kmeans-barbar', action='store_true',
                    help='only add the k-means|| fixtures (ref_kmeans_barbar.npz)')
    args = ap.parse_args()
    build_reference(args.reference)
    pca_mod, norm_mod = import_reference()
    if args.rng:
        return rng_fixture(args.reference)
    if args.calibrate:
        return calibrate_port(pca_mod, norm_mod)
    if args.normalize:
        return normalize_fixture(norm_mod)
    if args.kmeans_barbar:
        return kmeans_barbar_fixtures()
    if args.knn_large_barbar:
        return knn_large_barbar()
    if args.cells1m:
        return cells1m(pca_mod, norm_mod)
    if args.knn:
        return knn_fixtures()
    if args.knn_large:
        return knn_large()
    if args.pipeline1m:
        return pipeline1m(pca_mod, norm_mod)
    if args.knn_approx:
        return knn_approx()

    tiny = {}
    meta = {'tiny': [], 'large': None,
            'note': 'outputs of the reference brisc irlba/normalize_csr'}
    for (name, seed, n, g, T, dens, k, m, chunk) in TINY:
        w = synth.make_workload(seed, n, num_types=T, num_genes=g,
                                target_density=dens)
        indptr, indices, counts = oracle.generate_counts(w)
        data = ref_normalize(norm_mod, counts, indices, indptr)
        PCs, conv, _ = ref_irlba(pca_mod, data, indices, indptr, n, g, k, m,
                                 chunk)
        PCs1, conv1, _ = ref_irlba(pca_mod, data, indices, indptr, n, g, k, m,
                                   chunk, num_threads=1)
        assert np.array_equal(PCs, PCs1), 'reference not thread-invariant'
        tiny[f'{name}_data'] = data
        tiny[f'{name}_indices'] = indices
        tiny[f'{name}_indptr'] = indptr
        tiny[f'{name}_pcs'] = PCs
        meta['tiny'].append(dict(name=name, seed=seed, num_cells=n,
                                 num_genes=g, num_types=T, density=dens, k=k,
                                 subspace_size=m, chunk_size=chunk,
                                 converged=conv,
                                 csr_sha256=synth.csr_sha256(data, indices,
                                                             indptr)))
        print(name, 'converged', conv, flush=True)
    np.savez_compressed(os.path.join(HERE, 'ref_irlba_tiny.npz'), **tiny)

    wname, k = LARGE
    w = synth.make_workload(**synth.WORKLOADS[wname])
    indptr, indices, counts = oracle.generate_counts(w)
    data = ref_normalize(norm_mod, counts, indices, indptr)
    PCs, conv, t8 = ref_irlba(pca_mod, data, indices, indptr, w.num_cells,
                              w.num_genes, k, 100, 1024, num_threads=8)
    PCs1, conv1, t1 = ref_irlba(pca_mod, data, indices, indptr, w.num_cells,
                                w.num_genes, k, 100, 1024, num_threads=1)
    rows = np.random.Generator(np.random.PCG64(1)).choice(
        w.num_cells, ROW_SAMPLE, replace=False)
    rows.sort()
    np.savez_compressed(
        os.path.join(HERE, f'ref_irlba_{wname}.npz'),
        singular_values=np.linalg.norm(PCs.astype(np.float64), axis=0),
        rows=rows, pcs_rows=PCs[rows])
    meta['large'] = dict(
        workload=wname, k=k, subspace_size=100, chunk_size=1024,
        tolerance=1e-6, max_iterations=100, seed=0, converged=conv,
        thread_invariant=bool(np.array_equal(PCs, PCs1)),
        nnz=int(len(data)), scale=w.scale,
        csr_sha256=synth.csr_sha256(data, indices, indptr),
        counts_sha256=synth.csr_sha256(counts, indices, indptr),
        seconds_8_threads=round(t8, 3), seconds_1_thread=round(t1, 3),
        host_cpus=os.cpu_count())
    with open(os.path.join(HERE, 'ref_vectors.json'), 'w') as f:
        json.dump(meta, f, indent=1)
    print(json.dumps(meta['large'], indent=1))


if __name__ == '__main__':
    main()
