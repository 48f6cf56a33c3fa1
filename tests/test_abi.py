"""The C ABI: the library loads, exports every declared entry point, and
rejects bad arguments before touching a device (runs without a GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

from brisc_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fname in sorted(os.listdir(os.path.join(ROOT, 'include'))):
        if not fname.endswith('.h'):
            continue
        text = open(os.path.join(ROOT, 'include', fname)).read()
        text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
        text = re.sub(r'//[^\n]*', '', text)
        names += re.findall(r'\b(bpca_[a-z0-9_]+)\s*\(', text)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert 'bpca_irlba_f32' in names and 'bpca_irlba_device' in names
    L = _lib.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_version():
    assert _lib.lib().bpca_version() >= 10000


def _call_irlba(n=10, g=5, k=2, m=4, tol=1e-6, maxit=10, width=4,
                match_parallel=0, threads=1, chunk=1024):
    idx_t = np.int32 if width == 4 else np.int64
    indptr = np.zeros(n + 1, dtype=idx_t)
    PCs = np.zeros((n, max(k, 1)), dtype=np.float32)
    conv = ctypes.c_int(0)
    rc = _lib.lib().bpca_irlba_f32(
        np.zeros(1, np.float32).ctypes.data, np.zeros(1, idx_t).ctypes.data,
        indptr.ctypes.data, None, None, None, width, n, g, k, m, tol, maxit,
        0, match_parallel, 1, threads, chunk, PCs.ctypes.data,
        ctypes.byref(conv))
    return rc, _lib.last_error()


@pytest.mark.parametrize('kwargs,fragment', [
    (dict(k=0), 'k must be a positive integer'),
    (dict(k=5, m=4), 'subspace_size must be >= k'),
    (dict(n=1, m=1, k=1), 'only one cell'),
    (dict(g=1, m=1, k=1), 'only one gene'),
    (dict(m=6), 'subspace_size must be <= min'),
    (dict(tol=0.0), 'tolerance must be positive'),
    (dict(maxit=0), 'max_iterations'),
    (dict(width=2), 'index_width must be 4 or 8'),
    (dict(match_parallel=1, threads=2), 'match_parallel must be False unless num_threads is 1'),
    (dict(chunk=0), 'chunk_size'),
    (dict(g=2**31 - 1, m=4), 'num_genes too large'),
])
def test_argument_errors(kwargs, fragment):
    rc, msg = _call_irlba(**kwargs)
    assert rc == _lib.BPCA_ERR_INVALID_ARGUMENT
    assert fragment in msg


def test_python_check_maps_codes():
    with pytest.raises(ValueError, match='only one cell'):
        _lib.check(_call_irlba(n=1, m=1, k=1)[0])


def test_workspace_bytes_grows_with_cells():
    L = _lib.lib()
    a = L.bpca_workspace_bytes(10_000, 2000, 50, 100)
    b = L.bpca_workspace_bytes(20_000, 2000, 50, 100)
    assert b > a > 2 * 8 * 10_000 * 100


def test_single_hip_runtime_in_process():
    """torch and libbrisc_pca.so share one libamdhip64 (two runtimes in one
    process leave torch without a device)."""
    _lib.lib()
    maps = open('/proc/self/maps').read()
    runtimes = set(re.findall(r'\S*libamdhip64\S*', maps))
    assert len(runtimes) == 1, runtimes


def test_start_vector_matches_reference_draws():
    """The IRLBA start vector (brisc/pca.pyx:801-806) for several seeds
    against the reference's own srand + random_normal draws
    (tests/golden/ref_start_vector.npz, written by make_ref_vectors.py
    --rng): the FP32 draws bit for bit, and the FP64 normalisation."""
    from brisc_amd.irlba import start_vector
    ref = np.load(os.path.join(ROOT, 'tests', 'golden', 'ref_start_vector.npz'))
    assert len(ref.files) >= 5
    for name in ref.files:
        seed = int(name.split('_', 1)[1])
        want = ref[name]
        draws, v0 = start_vector(seed, len(want))
        assert np.array_equal(draws.view(np.uint32), want.view(np.uint32)), name
        x = want.astype(np.float64)
        assert np.allclose(v0, x / np.sqrt(np.sum(x * x)), rtol=0, atol=1e-15)
        # the draws are standard normal (mean, variance) and a planted
        # change of the seed gives a different stream
        if len(want) >= 2000:
            assert abs(want.mean()) < 0.1 and abs(want.var() - 1) < 0.1
    d1, _ = start_vector(1, 300)
    assert not np.array_equal(d1, ref['seed_0'][:300])
