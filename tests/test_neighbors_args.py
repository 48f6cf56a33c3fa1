"""SingleCell.neighbors() argument checks (host logic, no GPU): the
reference's checks and messages (brisc/single_cell.py:12336-12470)."""
import numpy as np
import pytest

from brisc_amd import SingleCell


def _sc(n=50, d=8, **obsm):
    X = np.random.default_rng(0).standard_normal((n, d)).astype(np.float32)
    return SingleCell(obs={'cell': np.arange(n),
                           'passed_QC': np.ones(n, bool)},
                      obsm={'pca': X, **obsm})


@pytest.mark.parametrize('kwargs,error,fragment', [
    (dict(PC_key='pcs'), ValueError, "PC_key 'pcs' is not a key of obsm"),
    (dict(num_neighbors=50), ValueError, 'less than the number of cells'),
    (dict(num_neighbors=0), ValueError, 'must be ≥ 1'),
    (dict(num_clusters=51), ValueError, 'num_clusters is 51'),
    (dict(num_clusters=4, num_clusters_searched=5), ValueError,
     'num_clusters_searched is 5'),
    (dict(num_init_iterations=3), ValueError,
     'num_init_iterations can only be specified'),
    (dict(oversampling_factor=2), ValueError,
     'oversampling_factor can only be specified'),
    (dict(kmeans_tolerance=0), ValueError, 'kmeans_tolerance'),
    (dict(num_kmeans_iterations=0), ValueError, 'num_kmeans_iterations'),
    (dict(chunk_size_search=51), ValueError, 'chunk_size_search is 51'),
    (dict(num_neighbors=2.0), TypeError, 'num_neighbors'),
    (dict(num_neighbors=49), ValueError, 'at most 256 are supported'),
])
def test_neighbors_argument_errors(kwargs, error, fragment):
    sc = _sc(n=50)
    if kwargs.get('num_neighbors') == 49:
        # the matrix-core tiles take at most 256 PCs
        sc = _sc(n=100, d=300)
        kwargs = dict(num_neighbors=65, num_clusters=1)
    with pytest.raises(error, match=fragment):
        sc.neighbors(**kwargs)


def test_neighbors_input_checks():
    sc = _sc()
    with pytest.raises(ValueError, match='already a key of obsm'):
        _sc(neighbors=np.zeros((50, 2))).neighbors()
    bad = SingleCell(obs={'cell': np.arange(50)},
                     obsm={'pca': np.zeros((50, 4), np.float64)})
    with pytest.raises(TypeError, match='must be float32'):
        bad.neighbors()
    nc = SingleCell(obs={'cell': np.arange(50)},
                    obsm={'pca': np.asfortranarray(
                        np.zeros((50, 4), np.float32))})
    with pytest.raises(ValueError, match='not C-contiguous'):
        nc.neighbors()
    with pytest.raises(ValueError, match='no cells pass QC'):
        sc.neighbors(QC_column=np.zeros(50, bool))
    one = SingleCell(obs={'cell': np.arange(1)},
                     obsm={'pca': np.zeros((1, 4), np.float32)})
    with pytest.raises(ValueError, match='only one cell'):
        one.neighbors()
