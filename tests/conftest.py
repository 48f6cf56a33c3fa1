import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (HIP device)')
    config.addinivalue_line('markers', 'slow: takes more than a few seconds')


@pytest.fixture(scope='session')
def golden():
    import json
    import numpy as np
    here = os.path.join(ROOT, 'tests', 'golden')
    with open(os.path.join(here, 'ref_vectors.json')) as f:
        meta = json.load(f)
    tiny = dict(np.load(os.path.join(here, 'ref_irlba_tiny.npz')))
    large = dict(np.load(os.path.join(here, 'ref_irlba_cells50k.npz')))
    return meta, tiny, large


@pytest.fixture(scope='session')
def cells50k():
    """The 50k-cell workload, generated on the CPU and log1pPF-normalised."""
    from brisc_amd import synth
    import oracle
    w = synth.make_workload(**synth.WORKLOADS['cells50k'])
    data, indices, indptr = oracle.generate_csr(w)
    return w, data, indices, indptr
