"""X^T u from the gene-sorted copy of X (problems with at most 4096 genes).

The gene-sorted path and the CSR path sum each gene's block partial in
different fixed orders, so they agree to FP32 rounding, not bit for bit;
both are checked against the FP64 oracle (``oracle.fp64_pca`` through the
gene Gram, or ``oracle.fp64_pca_cells`` when there are fewer cells than
genes).  The inputs are built to reach the copy's edge cases: empty rows,
a whole empty 256-row sub-block, genes present in every row (runs longer
than one 1024-entry window), genes present nowhere, a gene count that is
not a multiple of 8, the largest gene count the copy takes (4096) next to
the smallest it leaves to the CSR (4097), and fewer rows than one
sub-block.  The sorted path is bitwise reproducible run to run and alone
versus in a batch that mixes both paths.

Tolerances: singular values within 1e-6 of the oracle and 2e-6 between the
paths; PCs with a > 1% gap within 1e-4 up to sign.
"""
import numpy as np
import pytest

import oracle
from brisc_amd import Context, irlba_batched, irlba_device

import helpers

pytestmark = pytest.mark.gpu

SV_RTOL = 1e-6
PATH_RTOL = 2e-6
COL_RTOL = 1e-4


def _torch():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    return torch


def _dev(*arrays):
    torch = _torch()
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrays]


def _matrix(seed, n, g, empty_block=False, dense_genes=3):
    """A seeded CSR with log-normal-ish values, per-gene densities spread
    over three decades, `dense_genes` genes present in every row, one gene
    present in no row, every 37th row empty and, optionally, rows 256..511
    empty."""
    rng = np.random.default_rng(seed)
    p = np.clip(rng.lognormal(np.log(0.05), 1.2, g), 0.0, 0.9)
    p[:dense_genes] = 1.0
    p[g // 2] = 0.0
    mask = rng.random((n, g)) < p
    mask[::37] = False
    if empty_block and n > 512:
        mask[256:512] = False
    rows, cols = np.nonzero(mask)
    vals = np.log1p(rng.gamma(2.0, 2.0, len(rows))).astype(np.float32)
    indptr = np.zeros(n + 1, np.int64)
    np.add.at(indptr, rows + 1, 1)
    indptr = np.cumsum(indptr)
    return vals, cols.astype(np.int32), indptr


def _oracle(data, indices, indptr, n, g, k):
    if n >= g:
        return oracle.fp64_pca(data, indices, indptr, n, g, k)
    return oracle.fp64_pca_cells(data, indices, indptr, n, g, k)


@pytest.mark.parametrize('n,g,k,empty_block', [
    (3000, 1999, 20, True),
    (1795, 600, 15, True),
    (200, 37, 10, False),
    (2500, 4096, 15, False),
    (2500, 4097, 15, False),
])
def test_sorted_path_against_csr_path_and_oracle(n, g, k, empty_block):
    data, indices, indptr = _matrix(n * 7 + g, n, g, empty_block)
    print(f'case sorted X^T u n={n} g={g} k={k} nnz={len(data)}', flush=True)
    d, i, p = _dev(data, indices, indptr.astype(np.int32))
    ctx_s = Context(0)
    ctx_c = Context(0)
    ctx_c.set_sorted_xtu(False)
    assert ctx_s.sorted_xtu() and not ctx_c.sorted_xtu()
    Ps, svs, sts = irlba_device(d, i, p, n, g, k, context=ctx_s)
    Pc, svc, stc = irlba_device(d, i, p, n, g, k, context=ctx_c)
    assert sts['converged'] == 1 and stc['converged'] == 1
    Ps, Pc = Ps.cpu().numpy(), Pc.cpu().numpy()
    if g > 4096:   # the copy is not made: both contexts take the CSR path
        assert np.array_equal(Ps.view(np.uint32), Pc.view(np.uint32))
    out = _oracle(data, indices, indptr, n, g, k)
    ref = out['singular_values']
    ws = helpers.well_separated(ref, 1e-2)
    for name, P, sv in (('sorted', Ps, svs), ('csr', Pc, svc)):
        err = np.max(np.abs(sv - ref) / ref)
        col = helpers.column_errors_up_to_sign(P[:, ws], out['PCs'][:, ws]).max() if len(ws) else 0.0
        print(f'  {name}: sv rel err {err:.2e}, separated PCs ({len(ws)}) {col:.2e}', flush=True)
        assert err < SV_RTOL and col < COL_RTOL
    between = np.max(np.abs(svs - svc) / svc)
    print(f'  sorted vs csr: sv {between:.2e}', flush=True)
    assert between < PATH_RTOL


def test_sorted_path_is_reproducible_alone_and_in_a_mixed_batch():
    torch = _torch()
    shapes = [(3000, 1999, 20), (2000, 5000, 10), (1100, 300, 12)]
    probs, host = [], []
    for j, (n, g, k) in enumerate(shapes):
        data, indices, indptr = _matrix(500 + j, n, g, empty_block=j == 0)
        d, i, p = _dev(data, indices, indptr)
        probs.append(dict(data=d, indices=i, indptr=p, num_cells=n, num_genes=g, k=k))
        host.append((data, indices, indptr))
    print('case sorted X^T u batch vs solo', flush=True)
    ctx = Context(0)
    solo = []
    for pb in probs:
        P, sv, st = irlba_device(pb['data'], pb['indices'], pb['indptr'], pb['num_cells'],
                                 pb['num_genes'], pb['k'], context=ctx)
        P2, _, _ = irlba_device(pb['data'], pb['indices'], pb['indptr'], pb['num_cells'],
                                pb['num_genes'], pb['k'], context=ctx)
        assert torch.equal(P.view(torch.int32), P2.view(torch.int32))
        solo.append(P.cpu().numpy())
    batched = irlba_batched(probs, context=ctx)
    for j, (P, sv, st) in enumerate(batched):
        assert st['converged'] == 1
        assert np.array_equal(P.cpu().numpy().view(np.uint32), solo[j].view(np.uint32)), j
