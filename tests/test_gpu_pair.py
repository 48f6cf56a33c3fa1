"""X v and X^T u of a Lanczos step from one read of X, against the two
separate products and the FP64 oracle.

A one-pass problem with at most 4096 genes forms both sparse products in one
pass over X (``Context.set_pair_products``, on by default): X^T is applied
to the raw step vector less the mean of X vs, known before the pass from the
column sums of X, and the V side moves the result to X^T (u - mean of u) in
FP64.  Per gene the FP32 partial is the CSR path's chain over the rows in
order; the products agree with the separate ones to rounding, not bit for
bit.  Cases: a 2000-gene problem with a partial last data block, whole
blocks with few genes, and rows of three or more 256-element segments (the
path that re-reads a long row when its last segment is reached), mixed with
short rows, and empty rows (single and in a run).  The pair path is bitwise reproducible run to run and alone
versus in a batch beside a problem that takes the separate products.

Tolerances: singular values within 1e-6 of the oracle and 2e-6 between the
paths; PCs with a > 1% singular-value gap within 1e-4 up to sign.
"""
import numpy as np
import pytest

import oracle
from brisc_amd import Context, irlba_batched, irlba_device, synth

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


def _workload(seed, n, g, dense_every=0, empty_every=0):
    w = synth.make_workload(seed, n, num_types=8, num_genes=g, target_density=0.07)
    data, indices, indptr = oracle.generate_csr(w)
    if empty_every:
        # Every empty_every-th row emptied, and a run of 300 empty rows.
        keep = np.ones(n, bool)
        keep[::empty_every] = False
        keep[1000:1300] = False
        lens = np.diff(indptr) * keep
        mask = np.repeat(keep, np.diff(indptr))
        data, indices = data[mask], indices[mask]
        indptr = np.zeros(n + 1, np.int32)
        indptr[1:] = np.cumsum(lens)
    if not dense_every:
        return data, indices, indptr
    # Every dense_every-th row replaced by a row holding 40-100% of the genes
    # (3 to 12 segments of 256 elements).
    rng = np.random.default_rng(seed)
    rows_d, rows_i = [], []
    for r in range(n):
        if r % dense_every == 3:
            cnt = int(rng.integers(int(0.4 * g), g + 1))
            idx = np.sort(rng.choice(g, cnt, replace=False)).astype(np.int32)
            val = rng.uniform(0.05, 4.0, cnt).astype(np.float32)
        else:
            idx = indices[indptr[r]:indptr[r + 1]]
            val = data[indptr[r]:indptr[r + 1]]
        rows_i.append(idx)
        rows_d.append(val)
    lens = np.array([len(x) for x in rows_i], np.int64)
    indptr = np.zeros(n + 1, np.int32)
    indptr[1:] = np.cumsum(lens)
    return np.concatenate(rows_d), np.concatenate(rows_i), indptr


@pytest.mark.parametrize('n,g,k,m,dense_every,empty_every', [
    (70001, 2000, 20, 50, 0, 0),
    (65536, 400, 12, 40, 0, 0),
    (70000, 3000, 15, 40, 37, 0),
    (66001, 1000, 10, 40, 0, 13),
])
def test_pair_against_separate_products_and_oracle(n, g, k, m, dense_every, empty_every):
    data, indices, indptr = _workload(n + g, n, g, dense_every, empty_every)
    print(f'case pair products n={n} g={g} k={k} m={m} dense_every={dense_every} '
          f'empty_every={empty_every} nnz={len(data)}', flush=True)
    d, i, p = _dev(data, indices, indptr)
    ctx1, ctx2 = Context(0), Context(0)
    ctx2.set_pair_products(False)
    assert ctx1.pair_products() and not ctx2.pair_products()
    P1, sv1, st1 = irlba_device(d, i, p, n, g, k, subspace_size=m, context=ctx1)
    P2, sv2, st2 = irlba_device(d, i, p, n, g, k, subspace_size=m, context=ctx2)
    assert st1['pair_products'] == 1 and st2['pair_products'] == 0
    assert st1['reorth_passes'] == 1 and st2['reorth_passes'] == 1
    assert st1['converged'] == 1 and st2['converged'] == 1
    P1, P2 = P1.cpu().numpy(), P2.cpu().numpy()
    out = oracle.fp64_pca(data, indices, indptr, n, g, k)
    ref = out['singular_values']
    ws = helpers.well_separated(ref, 1e-2)
    for name, P, sv, st in (('pair', P1, sv1, st1), ('separate', P2, sv2, st2)):
        err = np.max(np.abs(sv - ref) / ref)
        err_cols = np.max(helpers.relative_sv_error(P, ref))
        col = helpers.column_errors_up_to_sign(P[:, ws], out['PCs'][:, ws]).max() if len(ws) else 0.0
        print(f'  {name}: restarts {st["restarts"]}, sv rel err {err:.2e}, PC norms {err_cols:.2e}, '
              f'separated PCs ({len(ws)}) {col:.2e}', flush=True)
        assert err < SV_RTOL and err_cols < SV_RTOL and col < COL_RTOL
    between = np.max(np.abs(sv1 - sv2) / sv2)
    print(f'  pair vs separate: sv {between:.2e}', flush=True)
    assert between < PATH_RTOL


def test_pair_reproducible_and_batch_versus_solo():
    torch = _torch()
    shapes = [(66000, 1200, 12, 29), (5000, 800, 10, 0), (68000, 2500, 10, 0)]
    probs = []
    for j, (n, g, k, dense) in enumerate(shapes):
        data, indices, indptr = _workload(500 + j, n, g, dense)
        d, i, p = _dev(data, indices, indptr)
        probs.append(dict(data=d, indices=i, indptr=p, num_cells=n, num_genes=g, k=k,
                          subspace_size=40))
    print('case pair batch vs solo', flush=True)
    ctx = Context(0)
    expect = [1, 0, 1]   # 5000 cells: two passes over U, separate products
    solo = []
    for pb, e in zip(probs, expect):
        P, sv, st = irlba_device(pb['data'], pb['indices'], pb['indptr'], pb['num_cells'],
                                 pb['num_genes'], pb['k'], subspace_size=40, context=ctx)
        P2, _, _ = irlba_device(pb['data'], pb['indices'], pb['indptr'], pb['num_cells'],
                                pb['num_genes'], pb['k'], subspace_size=40, context=ctx)
        assert st['pair_products'] == e
        assert torch.equal(P.view(torch.int32), P2.view(torch.int32))
        solo.append(P.cpu().numpy())
    batched = irlba_batched(probs, context=ctx)
    for j, (P, sv, st) in enumerate(batched):
        assert st['converged'] == 1 and st['pair_products'] == expect[j]
        assert np.array_equal(P.cpu().numpy().view(np.uint32), solo[j].view(np.uint32)), j
