"""One-pass reorthogonalisation against two passes and the FP64 oracle.

A one-pass problem (``Context.set_reorth_passes(1)``, or automatically at
n >= max(8 g, 65536)) reads U once per Lanczos step: the step's update of
the new column is made by the next step's pass, and X^T U[:, j] is formed
in gene space.  The two schemes perform the same arithmetic per element in
different orders, so they agree to rounding, not bit for bit; both are
checked against the FP64 oracle (``oracle.fp64_pca``, the gene Gram).
Cases: an odd cell count whose last data block is partial (the element
path of the one-pass kernel), whole 1024-row blocks (its 16-byte path),
the CSR X^T u path (``set_sorted_xtu(False)``), n = 8 g exactly (all with
X v and X^T u in separate passes, ``set_pair_products(False)``), and a
problem with too few cells (n < 8 g) that must stay on two passes.  A
one-pass problem is bitwise reproducible run to run and alone versus in a
batch beside a two-pass problem.

Tolerances: singular values within 1e-6 of the oracle and 2e-6 between the
schemes; PCs with a > 1% singular-value gap within 1e-4 up to sign.
"""
import numpy as np
import pytest

import oracle
from brisc_amd import Context, irlba_batched, irlba_device, synth

import helpers

pytestmark = pytest.mark.gpu

SV_RTOL = 1e-6
SCHEME_RTOL = 2e-6
COL_RTOL = 1e-4


def _torch():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    return torch


def _dev(*arrays):
    torch = _torch()
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrays]


def _workload(seed, n, g):
    w = synth.make_workload(seed, n, num_types=8, num_genes=g, target_density=0.07)
    data, indices, indptr = oracle.generate_csr(w)
    return data, indices, indptr


@pytest.mark.parametrize('n,g,k,m,sorted_xtu', [
    (40001, 500, 10, 30, True),
    (65536, 2000, 20, 50, True),
    (24000, 3000, 15, 40, False),
    (16000, 2000, 12, 40, True),
])
def test_one_pass_against_two_passes_and_oracle(n, g, k, m, sorted_xtu):
    data, indices, indptr = _workload(n + g, n, g)
    print(f'case one-pass reorth n={n} g={g} k={k} m={m} sorted={sorted_xtu} '
          f'nnz={len(data)}', flush=True)
    d, i, p = _dev(data, indices, indptr)
    ctx1, ctx2 = Context(0), Context(0)
    for c in (ctx1, ctx2):
        c.set_sorted_xtu(sorted_xtu)
        # The separate products (the pair products are tests/test_gpu_pair.py's).
        c.set_pair_products(False)
    ctx1.set_reorth_passes(1)
    ctx2.set_reorth_passes(2)
    assert ctx1.reorth_passes() == 1 and ctx2.reorth_passes() == 2
    P1, sv1, st1 = irlba_device(d, i, p, n, g, k, subspace_size=m, context=ctx1)
    P2, sv2, st2 = irlba_device(d, i, p, n, g, k, subspace_size=m, context=ctx2)
    assert st1['reorth_passes'] == 1 and st2['reorth_passes'] == 2
    assert st1['converged'] == 1 and st2['converged'] == 1
    P1, P2 = P1.cpu().numpy(), P2.cpu().numpy()
    out = oracle.fp64_pca(data, indices, indptr, n, g, k)
    ref = out['singular_values']
    ws = helpers.well_separated(ref, 1e-2)
    for name, P, sv, st in (('one pass', P1, sv1, st1), ('two passes', P2, sv2, st2)):
        err = np.max(np.abs(sv - ref) / ref)
        err_cols = np.max(helpers.relative_sv_error(P, ref))
        col = helpers.column_errors_up_to_sign(P[:, ws], out['PCs'][:, ws]).max() if len(ws) else 0.0
        print(f'  {name}: restarts {st["restarts"]}, sv rel err {err:.2e}, PC norms {err_cols:.2e}, '
              f'separated PCs ({len(ws)}) {col:.2e}', flush=True)
        assert err < SV_RTOL and err_cols < SV_RTOL and col < COL_RTOL
    between = np.max(np.abs(sv1 - sv2) / sv2)
    print(f'  one vs two passes: sv {between:.2e}', flush=True)
    assert between < SCHEME_RTOL


def test_too_few_cells_stay_on_two_passes():
    n, g, k = 7999, 1000, 10
    data, indices, indptr = _workload(77, n, g)
    print(f'case one-pass refused n={n} g={g}', flush=True)
    d, i, p = _dev(data, indices, indptr)
    ctx1, ctx2 = Context(0), Context(0)
    ctx1.set_reorth_passes(1)
    ctx2.set_reorth_passes(2)
    P1, _, st1 = irlba_device(d, i, p, n, g, k, subspace_size=30, context=ctx1)
    P2, _, st2 = irlba_device(d, i, p, n, g, k, subspace_size=30, context=ctx2)
    assert st1['reorth_passes'] == 2
    assert np.array_equal(P1.cpu().numpy().view(np.uint32), P2.cpu().numpy().view(np.uint32))


def test_automatic_choice_and_batch_versus_solo():
    torch = _torch()
    shapes = [(70001, 1500, 15), (3000, 1000, 10), (66000, 9000, 12)]
    probs = []
    for j, (n, g, k) in enumerate(shapes):
        data, indices, indptr = _workload(900 + j, n, g)
        d, i, p = _dev(data, indices, indptr)
        probs.append(dict(data=d, indices=i, indptr=p, num_cells=n, num_genes=g, k=k,
                          subspace_size=40))
    print('case one-pass batch vs solo', flush=True)
    ctx = Context(0)
    assert ctx.reorth_passes() == 0
    expect = [1, 2, 2]   # 3000 < 65536; 66000 < 8 * 9000
    solo = []
    for pb, e in zip(probs, expect):
        P, sv, st = irlba_device(pb['data'], pb['indices'], pb['indptr'], pb['num_cells'],
                                 pb['num_genes'], pb['k'], subspace_size=40, context=ctx)
        P2, _, _ = irlba_device(pb['data'], pb['indices'], pb['indptr'], pb['num_cells'],
                                pb['num_genes'], pb['k'], subspace_size=40, context=ctx)
        assert st['reorth_passes'] == e
        assert torch.equal(P.view(torch.int32), P2.view(torch.int32))
        solo.append(P.cpu().numpy())
    batched = irlba_batched(probs, context=ctx)
    for j, (P, sv, st) in enumerate(batched):
        assert st['converged'] == 1 and st['reorth_passes'] == expect[j]
        assert np.array_equal(P.cpu().numpy().view(np.uint32), solo[j].view(np.uint32)), j
