"""Failure handling and admission of the batched / served path on the GPU.

* a solve cancelled from a host timer thread stops (``KeyboardInterrupt``, the
  reference's signal check per step, ``brisc/pca.pyx:1173-1174``), and the
  next solve on the same context runs to the same bits as before;
* a cancel issued while nothing runs does not touch a later solve;
* a request larger than the HBM budget is refused (``BudgetError``), and with
  a budget that holds one request at a time the serving loop runs them one
  after another, each result bit-identical to its solo solve and within
  1e-6 of the FP64 oracle's singular values.
"""
import threading
import time

import numpy as np
import pytest

import oracle
from brisc_amd import _lib, irlba_device, serve, synth
from brisc_amd.irlba import Context

pytestmark = pytest.mark.gpu

SV_RTOL = 1e-6   # vs the FP64 oracle (achieved: at most 1.4e-7)


def test_cancel_from_timer_thread_then_next_solve_completes():
    import torch
    assert torch.cuda.is_available()
    from brisc_amd import irlba_batched
    ctx = Context(0)
    w = synth.make_workload(61, 1_000_000, num_types=20)
    d, i, p = synth.generate_device(w, context=ctx)
    n, g = w.num_cells, w.num_genes
    print('case cancel: idle cancel, then a solve', flush=True)
    ctx.cancel()                     # nothing running: no effect later
    P0, sv0, st0 = irlba_device(d, i, p, n, g, 20, context=ctx)
    assert st0['converged'] == 1
    # A long call: eight 1M-cell solves with 50 PCs in one batch (seconds).
    specs = [dict(data=d, indices=i, indptr=p, num_cells=n, num_genes=g, k=50,
                  seed=s_) for s_ in range(8)]
    t0 = time.perf_counter()
    irlba_batched(specs[:1], context=ctx)
    one = time.perf_counter() - t0
    print(f'case cancel: timer thread during a batch of 8 (one solve {one:.2f} s)', flush=True)
    timer = threading.Timer(min(0.5, one), ctx.cancel)
    t0 = time.perf_counter()
    timer.start()
    with pytest.raises(KeyboardInterrupt, match='cancelled'):
        irlba_batched(specs, context=ctx)
    elapsed = time.perf_counter() - t0
    timer.join()
    print(f'cancelled after {elapsed:.2f} s', flush=True)
    assert elapsed < 4 * one + 0.5
    P1, sv1, st1 = irlba_device(d, i, p, n, g, 20, context=ctx)
    assert st1['converged'] == 1
    assert np.array_equal(P0.cpu().numpy().view(np.uint32), P1.cpu().numpy().view(np.uint32))
    ctx.close()


def test_budget_refuses_a_request_that_cannot_fit():
    import torch
    ctx = Context(0, hbm_budget_bytes=1 << 20)
    w = synth.make_workload(62, 5_000, num_types=5)
    data, indices, indptr = oracle.generate_csr(w)
    d = torch.from_numpy(data).cuda()
    i = torch.from_numpy(indices).cuda()
    p = torch.from_numpy(indptr).cuda()
    print('case budget refusal', flush=True)
    with pytest.raises(_lib.BudgetError):
        irlba_device(d, i, p, w.num_cells, w.num_genes, 10, context=ctx)
    assert ctx.budget()['charged'] == 0
    ctx.reserve(1 << 19)
    with pytest.raises(_lib.BudgetError):
        ctx.reserve(1 << 20)
    assert ctx.budget()['charged'] == 1 << 19
    ctx.release(1 << 19)
    ctx.close()


def test_budget_for_one_request_serialises_the_loop():
    L = _lib.lib()
    stream = synth.request_stream(4, rate=500.0, low=8_000, high=12_000,
                                  k_range=(8, 12), seed=91)
    m = 40
    need = [synth.generate_device_peak_bytes(w)
            + L.bpca_workspace_bytes(w.num_cells, w.num_genes, k, m)
            for _, w, k in stream]
    budget = int(1.15 * max(need))
    assert budget < min(need) * 2 * 0.95     # never room for two at once
    ctx = Context(0, hbm_budget_bytes=budget)
    print('case serialised serving, budget', budget, [w.num_cells for _, w, _ in stream], flush=True)
    reqs = [serve.Request(rid, a, None, None, None, w.num_cells, w.num_genes, k,
                          subspace_size=m,
                          make=(lambda w=w: synth.generate_device(w, context=ctx)),
                          make_bytes=synth.generate_device_peak_bytes(w))
            for rid, (a, w, k) in enumerate(stream)]
    res, span = serve.serve(reqs, context=ctx, step_quantum=4, max_active=8)
    assert len(res) == len(stream)
    assert ctx.budget()['charged'] == 0
    spans = sorted((r.admitted, r.finished, r.rid) for r in res)
    for (a0, f0, _), (a1, f1, _) in zip(spans, spans[1:]):
        assert a1 >= f0 - 1e-6, spans        # one at a time
    for r in res:
        _, w, k = stream[r.rid]
        d, i, p = synth.generate_device(w, context=ctx)     # the request's matrix again
        Ps, svs, sts = irlba_device(d, i, p, w.num_cells, w.num_genes, k,
                                    subspace_size=m, context=ctx)
        assert np.array_equal(r.PCs.cpu().numpy().view(np.uint32), Ps.cpu().numpy().view(np.uint32))
        csr = (d.cpu().numpy(), i.cpu().numpy(), p.cpu().numpy())
        ref = oracle.fp64_pca(*csr, w.num_cells, w.num_genes, k, want_pcs=False)['singular_values']
        err = np.max(np.abs(r.singular_values - ref) / ref)
        print(f'request {r.rid}: n={w.num_cells} k={k} sv rel err vs FP64 oracle {err:.2e}', flush=True)
        assert err < SV_RTOL
    ctx.close()

