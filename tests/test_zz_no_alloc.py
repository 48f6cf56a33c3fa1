"""The allocation-free solve (bpca_irlba_device_noalloc) against the
allocating one (bpca_irlba_device), through the C ABI: the same PCs bit for
bit, singular values and statistics, on two calls in a row (the second
smaller, so stale scratch contents from the first would show), and a clear
refusal of a problem larger than the library's scratch array."""
import ctypes

import numpy as np
import pytest

import oracle
from brisc_amd import _lib, synth
from brisc_amd.irlba import Context

pytestmark = pytest.mark.gpu


def _solve(noalloc, ctx, d, i, p, n, g, k, m):
    """One solve through the C ABI; bpca_irlba_device takes one more
    argument (validate) before the outputs."""
    import torch
    L = _lib.lib()
    PCs = torch.empty((n, k), dtype=torch.float32, device='cuda')
    sv = np.zeros(k, dtype=np.float64)
    st = _lib.SolveStats()
    head = (ctx.handle, d.data_ptr(), i.data_ptr(), p.data_ptr(),
            p.element_size(), n, g, k, m, 1e-6, 100, 0)
    tail = (PCs.data_ptr(), sv.ctypes.data, ctypes.byref(st))
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    if noalloc:
        rc = L.bpca_irlba_device_noalloc(*head, *tail)
    else:
        rc = L.bpca_irlba_device(*head, 1, *tail)
    free1 = torch.cuda.mem_get_info()[0]
    return rc, PCs, sv, st, free0 - free1


def _device_csr(seed, n, types, wide):
    import torch
    w = synth.make_workload(seed, n, num_types=types)
    data, indices, indptr = oracle.generate_csr(w)
    return w, (torch.from_numpy(data).cuda(), torch.from_numpy(indices).cuda(),
               torch.from_numpy(indptr.astype(np.int64 if wide else np.int32)).cuda())


@pytest.mark.parametrize('passes', [0, 1])
def test_noalloc_matches_allocating_solve_twice(passes):
    import torch
    assert torch.cuda.is_available()
    ctx = Context(0)
    # passes = 1: one pass over U per step for n >= 8 g, so the 2000-gene
    # problems below also take X v and X^T u in one read of X (the padded
    # copy and segment table live in the scratch too).
    ctx.set_reorth_passes(passes)
    # First a larger problem, then a smaller one with other genes' counts.
    cases = [(501, 20_000, 20, 60, False), (502, 6_000, 12, 40, True)] if passes == 0 else \
        [(503, 20_000, 20, 60, False), (504, 17_000, 12, 40, True)]
    for seed, n, k, m, wide in cases:
        w, (d, i, p) = _device_csr(seed, n, 7, wide)
        print(f'case noalloc seed={seed} n={n} k={k} m={m} wide={wide}', flush=True)
        rc, P0, sv0, st0, used0 = _solve(True, ctx, d, i, p, n, w.num_genes, k, m)
        _lib.check(rc)
        rc, P1, sv1, st1, used1 = _solve(False, ctx, d, i, p, n, w.num_genes, k, m)
        _lib.check(rc)
        print('noalloc', st0.as_dict(), 'allocating', st1.as_dict(),
              'device bytes taken by the call: noalloc', used0, 'allocating', used1, flush=True)
        assert st0.converged == st1.converged == 1
        assert st0.pair_products == st1.pair_products == (1 if passes == 1 else 0)
        assert st0.restarts == st1.restarts and st0.lanczos_steps == st1.lanczos_steps
        assert np.array_equal(P0.cpu().numpy().view(np.uint32), P1.cpu().numpy().view(np.uint32))
        assert np.array_equal(sv0, sv1)
        # The call allocated no device memory: the workspace of this problem
        # (two FP64 n x m bases, 19 MB for the first) would have shown.  The
        # margin covers code objects the runtime loads at a first launch.
        assert used0 <= 4 << 20
    ctx.close()


def test_noalloc_refuses_problems_beyond_its_scratch():
    import torch
    ctx = Context(0)
    n, g = 300_000, 200
    indptr = torch.zeros(n + 1, dtype=torch.int64, device='cuda')
    d = torch.zeros(1, dtype=torch.float32, device='cuda')
    i = torch.zeros(1, dtype=torch.int32, device='cuda')
    print('case noalloc too large', flush=True)
    rc = _solve(True, ctx, d, i, indptr, n, g, 10, 100)[0]
    assert rc == _lib.BPCA_ERR_INVALID_ARGUMENT
    assert 'allocation-free solve needs' in _lib.last_error()
    ctx.close()
