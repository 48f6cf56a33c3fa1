"""Python entry points of the IRLBA path.

``irlba`` is the drop-in for ``brisc.pca.irlba`` (``brisc/pca.pyx:661-678``):
same keyword arguments, same in-place fill of the caller's ``PCs`` array, same
``bool`` return (converged).  ``SingleCell.pca`` calls it by keyword
(``brisc/single_cell.py:12082-12093``), so binding this function to the name
``irlba`` in ``brisc.single_cell`` routes the reference's ``pca()`` through
the GPU.

``irlba_device`` is the device-resident variant for CSR already in HBM (torch
tensors on the GPU); it returns the PCs as a device tensor, the singular
values and the per-solve statistics.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def _as_host(a, dtype=None, name='array'):
    arr = np.asarray(a)
    if dtype is not None and arr.dtype != dtype:
        raise TypeError(f'{name} must have dtype {np.dtype(dtype).name}, got '
                        f'{arr.dtype.name}')
    if arr.ndim != 1:
        raise ValueError(f'{name} must be one-dimensional')
    return np.ascontiguousarray(arr)


def irlba(data_matvec, indices_matvec, indptr_matvec, data_rmatvec,
          indices_rmatvec, indptr_rmatvec, num_cells, num_genes, k,
          subspace_size, tolerance, max_iterations, seed, match_parallel,
          is_csr, num_threads, chunk_size, PCs) -> bool:
    """Top-k PCs of the implicitly centred and scaled matrix, on the GPU.

    Mirrors ``brisc.pca.irlba`` (``brisc/pca.pyx:661-678``).  The matvec
    operand is the CSR of X, or its CSC on the reference's serial CSC path
    (``is_csr=False, num_threads=1, match_parallel=False``; transposed on the
    device); the rmatvec operand is accepted and not used, because
    ``X^T u`` is computed from the CSR.  ``PCs`` (float32,
    C-contiguous, ``(num_cells, k)``) is filled in place with ``U_k S_k``.
    Returns whether all ``k`` singular values converged.
    """
    data = _as_host(data_matvec, np.float32, 'data_matvec')
    indices = _as_host(indices_matvec, name='indices_matvec')
    indptr = _as_host(indptr_matvec, name='indptr_matvec')
    if indices.dtype != indptr.dtype or indices.dtype not in (np.int32,
                                                                np.int64):
        raise TypeError('indices and indptr must both be int32 or both int64')
    for name, arr in (('indices_rmatvec', indices_rmatvec),
                      ('indptr_rmatvec', indptr_rmatvec)):
        if arr is not None and np.asarray(arr).dtype != indices.dtype:
            raise TypeError(f'{name} must have the same dtype as '
                            f'indices_matvec')
    if not isinstance(PCs, np.ndarray) or PCs.dtype != np.float32 or \
            not PCs.flags.c_contiguous or PCs.shape != (num_cells, k):
        raise TypeError('PCs must be a C-contiguous float32 array of shape '
                        '(num_cells, k)')
    # On the reference's serial CSC path the matvec operand is the CSC of X.
    csc = not is_csr and num_threads == 1 and not match_parallel
    outer = num_genes if csc else num_cells
    if len(indptr) != outer + 1:
        raise ValueError(f'indptr_matvec must have '
                         f'{"num_genes" if csc else "num_cells"} + 1 entries')
    if int(indptr[-1]) > len(data) or len(indices) < int(indptr[-1]):
        raise ValueError('indptr_matvec[-1] exceeds the number of stored '
                         'elements')
    conv = ctypes.c_int(0)
    rc = _lib.lib().bpca_irlba_f32(
        data.ctypes.data, indices.ctypes.data, indptr.ctypes.data, None, None,
        None, indices.dtype.itemsize, int(num_cells), int(num_genes), int(k),
        int(subspace_size), float(tolerance), int(max_iterations),
        int(seed) & 0xFFFFFFFFFFFFFFFF, int(bool(match_parallel)),
        int(bool(is_csr)), int(num_threads), int(chunk_size),
        PCs.ctypes.data, ctypes.byref(conv))
    _lib.check(rc)
    return bool(conv.value)


class Context:
    """A device, stream and HBM budget for device-resident solves.

    The context queues its work on the caller's current torch stream so that
    tensors produced by torch are ready when the kernels read them.
    """

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0,
                 use_torch_stream: bool = True):
        L = _lib.lib()
        handle = ctypes.c_void_p()
        _lib.check(L.bpca_context_create(int(device), int(hbm_budget_bytes),
                                         ctypes.byref(handle)))
        self._h = handle
        self.device = device
        self.hbm_budget_bytes = hbm_budget_bytes
        self._use_torch_stream = use_torch_stream

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._h is None:
            raise RuntimeError('context is closed')
        if self._use_torch_stream:
            import torch
            stream = torch.cuda.current_stream(self.device).cuda_stream
            _lib.check(_lib.lib().bpca_context_set_stream(self._h, stream))
        return self._h

    def set_profiling(self, enable: bool = True, classes=None) -> None:
        """Record HIP events around each kernel class of later solves, or
        only around the named classes (names of _lib.KERNEL_CLASSES)."""
        mask = 0xFFFFFFFF
        if classes is not None:
            mask = 0
            for c in classes:
                mask |= 1 << _lib.KERNEL_CLASSES.index(c)
        _lib.check(_lib.lib().bpca_context_set_profiled_classes(self._h, mask))
        _lib.check(_lib.lib().bpca_context_set_profiling(self._h,
                                                         int(enable)))

    def reset_kernel_times(self) -> None:
        _lib.check(_lib.lib().bpca_context_reset_kernel_times(self._h))

    def kernel_times(self) -> dict:
        """{class: (total milliseconds, launches)} since the last reset."""
        n = len(_lib.KERNEL_CLASSES)
        ms = np.zeros(n, dtype=np.float64)
        cnt = np.zeros(n, dtype=np.int64)
        _lib.check(_lib.lib().bpca_context_kernel_times(
            self._h, ms.ctypes.data, cnt.ctypes.data))
        return {name: (float(ms[i]), int(cnt[i]))
                for i, name in enumerate(_lib.KERNEL_CLASSES)}

    def info(self) -> dict:
        """Device facts: max LDS per workgroup, CUs, graph replay on/off."""
        lds = ctypes.c_int64(0)
        cus = ctypes.c_int32(0)
        graph = ctypes.c_int32(0)
        _lib.check(_lib.lib().bpca_context_info(self._h, ctypes.byref(lds),
                                                ctypes.byref(cus),
                                                ctypes.byref(graph)))
        return {'max_lds_per_workgroup': lds.value, 'num_cus': cus.value,
                'graph_replay': bool(graph.value)}

    def set_kernel_config(self, spmv_group: int, spmvt_group: int) -> None:
        """Rows per load group of X v / X^T u (2, 4 or 8; speed only)."""
        _lib.check(_lib.lib().bpca_context_set_kernel_config(
            self._h, int(spmv_group), int(spmvt_group)))

    def kernel_config(self) -> tuple:
        a = ctypes.c_int32(0)
        b = ctypes.c_int32(0)
        _lib.check(_lib.lib().bpca_context_get_kernel_config(
            self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_sorted_xtu(self, enable: bool) -> None:
        """Compute X^T u of problems admitted from now on with at most 4096
        genes from a gene-sorted copy of X (the default) or from the CSR.
        The two orders of summation agree to rounding, not bit for bit."""
        _lib.check(_lib.lib().bpca_context_set_sorted_xtu(self._h, int(bool(enable))))

    def sorted_xtu(self) -> bool:
        a = ctypes.c_int32(0)
        _lib.check(_lib.lib().bpca_context_get_sorted_xtu(self._h, ctypes.byref(a)))
        return bool(a.value)

    def set_reorth_passes(self, passes: int) -> None:
        """Reorthogonalisation of problems admitted from now on: 2 = two
        passes over U per Lanczos step (dots, then the update), 1 = one pass
        wherever num_cells >= 8 num_genes, 0 (default) = one pass where
        num_cells >= max(8 num_genes, 65536).  Results agree to rounding."""
        _lib.check(_lib.lib().bpca_context_set_reorth_passes(self._h, int(passes)))

    def reorth_passes(self) -> int:
        a = ctypes.c_int32(0)
        _lib.check(_lib.lib().bpca_context_get_reorth_passes(self._h, ctypes.byref(a)))
        return int(a.value)

    def set_pair_products(self, enable: bool) -> None:
        """Form X v and X^T u of a Lanczos step from one read of X for
        problems admitted from now on that reorthogonalise in one pass and
        have at most 4096 genes (the default), or in two passes.  Results
        agree to rounding."""
        _lib.check(_lib.lib().bpca_context_set_pair_products(self._h, int(bool(enable))))

    def pair_products(self) -> bool:
        a = ctypes.c_int32(0)
        _lib.check(_lib.lib().bpca_context_get_pair_products(self._h, ctypes.byref(a)))
        return bool(a.value)

    def set_occupancy(self, spmv_wg: int, spmvt_wg: int) -> None:
        """At most this many workgroups per compute unit for X v / X^T u
        (0 = no cap; speed only)."""
        _lib.check(_lib.lib().bpca_context_set_occupancy(
            self._h, int(spmv_wg), int(spmvt_wg)))

    def occupancy(self) -> tuple:
        a = ctypes.c_int32(0)
        b = ctypes.c_int32(0)
        _lib.check(_lib.lib().bpca_context_get_occupancy(
            self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def autotune(self, data, indices, indptr, num_cells, num_genes,
                 subspace_size=100, reps=5) -> dict:
        """Time every load-group size of X v and X^T u on this matrix, then
        workgroup-per-CU caps at the fastest size, and keep the fastest
        (results are bit-identical for every configuration)."""
        import torch
        scratch = torch.empty((num_cells, 1), dtype=torch.float32,
                              device=data.device)
        pb, _ = _problem(data, indices, indptr, num_cells, num_genes, 1,
                         subspace_size, 1e-6, 100, 0, True, scratch)
        t = np.zeros(6, dtype=np.float64)
        _lib.check(_lib.lib().bpca_autotune(self.handle, ctypes.byref(pb),
                                            int(reps), t.ctypes.data))
        groups = (2, 4, 8)
        return {'spmv_ms': dict(zip(groups, t[:3].round(4).tolist())),
                'spmv_t_ms': dict(zip(groups, t[3:].round(4).tolist())),
                'chosen': self.kernel_config(),
                'workgroups_per_cu': self.occupancy()}

    def cancel(self) -> None:
        """Ask the solves running on this context to stop within their next
        few steps (``KeyboardInterrupt`` in the calls advancing them); a
        solve started later is not affected."""
        _lib.check(_lib.lib().bpca_context_cancel(self._h))

    def reserve(self, nbytes: int, force: bool = False) -> None:
        """Charge ``nbytes`` held outside the library (an input CSR) to the
        HBM budget; ``BudgetError`` if it does not fit now (unless
        ``force``)."""
        _lib.check(_lib.lib().bpca_context_reserve(self._h, int(nbytes),
                                                   int(bool(force))))

    def release(self, nbytes: int) -> None:
        """Return bytes charged by :meth:`reserve`."""
        _lib.check(_lib.lib().bpca_context_release(self._h, int(nbytes)))

    def budget(self) -> dict:
        """{'budget', 'charged', 'cached'} in bytes (budget 0 = none)."""
        b, c, k = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _lib.check(_lib.lib().bpca_context_budget(
            self._h, ctypes.byref(b), ctypes.byref(c), ctypes.byref(k)))
        return {'budget': b.value, 'charged': c.value, 'cached': k.value}

    def close(self) -> None:
        if self._h is not None:
            _lib.lib().bpca_context_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


_default_ctx = {}


def default_context(device: int = 0) -> Context:
    ctx = _default_ctx.get(device)
    if ctx is None:
        ctx = _default_ctx[device] = Context(device)
    return ctx


def _check_device_csr(data, indices, indptr, num_cells):
    import torch
    if not (data.is_cuda and indices.is_cuda and indptr.is_cuda):
        raise ValueError('data, indices and indptr must be device tensors')
    if data.dtype != torch.float32:
        raise TypeError('data must be float32')
    if indices.dtype != torch.int32:
        raise TypeError('indices must be int32')
    if indptr.dtype not in (torch.int32, torch.int64):
        raise TypeError('indptr must be int32 or int64')
    for t, name in ((data, 'data'), (indices, 'indices'), (indptr, 'indptr')):
        if t.dim() != 1 or not t.is_contiguous():
            raise ValueError(f'{name} must be a contiguous 1-D tensor')
    if indptr.numel() != num_cells + 1:
        raise ValueError('indptr must have num_cells + 1 entries')


def irlba_device(data, indices, indptr, num_cells: int, num_genes: int,
                 k: int = 50, subspace_size: int = 100,
                 tolerance: float = 1e-6, max_iterations: int = 100,
                 seed: int = 0, validate: bool = True, PCs=None,
                 context: Context | None = None):
    """IRLBA on a device-resident CSR (torch tensors on the GPU).

    Returns ``(PCs, singular_values, stats)``: a ``(num_cells, k)`` float32
    device tensor of ``U_k S_k``, a float64 NumPy array of the ``k`` singular
    values and a dict of solve statistics (converged, restarts, matvecs, ...).
    ``subspace_size`` is clipped to ``min(num_cells, num_genes)`` as
    ``SingleCell.pca`` does (``single_cell.py:11957-11958``).
    """
    import torch
    _check_device_csr(data, indices, indptr, num_cells)
    subspace_size = min(int(subspace_size), int(num_cells), int(num_genes))
    ctx = context or default_context(data.device.index or 0)
    if PCs is None:
        PCs = torch.empty((num_cells, k), dtype=torch.float32,
                          device=data.device)
    elif PCs.shape != (num_cells, k) or PCs.dtype != torch.float32 or \
            not PCs.is_contiguous():
        raise ValueError('PCs must be a contiguous float32 (num_cells, k) '
                         'tensor')
    sv = np.zeros(k, dtype=np.float64)
    st = _lib.SolveStats()
    rc = _lib.lib().bpca_irlba_device(
        ctx.handle, data.data_ptr(), indices.data_ptr(), indptr.data_ptr(),
        indptr.element_size(), int(num_cells), int(num_genes), int(k),
        int(subspace_size), float(tolerance), int(max_iterations),
        int(seed) & 0xFFFFFFFFFFFFFFFF, int(bool(validate)), PCs.data_ptr(),
        sv.ctypes.data, ctypes.byref(st))
    _lib.check(rc)
    return PCs, sv, st.as_dict()


def _problem(data, indices, indptr, num_cells, num_genes, k, subspace_size,
             tolerance, max_iterations, seed, validate, PCs):
    import torch
    _check_device_csr(data, indices, indptr, num_cells)
    subspace_size = min(int(subspace_size), int(num_cells), int(num_genes))
    if PCs is None:
        PCs = torch.empty((num_cells, k), dtype=torch.float32,
                          device=data.device)
    elif PCs.shape != (num_cells, k) or PCs.dtype != torch.float32 or \
            not PCs.is_contiguous():
        raise ValueError('PCs must be a contiguous float32 (num_cells, k) '
                         'tensor')
    pb = _lib.Problem(
        data=data.data_ptr(), indices=indices.data_ptr(),
        indptr=indptr.data_ptr(), indptr_width=indptr.element_size(),
        validate=int(bool(validate)), num_cells=int(num_cells),
        num_genes=int(num_genes), k=int(k), subspace_size=subspace_size,
        max_iterations=int(max_iterations), tolerance=float(tolerance),
        pad=0, seed=int(seed) & 0xFFFFFFFFFFFFFFFF, PCs=PCs.data_ptr())
    return pb, PCs


def irlba_batched(problems, context: Context | None = None,
                  tolerance: float = 1e-6, max_iterations: int = 100,
                  seed: int = 0, validate: bool = True):
    """Solve many independent PCAs in shared launches.

    ``problems`` is a sequence of dicts with keys ``data``, ``indices``,
    ``indptr`` (device tensors), ``num_cells``, ``num_genes``, ``k`` and
    optionally ``subspace_size`` (default 100), ``tolerance``,
    ``max_iterations``, ``seed``.  Every problem advances one Lanczos step per
    launch sequence (a replayed HIP graph); each result is bit-identical to
    solving that problem alone.  Returns a list of ``(PCs, singular_values,
    stats)`` in input order.
    """
    if not problems:
        return []
    ctx = context or default_context(problems[0]['data'].device.index or 0)
    arr = (_lib.Problem * len(problems))()
    outs = []
    kmax = 1
    for i, p in enumerate(problems):
        pb, PCs = _problem(p['data'], p['indices'], p['indptr'],
                           p['num_cells'], p['num_genes'], p['k'],
                           p.get('subspace_size', 100),
                           p.get('tolerance', tolerance),
                           p.get('max_iterations', max_iterations),
                           p.get('seed', seed), p.get('validate', validate),
                           p.get('PCs'))
        arr[i] = pb
        outs.append(PCs)
        kmax = max(kmax, int(p['k']))
    stats = (_lib.SolveStats * len(problems))()
    sv = np.zeros((len(problems), kmax), dtype=np.float64)
    _lib.check(_lib.lib().bpca_irlba_batched(ctx.handle, len(problems), arr,
                                            stats, sv.ctypes.data, kmax))
    return [(outs[i], sv[i, :int(problems[i]['k'])].copy(),
             stats[i].as_dict()) for i in range(len(problems))]


class Engine:
    """Continuous-batching IRLBA engine on one device.

    ``add`` admits a problem whenever its workspace fits the context's HBM
    budget (``BudgetError`` otherwise); ``step`` advances every admitted
    problem; ``collect`` returns finished ones.  Problems admitted while
    others run share the same launches, and each result is bit-identical to
    solving the problem alone.
    """

    def __init__(self, context: Context | None = None, device: int = 0):
        self.context = context or default_context(device)
        handle = ctypes.c_void_p()
        _lib.check(_lib.lib().bpca_engine_create(self.context.handle,
                                                 ctypes.byref(handle)))
        self._h = handle
        self._live = {}   # slot -> (tag, PCs, tensors kept alive, k)

    def add(self, data, indices, indptr, num_cells, num_genes, k,
            subspace_size=100, tolerance=1e-6, max_iterations=100, seed=0,
            validate=True, PCs=None, tag=None) -> int:
        pb, PCs = _problem(data, indices, indptr, num_cells, num_genes, k,
                           subspace_size, tolerance, max_iterations, seed,
                           validate, PCs)
        slot = ctypes.c_int32(-1)
        self.context.handle  # order after the caller's torch stream
        _lib.check(_lib.lib().bpca_engine_add(self._h, ctypes.byref(pb),
                                              ctypes.byref(slot)))
        self._live[slot.value] = (tag, PCs, (data, indices, indptr), int(k))
        return slot.value

    def step(self, max_steps: int = 1 << 30) -> int:
        """Advance up to max_steps Lanczos steps; stops early when a problem
        finishes.  Returns the number of finished, uncollected problems."""
        done = ctypes.c_int32(0)
        _lib.check(_lib.lib().bpca_engine_step(self._h, int(max_steps),
                                               ctypes.byref(done)))
        return done.value

    def collect(self):
        """[(tag, PCs, singular_values, stats)] of finished problems."""
        n = len(self._live)
        if n == 0:
            return []
        kmax = max(v[3] for v in self._live.values())
        slots = np.zeros(n, dtype=np.int32)
        stats = (_lib.SolveStats * n)()
        sv = np.zeros((n, kmax), dtype=np.float64)
        got = ctypes.c_int32(0)
        _lib.check(_lib.lib().bpca_engine_collect(
            self._h, n, slots.ctypes.data, stats, sv.ctypes.data, kmax,
            ctypes.byref(got)))
        out = []
        for i in range(got.value):
            tag, PCs, _, k = self._live.pop(int(slots[i]))
            out.append((tag, PCs, sv[i, :k].copy(), stats[i].as_dict()))
        return out

    @property
    def active(self) -> int:
        return len(self._live)

    def close(self) -> None:
        if self._h is not None:
            _lib.lib().bpca_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def start_vector(seed: int, num_genes: int):
    """The IRLBA start vector for ``seed``: ``(draws, v0)`` -- the FP32
    normal draws of ``brisc/pca.pyx:801-806`` and the vector normalised in
    FP64 (host computation, no device)."""
    draws = np.empty(num_genes, dtype=np.float32)
    v0 = np.empty(num_genes, dtype=np.float64)
    _lib.check(_lib.lib().bpca_start_vector(int(seed) & 0xFFFFFFFFFFFFFFFF,
                                            int(num_genes), draws.ctypes.data,
                                            v0.ctypes.data))
    return draws, v0


def tall_gemm(A, C, context: Context | None = None):
    """A (n x m) @ C (m x kout) through the device's MFMA restart product
    (float64; m, kout <= 128)."""
    A = np.asfortranarray(A, dtype=np.float64)
    C = np.asfortranarray(C, dtype=np.float64)
    n, m = A.shape
    kout = C.shape[1]
    ctx = context or default_context(0)
    out = np.zeros((n, kout), order='F')
    _lib.check(_lib.lib().bpca_test_gemm(ctx.handle, n, m, kout,
                                         A.ctypes.data, C.ctypes.data,
                                         out.ctypes.data))
    return out


def small_svd(B, context: Context | None = None):
    """The device's restart SVD on one m x (m+1) matrix (float64).

    Returns ``(Ub, Sb, Vb, sweeps)`` with ``B[:, :m] = Ub diag(Sb) Vb^T``.
    """
    B = np.asfortranarray(B, dtype=np.float64)
    m = B.shape[0]
    if B.shape != (m, m + 1):
        raise ValueError('B must have shape (m, m+1)')
    ctx = context or default_context(0)
    Ub = np.zeros((m, m), order='F')
    Sb = np.zeros(m)
    Vb = np.zeros((m, m), order='F')
    sweeps = ctypes.c_int32(0)
    _lib.check(_lib.lib().bpca_test_small_svd(
        ctx.handle, m, B.ctypes.data, Ub.ctypes.data, Sb.ctypes.data,
        Vb.ctypes.data, ctypes.byref(sweeps)))
    return Ub, Sb, Vb, sweeps.value
