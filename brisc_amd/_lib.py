"""ctypes binding of libbrisc_pca.so (the C ABI in include/bpca.h).

The library is built in-tree (``brisc_amd/libbrisc_pca.so``).  There is no
fallback: if it is missing, loading raises ``ImportError`` and every compute
entry point fails loudly.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libbrisc_pca.so')

BPCA_OK = 0
BPCA_ERR_INVALID_ARGUMENT = 1
BPCA_ERR_OUT_OF_MEMORY = 2
BPCA_ERR_DEVICE = 3
BPCA_ERR_NONCANONICAL = 4
BPCA_ERR_NUMERICAL = 5
BPCA_ERR_CANCELLED = 6
BPCA_ERR_BUDGET = 7
BPCA_ERR_NOT_FOUND = 8


class BpcaError(RuntimeError):
    """A device-side failure reported by the library."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class BudgetError(BpcaError, MemoryError):
    """A request does not fit the context's HBM budget."""


class Problem(ctypes.Structure):
    """struct bpca_problem (include/bpca.h)."""
    _fields_ = [('data', ctypes.c_void_p),
                ('indices', ctypes.c_void_p),
                ('indptr', ctypes.c_void_p),
                ('indptr_width', ctypes.c_int32),
                ('validate', ctypes.c_int32),
                ('num_cells', ctypes.c_uint64),
                ('num_genes', ctypes.c_uint32),
                ('k', ctypes.c_uint32),
                ('subspace_size', ctypes.c_uint32),
                ('max_iterations', ctypes.c_uint32),
                ('tolerance', ctypes.c_float),
                ('pad', ctypes.c_uint32),
                ('seed', ctypes.c_uint64),
                ('PCs', ctypes.c_void_p)]


class SolveStats(ctypes.Structure):
    _fields_ = [('converged', ctypes.c_int32),
                ('restarts', ctypes.c_int32),
                ('lanczos_steps', ctypes.c_int64),
                ('matvecs', ctypes.c_int64),
                ('seconds', ctypes.c_double),
                ('svd_sweeps', ctypes.c_int32),
                ('slot', ctypes.c_int32),
                ('smax', ctypes.c_double),
                ('reorth_passes', ctypes.c_int32),
                ('pair_products', ctypes.c_int32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


KERNEL_CLASSES = ('spmv', 'spmv_t', 'reorth', 'vside', 'restart', 'final',
                  'setup')

_lib = None


def lib() -> ctypes.CDLL:
    """Load libbrisc_pca.so once and declare the signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f'{LIB_PATH} is not built; build it with '
            f'`make -C brisc_amd/csrc` (or __graft_entry__.build())')
    # One HIP runtime per process.  PyTorch-ROCm bundles its own
    # libamdhip64.so (SONAME libamdhip64.so.7); loading it first makes this
    # library bind to that same runtime instead of pulling in a second copy
    # from /opt/rocm, which would leave torch without a device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    p = ctypes.c_void_p
    i32, u32, u64, i64 = ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int64
    f32 = ctypes.c_float
    sig = {
        'bpca_last_error': (ctypes.c_char_p, []),
        'bpca_version': (i32, []),
        'bpca_context_create': (i32, [i32, u64, ctypes.POINTER(p)]),
        'bpca_context_destroy': (i32, [p]),
        'bpca_context_set_stream': (i32, [p, p]),
        'bpca_workspace_bytes': (u64, [u64, u32, u32, u32]),
        'bpca_context_cancel': (i32, [p]),
        'bpca_context_reserve': (i32, [p, u64, i32]),
        'bpca_context_release': (i32, [p, u64]),
        'bpca_context_budget': (i32, [p, ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_uint64)]),
        'bpca_start_vector': (i32, [u64, u32, p, p]),
        'bpca_csr_transpose': (i32, [p, p, p, i32, p, i32, i64, i64, p, p, p]),
        'bpca_normalize': (i32, [p, p, i32, p, i32, p, i32, i32, i64, i64, p,
                                 i32, p, p]),
        'bpca_test_log1pf': (i32, [p, p, i64, p]),
        'bpca_context_set_profiling': (i32, [p, i32]),
        'bpca_context_set_profiled_classes': (i32, [p, ctypes.c_uint32]),
        'bpca_context_kernel_times': (i32, [p, p, p]),
        'bpca_context_reset_kernel_times': (i32, [p]),
        'bpca_irlba_f32': (i32, [p, p, p, p, p, p, i32, u64, u32, u32, u32,
                                 f32, u32, u64, i32, i32, u32, u64, p,
                                 ctypes.POINTER(i32)]),
        'bpca_irlba_device': (i32, [p, p, p, p, i32, u64, u32, u32, u32, f32,
                                    u32, u64, i32, p, p,
                                    ctypes.POINTER(SolveStats)]),
        'bpca_irlba_device_noalloc': (i32, [p, p, p, p, i32, u64, u32, u32,
                                            u32, f32, u32, u64, p, p,
                                            ctypes.POINTER(SolveStats)]),
        'bpca_synth_row_nnz': (i32, [p, u64, u64, u64, u32, p, p, p, p]),
        'bpca_synth_fill': (i32, [p, u64, u64, u64, u32, p, p, p, p, p, p]),
        'bpca_normalize_log1ppf': (i32, [p, p, p, u64, p]),
        'bpca_csr_subset_count': (i32, [p, p, i32, p, i32, i64, i64, i64, p, i64, p, p]),
        'bpca_csr_subset_fill': (i32, [p, p, p, i32, p, i32, p, i64, p, i32, p, i64, p, p]),
        'bpca_knn_exact': (i32, [p, p, i64, i32, i32, p, p]),
        'bpca_knn_ivf': (i32, [p, p, i64, i32, i32, i32, i32, i32,
                               ctypes.c_float, i64, ctypes.c_uint64, p, p, p]),
        'bpca_kmeans': (i32, [p, p, i64, i32, i32, i32, ctypes.c_float, i64,
                              ctypes.c_uint64, p, p, p, p]),
        'bpca_kmeans_init_cells': (i32, [i64, i32, ctypes.c_uint64, p]),
        'bpca_knn_ivf_ex': (i32, [p, p, i64, i32, i32, i32, i32, i32,
                                  ctypes.c_float, i64, ctypes.c_uint64, i32, i32,
                                  ctypes.c_float, p, p, p]),
        'bpca_kmeans_ex': (i32, [p, p, i64, i32, i32, i32, ctypes.c_float, i64,
                                 ctypes.c_uint64, i32, i32, ctypes.c_float,
                                 p, p, p, p]),
        'bpca_kmeans_barbar_init': (i32, [p, p, i64, i32, i32, i32, ctypes.c_float,
                                          i64, ctypes.c_uint64, p, p]),
        'bpca_test_small_svd': (i32, [p, i32, p, p, p, p, p]),
        'bpca_test_gemm': (i32, [p, i64, i32, i32, p, p, p]),
        'bpca_test_gemm_device': (i32, [p, i64, i32, i32, p, p, p]),
        'bpca_test_fp64_rate': (i32, [p, i32, i32, ctypes.POINTER(ctypes.c_double)]),
        'bpca_context_set_kernel_config': (i32, [p, i32, i32]),
        'bpca_context_get_kernel_config': (i32, [p, ctypes.POINTER(ctypes.c_int32),
                                                 ctypes.POINTER(ctypes.c_int32)]),
        'bpca_context_set_occupancy': (i32, [p, i32, i32]),
        'bpca_context_set_sorted_xtu': (i32, [p, i32]),
        'bpca_context_get_sorted_xtu': (i32, [p, ctypes.POINTER(i32)]),
        'bpca_context_set_reorth_passes': (i32, [p, i32]),
        'bpca_context_set_pair_products': (i32, [p, i32]),
        'bpca_context_get_pair_products': (i32, [p, ctypes.POINTER(i32)]),
        'bpca_context_get_reorth_passes': (i32, [p, ctypes.POINTER(i32)]),
        'bpca_context_get_occupancy': (i32, [p, ctypes.POINTER(ctypes.c_int32),
                                             ctypes.POINTER(ctypes.c_int32)]),
        'bpca_autotune': (i32, [p, ctypes.POINTER(Problem), i32, p]),
        'bpca_context_info': (i32, [p, ctypes.POINTER(ctypes.c_int64),
                                    ctypes.POINTER(ctypes.c_int32),
                                    ctypes.POINTER(ctypes.c_int32)]),
        'bpca_irlba_batched': (i32, [p, i32, ctypes.POINTER(Problem),
                                     ctypes.POINTER(SolveStats), p, i32]),
        'bpca_engine_create': (i32, [p, ctypes.POINTER(p)]),
        'bpca_engine_destroy': (i32, [p]),
        'bpca_engine_add': (i32, [p, ctypes.POINTER(Problem),
                                  ctypes.POINTER(ctypes.c_int32)]),
        'bpca_engine_step': (i32, [p, i32, ctypes.POINTER(ctypes.c_int32)]),
        'bpca_engine_collect': (i32, [p, i32, p, ctypes.POINTER(SolveStats),
                                      p, i32, ctypes.POINTER(ctypes.c_int32)]),
        'bpca_engine_active': (i32, [p, ctypes.POINTER(ctypes.c_int32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def last_error() -> str:
    msg = lib().bpca_last_error()
    return msg.decode() if msg else ''


def check(rc: int) -> None:
    """Raise the Python exception the reference would raise for ``rc``."""
    if rc == BPCA_OK:
        return
    msg = last_error()
    if rc == BPCA_ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    if rc == BPCA_ERR_NONCANONICAL:
        raise ValueError(msg)
    if rc == BPCA_ERR_OUT_OF_MEMORY:
        raise MemoryError(msg)
    if rc == BPCA_ERR_CANCELLED:
        raise KeyboardInterrupt(msg)
    if rc == BPCA_ERR_BUDGET:
        raise BudgetError(rc, msg)
    raise BpcaError(rc, msg)
