""".h5ad input and output for the PCA path (``libbrisc_h5ad.so``, the C ABI
in ``include/bpca_h5ad.h``).

* :func:`read_h5ad` -- a ``SingleCell`` from a file: X (``csr_matrix`` /
  ``csc_matrix`` group, or a dense dataset turned into a CSR, as
  ``brisc/single_cell.py:600-660`` does; ``X_key`` defaults to
  ``layers/UMIs``, then ``raw/X``, then ``X``), the obs and var dataframes
  (boolean, numeric, string and categorical columns; ``_index``) and obsm.
* :func:`stream_x_to_device` -- X straight to device tensors, read in slices
  into pinned host buffers whose copies to the device overlap the next read.
* :func:`write_h5ad` -- X, obs, var and obsm of a ``SingleCell`` in the same
  layout (``single_cell.py:3954``); :func:`write_obsm` -- one obsm array
  written back into an existing file (the PCs of ``pca()``).

The library is built in-tree; without it (or without the HDF5 library it
links) these functions raise ``ImportError``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libbrisc_h5ad.so')

INT, UINT, FLOAT, BOOL, STRING = 0, 1, 2, 3, 4
MISSING, GROUP, DATASET = 0, 1, 2

_L = None


def lib():
    global _L
    if _L is not None:
        return _L
    if not os.path.exists(LIB_PATH):
        raise ImportError(f'{LIB_PATH} is not built; build it with '
                          f'`make -C brisc_amd/csrc`')
    L = ctypes.CDLL(LIB_PATH)
    p, i32, i64, c = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_char_p
    P64 = ctypes.POINTER(ctypes.c_int64)
    PI = ctypes.POINTER(ctypes.c_int)
    sig = {
        'bh5_last_error': (c, []),
        'bh5_open': (i32, [c, i32, ctypes.POINTER(p)]),
        'bh5_create': (i32, [c, ctypes.POINTER(p)]),
        'bh5_close': (i32, [p]),
        'bh5_kind': (i32, [p, c]),
        'bh5_dataset_info': (i32, [p, c, PI, P64, PI, PI]),
        'bh5_read_rows': (i32, [p, c, i64, i64, p]),
        'bh5_read_strings': (i32, [p, c, p, i64, p, P64]),
        'bh5_attr_string': (i32, [p, c, c, p, i64]),
        'bh5_attr_int64': (i32, [p, c, c, p, i64, P64]),
        'bh5_list': (i32, [p, c, p, i64, P64]),
        'bh5_write_array': (i32, [p, c, p, i32, p, i32, i32, c]),
        'bh5_write_strings': (i32, [p, c, p, p, i64]),
        'bh5_make_group': (i32, [p, c, c]),
        'bh5_set_attr_string': (i32, [p, c, c, c]),
        'bh5_set_attr_strings': (i32, [p, c, c, p, p, i64]),
        'bh5_set_attr_int64': (i32, [p, c, c, p, i64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _L = L
    return L


def _check(rc: int) -> int:
    if rc < 0:
        raise OSError(lib().bh5_last_error().decode())
    return rc


_NP = {(INT, 1): np.int8, (INT, 2): np.int16, (INT, 4): np.int32, (INT, 8): np.int64,
       (UINT, 1): np.uint8, (UINT, 2): np.uint16, (UINT, 4): np.uint32, (UINT, 8): np.uint64,
       (FLOAT, 4): np.float32, (FLOAT, 8): np.float64, (BOOL, 1): np.bool_}


def _code(dtype) -> tuple:
    dtype = np.dtype(dtype)
    if dtype == np.bool_:
        return BOOL, 1
    kind = {'i': INT, 'u': UINT, 'f': FLOAT}.get(dtype.kind)
    if kind is None:
        raise TypeError(f'unsupported dtype {dtype}')
    return kind, dtype.itemsize


class H5File:
    """An open HDF5 file (read-only unless ``writable``; ``create`` makes a
    new, empty file)."""

    def __init__(self, path: str, writable: bool = False, create: bool = False):
        h = ctypes.c_void_p()
        if create:
            _check(lib().bh5_create(os.fsencode(path), ctypes.byref(h)))
        else:
            _check(lib().bh5_open(os.fsencode(path), int(writable), ctypes.byref(h)))
        self._h = h
        self.path = path

    def close(self) -> None:
        if self._h is not None:
            _check(lib().bh5_close(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- reading
    def kind(self, path: str) -> int:
        return _check(lib().bh5_kind(self._h, path.encode()))

    def __contains__(self, path: str) -> bool:
        return self.kind(path) != MISSING

    def info(self, path: str):
        """(shape tuple, numpy dtype or 'str')."""
        nd, cls, size = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        dims = (ctypes.c_int64 * 2)()
        _check(lib().bh5_dataset_info(self._h, path.encode(), ctypes.byref(nd), dims,
                                      ctypes.byref(cls), ctypes.byref(size)))
        shape = tuple(dims[:nd.value]) if nd.value else ()
        dtype = 'str' if cls.value == STRING else np.dtype(_NP[(cls.value, size.value)])
        return shape, dtype

    def read(self, path: str, first: int = 0, count: int | None = None,
             out: np.ndarray | None = None) -> np.ndarray:
        """Rows [first, first + count) of a numeric or boolean dataset (all
        rows by default), or a whole string dataset as an object array."""
        shape, dtype = self.info(path)
        if dtype == 'str':
            return self.read_strings(path)
        rows = shape[0] if shape else 1
        if count is None:
            count = rows - first
        tail = shape[1:] if len(shape) > 1 else ()
        if out is None:
            out = np.empty((count,) + tail if shape else (), dtype=dtype)
        _check(lib().bh5_read_rows(self._h, path.encode(), first, count,
                                   out.ctypes.data if out.size else None))
        return out

    def read_strings(self, path: str) -> np.ndarray:
        total = ctypes.c_int64()
        _check(lib().bh5_read_strings(self._h, path.encode(), None, 0, None,
                                      ctypes.byref(total)))
        shape, _ = self.info(path)
        n = shape[0] if shape else 1
        buf = ctypes.create_string_buffer(max(1, total.value))
        offs = np.zeros(n + 1, np.int64)
        _check(lib().bh5_read_strings(self._h, path.encode(), buf, total.value,
                                      offs.ctypes.data, ctypes.byref(total)))
        raw = buf.raw[:total.value]
        return np.array([raw[offs[i]:offs[i + 1]].decode('utf-8') for i in range(n)],
                        dtype=object)

    def attr(self, path: str, name: str):
        """A string attribute, or None when absent."""
        buf = ctypes.create_string_buffer(4096)
        rc = _check(lib().bh5_attr_string(self._h, path.encode(), name.encode(), buf, 4096))
        return None if rc == 1 else buf.value.decode('utf-8')

    def attr_ints(self, path: str, name: str):
        out = np.zeros(16, np.int64)
        count = ctypes.c_int64()
        rc = _check(lib().bh5_attr_int64(self._h, path.encode(), name.encode(),
                                         out.ctypes.data, 16, ctypes.byref(count)))
        return None if rc == 1 else out[:count.value]

    def members(self, path: str) -> list:
        need = ctypes.c_int64()
        _check(lib().bh5_list(self._h, path.encode(), None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value)
        _check(lib().bh5_list(self._h, path.encode(), buf, need.value, ctypes.byref(need)))
        return [s for s in buf.value.decode('utf-8').split('\n') if s]

    # -- writing
    def write(self, path: str, array, encoding_type: str | None = 'array') -> None:
        a = np.asarray(array)
        if a.dtype == object or a.dtype.kind in 'US':
            return self.write_strings(path, a)
        a = np.ascontiguousarray(a)
        if a.ndim not in (1, 2):
            raise ValueError('only 1-D and 2-D arrays are written')
        cls, size = _code(a.dtype)
        dims = np.array(a.shape, np.int64)
        _check(lib().bh5_write_array(self._h, path.encode(), a.ctypes.data if a.size else None,
                                     a.ndim, dims.ctypes.data, cls, size,
                                     encoding_type.encode() if encoding_type else None))

    @staticmethod
    def _pack(values):
        enc = [str(v).encode('utf-8') for v in values]
        offs = np.zeros(len(enc) + 1, np.int64)
        offs[1:] = np.cumsum([len(e) for e in enc])
        return b''.join(enc) or b'\0', offs

    def write_strings(self, path: str, values) -> None:
        buf, offs = self._pack(values)
        _check(lib().bh5_write_strings(self._h, path.encode(), buf, offs.ctypes.data,
                                       len(offs) - 1))

    def group(self, path: str, encoding_type: str | None = None) -> None:
        _check(lib().bh5_make_group(self._h, path.encode(),
                                    encoding_type.encode() if encoding_type else None))

    def set_attr(self, path: str, name: str, value) -> None:
        if isinstance(value, str):
            _check(lib().bh5_set_attr_string(self._h, path.encode(), name.encode(),
                                             value.encode('utf-8')))
        elif np.asarray(value).dtype.kind in 'iu':
            v = np.ascontiguousarray(value, np.int64).ravel()
            _check(lib().bh5_set_attr_int64(self._h, path.encode(), name.encode(),
                                            v.ctypes.data, len(v)))
        else:
            buf, offs = self._pack(value)
            _check(lib().bh5_set_attr_strings(self._h, path.encode(), name.encode(),
                                              buf, offs.ctypes.data, len(offs) - 1))


# ------------------------------------------------------------------ reading

def _default_x_key(f: H5File) -> str:
    """``layers/UMIs``, else ``raw/X``, else ``X`` (single_cell.py:609-621)."""
    has_umis, has_raw = 'layers/UMIs' in f, 'raw/X' in f
    if has_umis and has_raw:
        raise ValueError("both layers['UMIs'] and raw.X are present; this should "
                         "never happen in well-formed .h5ad files")
    return 'layers/UMIs' if has_umis else 'raw/X' if has_raw else 'X'


def x_layout(f: H5File, X_key: str):
    """('csr' | 'csc' | 'dense', shape) of the matrix at X_key."""
    if X_key not in f:
        raise ValueError(f'X_key {X_key!r} is not present in the .h5ad file')
    if f.kind(X_key) == DATASET:
        return 'dense', f.info(X_key)[0]
    enc = f.attr(X_key, 'encoding-type')
    if enc is None:
        fmt = f.attr(X_key, 'h5sparse_format')
        enc = None if fmt is None else fmt + '_matrix'
    if enc not in ('csr_matrix', 'csc_matrix'):
        raise ValueError(f"X has unsupported encoding-type {enc!r}, but should be "
                         f"'csr_matrix' or 'csc_matrix'")
    shape = f.attr_ints(X_key, 'shape')
    if shape is None:
        shape = f.attr_ints(X_key, 'h5sparse_shape')
    return enc[:3], tuple(int(s) for s in shape)


def read_x(f: H5File, X_key: str | None = None):
    """X as a scipy csr_array / csc_array (a dense X becomes a csr_array)."""
    import scipy.sparse as sp
    X_key = X_key or _default_x_key(f)
    layout, shape = x_layout(f, X_key)
    if layout == 'dense':
        return sp.csr_array(f.read(X_key))
    cls = sp.csr_array if layout == 'csr' else sp.csc_array
    return cls((f.read(f'{X_key}/data'), f.read(f'{X_key}/indices'),
                f.read(f'{X_key}/indptr')), shape=shape)


def read_dataframe(f: H5File, path: str, columns=None) -> dict:
    """A dataframe group as {'_index' name: index, column: values}: arrays,
    string arrays, categoricals (as category values) and nullable columns
    (masked entries None / NaN).  ``columns`` limits the columns read."""
    if path not in f:
        return {}
    index_name = f.attr(path, '_index') or '_index'
    names = [m for m in f.members(path) if m != index_name]
    out = {}
    if index_name in f.members(path):
        out[index_name] = f.read(f'{path}/{index_name}')
    for name in names:
        if columns is not None and name not in columns:
            continue
        sub = f'{path}/{name}'
        if f.kind(sub) == DATASET:
            out[name] = f.read(sub)
            continue
        enc = f.attr(sub, 'encoding-type')
        if enc == 'categorical':
            codes = f.read(f'{sub}/codes').astype(np.int64)
            cats = f.read(f'{sub}/categories')
            vals = np.asarray(cats)[np.clip(codes, 0, None)]
            if (codes < 0).any():
                vals = vals.astype(object)
                vals[codes < 0] = None
            out[name] = vals
        elif enc in ('nullable-integer', 'nullable-boolean'):
            vals = f.read(f'{sub}/values')
            mask = f.read(f'{sub}/mask').astype(bool)
            vals = vals.astype(object)
            vals[mask] = None
            out[name] = vals
        else:
            raise ValueError(f'{sub} has unsupported encoding-type {enc!r}')
    return out


def read_h5ad(path: str, X_key: str | None = None, obs_columns=None,
              var_columns=None, obsm: bool = True):
    """A ``SingleCell`` with X, obs, var and (optionally) obsm of the file."""
    from .single_cell import SingleCell
    with H5File(path) as f:
        X = read_x(f, X_key)
        obs = read_dataframe(f, 'obs', obs_columns)
        var = read_dataframe(f, 'var', var_columns)
        om = {}
        if obsm and 'obsm' in f:
            for key in f.members('obsm'):
                if f.kind(f'obsm/{key}') == DATASET:
                    om[key] = f.read(f'obsm/{key}')
    obs = obs or None
    var = var or None
    return SingleCell(X, obs=obs, var=var, obsm=om)


def stream_x_to_device(path: str, X_key: str | None = None, device: int = 0,
                       chunk_bytes: int = 256 << 20):
    """X of the file as device tensors (data in the file's type; indices and
    indptr as stored) with its layout and shape: ``(layout, shape, data,
    indices, indptr)``.  The values and indices are read in slices of about
    ``chunk_bytes`` into two pinned host buffers in turn; each slice's copy
    to the device runs on a side stream while the next slice is read."""
    import torch
    dev = torch.device('cuda', device)
    with H5File(path) as f:
        X_key = X_key or _default_x_key(f)
        layout, shape = x_layout(f, X_key)
        if layout == 'dense':
            raise ValueError('stream_x_to_device reads sparse X; a dense X is '
                             'read with read_x()')
        out = []
        stream = torch.cuda.Stream(device=dev)
        for name in ('data', 'indices', 'indptr'):
            dpath = f'{X_key}/{name}'
            (n,), dtype = f.info(dpath)
            tdtype = torch.from_numpy(np.empty(0, dtype)).dtype
            dst = torch.empty(n, dtype=tdtype, device=dev)
            step = max(1, chunk_bytes // np.dtype(dtype).itemsize)
            bufs = [torch.empty(min(step, max(n, 1)), dtype=tdtype).pin_memory()
                    for _ in range(2)]
            done = [None, None]
            for k, first in enumerate(range(0, n, step)):
                cnt = min(step, n - first)
                b = k & 1
                if done[b] is not None:
                    done[b].synchronize()       # the copy out of this buffer ended
                host = bufs[b][:cnt]
                f.read(dpath, first, cnt, out=host.numpy())
                with torch.cuda.stream(stream):
                    dst[first:first + cnt].copy_(host, non_blocking=True)
                    ev = torch.cuda.Event()
                    ev.record(stream)
                done[b] = ev
            stream.synchronize()
            out.append(dst)
    return (layout, shape) + tuple(out)


# ------------------------------------------------------------------ writing

def _write_dataframe(f: H5File, path: str, frame: dict, index_name: str) -> None:
    f.group(path, 'dataframe')
    cols = {k: np.asarray(v) for k, v in frame.items()}
    index = cols.pop(index_name, None)
    n = len(next(iter(cols.values()))) if cols else (len(index) if index is not None else 0)
    if index is None:
        index = np.arange(n).astype(str)
    f.set_attr(path, '_index', '_index')
    f.set_attr(path, 'column-order', list(cols) if cols else [])
    f.write_strings(f'{path}/_index', index)
    for name, values in cols.items():
        if values.dtype == object or values.dtype.kind in 'US':
            f.write_strings(f'{path}/{name}', values)
        else:
            f.write(f'{path}/{name}', values)


def write_h5ad(sc, path: str, overwrite: bool = False) -> None:
    """Write X (csr_matrix / csc_matrix group), obs, var and obsm of ``sc``
    to a new .h5ad file."""
    import scipy.sparse as sp
    from .single_cell import _columns
    if os.path.exists(path) and not overwrite:
        raise FileExistsError(f'{path} exists; set overwrite=True to replace it')
    with H5File(path, create=True) as f:
        f.set_attr('/', 'encoding-type', 'anndata')
        f.set_attr('/', 'encoding-version', '0.1.0')
        X = sc.X
        if X is not None:
            layout = 'csr' if isinstance(X, (sp.csr_array, sp.csr_matrix)) else 'csc'
            f.group('X', f'{layout}_matrix')
            f.set_attr('X', 'shape', np.array(X.shape, np.int64))
            f.write('X/data', X.data, None)
            f.write('X/indices', X.indices, None)
            f.write('X/indptr', X.indptr, None)
        obs = _columns(sc.obs)
        var = _columns(sc.var)
        obs_index = 'obs_names' if 'obs_names' in obs else next(iter(obs), None)
        var_index = next(iter(var), None)
        _write_dataframe(f, 'obs', obs, obs_index)
        _write_dataframe(f, 'var', var, var_index)
        f.group('obsm', 'dict')
        for key, value in sc.obsm.items():
            f.write(f'obsm/{key}', np.asarray(value))
        for name in ('varm', 'obsp', 'varp', 'layers', 'uns'):
            f.group(name, 'dict')


def write_obsm(path: str, key: str, array) -> None:
    """Write ``array`` as obsm[key] of an existing .h5ad file (replacing an
    array of that name): the PCs of pca() written back."""
    with H5File(path, writable=True) as f:
        if 'obsm' not in f:
            f.group('obsm', 'dict')
        f.write(f'obsm/{key}', np.ascontiguousarray(array))
