"""Synthetic single-cell count workloads for the sparse-PCA path.

A workload is a cells x genes matrix of negative-binomial counts:

* ``num_genes`` genes with base means ``mu_g ~ LogNormal(-1, 1.5)``;
* ``num_types`` cell types with equal priors; each type multiplies a random
  10% of the genes by ``LogNormal(0, 1)``;
* a per-cell library-size factor ``s_c ~ LogNormal(0, 0.4)``;
* counts ``~ NB(mean = c * mu_g * f_{t,g} * s_c, var = m + 0.2 m^2)``, i.e. NB
  with size ``r = 5``.

The literal parameters give ~38% nonzeros; the stated ~7% density is reached
with a global mean scale ``c`` that is calibrated here by bisection on the
exact expected density (Gauss-Hermite quadrature over the library factor), so
``c`` is a deterministic function of the seed.

The per-gene, per-type and per-cell parameters are drawn on the host with
NumPy's PCG64 stream.  The per-entry draw is a counter-based function of
``(seed, cell, gene)`` (SplitMix64 finaliser -> uniform double -> NB
inversion) so the device generator (``synth.hip``) and the CPU generator used
by the tests produce identical integer counts for any cell range.

The counts then go through log1pPF normalisation (``brisc/normalize.pyx:9-128``
semantics) to give the CSR float32 / int32 matrix the PCA path consumes.
"""
from __future__ import annotations

import dataclasses
import hashlib

import numpy as np

NB_SIZE = 5.0                # 1 / dispersion 0.2
TARGET_DENSITY = 0.07
DEFAULT_NUM_GENES = 2000


@dataclasses.dataclass
class Workload:
    """Host-side parameters of one synthetic count matrix."""
    seed: int
    num_cells: int
    num_genes: int
    num_types: int
    scale: float                 # calibrated global mean scale c
    type_gene_mean: np.ndarray   # (num_types, num_genes) float64: c*mu_g*f_tg
    cell_type: np.ndarray        # (num_cells,) int32
    cell_factor: np.ndarray      # (num_cells,) float64 library factor s_c

    @property
    def expected_density(self) -> float:
        return expected_density(self.type_gene_mean, 1.0)


def _quadrature(points: int = 48):
    x, w = np.polynomial.hermite_e.hermegauss(points)
    return np.exp(0.4 * x), w / w.sum()


def expected_density(base_means: np.ndarray, c: float,
                     points: int = 48) -> float:
    """Expected nonzero fraction of NB(c*base*s, size 5) over s~LogNormal(0,.4)."""
    s, w = _quadrature(points)
    m = c * base_means[..., None] * s
    q = NB_SIZE / (NB_SIZE + m)
    return float(((1.0 - q ** 5) * w).sum(-1).mean())


def calibrate_scale(base_means: np.ndarray, target: float = TARGET_DENSITY,
                    iterations: int = 80, points: int = 48) -> float:
    """Bisection on log c so the expected density equals ``target``."""
    lo, hi = -20.0, 5.0
    for _ in range(iterations):
        mid = 0.5 * (lo + hi)
        if expected_density(base_means, float(np.exp(mid)), points) < target:
            lo = mid
        else:
            hi = mid
    return float(np.exp(0.5 * (lo + hi)))


def make_workload(seed: int, num_cells: int, num_types: int = 20,
                  num_genes: int = DEFAULT_NUM_GENES,
                  target_density: float = TARGET_DENSITY,
                  quick: bool = False) -> Workload:
    """Draw the parameters of one workload from ``seed``.

    ``quick`` calibrates the density scale with 30 bisection steps on an
    8-point quadrature (for streams of many small problems); the default is
    80 steps on 48 points.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    mu = np.exp(rng.normal(-1.0, 1.5, num_genes))
    fold = np.ones((num_types, num_genes))
    num_marked = max(1, num_genes // 10)
    for t in range(num_types):
        genes = rng.choice(num_genes, num_marked, replace=False)
        fold[t, genes] = np.exp(rng.normal(0.0, 1.0, num_marked))
    base = mu[None, :] * fold
    c = (calibrate_scale(base, target_density, 30, 8) if quick
         else calibrate_scale(base, target_density))
    type_gene_mean = np.ascontiguousarray(c * base, dtype=np.float64)
    cell_type = rng.integers(0, num_types, num_cells).astype(np.int32)
    cell_factor = np.exp(rng.normal(0.0, 0.4, num_cells)).astype(np.float64)
    return Workload(seed=seed, num_cells=num_cells, num_genes=num_genes,
                    num_types=num_types, scale=c,
                    type_gene_mean=type_gene_mean, cell_type=cell_type,
                    cell_factor=cell_factor)


def csr_sha256(data: np.ndarray, indices: np.ndarray,
               indptr: np.ndarray) -> str:
    """SHA-256 over the raw bytes of a CSR triple (data, indices, indptr)."""
    h = hashlib.sha256()
    for a in (data, indices, indptr):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


# Named workloads.  Sizes follow the sparse-PCA benchmark configurations:
# 50k x 2000 (CPU-runnable), 1M x 2000, 10M x 2000 with 100 types, and the
# 256-problem subclustering batch / 1000-request serving stream with 5
# subtypes per problem.
WORKLOADS = {
    'cells50k': dict(seed=50_000, num_cells=50_000, num_types=20),
    'cells1m': dict(seed=1_000_000, num_cells=1_000_000, num_types=20),
    'cells10m': dict(seed=10_000_000, num_cells=10_000_000, num_types=100),
}


def generate_device(w: Workload, device: int = 0, context=None,
                    keep_counts: bool = False, row_chunk: int = 1 << 21):
    """Generate and log1pPF-normalise workload ``w`` on the GPU.

    Returns ``(data float32, indices int32, indptr int64)`` device tensors
    (plus the int32 counts when ``keep_counts``).  Rows are generated in
    chunks of ``row_chunk`` so the per-row parameters never need more than a
    few tens of MB of device memory at once.
    """
    import ctypes

    import torch

    from . import _lib
    from .irlba import default_context

    ctx = context or default_context(device)
    L = _lib.lib()
    dev = torch.device('cuda', device)
    n, g = w.num_cells, w.num_genes
    tgm = torch.from_numpy(np.ascontiguousarray(w.type_gene_mean)).to(dev)
    row_nnz = torch.empty(n, dtype=torch.int64, device=dev)
    for r0 in range(0, n, row_chunk):
        r1 = min(n, r0 + row_chunk)
        ct = torch.from_numpy(w.cell_type[r0:r1]).to(dev)
        cf = torch.from_numpy(w.cell_factor[r0:r1]).to(dev)
        _lib.check(L.bpca_synth_row_nnz(
            ctx.handle, w.seed, r0, r1 - r0, g, tgm.data_ptr(),
            ct.data_ptr(), cf.data_ptr(), row_nnz[r0:r1].data_ptr()))
    indptr = torch.zeros(n + 1, dtype=torch.int64, device=dev)
    torch.cumsum(row_nnz, 0, out=indptr[1:])
    del row_nnz
    nnz = int(indptr[-1].item())
    indices = torch.empty(nnz, dtype=torch.int32, device=dev)
    counts = torch.empty(nnz, dtype=torch.int32, device=dev)
    for r0 in range(0, n, row_chunk):
        r1 = min(n, r0 + row_chunk)
        ct = torch.from_numpy(w.cell_type[r0:r1]).to(dev)
        cf = torch.from_numpy(w.cell_factor[r0:r1]).to(dev)
        _lib.check(L.bpca_synth_fill(
            ctx.handle, w.seed, r0, r1 - r0, g, tgm.data_ptr(),
            ct.data_ptr(), cf.data_ptr(), indptr[r0:].data_ptr(),
            indices.data_ptr(), counts.data_ptr()))
    data = torch.empty(nnz, dtype=torch.float32, device=dev)
    _lib.check(L.bpca_normalize_log1ppf(ctx.handle, counts.data_ptr(),
                                        indptr.data_ptr(), n,
                                        data.data_ptr()))
    torch.cuda.synchronize(dev)
    if keep_counts:
        return data, indices, indptr, counts
    del counts
    return data, indices, indptr


def generate_device_peak_bytes(w: Workload, slack: float = 1.1) -> int:
    """Device bytes :func:`generate_device` holds at its peak for ``w``
    (row counts and offsets, then indices, counts and values of the expected
    number of nonzeros, with ``slack`` for the draw's spread)."""
    n = w.num_cells
    nnz = n * w.num_genes * w.expected_density * slack
    return int(12 * nnz + 16 * (n + 1) + 8 * w.type_gene_mean.size)


def batch_sizes(num_problems: int = 256, low: int = 500, high: int = 200_000,
                seed: int = 256) -> np.ndarray:
    """Cells per problem, log-uniform on [low, high] (subclustering batch)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.exp(rng.uniform(np.log(low), np.log(high),
                              num_problems)).astype(np.int64)


def subcluster_batch(num_problems: int = 256, low: int = 500,
                     high: int = 200_000, seed: int = 256):
    """Workloads of the subclustering batch: log-uniform cell counts, 2000
    genes, 5 subtypes per problem, one seed per problem."""
    sizes = batch_sizes(num_problems, low, high, seed)
    return [make_workload(seed * 100_003 + i, int(n), num_types=5, quick=True)
            for i, n in enumerate(sizes)]


def request_stream(num_requests: int = 1000, rate: float = 8.0,
                   low: int = 1_000, high: int = 2_000_000,
                   k_range=(20, 50), seed: int = 1000):
    """Serving stream: Poisson arrivals at ``rate`` per second, cells
    log-uniform on [low, high], num_PCs uniform on k_range (inclusive).
    Returns a list of (arrival seconds, Workload, k)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, num_requests))
    sizes = np.exp(rng.uniform(np.log(low), np.log(high), num_requests))
    ks = rng.integers(k_range[0], k_range[1] + 1, num_requests)
    return [(float(a), make_workload(seed * 100_003 + i, int(n), num_types=5,
                                     quick=True), int(kk))
            for i, (a, n, kk) in enumerate(zip(arrivals, sizes, ks))]


def pc_mixture(seed: int, num_cells: int, num_dims: int, centers: int,
               spread: float = 0.5) -> np.ndarray:
    """Synthetic PC-space points for the neighbour search: a seeded FP32
    Gaussian mixture (``centers`` centres drawn N(0, 9) per dimension, each
    cell a centre plus N(0, spread^2) noise), regenerated bit for bit from
    the seed on any host."""
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((centers, num_dims)) * 3.0
    lab = rng.integers(0, centers, num_cells)
    return (C[lab] + spread * rng.standard_normal((num_cells, num_dims))
            ).astype(np.float32)
