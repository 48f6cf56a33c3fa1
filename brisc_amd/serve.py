"""Request loop for many PCA requests on one GPU (continuous batching).

Requests arrive over time (a Poisson stream in the benchmark); each carries a
device-resident CSR and its ``num_PCs``.  The loop

* admits arrived requests in arrival order while they fit the context's HBM
  budget.  A request whose matrix the loop produces (``make``) is charged to
  the budget before it is materialised (its peak bytes during generation,
  then its resident CSR), and its solver workspace is charged at admission,
  so the budget bounds the device memory of every matrix the loop holds
  (``BudgetError`` means: wait until a running request finishes);
* advances all admitted requests together -- one Lanczos step per request
  per replay of the engine's macro-step graph -- so requests that arrive
  while others run join the same launches;
* retires finished requests as soon as they complete and records their
  latency (completion time minus arrival time).

Each request's result is bit-identical to solving it alone, because every
reduction in the kernels is keyed to the request's own data blocks.
``serve_sequential`` is the plain one-at-a-time GPU path over the same
stream, for comparison.
"""
from __future__ import annotations

import dataclasses
import time
from collections import deque

import numpy as np

from . import _lib
from .irlba import Context, Engine, irlba_device


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float            # seconds after the stream starts
    data: object              # device tensors (None until materialised)
    indices: object
    indptr: object
    num_cells: int
    num_genes: int
    k: int
    subspace_size: int = 100
    tolerance: float = 1e-6
    max_iterations: int = 100
    seed: int = 0
    # Optional producer of (data, indices, indptr) on the device, called when
    # the request arrives; the loop drops the request's matrix once it is
    # solved, so a long stream never holds every matrix at once.
    make: object = None
    # Device bytes ``make`` needs at its peak (an estimate is fine; the loop
    # charges the resident CSR's exact size once it exists).
    make_bytes: int = 0
    charged: int = 0          # bytes the loop has charged to the budget

    def materialise(self):
        if self.data is None and self.make is not None:
            self.data, self.indices, self.indptr = self.make()

    def resident_bytes(self) -> int:
        return sum(int(t.numel()) * t.element_size()
                   for t in (self.data, self.indices, self.indptr))

    def release(self):
        if self.make is not None:
            self.data = self.indices = self.indptr = None


@dataclasses.dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    finished: float
    PCs: object
    singular_values: np.ndarray
    stats: dict

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def summarize(results, makespan: float) -> dict:
    lat = np.array([r.latency for r in results]) * 1e3
    return {'requests': len(results),
            'solves_per_s': round(len(results) / makespan, 3),
            'makespan_s': round(makespan, 3),
            'p50_ms': round(float(np.percentile(lat, 50)), 2),
            'p99_ms': round(float(np.percentile(lat, 99)), 2),
            'mean_ms': round(float(lat.mean()), 2),
            'converged': int(sum(r.stats['converged'] for r in results))}


def serve(requests, context: Context | None = None, step_quantum: int = 16,
          max_active: int = 64, clock=time.perf_counter, progress=None):
    """Run the continuous-batching loop over ``requests`` (sorted by arrival).

    ``progress(done, total)`` is called as results complete.  Returns
    ``(results, makespan_seconds)``.
    """
    eng = Engine(context)
    ctx = eng.context
    queue = deque(sorted(requests, key=lambda r: r.arrival))
    total = len(queue)
    waiting = deque()
    results = []
    admitted_at = {}
    t0 = clock()
    try:
        while queue or waiting or eng.active:
            now = clock() - t0
            while queue and queue[0].arrival <= now:
                waiting.append(queue.popleft())
            while waiting and eng.active < max_active:
                r = waiting[0]
                if r.data is None and r.make is not None:
                    # Charge the matrix before it exists; only the head of
                    # the queue is ever materialised ahead of admission.
                    try:
                        ctx.reserve(r.make_bytes)
                    except _lib.BudgetError:
                        if not eng.active:
                            raise
                        break
                    r.charged = r.make_bytes
                    r.materialise()
                    ctx.release(r.charged)
                    r.charged = r.resident_bytes()
                    ctx.reserve(r.charged, force=True)
                if not _admit(eng, r):
                    break
                admitted_at[r.rid] = clock() - t0
                waiting.popleft()
            if eng.active:
                eng.step(step_quantum)
                done = eng.collect()
                t = clock() - t0
                for tag, PCs, sv, st in done:
                    results.append(Result(tag.rid, tag.arrival,
                                          admitted_at[tag.rid], t, PCs, sv,
                                          st))
                    tag.release()
                    if tag.charged:
                        ctx.release(tag.charged)
                        tag.charged = 0
                if done and progress:
                    progress(len(results), total)
            elif queue:
                delay = queue[0].arrival - (clock() - t0)
                if delay > 0:
                    time.sleep(delay)
        makespan = clock() - t0
    finally:
        eng.close()
        for r in list(waiting) + list(queue):
            if r.charged:
                ctx.release(r.charged)
                r.charged = 0
    return results, makespan


def _admit(eng, r) -> bool:
    """Add ``r`` to the engine; False when its workspace must wait for a
    running request to finish.  Device memory that torch keeps cached after
    finished requests is returned once before an allocation is given up."""
    for attempt in range(2):
        try:
            eng.add(r.data, r.indices, r.indptr, r.num_cells, r.num_genes,
                    r.k, subspace_size=r.subspace_size,
                    tolerance=r.tolerance, max_iterations=r.max_iterations,
                    seed=r.seed, validate=False, tag=r)
            return True
        except _lib.BudgetError:
            if not eng.active:
                raise
            return False
        except MemoryError:
            if attempt:
                if not eng.active:
                    raise
                return False
            import torch
            torch.cuda.empty_cache()
    return False


def serve_sequential(requests, context: Context | None = None,
                     clock=time.perf_counter, progress=None):
    """One request at a time, FIFO, on the plain single-solve GPU path.

    A request's matrix is materialised when it is taken up (after its
    arrival), so requests that queue up hold no device memory meanwhile.
    """
    queue = deque(sorted(requests, key=lambda r: r.arrival))
    total = len(queue)
    results = []
    t0 = clock()
    while queue:
        r = queue.popleft()
        now = clock() - t0
        if r.arrival > now:
            time.sleep(r.arrival - now)
        start = clock() - t0
        r.materialise()
        P, sv, st = irlba_device(r.data, r.indices, r.indptr, r.num_cells,
                                 r.num_genes, r.k,
                                 subspace_size=r.subspace_size,
                                 tolerance=r.tolerance,
                                 max_iterations=r.max_iterations, seed=r.seed,
                                 validate=False, context=context)
        results.append(Result(r.rid, r.arrival, start, clock() - t0, P, sv,
                              st))
        r.release()
        if progress:
            progress(len(results), total)
    return results, clock() - t0
