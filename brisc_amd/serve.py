"""Request loop for many PCA requests on one GPU (continuous batching).

Requests arrive over time (a Poisson stream in the benchmark); each carries a
device-resident CSR and its ``num_PCs``.  The loop

* admits every arrived request whose workspace fits the context's HBM budget
  (``BudgetError`` from the engine means: wait until a running request
  finishes and frees memory);
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

    def materialise(self):
        if self.data is None and self.make is not None:
            self.data, self.indices, self.indptr = self.make()

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
                r = queue.popleft()
                r.materialise()
                waiting.append(r)
            while waiting and eng.active < max_active:
                r = waiting[0]
                try:
                    eng.add(r.data, r.indices, r.indptr, r.num_cells,
                            r.num_genes, r.k, subspace_size=r.subspace_size,
                            tolerance=r.tolerance,
                            max_iterations=r.max_iterations, seed=r.seed,
                            validate=False, tag=r)
                except _lib.BudgetError:
                    if not eng.active:
                        raise
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
                if done and progress:
                    progress(len(results), total)
            elif queue:
                delay = queue[0].arrival - (clock() - t0)
                if delay > 0:
                    time.sleep(delay)
        makespan = clock() - t0
    finally:
        eng.close()
    return results, makespan


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
