#!/usr/bin/env python3
"""Seconds to 50 PCs of a synthetic cells x 2000-gene matrix on one MI355X.

Default mode (the headline): one step = one full IRLBA solve (k=50,
subspace 100, tol 1e-6, seed 0) of the device-resident CSR of the named
workload (default: 10M cells x 2000 genes, 100 cell types, ~1.4B nonzeros),
inputs generated on the device and resident in HBM before the timed region.

--mode batched: the subclustering batch (256 problems, cells log-uniform on
[500, 200k], 5 subtypes, 30 PCs) solved in shared launches, against the same
problems solved one after another; reports solves/s for both.
--mode serve: a Poisson stream of PCA requests (cells log-uniform on
[1k, 2M], 20-50 PCs) through the continuous-batching request loop, against
the one-at-a-time GPU path on the same stream; reports p50/p99 latency and
solves/s for both.

The JSON line carries the solve time, the roofline of the dominant kernel
(algorithmic bytes per launch / its average launch time, from HIP events
recorded on the kernels' own stream during the timed steps; beside it the
measured HBM bytes per launch over the same time) and a CPU baseline: the
C++/OpenMP restatement of the reference's parallel IRLBA run on the host
cores for two iterations on a row sample of the same matrix, its time per
Lanczos step scaled to the full matrix by nonzeros and by the full solve's
step count.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (matrix and vector), published


def algorithmic_bytes(nnz: int, n: int, g: int) -> int:
    """One sparse product X.v or X^T.u: read each operand once, write each
    result once: 8 z (values + column indices) + 4 (n+1) (row offsets) +
    4 n (cell vector) + 8 g (gene vector, read and written)."""
    return 8 * nnz + 4 * (n + 1) + 4 * n + 8 * g


def load_traffic(workload: str, kernel: str):
    path = os.path.join(ROOT, 'profiles', 'traffic.json')
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    entry = table.get(workload, {}).get(kernel)
    return None if entry is None else entry.get('bytes_per_launch')


def port_calibration():
    """Port / reference time ratio measured on the development host (both on
    the same 50k-cell input and thread count; make_ref_vectors.py
    --calibrate), or None."""
    path = os.path.join(ROOT, 'tests', 'golden', 'ref_vectors.json')
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get('port_calibration')


def cpu_baseline(data, indices, indptr, nnz_full, g, k, rows, steps_full):
    """The oracle's C++/OpenMP restatement of the reference's parallel IRLBA
    (oracle.irlba_port) on the host cores: two IRLBA iterations on the first
    ``rows`` rows of the same matrix, timed without the CSC copy that the
    reference receives ready-made from pca(); seconds per Lanczos step (each
    step is one X v, one X^T u and the reorthogonalisations, plus its share
    of the restart) scaled to the full matrix by nonzeros and multiplied by
    the Lanczos steps of the full solve (``steps_full``, counted by the
    device solve of the same algorithm)."""
    import oracle
    rows = min(rows, len(indptr) - 1)
    e = int(indptr[rows])
    d = np.ascontiguousarray(data[:e])
    ix = np.ascontiguousarray(indices[:e])
    ip = np.ascontiguousarray(indptr[:rows + 1]).astype(np.int64)
    threads = int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))
    threads = max(1, min(threads, os.cpu_count() or 1))
    t0 = time.perf_counter()
    _, conv, st = oracle.irlba_port(d, ix, ip, rows, g, k, max_iterations=2,
                                    num_threads=threads)
    dt = time.perf_counter() - t0 - st['csc_seconds']
    steps = max(1, st['matvecs'] // 2)
    scale = nnz_full / max(e, 1)
    value = dt / steps * scale * steps_full
    calib = port_calibration()
    out = {'value': round(value, 3),
           'unit': 's (seconds per Lanczos step on a row sample, scaled by '
                   'nonzeros, times the full solve\'s Lanczos steps)',
           'cores': threads, 'kind': 'port',
           'sample': f'first {rows} rows ({e} nnz, {100 * e / nnz_full:.0f}% of the '
                     f'nonzeros), 2 IRLBA iterations k={k}: {steps} Lanczos steps '
                     f'in {dt:.2f} s (+{st["csc_seconds"]:.2f} s CSC copy, not '
                     f'counted); x{scale:.2f} by nnz, x{steps_full} steps'}
    if calib:
        out['port_over_reference'] = calib['port_over_reference']
        out['calibration'] = (f'port / reference = {calib["port_over_reference"]} on '
                              f'{calib["workload"]} with {calib["threads"]} threads '
                              f'(development host, {calib["host_cpus"]} CPUs)')
    return out


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def bench_batched(args):
    import torch
    from brisc_amd import Context, irlba_batched, irlba_device, synth

    ctx = Context(0)
    t0 = time.perf_counter()
    works = synth.subcluster_batch(args.problems)
    mats = [synth.generate_device(w, context=ctx) for w in works]
    log(f'# batch of {len(works)} problems, cells {sum(w.num_cells for w in works)} '
        f'(min {min(w.num_cells for w in works)}, max {max(w.num_cells for w in works)}), '
        f'nnz {sum(int(m[0].numel()) for m in mats)}, prepared in {time.perf_counter() - t0:.1f} s')
    specs = [dict(data=d, indices=i, indptr=p, num_cells=w.num_cells,
                  num_genes=w.num_genes, k=args.batch_k, subspace_size=100,
                  validate=False)
             for w, (d, i, p) in zip(works, mats)]
    irlba_batched(specs[:4], context=ctx)          # warm-up (code objects, pool)
    torch.cuda.synchronize()
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        out = irlba_batched(specs, context=ctx)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        log(f'# batched: {times[-1]:.3f} s')
    batched_s = min(times)
    if args.no_sequential:
        print(json.dumps({'metric': 'batched PCA solves/s', 'value': round(len(specs) / batched_s, 3),
                          'unit': 'solves/s', 'batched_seconds': round(batched_s, 4),
                          'problems': len(specs)}), flush=True)
        return
    t0 = time.perf_counter()
    seq = [irlba_device(sp['data'], sp['indices'], sp['indptr'], sp['num_cells'],
                        sp['num_genes'], sp['k'], subspace_size=100,
                        validate=False, context=ctx) for sp in specs]
    torch.cuda.synchronize()
    seq_s = time.perf_counter() - t0
    log(f'# sequential: {seq_s:.3f} s')
    same = sum(int(np.array_equal(a[0].cpu().numpy().view(np.uint32),
                                  b[0].cpu().numpy().view(np.uint32)))
               for a, b in zip(out, seq))
    conv = sum(int(o[2]['converged']) for o in out)
    steps = [o[2]['lanczos_steps'] for o in out]
    line = {
        'metric': 'batched PCA solves/s (256 subclustering problems)',
        'value': round(len(specs) / batched_s, 3), 'unit': 'solves/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': 1,
        'ms_per_step': round(batched_s * 1e3, 2), 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32 spmv / f64 lanczos', 'data': 'synthetic',
        'config': {'workload': 'subcluster_batch', 'problems': len(specs),
                   'cells_low': 500, 'cells_high': 200_000, 'num_types': 5,
                   'k': args.batch_k, 'subspace_size': 100},
        'batched_seconds': round(batched_s, 4),
        'sequential_seconds': round(seq_s, 4),
        'sequential_solves_per_s': round(len(specs) / seq_s, 3),
        'speedup_vs_sequential': round(seq_s / batched_s, 3),
        'bitwise_equal_to_sequential': f'{same}/{len(specs)}',
        'converged': f'{conv}/{len(specs)}',
        'lanczos_steps_max': int(max(steps)), 'lanczos_steps_sum': int(sum(steps)),
    }
    print(json.dumps(line), flush=True)


def bench_serve(args):
    """The serving stream: continuous batching against the one-at-a-time GPU
    path on the same Poisson stream, and again on a prefix of the stream at
    an arrival rate the one-at-a-time path can sustain."""
    import hashlib

    import torch
    from brisc_amd import Context, serve, synth

    free = torch.cuda.mem_get_info()[0]
    nominal = int(args.hbm_budget_gb * 1e9)
    reserve = 4 << 30   # torch's cached blocks and the engine's small tables
    budget = min(nominal, free - reserve)
    ctx = Context(0, hbm_budget_bytes=budget)
    t0 = time.perf_counter()
    stream = synth.request_stream(args.requests, args.rate, high=args.max_cells)

    def maker(w):
        return lambda: synth.generate_device(w, context=ctx)

    def requests(items, rate_scale=1.0):
        # Each request's CSR is generated on the device when it arrives and
        # dropped once it is solved (the whole stream would not fit in HBM).
        return [serve.Request(rid, arrival * rate_scale, None, None, None,
                              w.num_cells, w.num_genes, k, make=maker(w),
                              make_bytes=synth.generate_device_peak_bytes(w))
                for rid, (arrival, w, k) in enumerate(items)]

    cells = [w.num_cells for _, w, _ in stream]
    log(f'# {len(stream)} requests, cells {sum(cells)} (max {max(cells)}), '
        f'rate {args.rate}/s, budget {budget / 1e9:.1f} GB, parameters drawn in '
        f'{time.perf_counter() - t0:.1f} s')
    serve.serve(requests(stream[:2]), context=ctx)                 # warm-up

    def progress(tag):
        t_start = time.perf_counter()

        def report(done, total):
            if done % 100 == 0 or done == total:
                log(f'#   {tag}: {done}/{total} solved, {time.perf_counter() - t_start:.1f} s')
        return report

    def digests(results):
        out = {}
        for r in results:
            out[r.rid] = hashlib.sha256(r.PCs.cpu().numpy().tobytes()).hexdigest()
            r.PCs = None
        return out

    def both(items, rate_scale, tag):
        res_b, span_b = serve.serve(requests(items, rate_scale), context=ctx,
                                    progress=progress(f'{tag} continuous batching'))
        sb = serve.summarize(res_b, span_b)
        log(f'# {tag} continuous batching: {sb}')
        nonconv = [dict(rid=r.rid, num_cells=items[r.rid][1].num_cells,
                        k=items[r.rid][2], restarts=r.stats['restarts'])
                   for r in res_b if not r.stats['converged']]
        dig_b = digests(res_b)
        torch.cuda.empty_cache()
        if args.no_sequential:
            return sb, None, None, nonconv
        res_s, span_s = serve.serve_sequential(requests(items, rate_scale), context=ctx,
                                               progress=progress(f'{tag} one at a time'))
        ss = serve.summarize(res_s, span_s)
        log(f'# {tag} one at a time: {ss}')
        dig_s = digests(res_s)
        torch.cuda.empty_cache()
        same = sum(int(dig_b[rid] == dig_s.get(rid)) for rid in dig_b)
        return sb, ss, f'{same}/{len(dig_b)} (SHA-256 of each PCs array)', nonconv

    sb, ss, same, nonconv = both(stream, 1.0, 'full stream')
    line = {
        'metric': 'serving PCA requests: p50/p99 ms and solves/s',
        'value': sb['solves_per_s'], 'unit': 'solves/s', 'n_gpus': 1,
        'steps': 1, 'warmup': 1, 'ms_per_step': sb['makespan_s'] * 1e3,
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32 spmv / f64 lanczos', 'data': 'synthetic',
        'config': {'workload': 'request_stream', 'requests': len(stream),
                   'rate_per_s': args.rate, 'cells_low': 1000,
                   'cells_high': args.max_cells, 'k_range': [20, 50],
                   'hbm_budget_gb_nominal': args.hbm_budget_gb,
                   'hbm_budget_bytes': budget,
                   'budget_note': f'nominal budget, less what is not free on the device and a '
                                  f'{reserve >> 30} GiB reserve'},
        'continuous_batching': sb, 'one_at_a_time': ss,
        'bitwise_equal_to_one_at_a_time': same,
        'not_converged': nonconv,
    }
    if args.compare_requests > 0 and not args.no_sequential:
        items = stream[:args.compare_requests]
        scale = args.rate / args.compare_rate
        cb, cs, csame, _ = both(items, scale, f'{args.compare_rate}/s prefix')
        line['sustainable_rate_comparison'] = {
            'requests': len(items), 'rate_per_s': args.compare_rate,
            'continuous_batching': cb, 'one_at_a_time': cs,
            'bitwise_equal_to_one_at_a_time': csame}
    print(json.dumps(line), flush=True)


def bench_subset(args):
    """Device QC x HVG subset (operand preparation of SingleCell.pca) on the
    10M-cell matrix: 90% of cells pass QC, 1000 of the 2000 genes are HVGs.
    Algorithmic bytes: the selected rows' values, indices and offsets read
    once, every column index looked up once, the subset written once."""
    import scipy.sparse as sp
    import torch
    from brisc_amd import Context, synth
    from brisc_amd.subset import csr_subset

    spec = synth.WORKLOADS[args.workload]
    w = synth.make_workload(**spec)
    ctx = Context(0)
    data, indices, indptr = synth.generate_device(w, context=ctx)
    n, g, nnz = w.num_cells, w.num_genes, int(data.numel())
    rng = np.random.default_rng(7)
    qc = rng.random(n) < 0.9
    hvg = np.zeros(g, bool)
    hvg[rng.choice(g, g // 2, replace=False)] = True
    rows = np.flatnonzero(qc)
    d_qc = torch.from_numpy(qc).cuda()       # QC mask resident on the device
    out = csr_subset(data, indices, indptr, g, rows=d_qc, genes=hvg, context=ctx)
    torch.cuda.synchronize()
    times = []
    for _ in range(args.steps):
        del out
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = csr_subset(data, indices, indptr, g, rows=d_qc, genes=hvg, context=ctx)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    sec = min(times)
    out_nnz = int(out[0].numel())
    sel_nnz = int((indptr[1:] - indptr[:-1])[torch.from_numpy(rows).cuda()].sum().item())
    # read: selected rows' data + indices (4+4 B) for the fill, indices again
    # for the count, offsets and row list; write: subset data + indices
    # (4+4 B), offsets and per-row counts.
    alg = (8 * sel_nnz + 4 * sel_nnz + 16 * rows.size + 8 * rows.size
           + 8 * out_nnz + 16 * (rows.size + 1))
    # CPU: SciPy's X[rows][:, genes] on a 1M-row sample of the same matrix.
    sample = min(1_000_000, n)
    e = int(indptr[sample].item())
    Xs = sp.csr_array((data[:e].cpu().numpy(), indices[:e].cpu().numpy(),
                       indptr[:sample + 1].cpu().numpy()), shape=(sample, g))
    rs = np.flatnonzero(qc[:sample])
    t0 = time.perf_counter()
    Ys = Xs[rs][:, np.flatnonzero(hvg)]
    Ys.sort_indices()
    cpu_s = time.perf_counter() - t0
    line = {
        'metric': 'device QC x HVG subset of the CSR (operand preparation)',
        'value': round(sec * 1e3, 3), 'unit': 'ms', 'n_gpus': 1,
        'steps': args.steps, 'higher_is_better': False, 'dtype': 'u32/f32 copy',
        'data': 'synthetic',
        'config': {'workload': args.workload, 'num_cells': n, 'num_genes': g,
                   'nnz': nnz, 'qc_rows': int(rows.size), 'hvg_genes': int(hvg.sum()),
                   'subset_nnz': out_nnz},
        'algorithmic_bytes': alg,
        'achieved_gbs': round(alg / sec / 1e9, 1),
        'cpu_baseline': {'value': round(cpu_s * n / sample * 1e3, 1), 'unit': 'ms (scaled by rows)',
                         'cores': 1, 'kind': 'scipy X[rows][:, genes]',
                         'sample': f'first {sample} rows, {cpu_s:.2f} s'},
    }
    print(json.dumps(line), flush=True)


def bench_knn(args):
    """k nearest neighbours of every cell (SingleCell.neighbors) on
    num_cells cells in 50-dimensional PC space: the seeded mixture
    synth.pc_mixture(34, n, 50, 100 centres, spread 1), generated on the host
    and resident on the device before the timed region.  --knn-search ivf
    times the default search (k-means + inverted file, as the reference);
    exact times the exhaustive search, whose dominant work is the FP32 dot
    products, 2 * n^2 * d flops per call, priced against the FP32 peak.  The
    CPU baseline is the exhaustive search in NumPy (FP32 sgemm + partial
    sort) on a sample of queries against all cells, scaled to all queries."""
    import torch
    from brisc_amd import Context, synth
    from brisc_amd.knn import knn_exact, knn_ivf

    n, d, k = args.knn_cells, 50, args.knn_k
    ivf = args.knn_search == 'ivf'
    barbar = bool(args.kmeans_barbar)
    search = (lambda X, k, context: knn_ivf(X, k, context=context,
                                            kmeans_barbar=barbar)[:2]) if ivf \
        else knn_exact
    Xh = synth.pc_mixture(34, n, d, 100, 1.0)
    X = torch.from_numpy(Xh).cuda()
    ctx = Context(0)
    nb, dist = search(X, k, context=ctx)
    torch.cuda.synchronize()
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        nb, dist = search(X, k, context=ctx)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    sec = sum(times) / len(times)
    flops = 2.0 * n * n * d
    if ivf:
        _, _, info = knn_ivf(X, k, context=ctx, kmeans_barbar=barbar)
        ex, _ = knn_exact(X, k, context=ctx)
        sample = torch.arange(0, n, max(1, n // 20000), device='cuda')
        a, b = nb[sample].cpu().numpy(), ex[sample].cpu().numpy()
        recall = float(np.mean([len(set(x) & set(y)) / k for x, y in zip(a.tolist(), b.tolist())]))
        line = {
            'metric': 'k nearest neighbours of every cell (neighbors() default inverted-file search'
                      + (', k-means|| initialisation)' if barbar else ')'),
            'value': round(sec, 4), 'unit': 's', 'n_gpus': 1, 'steps': args.steps,
            'higher_is_better': False, 'dtype': 'f32', 'data': 'synthetic',
            'config': {'workload': f'pcs{n}', 'num_cells': n, 'num_dims': d,
                       'num_neighbors': k, **info},
            'recall_vs_exact': round(recall, 5),
        }
        # The reference's own default search on the same points (fixture
        # written on the development host): its time and a row sample.
        ref_meta = os.path.join(ROOT, 'tests', 'golden', 'ref_vectors.json')
        ref_rows = os.path.join(ROOT, 'tests', 'golden', 'ref_knn_pcs1m.npz')
        if os.path.exists(ref_meta) and os.path.exists(ref_rows):
            with open(ref_meta) as f:
                big = json.load(f).get('knn_large')
            if big and big['num_cells'] == n and big['k'] == k:
                r = np.load(ref_rows)
                mine = nb.cpu().numpy()[r['rows']].astype(np.int64)
                same = float((np.sort(mine, 1) == np.sort(r['neighbors'].astype(np.int64), 1)).all(1).mean())
                if barbar:
                    line['reference_recall_rows'] = big['recall_kmeans_barbar_rows']
                line['reference'] = {
                    'seconds_8_threads': big['seconds_kmeans_barbar_8_threads'] if barbar
                    else big['seconds_default_8_threads'],
                    'host': 'development container (8 cores); the reference does not run here',
                    'rows_with_same_neighbour_set': round(same, 4),
                    'rows_compared': int(len(r['rows']))}
        print(json.dumps(line), flush=True)
        return
    line = {
        'metric': 'exact kNN of every cell (exhaustive neighbors search)',
        'value': round(sec, 4), 'unit': 's', 'n_gpus': 1, 'steps': args.steps,
        'higher_is_better': False, 'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': f'pcs{n}', 'num_cells': n, 'num_dims': d,
                   'num_neighbors': k},
        'roofline': {'bound': 'mfma', 'achieved': round(flops / sec / 1e12, 2),
                     'peak': 157.3, 'unit': 'TFLOP/s',
                     'frac': round(flops / sec / 1e12 / 157.3, 4),
                     'traffic': None},
    }
    if not args.no_cpu_baseline:
        norms = np.einsum('ij,ij->i', Xh, Xh)
        sample = min(n, 2048)
        t0 = time.perf_counter()
        for q0 in range(0, sample, 256):
            Q = Xh[q0:q0 + 256]
            D = norms[q0:q0 + 256, None] + (norms[None, :] - 2.0 * (Q @ Xh.T))
            D[np.arange(len(Q)), np.arange(q0, q0 + len(Q))] = np.inf
            np.argpartition(D, k, axis=1)
        cpu_s = time.perf_counter() - t0
        line['cpu_baseline'] = {
            'value': round(cpu_s * n / sample, 2), 'unit': 's (scaled by queries)',
            'cores': int(os.environ.get('OMP_NUM_THREADS', os.cpu_count())),
            'kind': 'port', 'sample': f'{sample} queries against all {n} cells, {cpu_s:.2f} s'}
    print(json.dumps(line), flush=True)


def bench_pipeline(args):
    """The device chain SingleCell.normalize (log1pPF) -> pca (50 PCs) ->
    neighbors (k = 20, default search) on the 1M-cell workload: counts
    generated on the device, every stage device-resident, PCs never leave
    HBM.  Stage times from the host clock around each synchronous stage;
    the reference's own chain on the same workload (fixture written on the
    development host, 8 threads) is reported beside it, with the share of
    sampled cells whose neighbour set equals the reference's."""
    import torch
    from brisc_amd import Context, irlba_device, synth
    from brisc_amd.knn import knn_ivf

    w = synth.make_workload(**synth.WORKLOADS[args.workload])
    ctx = Context(0)
    runs = []
    for rep in range(max(1, args.steps) + 1):       # the first is a warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        data, indices, indptr = synth.generate_device(w, context=ctx)
        t1 = time.perf_counter()
        P, sv, stats = irlba_device(data, indices, indptr, w.num_cells,
                                    w.num_genes, 50, context=ctx)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        nb, dist, info = knn_ivf(P, 20, context=ctx)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        runs.append((t1 - t0, t2 - t1, t3 - t2))
        del data, indices, indptr
    gen, pca, knn = (min(r[i] for r in runs[1:]) for i in range(3))
    line = {
        'metric': 'device chain generate+normalize -> pca -> neighbors',
        'value': round(pca + knn, 4), 'unit': 's (pca + neighbors)',
        'n_gpus': 1, 'steps': args.steps, 'higher_is_better': False,
        'dtype': 'f32 products / f64 Lanczos', 'data': 'synthetic',
        'config': {'workload': args.workload, 'num_cells': w.num_cells,
                   'num_genes': w.num_genes, 'num_PCs': 50, 'num_neighbors': 20,
                   **info},
        'seconds': {'generate_normalize': round(gen, 4), 'pca': round(pca, 4),
                    'neighbors': round(knn, 4)},
        'pca_converged': bool(stats['converged']),
    }
    ref_meta = os.path.join(ROOT, 'tests', 'golden', 'ref_vectors.json')
    ref_rows = os.path.join(ROOT, 'tests', 'golden', 'ref_pipeline_cells1m.npz')
    if args.workload == 'cells1m' and os.path.exists(ref_rows):
        with open(ref_meta) as f:
            ref = json.load(f).get('pipeline_cells1m')
        r = np.load(ref_rows)
        mine = nb.cpu().numpy()[r['rows']].astype(np.int64)
        same = float((np.sort(mine, 1) == np.sort(r['neighbors'].astype(np.int64), 1)).all(1).mean())
        line['reference'] = {
            'seconds_8_threads': {'normalize': ref['seconds_normalize_8_threads'],
                                  'pca': ref['seconds_pca_8_threads'],
                                  'neighbors': ref['seconds_neighbors_8_threads']},
            'host': 'development container (8 cores); the reference does not run here',
            'rows_with_same_neighbour_set': round(same, 4),
            'rows_compared': int(len(r['rows']))}
    print(json.dumps(line), flush=True)


def bench_normalize(args):
    """SingleCell.normalize() on the device: the workload's int32 counts
    (generated on the device) normalised by each method, CSR with 90% of
    cells passing QC, and log1pPF on the same counts in CSC.  Algorithmic
    bytes per CSR launch: counts read twice (row sums, then the values) and
    the float32 values written (12 z), the offsets read twice, the row sums
    written; PFlog1pPF also reads and writes the values once more (8 z).
    CPU baseline: the oracle's log1pPF (the reference's loop in C, one core)
    on a 1M-row sample, scaled by nonzeros."""
    import torch
    from brisc_amd import Context, synth
    from brisc_amd.normalize import normalize_device
    from brisc_amd.subset import csr_transpose

    spec = synth.WORKLOADS[args.workload]
    w = synth.make_workload(**spec)
    ctx = Context(0)
    _, indices, indptr, counts = synth.generate_device(w, context=ctx, keep_counts=True)
    n, g, nnz = w.num_cells, w.num_genes, int(counts.numel())
    qc = torch.from_numpy(np.random.default_rng(7).random(n) < 0.9).cuda()
    out = torch.empty(nnz, dtype=torch.float32, device='cuda')
    results = {}

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        best = None
        for _ in range(max(args.steps, 1)):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return best

    for method in ('logCP10k', 'log1pPF', 'PFlog1pPF'):
        sec = timed(lambda: normalize_device(counts, indices, indptr, n, g, QC=qc,
                                             method=method, out=out, context=ctx))
        alg = 12 * nnz + 16 * (n + 1) + 8 * n + (8 * nnz + 8 * n if method == 'PFlog1pPF' else 0)
        results[f'csr_{method}'] = {'ms': round(sec * 1e3, 3),
                                    'gbs': round(alg / sec / 1e9, 1)}
    # The CSC layout of the same counts (column offsets over genes), the
    # counts held as float32 (a count type the reference accepts).
    t_data, t_idx, t_ptr = csr_transpose(counts.float(), indices, indptr, g, context=ctx)
    sec = timed(lambda: normalize_device(t_data, t_idx, t_ptr, n, g, is_csr=False, QC=qc,
                                         method='log1pPF', out=out, context=ctx))
    results['csc_log1pPF'] = {'ms': round(sec * 1e3, 3)}
    del t_data, t_idx, t_ptr
    import oracle
    sample = min(1_000_000, n)
    e = int(indptr[sample].item())
    c_s = counts[:e].cpu().numpy()
    p_s = indptr[:sample + 1].cpu().numpy()
    t0 = time.perf_counter()
    oracle.log1ppf(c_s, p_s)
    cpu_s = time.perf_counter() - t0
    line = {
        'metric': 'device SingleCell.normalize() of the counts (log1pPF, CSR)',
        'value': results['csr_log1pPF']['ms'], 'unit': 'ms', 'n_gpus': 1,
        'steps': args.steps, 'higher_is_better': False, 'dtype': 'i32 counts -> f32',
        'data': 'synthetic',
        'config': {'workload': args.workload, 'num_cells': n, 'num_genes': g, 'nnz': nnz,
                   'qc_fraction': 0.9},
        'methods': results,
        'cpu_baseline': {'value': round(cpu_s * nnz / max(e, 1) * 1e3, 1),
                         'unit': 'ms (scaled by nonzeros)', 'cores': 1, 'kind': 'port',
                         'sample': f'oracle log1pPF on the first {sample} rows, {cpu_s:.2f} s'},
    }
    print(json.dumps(line), flush=True)


def bench_copy(args):
    """A plain device copy of known size (calibrates HBM byte counters and
    gives the achievable copy rate)."""
    import torch
    nbytes = int(args.copy_gib * (1 << 30))
    src = torch.ones(nbytes // 4, dtype=torch.float32, device='cuda')
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / max(args.steps, 1)
    print(json.dumps({'metric': 'device copy', 'bytes_read_per_launch': nbytes,
                      'bytes_written_per_launch': nbytes,
                      'launches': args.steps + 1,
                      'ms_per_copy': round(sec * 1e3, 4),
                      'copy_gbs_read_plus_write': round(2 * nbytes / sec / 1e9, 1)}),
          flush=True)


def bench_fp64(args):
    """The FP64 products of a restart (U Ub[:, :kc], kc = 99 and 92) and of
    the final step (kout = 50) at the headline shape (10M cells, m = 100) on
    device arrays, priced against the published FP64 matrix peak and against
    the FP64 rates this device sustains with register operands only
    (bpca_test_fp64_rate: v_mfma_f64_16x16x4f64 and v_fma_f64)."""
    import ctypes

    import torch
    from brisc_amd import _lib, default_context

    ctx = default_context(0)
    L = _lib.lib()
    rates = {}
    for which, name in ((0, 'mfma_f64_16x16x4'), (1, 'valu_fma_f64')):
        t = ctypes.c_double(0.0)
        _lib.check(L.bpca_test_fp64_rate(ctx.handle, which, 5, ctypes.byref(t)))
        rates[name] = round(t.value, 1)
    log(f'# FP64 rates with register operands: {rates}')
    n, m = 10_000_000, 100
    A = torch.randn(n * m, dtype=torch.float64, device='cuda')
    products = {}
    for kout in (99, 92, 50):
        C = torch.randn(m * kout, dtype=torch.float64, device='cuda')
        O = torch.empty(n * kout, dtype=torch.float64, device='cuda')
        call = lambda: _lib.check(L.bpca_test_gemm_device(ctx.handle, n, m, kout, A.data_ptr(),
                                                          C.data_ptr(), O.data_ptr()))
        call()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            call()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / args.steps * 1e3
        tf = 2.0 * n * m * kout / (ms * 1e-3) / 1e12
        products[f'kout{kout}'] = {
            'ms': round(ms, 3), 'tflops': round(tf, 1),
            'frac_of_published_78.6': round(tf / FP64_PEAK_TFLOPS, 3),
            'frac_of_measured_mfma': round(tf / rates['mfma_f64_16x16x4'], 3),
            'hbm_bytes': 8 * n * m + 8 * n * kout}
        log(f'# product n={n} m={m} kout={kout}: {products[f"kout{kout}"]}')
        del C, O
    print(json.dumps({'metric': 'FP64 restart / final products (10M x 100)',
                      'unit': 'TFLOP/s', 'fp64_published_peak_tflops': FP64_PEAK_TFLOPS,
                      'measured_register_rates_tflops': rates, 'products': products,
                      'steps': args.steps, 'note': 'host-timed per product, synchronised'}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--mode', choices=('single', 'batched', 'serve', 'copy', 'subset', 'knn',
                                       'pipeline', 'normalize', 'fp64'),
                    default='single')
    ap.add_argument('--workload', default='cells10m')
    ap.add_argument('--kmeans-barbar', action='store_true',
                    help='--mode knn --knn-search ivf: k-means|| initialisation')
    ap.add_argument('--k', type=int, default=50)
    ap.add_argument('--cpu-sample-rows', type=int, default=2_500_000)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--problems', type=int, default=256)
    ap.add_argument('--batch-k', type=int, default=30)
    ap.add_argument('--no-sequential', action='store_true',
                    help='batched / serve modes: skip the one-at-a-time comparison')
    ap.add_argument('--requests', type=int, default=1000)
    ap.add_argument('--rate', type=float, default=8.0)
    ap.add_argument('--max-cells', type=int, default=2_000_000)
    ap.add_argument('--hbm-budget-gb', type=float, default=288.0)
    ap.add_argument('--compare-requests', type=int, default=200,
                    help='serve mode: prefix of the stream run again at --compare-rate')
    ap.add_argument('--compare-rate', type=float, default=4.0)
    ap.add_argument('--copy-gib', type=float, default=4.0)
    ap.add_argument('--knn-cells', type=int, default=1_000_000)
    ap.add_argument('--knn-k', type=int, default=20)
    ap.add_argument('--knn-search', choices=('exact', 'ivf'), default='exact')
    ap.add_argument('--xtu', choices=('pair', 'sorted', 'csr'), default='pair',
                    help='X^T u in the pass of X v (default), from the gene-sorted '
                         'copy of X, or from the CSR in a pass of its own')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py runs on one GPU only (asked for {args.gpus}); the '
              f'multi-GPU sharding of this path is described in DESIGN.md and '
              f'not built', file=sys.stderr)
        sys.exit(2)
    if args.mode == 'batched':
        return bench_batched(args)
    if args.mode == 'serve':
        return bench_serve(args)
    if args.mode == 'copy':
        return bench_copy(args)
    if args.mode == 'fp64':
        return bench_fp64(args)
    if args.mode == 'subset':
        return bench_subset(args)
    if args.mode == 'knn':
        return bench_knn(args)
    if args.mode == 'normalize':
        return bench_normalize(args)
    if args.mode == 'pipeline':
        if args.workload == 'cells10m' and '--workload' not in sys.argv:
            args.workload = 'cells1m'
        return bench_pipeline(args)

    import torch
    from brisc_amd import Context, irlba_device, synth

    spec = synth.WORKLOADS[args.workload]
    w = synth.make_workload(**spec)
    ctx = Context(0)
    ctx.set_pair_products(args.xtu == 'pair')
    ctx.set_sorted_xtu(args.xtu != 'csr')
    t0 = time.perf_counter()
    data, indices, indptr = synth.generate_device(w, context=ctx)
    gen_s = time.perf_counter() - t0
    n, g, nnz = w.num_cells, w.num_genes, int(data.numel())
    print(f'# workload {args.workload}: n={n} g={g} nnz={nnz} '
          f'(generated on device in {gen_s:.2f} s)', file=sys.stderr, flush=True)
    PCs = torch.empty((n, args.k), dtype=torch.float32, device='cuda')
    tuned = ctx.autotune(data, indices, indptr, n, g, reps=3)
    log(f'# autotune: {tuned}')

    for i in range(args.warmup):
        _, sv, st = irlba_device(data, indices, indptr, n, g, args.k,
                                 PCs=PCs, context=ctx)
        print(f'# warmup {i}: {st["seconds"]:.3f} s restarts={st["restarts"]} '
              f'converged={st["converged"]}', file=sys.stderr, flush=True)
    # The timed solves record HIP events around the sparse products only
    # (each event record waits for the launches before it, so timing every
    # kernel class would slow the solve); the per-class breakdown comes from
    # one more solve after the timed region.
    ctx.set_profiling(True, classes=('spmv', 'spmv_t'))
    ctx.reset_kernel_times()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for i in range(args.steps):
        _, sv, st = irlba_device(data, indices, indptr, n, g, args.k,
                                 PCs=PCs, context=ctx)
        stats.append(st)
        print(f'# step {i}: {st["seconds"]:.3f} s restarts={st["restarts"]} '
              f'matvecs={st["matvecs"]}', file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ktimes = ctx.kernel_times()
    ctx.set_profiling(True)
    ctx.reset_kernel_times()
    _, _, st_b = irlba_device(data, indices, indptr, n, g, args.k, PCs=PCs, context=ctx)
    ctx.set_profiling(False)
    btimes = ctx.kernel_times()

    sec = elapsed / args.steps
    bytes_per = algorithmic_bytes(nnz, n, g)
    # Every macro step launches both products, but a problem that is
    # restarting returns at once; the average is over the products actually
    # computed (one X v and one X^T u per Lanczos step).
    products = sum(int(st['lanczos_steps']) for st in stats)
    # With the pair products (X v and X^T u in one read of X, k_spmv_pair)
    # the 'spmv' class times that kernel and the 'spmv_t' class is empty;
    # its algorithmic bytes are those of one read of X: the same formula.
    pair = bool(stats[-1].get('pair_products', 0))
    kern = {}
    for name in (('spmv',) if pair else ('spmv', 'spmv_t')):
        ms, cnt = ktimes[name]
        avg = ms / max(products, 1)
        key = 'pair' if pair else name
        kern[key] = {'avg_ms': round(avg, 4), 'products': products,
                     'launches': cnt,
                     'gbs': round(bytes_per / (avg * 1e-3) / 1e9, 1)}
    if pair:
        dominant, dclass = 'pair', 'spmv'
    else:
        dclass = max(('spmv', 'spmv_t'), key=lambda k_: ktimes[k_][0])
        dominant = dclass
    achieved = kern[dominant]['gbs']
    # X^T u reads the gene-sorted copy (5 bytes a nonzero) for g <= 4096.
    sorted_xtu = args.xtu == 'sorted' and g <= 4096
    tkey = dominant + ('_sorted' if dominant == 'spmv_t' and sorted_xtu else '')
    traffic = load_traffic(args.workload, tkey)
    kname = {'pair': 'k_spmv_pair', 'spmv': 'k_spmv', 'spmv_t': 'k_spmv_t'}[dominant]
    if tkey == 'spmv_t_sorted':
        kname = 'k_spmv_t_sorted'
    roofline = {'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS,
                'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4),
                'traffic': traffic,
                'kernel': kname, 'bytes_per_launch': bytes_per,
                'products_per_launch': 2 if pair else 1}
    if pair:
        # The two products' algorithmic bytes as two separate passes would
        # move them, over the same launch time.
        roofline['two_pass_equivalent_gbs'] = round(2 * achieved, 1)
    if traffic:
        # Measured HBM bytes per launch over the same average launch time.
        phys = traffic / (kern[dominant]['avg_ms'] * 1e-3) / 1e9
        roofline['achieved_physical'] = round(phys, 1)
        roofline['frac_physical'] = round(phys / HBM_PEAK_GBS, 4)
    total_ms = sum(v[0] for v in btimes.values())
    breakdown = {k_: {'ms_per_solve': round(v[0], 2),
                      'share': round(v[0] / total_ms, 4) if total_ms else 0}
                 for k_, v in btimes.items()}
    breakdown['note'] = (f'one solve after the timed region with every kernel class timed '
                         f'({st_b["seconds"]:.3f} s with those events)')
    spmv_ms = ktimes['spmv'][0] + ktimes['spmv_t'][0]
    spmv_launches = 2 * products

    baseline = None
    if not args.no_cpu_baseline:
        rows = min(args.cpu_sample_rows, n)
        e = int(indptr[rows].item())
        baseline = cpu_baseline(data[:e].cpu().numpy(), indices[:e].cpu().numpy(),
                                indptr[:rows + 1].cpu().numpy(), nnz, g,
                                args.k, rows, int(stats[-1]['lanczos_steps']))

    line = {
        'metric': 'sec to 50 PCs @10M cells (HBM GB/s % peak); batched PCA solves/s; p50/p99 ms',
        'value': round(sec, 4), 'unit': 's', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': round(sec * 1e3, 2),
        'higher_is_better': False, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32 spmv / f64 lanczos', 'data': 'synthetic',
        'config': {'workload': args.workload, 'num_cells': n, 'num_genes': g,
                   'nnz': nnz, 'num_types': spec['num_types'], 'k': args.k,
                   'subspace_size': 100, 'tolerance': 1e-6, 'seed': 0},
        'roofline': roofline,
        'cpu_baseline': baseline,
        'solve': {'restarts': stats[-1]['restarts'],
                  'matvecs': stats[-1]['matvecs'],
                  'converged': bool(stats[-1]['converged']),
                  'svd_sweeps_last': stats[-1]['svd_sweeps'],
                  'top_singular_values': [round(float(x), 4) for x in sv[:5]]},
        'spmv_products_gbs': round(bytes_per * spmv_launches / (spmv_ms * 1e-3) / 1e9, 1) if spmv_ms else None,
        'kernels': kern, 'breakdown': breakdown,
        'xtu_layout': 'one pass with X v' if pair else ('gene-sorted copy' if sorted_xtu else 'csr'),
        'generate_seconds': round(gen_s, 3),
        'autotune': tuned,
    }
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
