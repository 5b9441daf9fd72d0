"""Benchmark: T-matrix columns/s of the JVIE solve on one MI355X, mixed precision
against FP64, at a T-matrix accuracy well inside 1e-6.

Workload (default): two-layer sphere, 64 cells across (centre-in, jitter 0.1 h,
seed 0), cube-split to 824,256 tets; eps 3.5+0.2i for centroid r < 0.6 a,
2.0+0.02i outside; ka 15; nmax 20, i.e. 880 T-matrix columns.

Both legs solve the same columns in the same seeded random order through the
library's continuously refilled batch (scadyn_tjob_*), each to a matched
accuracy:
  * FP64: the reference's GMRES (x0 = rhs, MGS, restart length 50) on the FP64
    matvec, tol 1e-8 on |g(i+1)|/||b|| and up to 60 restart cycles, so that
    every column converges (the reference's default of 4 cycles stops all of
    them at the 200-iteration cap on this workload);
  * mixed: FP32 matvec, Krylov basis stored in binary16 (FP64 arithmetic;
    --basis), inside FP64 iterative refinement, done when the FP64 true
    residual is <= 1e-8 ||b||.
The relative Frobenius difference of the two legs' T columns is reported.

A step is K converged columns (default K = 4): the batch keeps every slot busy
and refills a slot as soon as its column converges, so a step is the time in
which K more columns finish.  The first fills of the batch's slots are spread
over 800 solver steps (about one column's iterations), so that columns finish
at an even rate rather than in batch-wide bursts.  --warmup steps run first
(the first also fills the pipeline: no column converges before about 700
solver iterations), then --steps steps are timed, with a device
synchronisation on both sides.  `value`
is the mixed leg's converged columns per second over the timed steps.  The
FP64 leg runs --fp64-warmup (5) and --fp64-steps (10) steps of the same kind.

Per-kernel times come from the library's HIP-event profiler on the operator's
stream over the timed steps; `roofline` names the matvec kernel with the
largest total time in the headline precision.  The CPU baseline runs a
multi-threaded port of the reference's matvec on the host (oracle/cpu_port.py)
on the same operator and scales its time by the FP64 leg's matvecs per column.

Other modes:
  --workload serving | serving-small | bars16   request-loop workloads
  --matvecs N [--batch B] [--precision P]       N batched matvecs (counter runs)
  --copy-mib M                                  one device copy of M MiB (counter
                                                calibration)
  --gmres NMAX                                  iteration counts of the columns of
                                                a smaller expansion order
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'T-matrix columns/s, mixed vs FP64 pFFT-GMRES at 1e-6 rel err; p99 latency'
HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6
# scadyn_opts.basis values by name (storage of the mixed-precision Krylov basis)
BASES = {'auto': 0, 'c64': 1, 'f16': 2}
MV_CLASSES = ['scatter', 'fftx', 'ffty', 'fftz_conv', 'iffty', 'ifftx', 'gather_near']

from scadyn_amd.workloads import make_workload, request_stream  # noqa: E402


def say(msg):
    print(f'[bench] {msg}', file=sys.stderr, flush=True)


# HBM bytes per launch of a matvec kernel, from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE runs of `bench.py --matvecs` on the same workload and batch, each
# scaled by the copy-kernel calibration factor (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0):
# (workload, kernel class, precision, batch) -> bytes
MEASURED_TRAFFIC = {
    # profiles/pmc_traffic_matvec_layered824k.md (the bench batches: mixed 64, FP64 43)
    ('layered824k', 'fftz_conv', 'mixed', 64): 2.0 * 9662131 * 1024 + 19245600 * 1024,
    ('layered824k', 'gather_near', 'mixed', 64): 2.0 * 17700921 * 1024 + 2472768 * 1024,
    ('layered824k', 'fftz_conv', 'fp64', 43): 2.0 * 13006873 * 1024 + 25861275 * 1024,
    ('layered824k', 'gather_near', 'fp64', 43): 2.0 * 20667058 * 1024 + 1663739 * 1024,
    # profiles/pmc_traffic_matvec_ellipsoid200k_b112.md
    ('ellipsoid200k', 'gather_near', 'fp64', 112): 28.08e9,
    ('ellipsoid200k', 'fftz_conv', 'fp64', 112): 39.17e9,
    ('ellipsoid200k', 'gather_near', 'mixed', 112): 15.92e9,
    ('ellipsoid200k', 'fftz_conv', 'mixed', 112): 19.57e9,
}


def t_rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def kernel_table(tim):
    out = {}
    for n in MV_CLASSES:
        t = tim[n]
        if not t['launches']:
            continue
        avg_ms = t['ms'] / t['launches']
        out[n] = {'avg_us': 1000.0 * avg_ms, 'launches': t['launches'],
                  'bytes_per_launch': t['bytes'] / t['launches'],
                  'GBps': t['bytes'] / t['launches'] / (avg_ms * 1e-3) / 1e9}
    out['krylov_ms_total'] = tim['krylov']['ms']
    return out


def run_leg(name, op, nmax, order, K, warmup, steps, prec_key, spread):
    """warmup + timed steps of K converged columns on one operator."""
    from scadyn_amd import native
    job = native.TJob(op, nmax, order, spread=spread)
    t0 = time.time()
    done = 0
    for w in range(warmup):
        done = job.advance(K)
        say(f'{name} warmup {w}: {done} columns converged, {time.time() - t0:.1f} s')
    op.profile(True)
    d0 = done
    t0 = time.time()
    for s in range(steps):
        done = job.advance(K)
        say(f'{name} step {s}: {done - d0} columns converged in {time.time() - t0:.1f} s')
    elapsed = time.time() - t0
    tim = op.timings(prec_key)
    op.profile(False)
    T, it = job.result()
    job.close()
    ncols = done - d0
    fin = it >= 0
    cap = op.opts.maxit * op.opts.restart
    info = op.info()
    r = {'columns': int(ncols), 'seconds': elapsed, 'columns_per_s': ncols / elapsed,
         'ms_per_step': 1000.0 * elapsed / max(steps, 1), 'steps': steps, 'warmup': warmup,
         'batch': info['batch'], 'mean_iterations': float(it[fin].mean()) if fin.any() else None,
         'max_iterations': int(it[fin].max()) if fin.any() else None,
         'columns_at_iteration_cap': int(np.sum(it[fin] >= cap)),
         # the reference's re-orthogonalisation test (its extra pass is not applied)
         'reorth_test_fired': info['reorth_fired'], 'kernels': kernel_table(tim),
         'fft_tiles': info['fft_tiles' if prec_key == 'fp64' else 'fft_tiles_f32']}
    say(f"{name}: {r['columns_per_s']:.3f} columns/s ({ncols} columns in {elapsed:.1f} s), batch {r['batch']}, "
        f"mean iterations {r['mean_iterations']}")
    return r, T, it, tim


def roofline(tim, workload, prec, batch):
    dom = max(MV_CLASSES, key=lambda n: tim[n]['ms'])
    d = tim[dom]
    avg_ms = d['ms'] / max(d['launches'], 1)
    bpl = d['bytes'] / max(d['launches'], 1)
    achieved = bpl / (avg_ms * 1e-3) / 1e9
    return {'bound': 'hbm', 'kernel': dom, 'dtype': 'f32' if prec == 'mixed' else 'f64', 'achieved': achieved,
            'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
            'traffic': MEASURED_TRAFFIC.get((workload, dom, prec, batch)), 'avg_launch_us': 1000.0 * avg_ms,
            'bytes_per_launch': bpl}


def headline(args):
    from scadyn_amd import native
    t = time.time()
    c, e, param, k, nmax, desc = make_workload(args.workload)
    if args.nmax:
        nmax = args.nmax
        desc += f' (nmax overridden to {nmax})'
    ncol = 2 * ((nmax + 1) ** 2 - 1)
    order = np.random.default_rng(0).permutation(ncol)
    say(f'workload {args.workload}: {e.shape[0]} tets, nmax {nmax}, {ncol} columns ({time.time() - t:.1f} s mesh)')
    common = dict(maxit=args.maxit, restart=args.restart, batch=args.batch, mgs_block=args.mgs_block)
    legs, Ts = {}, {}
    cpu = None
    for name in ('mixed', 'fp64'):
        t = time.time()
        if name == 'mixed':
            op = native.Operator(c, e, param, k, precision=native.PREC_MIXED, tol=args.inner_tol,
                                 refine_tol=args.tol, basis=BASES[args.basis], **common)
            warm, steps = args.warmup, args.steps
        else:
            op = native.Operator(c, e, param, k, precision=native.PREC_FP64, tol=args.tol, **common)
            warm, steps = args.fp64_warmup, args.fp64_steps
        setup = time.time() - t
        say(f'{name} operator set up in {setup:.1f} s')
        r, T, it, tim = run_leg(name, op, nmax, order, args.cols_per_step, warm, steps, name, args.spread)
        r['setup_s'] = setup
        info = op.info()
        r['roofline'] = roofline(tim, args.workload, name, info['batch'])
        tg = tim['tgemm']
        if tg['launches']:
            tf = tg['bytes'] / (tg['ms'] * 1e-3) / 1e12
            r['assembly_gemm'] = {'kernel': 'tgemm (FP64 MFMA 16x16x4)', 'avg_us': 1000.0 * tg['ms'] / tg['launches'],
                                  'achieved': tf, 'peak': FP64_MFMA_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                                  'frac': tf / FP64_MFMA_PEAK_TFLOPS}
        legs[name], Ts[name] = r, (T, it)
        if name == 'fp64' and not args.no_cpu:
            # matvecs per column of the FP64 leg: `ite` counts the unconverged
            # Arnoldi steps; add the converged one and one restart matvec per cycle
            fin = it >= 0
            mvpc = float(np.mean(it[fin] + 1 + (it[fin] // args.maxit + 1)))
            from oracle import cpu_port
            cpu = cpu_port.cpu_baseline(op, c, e, param, k, mvpc, budget_s=args.cpu_budget)
            say(f"cpu baseline {cpu['value']:.4g} columns/s ({cpu['sample']})")
        op.close()
    both = (Ts['mixed'][1] >= 0) & (Ts['fp64'][1] >= 0)
    err = t_rel(Ts['mixed'][0][:, both], Ts['fp64'][0][:, both]) if both.any() else None
    head = legs['mixed']
    cfg = {'workload': desc, 'n_tet': int(e.shape[0]), 'nmax': nmax, 'columns_total': ncol,
           'columns_per_step': args.cols_per_step, 'column_order': 'seeded random permutation (seed 0)',
           'gmres': {'tol': args.tol, 'mixed_cycle_tol': args.inner_tol, 'maxit': args.maxit,
                     'restart': args.restart, 'mixed_basis': args.basis},
           'mixed': {kk: v for kk, v in head.items() if kk != 'roofline'},
           'fp64': legs['fp64'],
           'speedup_mixed_over_fp64': head['columns_per_s'] / legs['fp64']['columns_per_s'],
           'T_rel_err_mixed_vs_fp64': err, 'T_columns_compared': int(both.sum())}
    out = {'metric': METRIC, 'value': head['columns_per_s'], 'unit': 'columns/s', 'n_gpus': 1,
           'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': head['ms_per_step'],
           'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32',
           'data': 'synthetic (seeded mesh generator)', 'config': cfg, 'roofline': head['roofline'],
           'cpu_baseline': cpu}
    print(json.dumps(out), flush=True)


def run_requests(name, args):
    """Request-loop workloads: throughput and per-request latency."""
    from scadyn_amd import native
    rng = np.random.default_rng(0)
    reqs = request_stream(name, rng)
    prec = native.PREC_FP64 if args.precision == 'fp64' else native.PREC_MIXED
    cols = [2 * ((r[4] + 1) ** 2 - 1) for r in reqs]
    opts = dict(precision=prec, tol=args.inner_tol if prec == native.PREC_MIXED else args.tol,
                refine_tol=args.tol, maxit=args.maxit, restart=args.restart)
    if name.startswith('serving'):
        # arrival rate for ~80% utilisation from a calibration solve of the first request
        c, e, p, k, nmax, _ = reqs[0]
        t = time.time()
        native.tmat_solve(c, e, p, k, nmax, **opts)
        work0 = cols[0] * e.shape[0]
        rate_work = work0 / (time.time() - t)          # (columns x tets) per second
        mean_service = np.mean([cc * r[1].shape[0] for cc, r in zip(cols, reqs)]) / rate_work
        gaps = rng.exponential(mean_service / 0.8, len(reqs))
        arrivals = np.concatenate([[0.0], np.cumsum(gaps[1:])])
    else:
        arrivals = np.zeros(len(reqs))
    say(f'{name}: {len(reqs)} requests, {sum(cols)} columns')
    sv = native.Server(max_active=args.max_active)
    t0 = time.time()
    ids = []
    for (c, e, p, k, nmax, _), at in zip(reqs, arrivals):
        while time.time() - t0 < at:
            time.sleep(0.005)
        ids.append(sv.submit(c, e, p, k, nmax, **opts))
    lat = []
    for rid in ids:
        _, it, times = sv.wait(rid)
        lat.append(times[3] - times[0])
        say(f'request {rid} done: latency {lat[-1]:.2f} s')
    makespan = time.time() - t0
    sv.close()
    lat = np.array(lat)
    res = {'columns': int(sum(cols)), 'requests': len(reqs), 'makespan_s': makespan,
           'columns_per_s': sum(cols) / makespan, 'p50_latency_s': float(np.percentile(lat, 50)),
           'p99_latency_s': float(np.percentile(lat, 99)), 'max_active': args.max_active}
    out = {'metric': METRIC, 'value': res['columns_per_s'], 'unit': 'columns/s', 'n_gpus': 1,
           'steps': 1, 'warmup': 0, 'ms_per_step': 1000.0 * res['makespan_s'], 'higher_is_better': True,
           'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64' if prec == native.PREC_FP64 else 'f32',
           'data': 'synthetic (seeded mesh generator)', 'config': dict(workload=name, **res)}
    print(json.dumps(out), flush=True)


def copy_calibration(mib):
    """One device-to-device copy of known size: the yardstick that scales the
    FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device='cuda')
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    best = None
    for rep in range(8):  # the first copy warms up; the best of the rest is reported
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        dst.copy_(src)
        b.record()
        torch.cuda.synchronize()
        if rep:
            best = a.elapsed_time(b) if best is None else min(best, a.elapsed_time(b))
    say(f'copy {mib} MiB read + {mib} MiB write: best {best:.3f} ms, '
        f'{2.0 * mib * (1 << 20) / (best * 1e-3) / 1e9:.0f} GB/s')


def probe(args):
    """A short, fixed amount of device work for hardware-counter runs (which
    serialise kernels) and solver diagnostics."""
    from scadyn_amd import native
    c, e, param, k, nmax, desc = make_workload(args.workload)
    prec = native.PREC_MIXED if args.precision == 'mixed' else native.PREC_FP64
    say(f'{args.workload}: {e.shape[0]} tets ({desc})')
    op = native.Operator(c, e, param, k, batch=args.batch or 64, precision=prec, tol=args.tol,
                         maxit=args.maxit, restart=args.restart, basis=BASES[args.basis])
    if args.matvecs:
        B = op.opts.batch
        op.profile(True)
        rng = np.random.default_rng(0)
        x = rng.standard_normal((B, 3 * op.N)) + 1j * rng.standard_normal((B, 3 * op.N))
        for i in range(args.matvecs):
            t = time.time()
            x = op.matvec(x, precision=prec)
            x /= np.linalg.norm(x, axis=1, keepdims=True)
            say(f'matvec {i}: {time.time() - t:.3f} s (host copies included)')
        tm = op.timings(args.precision)
        say('kernel averages (us per launch): ' +
            str({kk: round(v['ms'] * 1e3 / v['launches']) for kk, v in tm.items() if v['launches']}))
    if args.gmres:
        t = time.time()
        _, it = op.tmatrix(args.gmres)
        cap = args.maxit * args.restart
        say(f'T-matrix nmax {args.gmres}: {len(it)} columns in {time.time() - t:.2f} s; '
            f'maxit {args.maxit} restart {args.restart}: iterations {it.tolist()}; '
            f'{int(np.sum(it >= cap))} columns at the cap {cap}')
    op.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=5)
    ap.add_argument('--workload', default='layered824k')
    ap.add_argument('--cols-per-step', type=int, default=4, help='converged columns per step')
    ap.add_argument('--fp64-steps', type=int, default=10)
    ap.add_argument('--fp64-warmup', type=int, default=5)
    ap.add_argument('--spread', type=int, default=800,
                    help='solver steps over which the first fills of the batch are spread (about the '
                         'iterations of one column): columns then finish at an even rate')
    ap.add_argument('--batch', type=int, default=0)
    ap.add_argument('--nmax', type=int, default=0, help='override the workload expansion order')
    ap.add_argument('--tol', type=float, default=1e-8,
                    help='FP64 GMRES tolerance, and the FP64 true-residual target of the mixed solve')
    ap.add_argument('--inner-tol', type=float, default=1e-4,
                    help='mixed: residual reduction that ends a restart cycle early')
    ap.add_argument('--maxit', type=int, default=50, help='Krylov dimension per cycle (reference default 50)')
    ap.add_argument('--basis', choices=sorted(BASES), default='auto',
                    help='mixed: storage of the Krylov basis (auto: binary16 for 2^20 to 2^26 rows)')
    ap.add_argument('--mgs-block', type=int, default=0,
                    help='MGS coefficients per pass over the rows (0 automatic; 1, 2, 4, 6 or 8)')
    ap.add_argument('--restart', type=int, default=60, help='restart cycles (enough for every column to converge)')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--cpu-budget', type=float, default=25.0, help='seconds of CPU-baseline work')
    ap.add_argument('--precision', default='mixed', choices=['fp64', 'mixed'],
                    help='request-loop and probe modes')
    ap.add_argument('--max-active', type=int, default=8)
    ap.add_argument('--matvecs', type=int, default=0)
    ap.add_argument('--copy-mib', type=int, default=0)
    ap.add_argument('--gmres', type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU path is not built.',
              file=sys.stderr)
        sys.exit(2)
    if args.copy_mib:
        copy_calibration(args.copy_mib)
    elif args.matvecs or args.gmres:
        probe(args)
    elif args.workload in ('bars16', 'serving', 'serving-small'):
        run_requests(args.workload, args)
    else:
        headline(args)


if __name__ == '__main__':
    main()
