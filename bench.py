"""Benchmark: T-matrix columns/s of the JVIE solve on one MI355X.

Workload (the headline configuration): Gaussian random ellipsoid, semi-axes
1 : 0.8 : 0.6, log-normal radial perturbation sigma = 0.1, correlation angle
20 deg, real spherical harmonics l <= 10 (seed 0), cube-split to 198,390 tets,
eps = 3.0 + 0.1i, size parameter ka = 8 (a = volume-equivalent radius),
nmax = 12, i.e. 336 right-hand sides solved as one batched launch, FP64.
The reference's GMRES settings apply (tol 1e-5, Krylov dimension 50, 4
restart cycles).  The same solve in mixed precision (FP32 matvecs in the
Arnoldi steps, FP64 true residual at every restart, same iteration budget and
tolerance) follows by default; its columns/s, the ratio to FP64 and the
relative difference of the two T-matrices are reported beside the FP64 value.

A step is one full T-matrix solve (compute_T_matrix: VSWF right-hand sides,
336 restarted-GMRES systems, projection GEMM) on the device-resident operator;
the operator setup (projections, FFT(G), near zone) is built before the timed
region and reported separately.

The CPU baseline times the oracle (numpy restatement, one core) on a bounded
sample: its matvec on the full-size operator, scaled by the matvecs per column
that the GPU solve needed.
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
ORTHO_NAMES = {0: 'mgs', 1: 'cgs+dgks', 2: 'block-mgs'}
DTYPES = {'fp64': 'f64', 'mixed': 'f32 matvec + f64 refinement'}

from scadyn_amd.workloads import make_workload, request_stream  # noqa: E402


def run_requests(name, args):
    """Request-loop workloads: throughput and per-request latency."""
    from scadyn_amd import native
    rng = np.random.default_rng(0)
    reqs = request_stream(name, rng)
    prec = native.PREC_MIXED if args.precision != 'fp64' else native.PREC_FP64
    cols = [2 * ((r[4] + 1) ** 2 - 1) for r in reqs]
    if name.startswith('serving'):
        # arrival rate for ~80% utilisation from a calibration solve of the first request
        c, e, p, k, nmax, _ = reqs[0]
        t = time.time()
        native.tmat_solve(c, e, p, k, nmax, precision=prec, tol=args.tol, refine_tol=args.refine_tol)
        work0 = cols[0] * e.shape[0]
        rate_work = work0 / (time.time() - t)          # (columns x tets) per second
        mean_service = np.mean([cc * r[1].shape[0] for cc, r in zip(cols, reqs)]) / rate_work
        gaps = rng.exponential(mean_service / 0.8, len(reqs))
        arrivals = np.concatenate([[0.0], np.cumsum(gaps[1:])])
    else:
        arrivals = np.zeros(len(reqs))
    print(f'[bench] {name}: {len(reqs)} requests, {sum(cols)} columns', file=sys.stderr, flush=True)
    sv = native.Server(max_active=args.max_active)
    t0 = time.time()
    ids = []
    for (c, e, p, k, nmax, _), at in zip(reqs, arrivals):
        while time.time() - t0 < at:
            time.sleep(0.005)
        ids.append(sv.submit(c, e, p, k, nmax, precision=prec, tol=args.tol, refine_tol=args.refine_tol))
    lat = []
    for rid in ids:
        _, it, times = sv.wait(rid)
        lat.append(times[3] - times[0])
    makespan = time.time() - t0
    sv.close()
    lat = np.array(lat)
    return {'columns': int(sum(cols)), 'requests': len(reqs), 'makespan_s': makespan,
            'columns_per_s': sum(cols) / makespan, 'p50_latency_s': float(np.percentile(lat, 50)),
            'p99_latency_s': float(np.percentile(lat, 99)), 'max_active': args.max_active}


def cpu_baseline(op_gpu, c, e, param, k, matvecs_per_col, budget_s=25.0):
    """Oracle matvec on the full-size operator (operator arrays exported from the
    device build), scaled to columns/s with the measured matvecs per column."""
    from oracle import pfft_ref
    t0 = time.time()
    o = pfft_ref.Operator()
    o.g = pfft_ref.build_geometry(c, e)
    o.k = k
    o.param = param
    S = op_gpu.export('S')
    o.S, o.Sx, o.Sy, o.Sz = S[0], S[1], S[2], S[3]
    o.indS = np.stack([pfft_ref.box_nodes(o.g, cc)[0] for cc in o.g.tet_cube])
    o.Fg = np.fft.fftn(pfft_ref.green_grid(o.g, k))
    o.sp_mat = op_gpu.export('blocks')
    o.sp_ind = op_gpu.export('blkind').astype(np.int64) - 1
    prep = time.time() - t0
    rng = np.random.default_rng(0)
    x = rng.standard_normal(3 * o.g.N) + 1j * rng.standard_normal(3 * o.g.N)
    n = 0
    from threadpoolctl import threadpool_limits
    with threadpool_limits(1):  # the scalar port on one core
        t0 = time.time()
        while True:
            x = pfft_ref.matvec(o, x)
            x /= np.linalg.norm(x)
            n += 1
            if time.time() - t0 > budget_s or n >= 4:
                break
        per_mv = (time.time() - t0) / n
    cols = 1.0 / (per_mv * matvecs_per_col)
    return {'value': cols, 'unit': 'columns/s', 'cores': 1, 'kind': 'port',
            'sample': (f'{n} oracle matvecs (numpy, 1 thread) on the full {o.g.N}-tet operator, '
                       f'{per_mv:.2f} s each, x {matvecs_per_col:.1f} matvecs/column measured on the '
                       f'GPU; operator arrays exported from the device build ({prep:.0f} s prep)')}


def t_rel(Ta, Tb):
    """Relative Frobenius difference of two T-matrices (four nm x nm blocks each)."""
    num = np.sqrt(sum(np.linalg.norm(a - b) ** 2 for a, b in zip(Ta, Tb)))
    return float(num / np.sqrt(sum(np.linalg.norm(b) ** 2 for b in Tb)))


# HBM bytes per launch of the dominant matvec kernel, from rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE runs of `python -m scadyn_amd.probe` on the same
# workload and batch, each scaled by the copy-kernel calibration factor:
# (workload, kernel, precision, batch) -> bytes
MEASURED_TRAFFIC = {
    # profiles/pmc_traffic_matvec_ellipsoid200k_b112.md
    ('ellipsoid200k', 'gather_near', 'fp64', 112): 28.08e9,
    ('ellipsoid200k', 'fftz_conv', 'fp64', 112): 39.17e9,
    ('ellipsoid200k', 'gather_near', 'mixed', 112): 15.92e9,
    ('ellipsoid200k', 'fftz_conv', 'mixed', 112): 19.57e9,
}


def measured_traffic(workload, kernel, precision, batch):
    return MEASURED_TRAFFIC.get((workload, kernel, precision, batch))


def compare(args, c, e, param, k, nmax, desc):
    """Mixed precision vs FP64 on the same workload at a matched accuracy.

    FP64: the reference's GMRES to --tol64.  Mixed: GMRES with the FP32 matvec
    in the Arnoldi steps and the FP64 true residual at every restart, done when
    that residual is <= --tol64 ||b|| (a cycle ends early once it has reduced
    the residual by --inner-tol).  Reports the columns/s of both (operator
    setup excluded), their ratio and the relative Frobenius difference of the
    two T-matrices."""
    from scadyn_amd import native
    ncol = 2 * ((nmax + 1) ** 2 - 1)
    common = dict(batch=args.batch, ortho=args.ortho, maxit=args.maxit, restart=args.restart)
    res = {}
    for name, prec, tol in (('fp64', native.PREC_FP64, args.tol64), ('mixed', native.PREC_MIXED, args.inner_tol)):
        op = native.Operator(c, e, param, k, precision=prec, tol=tol, refine_tol=args.tol64, **common)
        t = time.time()
        T, it = op.tmatrix(nmax)
        el = time.time() - t
        inf = op.info()
        res[name] = {'T': T, 'seconds': el, 'columns_per_s': ncol / el, 'mean_iters': float(it.mean()),
                     'max_iters': int(it.max()), 'max_restart_cycles': inf['outer_rounds']}
        print(f"[bench] {name}: {el:.2f} s, {ncol / el:.3f} columns/s, mean iterations {it.mean():.1f}, "
              f"max restart cycles {inf['outer_rounds']}", file=sys.stderr, flush=True)
        op.close()
    err = t_rel(res['mixed']['T'], res['fp64']['T'])
    for r in res.values():
        r.pop('T')
    out = {'metric': METRIC, 'value': res['mixed']['columns_per_s'], 'unit': 'columns/s', 'n_gpus': 1,
           'steps': 1, 'warmup': 0, 'ms_per_step': 1000.0 * res['mixed']['seconds'], 'higher_is_better': True,
           'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32 matvec + f64 refinement',
           'data': 'synthetic (seeded mesh generator)',
           'config': {'workload': desc, 'n_tet': int(e.shape[0]), 'columns': ncol, 'nmax': nmax,
                      'fp64': res['fp64'], 'mixed': res['mixed'],
                      'speedup_mixed_over_fp64': res['mixed']['columns_per_s'] / res['fp64']['columns_per_s'],
                      'T_rel_frobenius_mixed_vs_fp64': err,
                      'gmres': {'tol64': args.tol64, 'inner_tol': args.inner_tol, 'refine_tol': args.tol64,
                                'maxit': args.maxit, 'restart': args.restart,
                                'ortho': ORTHO_NAMES[args.ortho]}}}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    # a step is one whole 336-column T-matrix solve (tens of seconds); kernels are
    # compiled ahead of time, so no warm-up step is needed for steady timings
    ap.add_argument('--steps', type=int, default=1)
    ap.add_argument('--warmup', type=int, default=0)
    ap.add_argument('--workload', default='ellipsoid200k')
    ap.add_argument('--batch', type=int, default=0)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--precision', default='both', choices=['fp64', 'mixed', 'both'],
                    help='both: the FP64 solve (the reported value) and the mixed-precision one, with '
                         'their ratio and the relative difference of the two T-matrices')
    ap.add_argument('--tol', type=float, default=1e-5)
    ap.add_argument('--refine-tol', type=float, default=1e-9, help='request loop: mixed-precision target')
    ap.add_argument('--max-active', type=int, default=8)
    ap.add_argument('--ortho', type=int, default=0, help='0 MGS (reference), 1 CGS + DGKS, 2 block MGS')
    ap.add_argument('--mgs-block', type=int, default=0, help='MGS coefficients per pass (0 automatic, 1, 2, 4, 6, 8)')
    ap.add_argument('--maxit', type=int, default=50, help='Krylov dimension per cycle (reference default 50)')
    ap.add_argument('--restart', type=int, default=4, help='restart cycles (reference default 4)')
    ap.add_argument('--compare', action='store_true',
                    help='mixed precision vs FP64 at a matched accuracy: both T-matrices, their '
                         'relative difference and the columns/s of each')
    ap.add_argument('--tol64', type=float, default=1e-8,
                    help='--compare: FP64 GMRES tolerance, and the true-residual target of the mixed solve')
    ap.add_argument('--inner-tol', type=float, default=1e-4,
                    help='--compare: residual reduction that ends a mixed-precision restart cycle early')
    ap.add_argument('--nmax', type=int, default=0, help='override the workload expansion order')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU path is not built.',
              file=sys.stderr)
        sys.exit(2)
    from scadyn_amd import native
    if args.workload in ('bars16', 'serving', 'serving-small'):
        res = run_requests(args.workload, args)
        out = {'metric': METRIC, 'value': res['columns_per_s'], 'unit': 'columns/s', 'n_gpus': 1,
               'steps': 1, 'warmup': 0, 'ms_per_step': 1000.0 * res['makespan_s'], 'higher_is_better': True,
               'scaling': 'weak', 'vs_baseline': None,
               'dtype': DTYPES['fp64' if args.precision == 'fp64' else 'mixed'],
               'data': 'synthetic (seeded mesh generator)', 'config': dict(workload=args.workload, **res)}
        print(json.dumps(out), flush=True)
        return
    c, e, param, k, nmax, desc = make_workload(args.workload)
    if args.nmax:
        nmax = args.nmax
        desc += f' (nmax overridden to {nmax})'
    print(f'[bench] workload {args.workload}: {e.shape[0]} tets, nmax {nmax}', file=sys.stderr, flush=True)
    if args.compare:
        compare(args, c, e, param, k, nmax, desc)
        return
    legs = ['fp64', 'mixed'] if args.precision == 'both' else [args.precision]
    runs, cpu = {}, None
    for name in legs:
        r = runs[name] = measure(args, name, c, e, param, k, nmax)
        if name == legs[0] and not args.no_cpu:
            cpu = cpu_baseline(r['op'], c, e, param, k, r['matvecs_per_column'])
            print(f"[bench] cpu baseline {cpu['value']:.4g} columns/s", file=sys.stderr, flush=True)
        r.pop('op').close()  # one operator resident at a time
    head = runs[legs[0]]
    ncol = head['columns']
    cfg = {'workload': desc, 'n_tet': int(e.shape[0]), 'columns': ncol, 'nmax': nmax,
           'grid': head['grid'], 'fft': head['fft'], 'near_blocks': head['near_blocks'],
           'gmres': {'tol': args.tol, 'maxit': args.maxit, 'restart': args.restart,
                     'ortho': ORTHO_NAMES[args.ortho], 'mgs_block': args.mgs_block}}
    for name, r in runs.items():
        cfg[name] = {kk: v for kk, v in r.items() if kk not in ('T', 'roofline', 'kernels', 'assembly_gemm')}
        cfg[name]['kernels'] = r['kernels']
    if len(runs) == 2:
        cfg['speedup_mixed_over_fp64'] = runs['mixed']['columns_per_s'] / runs['fp64']['columns_per_s']
        cfg['T_rel_frobenius_mixed_vs_fp64'] = t_rel(runs['mixed']['T'], runs['fp64']['T'])
    out = {
        'metric': METRIC, 'value': head['columns_per_s'], 'unit': 'columns/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': head['ms_per_step'],
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': DTYPES[legs[0]], 'data': 'synthetic (seeded mesh generator)', 'config': cfg,
        'roofline': head['roofline'], 'cpu_baseline': cpu, 'assembly_gemm': head['assembly_gemm'],
    }
    print(json.dumps(out), flush=True)


def measure(args, name, c, e, param, k, nmax):
    """Operator setup, then --warmup untimed and --steps timed T-matrix solves in
    one precision; per-kernel timings from the library's HIP-event profiler."""
    from scadyn_amd import native
    t0 = time.time()
    prec = native.PREC_MIXED if name == 'mixed' else native.PREC_FP64
    # mixed: a restart cycle ends once the Arnoldi estimate has fallen by --tol,
    # and a system is done when its FP64 true residual is <= --tol ||b||, the
    # criterion of the FP64 GMRES
    op = native.Operator(c, e, param, k, batch=args.batch, precision=prec, tol=args.tol,
                         refine_tol=args.tol, ortho=args.ortho, mgs_block=args.mgs_block, maxit=args.maxit,
                         restart=args.restart)
    setup_s = time.time() - t0
    ncol = 2 * ((nmax + 1) ** 2 - 1)
    for w in range(args.warmup):
        t = time.time()
        _, it = op.tmatrix(nmax)
        print(f'[bench] {name} warmup {w}: {time.time() - t:.2f} s, mean iters {it.mean():.1f}', file=sys.stderr,
              flush=True)
    op.profile(True)
    t0 = time.time()
    iters = []
    for s in range(args.steps):
        T, it = op.tmatrix(nmax)
        iters.append(it)
        print(f'[bench] {name} step {s}: {time.time() - t0:.2f} s cumulative', file=sys.stderr, flush=True)
    elapsed = time.time() - t0
    tim = op.timings()
    op.profile(False)
    info = op.info()
    it = np.concatenate(iters)
    # matvecs per column: `ite` counts unconverged inner iterations; add the
    # converged one and one restart matvec per cycle
    mv_per_col = float(np.mean(it + 1 + (it // op.opts.maxit + 1)))
    mv_classes = ['scatter', 'fftx', 'ffty', 'fftz_conv', 'iffty', 'ifftx', 'gather_near']
    dom = max(mv_classes, key=lambda n: tim[n]['ms'])
    d = tim[dom]
    avg_ms = d['ms'] / max(d['launches'], 1)
    bytes_per_launch = d['bytes'] / max(d['launches'], 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    per_kernel = {n: {'avg_us': 1000.0 * tim[n]['ms'] / max(tim[n]['launches'], 1),
                      'launches': tim[n]['launches'],
                      'GBps': (tim[n]['bytes'] / max(tim[n]['launches'], 1))
                      / max(1e-12, tim[n]['ms'] / max(tim[n]['launches'], 1) * 1e-3) / 1e9}
                  for n in mv_classes}
    per_kernel['krylov'] = {'ms_total': tim['krylov']['ms']}
    tg = tim['tgemm']
    gemm = None
    if tg['launches']:
        tflops = tg['bytes'] / (tg['ms'] * 1e-3) / 1e12
        gemm = {'kernel': 'tgemm (FP64 MFMA 16x16x4)', 'avg_us': 1000.0 * tg['ms'] / tg['launches'],
                'achieved': tflops, 'peak': FP64_MFMA_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                'frac': tflops / FP64_MFMA_PEAK_TFLOPS}
        per_kernel['tgemm'] = gemm
    print(f'[bench] {name} kernels ' + json.dumps(per_kernel), file=sys.stderr, flush=True)
    value = ncol * args.steps / elapsed
    print(f"[bench] {name}: {value:.3f} columns/s, mean iterations {it.mean():.1f}, batch {info['batch']}, "
          f"fft tiles {info['fft_tiles']} / {info['fft_tiles_f32']}", file=sys.stderr, flush=True)
    dt = 'f32' if name == 'mixed' else 'f64'
    return {
        'op': op, 'T': T, 'columns': ncol, 'columns_per_s': value, 'ms_per_step': 1000.0 * elapsed / args.steps,
        'seconds': elapsed, 'setup_s': setup_s, 'columns_per_s_with_setup': ncol * args.steps / (
            elapsed + setup_s * args.steps),
        'grid': [info['Nx'], info['Ny'], info['Nz']], 'fft': [info['Lx'], info['Ly'], info['Lz']],
        'near_blocks': info['n_blocks'], 'batch': info['batch'], 'fft_tiles': info['fft_tiles'],
        'fft_tiles_f32': info['fft_tiles_f32'], 'mean_gmres_iters': float(it.mean()),
        'matvecs_per_column': mv_per_col, 'max_restart_cycles': info['outer_rounds'],
        'columns_at_iteration_cap': int(np.sum(it >= args.maxit * args.restart)),
        'kernels': per_kernel, 'assembly_gemm': gemm,
        'roofline': {'bound': 'hbm', 'kernel': dom, 'dtype': dt, 'achieved': achieved, 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
                     'traffic': measured_traffic(args.workload, dom, name, info['batch']),
                     'avg_launch_us': avg_ms * 1000.0, 'bytes_per_launch': bytes_per_launch},
    }


if __name__ == '__main__':
    main()
