"""Profiling driver: a short, fixed amount of device work on one workload, for
hardware-counter runs (which serialise kernels and would stretch a full
T-matrix solve to many minutes) and for solver diagnostics.

    python -m scadyn_amd.probe --workload ellipsoid200k --matvecs 3 --batch 64
    python -m scadyn_amd.probe --copy-mib 1024          # copy kernel of known size
    python -m scadyn_amd.probe --workload ellipsoid200k --gmres 3 --maxit 100 --restart 10

--matvecs runs that many batched operator applications (all matvec kernels,
one launch each per application); --copy-mib runs one device-to-device copy
kernel of the given size, the yardstick that scales FETCH_SIZE / WRITE_SIZE
readings to bytes; --gmres solves the 2 nmax (nmax + 2) T-matrix columns of
a smaller expansion order and prints the iteration count of every column.
"""
import argparse
import sys
import time

import numpy as np


def say(msg):
    print(f'[probe] {msg}', file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--workload', default='ellipsoid200k')
    ap.add_argument('--matvecs', type=int, default=0)
    ap.add_argument('--batch', type=int, default=64)
    ap.add_argument('--precision', default='fp64', choices=['fp64', 'mixed'])
    ap.add_argument('--copy-mib', type=int, default=0)
    ap.add_argument('--gmres', type=int, default=0, help='nmax of the T-matrix columns to solve')
    ap.add_argument('--maxit', type=int, default=50)
    ap.add_argument('--restart', type=int, default=4)
    ap.add_argument('--tol', type=float, default=1e-5)
    ap.add_argument('--ortho', type=int, default=0)
    ap.add_argument('--kernel-times', action='store_true', help='print per-kernel HIP-event averages of the matvecs')
    args = ap.parse_args(argv)

    if args.copy_mib:
        import torch
        n = args.copy_mib * (1 << 20) // 8
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
        gbs = 2.0 * args.copy_mib * (1 << 20) / (best * 1e-3) / 1e9
        say(f'copy {args.copy_mib} MiB read + {args.copy_mib} MiB write: best {best:.3f} ms, {gbs:.0f} GB/s')
        return

    from scadyn_amd import native
    from scadyn_amd.workloads import make_workload
    c, e, param, k, nmax, desc = make_workload(args.workload)
    prec = native.PREC_MIXED if args.precision == 'mixed' else native.PREC_FP64
    say(f'{args.workload}: {e.shape[0]} tets ({desc})')
    op = native.Operator(c, e, param, k, batch=args.batch, precision=prec, tol=args.tol, maxit=args.maxit,
                         restart=args.restart, ortho=args.ortho)
    say(f'operator: {op.info()}')
    if args.kernel_times:
        op.profile(True)
    if args.matvecs:
        rng = np.random.default_rng(0)
        x = rng.standard_normal((args.batch, 3 * op.N)) + 1j * rng.standard_normal((args.batch, 3 * op.N))
        for i in range(args.matvecs):
            t = time.time()
            x = op.matvec(x, precision=prec)
            x /= np.linalg.norm(x, axis=1, keepdims=True)
            say(f'matvec {i}: {time.time() - t:.3f} s (host copies included)')
        if args.kernel_times:
            tm = op.timings()
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


if __name__ == '__main__':
    main()
