#!/usr/bin/env python3
"""CCSD seconds per iteration at o=48, v=320 on one MI355X: mixed vs fp64.

One "step" is one CC cycle of the closed-shell CCSD amplitude path (the xvcc
loop body: increments, Jacobi update, energy, DIIS) on the synthetic RHF/C1
system o=48, v=320, n_aux=4*(o+v), seed 20260.  Inputs are generated on the GPU
and are resident in HBM before timing starts.

Headline (``value``): seconds per cycle of the mixed-precision form (ladder,
W(mbej) build and ring contractions as bf16x3 split products on MFMA with fp32
accumulation; everything else fp64).  Outside the timed region the convergence
study runs both forms from the MP1 start: fp64 to |RESMITER| < 1e-9, mixed to
1e-6 followed by fp64 polish to 1e-9, and reports |dE|, the fp64 seconds per
cycle and the speedup.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, 'aces-ii_amd')):
    if _p not in sys.path:
        sys.path.insert(0, _p)

FP64_PEAK = 78.6e12
BF16_PEAK = 2.5e15
FP32_PEAK = 157.3e12


def f_ref(o, v):
    """Reference-equivalent flops per CC cycle (RHF, C1; the reference's XGEMM shapes)."""
    return o * o * v ** 3 * (v + 1) + 10 * o ** 3 * v ** 3 + 4 * o ** 4 * v ** 2 + 20 * o * o * v ** 3


def f_exec(o, v):
    """Flops this implementation executes in its MFMA contractions per cycle."""
    ov = o * v
    nvp, nvm = v * (v + 1) // 2, v * (v - 1) // 2
    nop, nom = o * (o + 1) // 2, o * (o - 1) // 2
    ladder = 2 * (nvp + ov) * nop * nvp + 2 * (nvm + ov) * nom * nvm
    rings = 4 * 2 * ov ** 3
    return ladder, rings


def cpu_baseline(o, v, seed, budget_s=20.0):
    """Time the oracle's closed-shell increments on the host on a bounded sample.

    The sample is the same workload family at a reduced size (o_s, v_s) chosen
    to fit the time budget; seconds per cycle are scaled to o=48, v=320 by the
    ratio of reference-equivalent flops.
    """
    import numpy as np
    from oracle import rccsd
    from aces2 import synth
    threads = os.environ.get('OMP_NUM_THREADS') or str(min(16, os.cpu_count() or 1))
    best = None
    for (os_, vs_) in [(8, 48), (12, 72), (16, 96)]:
        eps, B = synth.draw(os_, vs_, seed, sigma=synth.scaled_sigma(os_, vs_))
        blk, eri = synth.blocks_numpy(eps, B, os_)
        g = rccsd.blocks(eri, os_)
        t1 = np.zeros((os_, vs_))
        _, t2 = (None, g['ovov'].transpose(0, 2, 1, 3) / rccsd.denominators(blk['fock'], os_)[1])
        rccsd.increments(t1, t2, blk['fock'], g, os_)
        t0 = time.perf_counter()
        rccsd.increments(t1 + 1e-3, t2, blk['fock'], g, os_)
        dt = time.perf_counter() - t0
        best = (os_, vs_, dt)
        if dt * (f_ref(os_ + 4, vs_ + 24) / f_ref(os_, vs_)) > budget_s:
            break
    os_, vs_, dt = best
    scale = f_ref(o, v) / f_ref(os_, vs_)
    return {'value': dt * scale, 'unit': 's/iter', 'cores': int(threads), 'kind': 'port',
            'sample': 'oracle.rccsd increments (numpy/BLAS, %s threads) at o=%d, v=%d: %.3f s, scaled x%.1f '
                      'by reference-equivalent flops to o=%d, v=%d' % (threads, os_, vs_, dt, scale, o, v)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--o', type=int, default=48)
    ap.add_argument('--v', type=int, default=320)
    ap.add_argument('--seed', type=int, default=20260)
    ap.add_argument('--prec', default='bf16x3', choices=['bf16x3', 'fp32', 'fp64'])
    ap.add_argument('--no-convergence', action='store_true', help='skip the convergence study')
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline')
    ap.add_argument('--maxcyc', type=int, default=60)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) > 1:
        print('bench.py: this build measures one GPU only (the CCSD path is benchmarked on a single '
              'MI355X); --gpus %d is not supported' % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from aces2 import synth
    from aces2.ccsd import CCSD

    o, v = args.o, args.v
    log = lambda *a: print('[bench]', *a, file=sys.stderr, flush=True)
    t_setup = time.perf_counter()
    eps, B = synth.draw(o, v, args.seed, sigma=synth.scaled_sigma(o, v))
    blocks = synth.blocks_device(eps, B, o)
    del B
    log('integrals generated (%.1f s)' % (time.perf_counter() - t_setup))
    cc = CCSD(o, v, blocks)
    del blocks
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    log('engine set up: %.1f GB device, %.1f s' % (cc.device_bytes / 1e9, time.perf_counter() - t_setup))

    # ---- timed region: K cycles of the headline form after W warmup cycles ----
    cc.set_diis_order(6)
    cc.reset_diis()
    cc.start_mp1()
    for icycle in range(1, args.warmup + 1):
        cc.cycle(icycle, args.prec, 1e-30)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for icycle in range(args.warmup + 1, args.warmup + args.steps + 1):
        cc.cycle(icycle, args.prec, 1e-30)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    s_per_cycle = (t1 - t0) / args.steps
    log('%s: %.4f s/cycle' % (args.prec, s_per_cycle))

    # ---- kernel-level timing (hipEvents around each MFMA launch, same stream) ----
    cc.set_profiling(True)
    cc.profile(reset=True)
    nprof = 2
    for icycle in range(args.warmup + args.steps + 1, args.warmup + args.steps + 1 + nprof):
        cc.cycle(icycle, args.prec, 1e-30)
    prof = cc.profile(reset=True)
    cc.set_profiling(False)
    ms = sum(p['ms'] for p in prof.values())
    fl = sum(p['flops'] for p in prof.values())
    nl = sum(p['launches'] for p in prof.values())
    execf = 3.0 if args.prec == 'bf16x3' else 1.0
    peak = {'bf16x3': BF16_PEAK, 'fp32': FP32_PEAK, 'fp64': FP64_PEAK}[args.prec]
    achieved = execf * fl / (ms * 1e-3)
    roofline = {'bound': 'mfma', 'kernel': {'bf16x3': 'gemm_bf16x3_kernel', 'fp32': 'gemm16_kernel<float>',
                                             'fp64': 'gemm16_kernel<double>'}[args.prec],
                'achieved': achieved / 1e12, 'peak': peak / 1e12, 'unit': 'TFLOP/s',
                'frac': achieved / peak, 'traffic': None,
                'avg_launch_ms': ms / nl, 'launches_per_cycle': nl / nprof,
                'gemm_share_of_cycle': (ms / nprof) / (s_per_cycle * 1e3),
                'per_class': {k: {'TFLOP/s': execf * p['flops'] / max(p['ms'], 1e-9) / 1e9, 'ms': p['ms'] / nprof}
                              for k, p in prof.items()}}
    log('roofline', json.dumps(roofline))

    result = {
        'metric': 'CCSD s/iter at o=48,v=320: mixed vs FP64 on GPU at |dE|<=1e-7 Eh; MFMA % peak',
        'value': s_per_cycle,
        'unit': 's/iter',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': s_per_cycle * 1e3,
        'higher_is_better': False,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': {'bf16x3': 'bf16x3/fp32-acc + f64', 'fp32': 'f32 + f64', 'fp64': 'f64'}[args.prec],
        'data': 'synthetic (DF integrals, seed %d, sigma=0.05*sqrt(95/(o*v)))' % args.seed,
        'config': {'workload': 'closed-shell CCSD cycle, RHF C1, o=%d v=%d n_aux=%d' % (o, v, 4 * (o + v)),
                   'o': o, 'v': v, 'form': args.prec, 'diis': 6},
        'effective_TFLOPs_ref': f_ref(o, v) / s_per_cycle / 1e12,
        'roofline': roofline,
    }

    # ---- convergence study (untimed for `value`) ----
    if not args.no_convergence:
        def run(form, conv, polish_conv=None):
            cc.set_diis_order(6)
            cc.reset_diis()
            cc.start_mp1()
            times, hist = [], []
            icycle, e_final, phase = 0, None, form
            tol = 10.0 ** (-conv)
            while icycle < args.maxcyc:
                icycle += 1
                ta = time.perf_counter()
                st = cc.cycle(icycle, phase, tol)
                times.append((phase, time.perf_counter() - ta))
                hist.append((phase, st['energy'], st['resmiter']))
                if st['converged']:
                    if phase != 'fp64' and polish_conv is not None:
                        # iterative refinement: continue with fp64 cycles to the tight threshold
                        phase = 'fp64'
                        tol = 10.0 ** (-polish_conv)
                        continue
                    e_final = st['final_energy']
                    break
            return e_final, times, hist, icycle

        e64, times64, hist64, n64 = run('fp64', 9)
        log('fp64: E=%r in %d cycles' % (e64, n64))
        emx, timesmx, histmx, nmx = run(args.prec, 6, polish_conv=9) if args.prec != 'fp64' else (e64, times64, hist64, n64)
        log('%s+polish: E=%r in %d cycles' % (args.prec, emx, nmx))
        med = lambda xs: float(np.median(xs)) if xs else None
        s64 = med([t for p, t in times64[2:] if p == 'fp64'])
        slo = med([t for p, t in timesmx[2:] if p == args.prec])
        result['convergence'] = {
            'fp64': {'energy': e64, 'cycles': n64, 's_per_cycle_median': s64,
                     'total_s': sum(t for _, t in times64)},
            'mixed': {'energy': emx, 'cycles': nmx,
                      'lowp_cycles': sum(1 for p, _ in timesmx if p != 'fp64'),
                      'polish_cycles': sum(1 for p, _ in timesmx if p == 'fp64'),
                      's_per_cycle_median_lowp': slo, 'total_s': sum(t for _, t in timesmx)},
            'abs_dE': abs(emx - e64) if (emx is not None and e64 is not None) else None,
            'speedup_per_cycle': (s64 / slo) if (s64 and slo) else None,
            'speedup_total': (sum(t for _, t in times64) / sum(t for _, t in timesmx)) if timesmx else None,
        }
        result['fp64_s_per_iter'] = s64

    if not args.no_cpu:
        log('cpu baseline ...')
        result['cpu_baseline'] = cpu_baseline(o, v, args.seed)
    else:
        result['cpu_baseline'] = None
    print(json.dumps(result), flush=True)


if __name__ == '__main__':
    main()
