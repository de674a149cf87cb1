#!/usr/bin/env python3
"""CCSD seconds per iteration at o=48, v=320 on one MI355X: mixed vs fp64.

One "step" is one CC cycle of the closed-shell CCSD amplitude path (the xvcc
loop body: increments, Jacobi update, energy, DIIS) on the synthetic RHF/C1
system o=48, v=320, n_aux=4*(o+v), seed 20260.  Inputs are generated on the GPU
and are resident in HBM before timing starts.

Headline (``value``): seconds per cycle of the mixed-precision form (ladders,
W(mbej) build, ring and the read-bound T1-dressing contractions as bf16x3 split
products on MFMA with fp32 accumulation; the rest fp64).  Outside the timed
region the convergence study runs from the MP1 start: fp64 to |RESMITER| < 1e-9,
and mixed to 1e-7 followed (DIIS restarted at the switch) by a polish to 1e-9,
once with refinement cycles (the large products as bf16x3 products of the
operands' differences to an fp64 anchor taken at the switch) and once with fp64
cycles; it reports |dE|, the seconds per cycle of each form and the speedups.

With ``--workload lambda`` the step is one cycle of the Lambda equations (the
xlambda loop body: the engine's reverse pass, Jacobi step, Lambda DIIS) at the
converged fp64 amplitudes of o=32, v=256; the same run checks dE/dmu from the
response density and from the Lambda-weighted increments against central
finite differences (h = 1e-4) for a Fock perturbation F + mu P and a rank-1
integral perturbation (pq|rs) + mu x_pq x_rs, at fp64 amplitudes and at the
amplitudes of the mixed form after its refinement polish.

With ``--workload sym`` the system is the symmetry-blocked form: D2h (8
irreps), o=64, v=384 split over the irreps by a seeded multinomial (at least
2 occupied and 16 virtual orbitals per irrep), integrals from the same DF
generator with symmetry-forbidden elements zeroed; the step is one CC cycle of
the irrep-blocked engine (one grouped MFMA launch per contraction term) in the
mixed form, followed outside the timed region by the refinement polish to 1e-9 and
the Lambda equations on the same lists.

With ``--workload triples`` the step is one evaluation of the non-iterative
triples correction (T) at o=32, v=256 on the amplitudes of a 20-cycle fp64
CCSD run from the MP1 start (the o=32, v=256 configuration's run): the W
products in fp64, and outside the timed region the same evaluation with bf16x3
W products and its deviation from fp64.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ccsd|lambda|sym|triples]
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


def f_ref_classes(o, v):
    """Reference-equivalent flops (2mnk of the reference's XGEMM shapes, RHF C1)
    of the three product classes the engine times separately:
      ladder  DRLAD/LADAB1 particle ladder on the A<=B list (o^2 v^3 (v+1)), the
              T1INW1 product that the engine folds into its rows (2 o^3 v^3), and
              the hole ladder + QUAD1 W(mnij) (4 o^4 v^2)
      wbuild  DWMBEJ, two (ov)^3 products (4 o^3 v^3)
      ring    DRRNG, two (ov)^3 products (4 o^3 v^3)"""
    return {'ladder': o * o * v ** 3 * (v + 1) + 2 * o ** 3 * v ** 3 + 4 * o ** 4 * v ** 2,
            'wbuild': 4 * o ** 3 * v ** 3, 'ring': 4 * o ** 3 * v ** 3}


def cpu_baseline(o, v, seed):
    """The reference algorithm's fp64 work per CC cycle at the full (o, v), timed
    on the host with numpy/BLAS on a bounded sample of each product class.

    The reference's CPU cycle is its XGEMM calls (BLAS dgemm) plus O(o^2 v^2)
    list passes.  Each class is timed at its real shape with one dimension
    sampled, and extrapolated linearly in that dimension (GEMM time is linear in
    each dimension at these sizes):
      particle ladder  <Ab|Cd>(A<=B, Cd) x tau(Cd, Ij): 768 of v(v+1)/2 rows
      (ov)^3 products  DWMBEJ x2, DRRNG x2, T1INW1: 1536 of ov columns, x5
      hole ladder      o^2 x v^2 x o^2, twice (W(mnij) and the ladder)
      o^2 v^3 products T1RING, QUAD3, FEACONT, T1INT2B, T2T1AB1 (~20 o^2 v^3 flops):
                       one (o^2 v) x v x v product, x10
      list passes      NEWT2, CMPENG, DODIIS0 (6 vectors), SYMETRIZE_T2, DRMOVE:
                       one a*x+y pass over o^2 v^2, x16
    Operand values do not change BLAS time; random operands are used."""
    import numpy as np
    threads = int(os.environ.get('OMP_NUM_THREADS') or min(16, os.cpu_count() or 1))
    rng = np.random.default_rng(seed)
    ov, oo, vv, nvp = o * v, o * o, v * v, v * (v + 1) // 2

    def timed(fn, reps=1):
        fn()
        best = None
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return best

    parts = {}
    rows = min(768, nvp)
    A = rng.standard_normal((rows, vv))
    T = rng.standard_normal((vv, oo))
    parts['particle_ladder'] = timed(lambda: A @ T) * nvp / rows
    Y = rng.standard_normal((oo, vv))
    parts['hole_ladder'] = 2 * timed(lambda: Y @ T, reps=1)
    del A, Y
    cols = min(1536, ov)
    W = rng.standard_normal((ov, ov))
    X = rng.standard_normal((ov, cols))
    parts['ov3_products'] = 5 * timed(lambda: W @ X) * ov / cols
    del X
    Z = rng.standard_normal((oo * v, v))
    Q = rng.standard_normal((v, v))
    parts['o2v3_products'] = 10 * timed(lambda: Z @ Q)
    del Z, Q
    x = W.reshape(-1)[:oo * vv]
    y = np.empty_like(x)
    parts['list_passes'] = 16 * timed(lambda: np.multiply(x, 0.5, out=y))
    del W, T
    total = sum(parts.values())
    return {'value': total, 'unit': 's/iter', 'cores': threads, 'kind': 'port',
            'sample': 'the reference algorithm\'s fp64 products per CC cycle at o=%d, v=%d on the host (numpy/'
                      'OpenBLAS, %d threads), each class at its real shape with one dimension sampled and '
                      'extrapolated linearly in it: particle ladder %d of %d rows, (ov)^3 products %d of %d '
                      'columns (x5), hole ladder in full (x2), one o^2v^3 product (x10), one o^2v^2 list pass '
                      '(x16); seconds per class %s' % (o, v, threads, rows, nvp, cols, ov,
                                                       {k: round(t, 3) for k, t in parts.items()})}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--workload', default='ccsd', choices=['ccsd', 'lambda', 'sym', 'triples'])
    ap.add_argument('--o', type=int, default=None)
    ap.add_argument('--v', type=int, default=None)
    ap.add_argument('--seed', type=int, default=20260)
    ap.add_argument('--prec', default='bf16x3', choices=['bf16x3', 'fp32', 'fp64'])
    ap.add_argument('--no-convergence', action='store_true', help='skip the convergence study')
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline')
    ap.add_argument('--maxcyc', type=int, default=60)
    ap.add_argument('--calib-copy', action='store_true',
                    help='copy 1 GiB device to device first (the reference dispatch that scales '
                         'FETCH_SIZE/WRITE_SIZE counters to bytes)')
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) > 1:
        print('bench.py: this build measures one GPU only (the CCSD path is benchmarked on a single '
              'MI355X); --gpus %d is not supported' % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.workload == 'lambda':
        return bench_lambda(args)
    if args.workload == 'sym':
        return bench_sym(args)
    if args.workload == 'triples':
        return bench_triples(args)
    import numpy as np
    import torch
    from aces2 import synth
    from aces2.ccsd import CCSD

    o = args.o or 48
    v = args.v or 320
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
    if args.calib_copy:
        src = torch.ones(1 << 27, dtype=torch.float64, device='cuda')
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log('calibration copy: %d bytes read, %d bytes written' % (src.numel() * 8, dst.numel() * 8))
        del src, dst
        torch.cuda.empty_cache()

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
    prof_all = cc.profile(reset=True)
    cc.set_profiling(False)
    prof = {k: prof_all[k] for k in ('ladder', 'wbuild', 'ring')}
    ms = sum(p['ms'] for p in prof.values())
    fl = sum(p['flops'] for p in prof.values())
    nl = sum(p['launches'] for p in prof.values())
    execf = 3.0 if args.prec == 'bf16x3' else 1.0
    peak = {'bf16x3': BF16_PEAK, 'fp32': FP32_PEAK, 'fp64': FP64_PEAK}[args.prec]
    # achieved: reference-equivalent flops of these products per second (their
    # F_ref share); executed MFMA flops (3 x 2mnk for bf16x3) give mfma_pct_peak
    fr = f_ref_classes(o, v)
    fref = nprof * sum(fr.values())
    achieved = fref / (ms * 1e-3)
    executed = execf * fl / (ms * 1e-3)
    roofline = {'bound': 'mfma', 'kernel': {'bf16x3': 'gemm_bf16x3_dl_kernel<4,8,2>', 'fp32': 'gemm16_kernel<float>',
                                             'fp64': 'gemm_f64_dl_kernel'}[args.prec],
                'achieved': achieved / 1e12, 'peak': peak / 1e12, 'unit': 'TFLOP/s',
                'frac': achieved / peak, 'traffic': None,
                'flops_basis': 'reference-equivalent flops (2mnk of the reference XGEMM shapes) of the particle '
                               'and hole ladders, W(mbej) build and ring products: %.4g per cycle' % (fref / nprof),
                'algorithmic_TFLOPs': fl / (ms * 1e-3) / 1e12,
                'mfma_pct_peak': 100.0 * executed / peak, 'executed_TFLOPs': executed / 1e12,
                'avg_launch_ms': ms / nl, 'launches_per_cycle': nl / nprof,
                'gemm_share_of_cycle': (ms / nprof) / (s_per_cycle * 1e3),
                'per_class': {k: {'ref_TFLOP/s': nprof * fr[k] / max(p['ms'], 1e-9) / 1e9,
                                  'executed_TFLOP/s': execf * p['flops'] / max(p['ms'], 1e-9) / 1e9,
                                  'ms': p['ms'] / nprof}
                              for k, p in prof.items()},
                'fp64_other': {'TFLOP/s': prof_all['other']['flops'] / max(prof_all['other']['ms'], 1e-9) / 1e9,
                               'ms': prof_all['other']['ms'] / nprof,
                               'launches': prof_all['other']['launches'] / nprof}}
    tr = measured_traffic('ccsd_cycle_o%d_v%d_%s' % (o, v, args.prec))
    if tr is not None:
        roofline['traffic'] = tr['bytes_per_launch']
        roofline['traffic_detail'] = tr
    log('roofline', json.dumps(roofline))

    # ---- the plain form's MFMA rate: one fp64 cycle with the same per-launch timing ----
    cc.set_profiling(True)
    cc.profile(reset=True)
    cc.cycle(args.warmup + args.steps + 1 + nprof, 'fp64', 1e-30)
    p64 = cc.profile(reset=True)
    cc.set_profiling(False)
    big = {k: p64[k] for k in ('ladder', 'wbuild', 'ring')}
    ms64 = sum(p['ms'] for p in big.values())
    fl64 = sum(p['flops'] for p in big.values())
    ms64all = ms64 + p64['other']['ms']
    fl64all = fl64 + p64['other']['flops']
    fp64_roofline = {'bound': 'mfma', 'kernel': 'gemm_f64_dl_kernel',
                     'achieved': fl64 / (ms64 * 1e-3) / 1e12,
                     'peak': FP64_PEAK / 1e12, 'unit': 'TFLOP/s', 'frac': fl64 / (ms64 * 1e-3) / FP64_PEAK,
                     'flops_basis': 'executed 2mnk (fp64 MFMA utilisation)',
                     'ref_equivalent_TFLOPs': sum(fr.values()) / (ms64 * 1e-3) / 1e12,
                     'scope': 'ladder + W build + ring contractions of one fp64 cycle',
                     'all_fp64_gemms': {'TFLOP/s': fl64all / (ms64all * 1e-3) / 1e12,
                                        'frac': fl64all / (ms64all * 1e-3) / FP64_PEAK, 'ms': ms64all},
                     'per_class': {k: {'TFLOP/s': p['flops'] / max(p['ms'], 1e-9) / 1e9, 'ms': p['ms']}
                                   for k, p in p64.items() if k != 'adjoint' and p['launches'] > 0}}
    log('fp64 roofline', json.dumps(fp64_roofline))

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
        'fp64_roofline': fp64_roofline,
    }

    # ---- convergence study (untimed for `value`) ----
    if not args.no_convergence:
        def run(form, conv, polish_conv=None, polish='refine'):
            cc.set_diis_order(6)
            cc.reset_diis()
            cc.start_mp1()
            times, hist = [], []
            icycle, e_final, phase, ic0 = 0, None, form, 0
            tol = 10.0 ** (-conv)
            while icycle < args.maxcyc:
                icycle += 1
                ta = time.perf_counter()
                st = cc.cycle(icycle - ic0, phase, tol)
                times.append((phase, time.perf_counter() - ta))
                hist.append((phase, st['energy'], st['resmiter']))
                if st['converged']:
                    if phase == form and polish_conv is not None:
                        # iterative refinement: continue with `polish` cycles to the
                        # tight threshold, DIIS restarted (the low-precision error
                        # vectors would hold the extrapolation at their accuracy);
                        # 'refine' anchors its difference products here
                        phase = polish
                        tol = 10.0 ** (-polish_conv)
                        cc.reset_diis()
                        cc.reset_anchor()
                        ic0 = icycle
                        continue
                    e_final = st['final_energy']
                    break
            return e_final, times, hist, icycle

        e64, times64, hist64, n64 = run('fp64', 9)
        log('fp64: E=%r in %d cycles' % (e64, n64))
        med = lambda xs: float(np.median(xs)) if xs else None
        s64 = med([t for p, t in times64[2:] if p == 'fp64'])
        result['convergence'] = {
            'fp64': {'energy': e64, 'cycles': n64, 's_per_cycle_median': s64,
                     'total_s': sum(t for _, t in times64)}}
        if args.prec != 'fp64':
            # headline: refinement polish; beside it the plain fp64 polish
            for key, pol in (('mixed', 'refine'), ('mixed_fp64_polish', 'fp64')):
                emx, timesmx, histmx, nmx = run(args.prec, 7, polish_conv=9, polish=pol)
                log('%s+%s polish: E=%r in %d cycles' % (args.prec, pol, emx, nmx))
                slo = med([t for p, t in timesmx[2:] if p == args.prec])
                spol = med([t for p, t in timesmx[1:] if p == pol])
                result['convergence'][key] = {
                    'energy': emx, 'cycles': nmx, 'polish': pol,
                    'lowp_cycles': sum(1 for p, _ in timesmx if p == args.prec),
                    'polish_cycles': sum(1 for p, _ in timesmx if p == pol),
                    's_per_cycle_median_lowp': slo, 's_per_cycle_median_polish': spol,
                    'total_s': sum(t for _, t in timesmx),
                    'abs_dE': abs(emx - e64) if (emx is not None and e64 is not None) else None,
                    'speedup_total': sum(t for _, t in times64) / sum(t for _, t in timesmx)}
            m = result['convergence']['mixed']
            result['convergence'].update({
                'abs_dE': m['abs_dE'],
                'speedup_per_cycle': (s64 / m['s_per_cycle_median_lowp'])
                if (s64 and m['s_per_cycle_median_lowp']) else None,
                'speedup_total': m['speedup_total']})
        result['fp64_s_per_iter'] = s64

    if not args.no_cpu:
        log('cpu baseline ...')
        result['cpu_baseline'] = cpu_baseline(o, v, args.seed)
    else:
        result['cpu_baseline'] = None
    print(json.dumps(result), flush=True)


def measured_traffic(name):
    """HBM bytes per launch of the dominant kernel, from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary profiles/hbm_traffic_<name>.md (the line
    'traffic-json: {...}'), or None when that workload has no such summary."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles', 'hbm_traffic_%s.md' % name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith('traffic-json:'):
                    return json.loads(line.split(':', 1)[1])
    except OSError:
        return None
    return None


def cpu_baseline_lambda(o, v, seed, budget_s=20.0):
    """Oracle reverse pass (torch autograd of the restated increments, CPU) on a
    bounded sample, scaled to (o, v) by reference-equivalent flops."""
    import numpy as np
    import torch
    from oracle import lambda_, rccsd
    from aces2 import synth
    threads = torch.get_num_threads()
    best = None
    for (os_, vs_) in [(8, 48), (12, 72), (16, 96)]:
        eps, B = synth.draw(os_, vs_, seed, sigma=synth.scaled_sigma(os_, vs_))
        blk, eri = synth.blocks_numpy(eps, B, os_)
        g = rccsd.blocks(eri, os_)
        f = blk['fock']
        rng = np.random.default_rng(0)
        t1 = rng.normal(0, 0.01, (os_, vs_))
        t2 = rng.normal(0, 0.01, (os_, os_, vs_, vs_))
        lambda_.lagrangian_grads(t1, t2, t1, t2, f, g, os_)
        t0 = time.perf_counter()
        lambda_.lagrangian_grads(t1, t2, t1, t2, f, g, os_)
        dt = time.perf_counter() - t0
        best = (os_, vs_, dt)
        if dt * (f_ref(os_ + 4, vs_ + 24) / f_ref(os_, vs_)) > budget_s:
            break
    os_, vs_, dt = best
    scale = f_ref(o, v) / f_ref(os_, vs_)
    return {'value': dt * scale, 'unit': 's/iter', 'cores': int(threads), 'kind': 'port',
            'sample': 'oracle.lambda_.lagrangian_grads (torch CPU autograd, %d threads; forward + reverse of the '
                      'increments) at o=%d, v=%d: %.3f s, scaled x%.1f by reference-equivalent flops to o=%d, v=%d'
                      % (threads, os_, vs_, dt, scale, o, v)}


def bench_lambda(args):
    """Lambda equations and response density at o=32, v=256 (differentiation form)."""
    import numpy as np
    import torch
    from aces2 import synth
    from aces2.ccsd import CCSD

    o = args.o or 32
    v = args.v or 256
    n = o + v
    log = lambda *a: print('[bench]', *a, file=sys.stderr, flush=True)
    conv = 10
    t_setup = time.perf_counter()
    sigma = synth.scaled_sigma(o, v)
    eps, B = synth.draw(o, v, args.seed, sigma=sigma)
    blocks = synth.blocks_device(eps, B, o)
    F0 = blocks['fock'].clone()
    cc = CCSD(o, v, blocks)
    del blocks
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    log('engine set up: %.1f GB device, %.1f s' % (cc.device_bytes / 1e9, time.perf_counter() - t_setup))

    def converge(c, prec='fp64', start='mp1', polish=False):
        c.set_diis_order(6)
        c.reset_diis()
        if start == 'mp1':
            c.start_mp1()
        tol = 10.0 ** (-(7 if prec != 'fp64' else conv))
        phase, icycle, ic0, ncyc = prec, 0, 0, {'lowp': 0, 'refine': 0, 'fp64': 0}
        while icycle < args.maxcyc:
            icycle += 1
            st = c.cycle(icycle - ic0, phase, tol)
            ncyc[phase if phase in ('fp64', 'refine') else 'lowp'] += 1
            if st['converged']:
                if phase == prec and prec != 'fp64':
                    # polish with refinement cycles anchored here (DIIS restarted)
                    phase, tol = 'refine', 10.0 ** (-conv)
                    c.reset_diis()
                    c.reset_anchor()
                    ic0 = icycle
                    continue
                return st['final_energy'], ncyc
        raise RuntimeError('CCSD did not converge in %d cycles' % args.maxcyc)

    ta = time.perf_counter()
    E0, ncyc0 = converge(cc)
    t_ccsd = time.perf_counter() - ta
    T0 = cc.get_amplitudes()
    log('fp64 CCSD: E=%r, %d cycles, %.1f s' % (E0, ncyc0['fp64'], t_ccsd))

    # ---- timed region: K Lambda cycles after W warmup cycles ----
    tol = 10.0 ** (-conv)
    cc.set_lambda_diis_order(6)
    cc.lambda_start()
    icycle = 0
    for _ in range(args.warmup):
        icycle += 1
        cc.lambda_cycle(icycle, tol)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        icycle += 1
        cc.lambda_cycle(icycle, tol)
    torch.cuda.synchronize()
    s_per_cycle = (time.perf_counter() - t0) / args.steps
    log('lambda: %.4f s/cycle' % s_per_cycle)
    cc.set_profiling(True)
    cc.profile(reset=True)
    nprof = 2
    converged = False
    for _ in range(nprof):
        icycle += 1
        converged = cc.lambda_cycle(icycle, tol)['converged'] or converged
    prof = cc.profile(reset=True)
    cc.set_profiling(False)
    while not converged and icycle < args.maxcyc:
        icycle += 1
        converged = cc.lambda_cycle(icycle, tol)['converged']
    if not converged:
        raise RuntimeError('Lambda equations did not converge')
    lam_cycles = icycle
    adj = prof['adjoint']
    achieved = adj['flops'] / (adj['ms'] * 1e-3)
    fp64_roofline = {'bound': 'mfma',
                     'kernel': 'gemm16_kernel<double> / gemm_f64_dl_kernel (reverse-pass contractions)',
                     'achieved': achieved / 1e12, 'peak': FP64_PEAK / 1e12, 'unit': 'TFLOP/s',
                     'frac': achieved / FP64_PEAK, 'avg_launch_ms': adj['ms'] / adj['launches'],
                     'launches_per_cycle': adj['launches'] / nprof,
                     'gemm_share_of_cycle': (adj['ms'] / nprof) / (s_per_cycle * 1e3)}
    log('fp64 roofline', json.dumps(fp64_roofline))

    # ---- whole Lambda solves: fp64, and mixed (fp64 anchors, the linear part
    # on the difference to the anchor with its ring/W-build/ladder products in
    # bf16x3); the mixed solve's seconds per cycle (anchors included) is `value`
    solves = {}
    for form in ('fp64', 'refine'):
        torch.cuda.synchronize()
        ta = time.perf_counter()
        h = cc.solve_lambda(maxcyc=args.maxcyc, conv=conv, order=6, prec=form)
        torch.cuda.synchronize()
        solves[form] = {'cycles': h['cycles'], 'total_s': time.perf_counter() - ta}
        solves[form]['s_per_cycle'] = solves[form]['total_s'] / h['cycles']
        log('lambda solve %s: %d cycles, %.2f s' % (form, h['cycles'], solves[form]['total_s']))
    solves['speedup_total'] = solves['fp64']['total_s'] / solves['refine']['total_s']
    # timed region of the mixed form: K cycles after W warmup cycles from the
    # Lambda start (the first warmup cycle is the fp64 anchor)
    s_fp64 = s_per_cycle
    if args.prec != 'fp64':
        cc.set_lambda_precision('refine')
        cc.lambda_start()
        icycle = 0
        for _ in range(args.warmup):
            icycle += 1
            cc.lambda_cycle(icycle, tol)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            icycle += 1
            cc.lambda_cycle(icycle, tol)
        torch.cuda.synchronize()
        s_per_cycle = (time.perf_counter() - t0) / args.steps
        cc.set_lambda_precision('fp64')
        log('lambda mixed: %.4f s/cycle (fp64 %.4f)' % (s_per_cycle, s_fp64))
    cc.set_profiling(True)
    cc.profile(reset=True)
    hp = cc.solve_lambda(maxcyc=args.maxcyc, conv=conv, order=6, prec='refine')
    pm = cc.profile(reset=True)
    cc.set_profiling(False)
    lp = pm['adjoint_lp']
    achieved = lp['flops'] / (lp['ms'] * 1e-3)
    roofline = {'bound': 'mfma', 'kernel': 'gemm_bf16x3_dl_kernel (reverse-pass products of the mixed cycles)',
                'achieved': achieved / 1e12, 'peak': BF16_PEAK / 1e12, 'unit': 'TFLOP/s',
                'frac': achieved / BF16_PEAK, 'traffic': None,
                'flops_basis': 'algorithmic 2mnk of the reverse-pass products',
                'mfma_pct_peak': 100.0 * 3.0 * achieved / BF16_PEAK, 'avg_launch_ms': lp['ms'] / lp['launches'],
                'launches_per_cycle': lp['launches'] / hp['cycles'],
                'fp64_anchor_ms_per_cycle': pm['adjoint']['ms'] / hp['cycles']}
    log('roofline', json.dumps(roofline))

    # ---- dE/dmu checks ----
    rng = np.random.default_rng(args.seed + 1)
    P = rng.normal(0.0, 0.01, (n, n))
    P = torch.tensor(P + P.T, dtype=torch.float64, device='cuda')
    x = rng.normal(0.0, sigma, (n, n))
    x = np.triu(x) + np.triu(x, 1).T
    hh = 1e-4

    def analytic(c, t1, t2):
        D = c.density()
        l1, l2 = c.get_lambda()
        an_f = float(torch.sum(D * P))
        bd = synth.blocks_device(np.zeros(n), x[None], o, fock_matrix=np.zeros((n, n)))
        cd = CCSD(o, v, bd)
        del bd
        cd.set_amplitudes(t1, t2)
        i1, i2 = cd.increments('fp64')
        an_g = cd.energy(t1, t2) + float(torch.sum(l1 * i1) + torch.sum(l2 * i2))
        cd.close()
        del i1, i2, l1, l2
        torch.cuda.empty_cache()
        return an_f, an_g

    an_f, an_g = analytic(cc, *T0)
    log('analytic fp64: fock %.12e integral %.12e' % (an_f, an_g))
    e_f = []
    for mu in (hh, -hh, 2 * hh, -2 * hh):
        cc.set_fock(F0 + mu * P)
        cc.set_amplitudes(*T0)
        e_f.append(converge(cc, start='current')[0])
    cc.set_fock(F0)
    fd_f = (e_f[0] - e_f[1]) / (2 * hh)
    fd_f2 = (e_f[2] - e_f[3]) / (4 * hh)
    # mixed form: bf16x3 to 1e-7, refinement polish (DIIS restarted) to 1e-10, Lambda in fp64 at those amplitudes
    Em, ncm = converge(cc, prec=args.prec if args.prec != 'fp64' else 'bf16x3')
    Tm = cc.get_amplitudes()
    cc.set_lambda_diis_order(6)
    lm = cc.solve_lambda(maxcyc=args.maxcyc, conv=conv, order=6, prec='refine')
    an_fm, an_gm = analytic(cc, *Tm)
    cc.close()
    del cc
    torch.cuda.empty_cache()
    e_g = []
    Baug = np.concatenate([B, x[None]], axis=0)
    for mu in (hh, -hh, 2 * hh, -2 * hh):
        w = np.ones(Baug.shape[0])
        w[-1] = mu
        bp = synth.blocks_device(eps, Baug, o, weights=w)
        c2 = CCSD(o, v, bp)
        del bp
        c2.set_amplitudes(*T0)
        e_g.append(converge(c2, start='current')[0])
        c2.close()
        torch.cuda.empty_cache()
    fd_g = (e_g[0] - e_g[1]) / (2 * hh)
    fd_g2 = (e_g[2] - e_g[3]) / (4 * hh)
    # Richardson: the O(h^2) truncation of the central difference cancels in (4 FD(h) - FD(2h)) / 3
    rf, rg = (4 * fd_f - fd_f2) / 3, (4 * fd_g - fd_g2) / 3
    fd = {'h': hh, 'conv': conv,
          'fock': {'analytic_fp64': an_f, 'analytic_mixed_polished': an_fm, 'central_fd': fd_f,
                   'central_fd_2h': fd_f2, 'richardson': rf,
                   'abs_err_fp64': abs(an_f - fd_f), 'abs_err_mixed': abs(an_fm - fd_f),
                   'abs_err_fp64_richardson': abs(an_f - rf), 'abs_err_mixed_richardson': abs(an_fm - rf)},
          'integral_rank1': {'analytic_fp64': an_g, 'analytic_mixed_polished': an_gm, 'central_fd': fd_g,
                             'central_fd_2h': fd_g2, 'richardson': rg,
                             'abs_err_fp64': abs(an_g - fd_g), 'abs_err_mixed': abs(an_gm - fd_g),
                             'abs_err_fp64_richardson': abs(an_g - rg), 'abs_err_mixed_richardson': abs(an_gm - rg)},
          'ccsd_energy_fp64': E0, 'ccsd_energy_mixed': Em, 'abs_dE_mixed': abs(Em - E0),
          'ccsd_cycles_fp64': ncyc0['fp64'], 'ccsd_cycles_mixed': ncm,
          'lambda_cycles': lam_cycles, 'lambda_cycles_mixed': lm.get('cycles')}
    log('fd', json.dumps(fd))
    result = {
        'metric': 'Lambda (xlambda) s/iter at o=%d,v=%d on GPU; dE/dmu vs central FD' % (o, v),
        'value': s_per_cycle,
        'unit': 's/iter',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': s_per_cycle * 1e3,
        'higher_is_better': False,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'bf16x3/fp32-acc + f64' if args.prec != 'fp64' else 'f64',
        'data': 'synthetic (DF integrals, seed %d, sigma=0.05*sqrt(95/(o*v)))' % args.seed,
        'config': {'workload': 'closed-shell CCSD Lambda cycle (reverse pass + Jacobi + DIIS), RHF C1, o=%d v=%d '
                               'n_aux=%d' % (o, v, 4 * n), 'o': o, 'v': v, 'diis': 6,
                   'form': 'mixed (fp64 anchor + bf16x3 linear part)' if args.prec != 'fp64' else 'fp64'},
        'roofline': roofline if args.prec != 'fp64' else dict(fp64_roofline, traffic=None),
        'fp64_roofline': fp64_roofline,
        'fp64_s_per_iter': s_fp64,
        'lambda_solves': solves,
        'fd_checks': fd,
    }
    if not args.no_cpu:
        log('cpu baseline ...')
        result['cpu_baseline'] = cpu_baseline_lambda(o, v, args.seed)
    else:
        result['cpu_baseline'] = None
    print(json.dumps(result), flush=True)


def bench_sym(args):
    """Symmetry-blocked grouped form (D2h, o=64, v=384)."""
    import numpy as np
    import torch
    from aces2 import synth
    from aces2.ccsd import SymCCSD

    o = args.o or 64
    v = args.v or 384
    nirr = 8
    log = lambda *a: print('[bench]', *a, file=sys.stderr, flush=True)
    rng = np.random.default_rng(args.seed + 5)
    occ = synth.irrep_split(o, nirr, 2, rng)
    vir = synth.irrep_split(v, nirr, 16, rng)
    t0 = time.perf_counter()
    eps, B, irr = synth.sym_draw(occ, vir, args.seed)
    log('irreps occ %s vir %s; factors drawn (%.1f s)' % (occ, vir, time.perf_counter() - t0))
    blocks = synth.blocks_device(eps, B, o)
    del B
    torch.cuda.synchronize()
    log('dense integral blocks on the GPU (%.1f s)' % (time.perf_counter() - t0))
    cc = SymCCSD(occ, vir, blocks)
    del blocks
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    log('irrep-blocked engine set up: %.1f GB, packed T2 %d of %d dense elements (%.1f s)' %
        (cc.device_bytes / 1e9, cc.packed_t2, o * o * v * v, time.perf_counter() - t0))
    prec = args.prec if args.prec != 'fp64' else 'bf16x3'

    def timed_cycles(form):
        cc.set_diis_order(6)
        cc.reset_diis()
        cc.start_mp1()
        for ic in range(1, args.warmup + 1):
            cc.cycle(ic, form, 1e-30)
        torch.cuda.synchronize()
        ta = time.perf_counter()
        for ic in range(args.warmup + 1, args.warmup + args.steps + 1):
            cc.cycle(ic, form, 1e-30)
        torch.cuda.synchronize()
        return (time.perf_counter() - ta) / args.steps

    s_mixed = timed_cycles(prec)
    log('%s: %.4f s/cycle' % (prec, s_mixed))
    s64 = timed_cycles('fp64')
    log('fp64: %.4f s/cycle' % s64)
    cc.set_profiling(True)
    cc.profile(reset=True)
    cc.start_mp1()
    cc.reset_diis()
    for ic in (1, 2):
        cc.cycle(ic, prec, 1e-30)
    prof = cc.profile(reset=True)
    cc.set_profiling(False)
    heavy = {k: prof[k] for k in ('ladder', 'wbuild', 'ring')}
    ms = sum(p['ms'] for p in heavy.values())
    fl = sum(p['flops'] for p in heavy.values())
    nl = sum(p['launches'] for p in heavy.values())
    execf = 3.0 if prec == 'bf16x3' else 1.0
    peak = BF16_PEAK if prec == 'bf16x3' else FP32_PEAK
    roofline = {'bound': 'mfma', 'kernel': 'grouped %s GEMM (8 irrep blocks per launch)' % prec,
                'achieved': fl / (ms * 1e-3) / 1e12, 'peak': peak / 1e12, 'unit': 'TFLOP/s',
                'frac': fl / (ms * 1e-3) / peak, 'traffic': None,
                'flops_basis': 'algorithmic 2mnk of the irrep-block products',
                'mfma_pct_peak': 100.0 * execf * fl / (ms * 1e-3) / peak, 'avg_launch_ms': ms / max(nl, 1),
                'launches_per_cycle': nl / 2, 'gemm_share_of_cycle': (ms / 2) / (s_mixed * 1e3),
                'per_class': {k: {'TFLOP/s': execf * p['flops'] / max(p['ms'], 1e-9) / 1e9, 'ms': p['ms'] / 2}
                              for k, p in heavy.items()},
                'fp64_other': {'TFLOP/s': prof['other']['flops'] / max(prof['other']['ms'], 1e-9) / 1e9,
                               'ms': prof['other']['ms'] / 2}}
    log('roofline', json.dumps(roofline))

    # convergence: mixed to 1e-7, refinement polish (DIIS restarted) to 1e-9; fp64 reference run
    def converge(form, conv_lo, conv_hi):
        cc.set_diis_order(6)
        cc.reset_diis()
        cc.start_mp1()
        phase, tol, ic, ic0, n = form, 10.0 ** (-conv_lo), 0, 0, {'lowp': 0, 'refine': 0, 'fp64': 0}
        while ic < args.maxcyc:
            ic += 1
            st = cc.cycle(ic - ic0, phase, tol)
            n[phase if phase in ('fp64', 'refine') else 'lowp'] += 1
            if st['converged']:
                if phase == form and form != 'fp64':
                    # polish with refinement cycles anchored here (DIIS restarted)
                    phase, tol = 'refine', 10.0 ** (-conv_hi)
                    cc.reset_diis()
                    cc.reset_anchor()
                    ic0 = ic
                    continue
                return st['final_energy'], n
        raise RuntimeError('CCSD did not converge')

    ta = time.perf_counter()
    e64, n64 = converge('fp64', 9, 9)
    t64 = time.perf_counter() - ta
    ta = time.perf_counter()
    emx, nmx = converge(prec, 7, 9)
    tmx = time.perf_counter() - ta
    log('fp64 %r (%s) %.1f s; mixed+polish %r (%s) %.1f s' % (e64, n64, t64, emx, nmx, tmx))
    # Lambda on the same lists at the polished amplitudes: fp64 cycles, then
    # mixed cycles (fp64 anchor + bf16x3 linear part) to the same multipliers
    ta = time.perf_counter()
    lh = cc.solve_lambda(maxcyc=args.maxcyc, conv=9, order=6)
    tl = time.perf_counter() - ta
    s_lam = float(np.median(lh['ms_cycle'])) / 1e3
    D = cc.density()
    trace = float(torch.trace(D))
    log('lambda fp64: %d cycles, %.1f s, median %.4f s/cycle; trace(D_corr) = %.2e' %
        (lh['cycles'], tl, s_lam, trace))
    ta = time.perf_counter()
    lm = cc.solve_lambda(maxcyc=args.maxcyc, conv=9, order=6, prec='refine')
    tlm = time.perf_counter() - ta
    s_lam_mx = float(np.median(lm['ms_cycle'])) / 1e3
    dD = float((cc.density() - D).abs().max())
    log('lambda mixed: %d cycles, %.1f s, median %.4f s/cycle; |D_mixed - D_fp64| = %.1e' %
        (lm['cycles'], tlm, s_lam_mx, dD))
    cc.close()
    result = {
        'metric': 'CCSD s/iter, D2h irrep-blocked o=%d,v=%d (grouped MFMA), mixed vs FP64; Lambda s/iter' % (o, v),
        'value': s_mixed,
        'unit': 's/iter',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': s_mixed * 1e3,
        'higher_is_better': False,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'bf16x3/fp32-acc + f64' if prec == 'bf16x3' else 'f32 + f64',
        'data': 'synthetic (DF integrals with D2h symmetry, seed %d, sigma=0.05*sqrt(95/(o*v)))' % args.seed,
        'config': {'workload': 'closed-shell CCSD cycle, RHF D2h (8 irreps), o=%d v=%d' % (o, v), 'o': o, 'v': v,
                   'occ_per_irrep': occ, 'vir_per_irrep': vir, 'form': prec, 'diis': 6},
        'roofline': roofline,
        'fp64_s_per_iter': s64,
        'speedup_mixed_vs_fp64': s64 / s_mixed,
        'convergence': {'fp64': {'energy': e64, 'cycles': n64, 'total_s': t64},
                        'mixed_polish': {'energy': emx, 'cycles': nmx, 'total_s': tmx},
                        'abs_dE': abs(emx - e64)},
        'lambda': {'cycles': lh['cycles'], 's_per_cycle_median': s_lam, 'total_s': tl,
                   'density_trace_corr': trace,
                   'mixed': {'cycles': lm['cycles'], 's_per_cycle_median': s_lam_mx, 'total_s': tlm,
                             'max_abs_density_diff_vs_fp64': dD}},
        'cpu_baseline': None,
    }
    print(json.dumps(result), flush=True)



def cpu_baseline_triples(o, v, eps, B, t1, t2, nsample=2):
    """The oracle's per-triple (T) work (oracle.triples.triple_energy, numpy/BLAS)
    timed on the host for a sample of occupied triples with three distinct
    indices at the full (o, v), scaled to all o(o+1)(o+2)/6 triples."""
    import numpy as np
    from oracle import triples
    threads = int(os.environ.get('OMP_NUM_THREADS') or min(16, os.cpu_count() or 1))
    naux = B.shape[0]
    Bov = np.ascontiguousarray(B[:, :o, o:])
    Bvv = np.ascontiguousarray(B[:, o:, o:]).reshape(naux, v * v)
    Boo = np.ascontiguousarray(B[:, :o, :o])
    ovoo = (Bov.reshape(naux, o * v).T @ Boo.reshape(naux, o * o)).reshape(o, v, o, o)
    ovov = (Bov.reshape(naux, o * v).T @ Bov.reshape(naux, o * v)).reshape(o, v, o, v)
    eo, ev = eps[:o], eps[o:]
    sample = [(o - 1, o // 2, 0), (o - 2, o // 3, 1)][:nsample]
    ovvv = {}
    for trip in sample:
        for x in trip:
            if x not in ovvv:
                ovvv[x] = (Bov[:, x, :].T @ Bvv).reshape(v, v, v)
    g = {'ovvv': ovvv, 'ovoo': ovoo, 'ovov': ovov}
    triples.triple_energy(t1, t2, g, eo, ev, *sample[0])
    t0 = time.perf_counter()
    for trip in sample:
        triples.triple_energy(t1, t2, g, eo, ev, *trip)
    per = (time.perf_counter() - t0) / len(sample)
    ntrip = o * (o + 1) * (o + 2) // 6
    return {'value': per * ntrip, 'unit': 's/eval', 'cores': threads, 'kind': 'port',
            'sample': 'oracle.triples.triple_energy (numpy/OpenBLAS, %d threads; W of one occupied triple from '
                      'six v^3 particle and six hole products, then the energy sums) for %d triples with distinct '
                      'indices at o=%d, v=%d: %.3f s per triple, x%d triples' % (threads, len(sample), o, v, per,
                                                                                ntrip)}


def bench_triples(args):
    """(T) of CCSD(T) at o=32, v=256 after the 20-cycle fp64 CCSD run."""
    import numpy as np
    import torch
    from aces2 import synth
    from aces2.ccsd import CCSD

    o = args.o or 32
    v = args.v or 256
    log = lambda *a: print('[bench]', *a, file=sys.stderr, flush=True)
    t_setup = time.perf_counter()
    eps, B = synth.draw(o, v, args.seed, sigma=synth.scaled_sigma(o, v))
    blocks = synth.blocks_device(eps, B, o)
    cc = CCSD(o, v, blocks)
    del blocks
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    log('engine set up: %.1f GB device, %.1f s' % (cc.device_bytes / 1e9, time.perf_counter() - t_setup))
    h = cc.run(maxcyc=20, conv=14, order=6, start='mp1', raise_on_fail=False)
    e_cc = h['energy'][-1]
    log('CCSD: 20 fp64 cycles, E = %.12f' % e_cc)
    steps, warmup = max(1, args.steps), max(0, args.warmup)
    for _ in range(warmup):
        cc.triples(prec='fp64')
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r = cc.triples(prec='fp64')
    torch.cuda.synchronize()
    s_eval = (time.perf_counter() - t0) / steps
    log('(T) fp64: %.3f s per evaluation, E[4]T %.12f E[5]ST %.12f' % (s_eval, r['e4t'], r['e5st']))
    # per-launch timing of the W products (hipEvents on the engine's stream)
    cc.set_profiling(True)
    cc.profile(reset=True)
    cc.triples(prec='fp64')
    p = cc.profile(reset=True)['triples']
    rb = cc.triples(prec='bf16x3')
    pb = cc.profile(reset=True)['triples']
    cc.set_profiling(False)
    t0 = time.perf_counter()
    rb = cc.triples(prec='bf16x3')
    s_bf = time.perf_counter() - t0
    ntrip = r['ntriples']
    algo = ntrip * 6 * 2.0 * v ** 3 * (v + o)
    kp64 = (v + o + 15) // 16 * 16
    exec64 = ntrip * 6 * 2.0 * v ** 3 * kp64
    achieved = algo / (p['ms'] * 1e-3)
    roofline = {'bound': 'mfma', 'kernel': 'gemm_f64_dl_kernel',
                'achieved': achieved / 1e12, 'peak': FP64_PEAK / 1e12, 'unit': 'TFLOP/s',
                'frac': achieved / FP64_PEAK, 'traffic': None,
                'flops_basis': 'algorithmic 2 v^3 (v + o) per permutation, 6 per occupied triple: %.4g per (T)' % algo,
                'executed_TFLOPs': exec64 / (p['ms'] * 1e-3) / 1e12,
                'avg_launch_ms': p['ms'] / max(p['launches'], 1), 'launches_per_eval': p['launches'],
                'gemm_share_of_eval': p['ms'] * 1e-3 / s_eval}
    tr = measured_traffic('triples_o%d_v%d_fp64' % (o, v))
    if tr is not None:
        roofline['traffic'] = tr['bytes_per_launch']
        roofline['traffic_detail'] = tr
    log('roofline', json.dumps(roofline))
    dE = max(abs(rb[k] - r[k]) for k in ('e4t', 'e5st', 'e4t_aaa', 'e4t_baa'))
    result = {
        'metric': '(T) s per evaluation at o=32,v=256 (fp64 W products)',
        'value': s_eval,
        'unit': 's/eval',
        'n_gpus': 1,
        'steps': steps,
        'warmup': warmup,
        'ms_per_step': s_eval * 1e3,
        'higher_is_better': False,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic (DF integrals, seed %d, sigma=0.05*sqrt(95/(o*v))), amplitudes of 20 fp64 CCSD cycles'
                % args.seed,
        'config': {'workload': 'closed-shell (T) after CCSD, RHF C1, o=%d v=%d' % (o, v), 'o': o, 'v': v,
                   'triples': ntrip},
        'energies': {'ccsd_corr': e_cc, 'e4t': r['e4t'], 'e5st': r['e5st'], 'e4t_aaa': r['e4t_aaa'],
                     'e4t_baa': r['e4t_baa']},
        'algorithmic_TFLOPs_eval': algo / s_eval / 1e12,
        'roofline': roofline,
        'bf16x3': {'s_per_eval': s_bf, 'abs_dE_vs_fp64': dE, 'rel_dE_vs_fp64': dE / abs(r['e4t']),
                   'algorithmic_TFLOPs_eval': algo / s_bf / 1e12,
                   'W_products_ms': pb['ms'], 'W_products_executed_TFLOPs':
                       3.0 * ntrip * 6 * 2.0 * v ** 3 * ((v + o + 31) // 32 * 32) / (pb['ms'] * 1e-3) / 1e12},
    }
    if not args.no_cpu:
        t1, t2 = cc.get_amplitudes()
        t1h, t2h = t1.cpu().numpy(), t2.cpu().numpy()
        cc.close()
        result['cpu_baseline'] = cpu_baseline_triples(o, v, eps, B, t1h, t2h)
    else:
        result['cpu_baseline'] = None
    print(json.dumps(result), flush=True)

if __name__ == '__main__':
    main()
