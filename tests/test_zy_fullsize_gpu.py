"""The GPU path at the full sizes of the measured workloads.

Every case runs the HIP library through the C ABI at its real shape and checks
it with size-independent properties and with the element-sampled oracle
(``oracle/slices.py``, pinned against ``oracle.rccsd`` in ``test_oracle.py``):

  o=32, v=256   fp64 CCSD: MP1 / closed-form MP2 from T = 0, the reference's
                20-cycle run from the MP1 start, fp64 increments against the
                oracle's inc1 and inc2[i, :, a, :] slices, (T) of sampled
                occupied triples against the oracle's, bf16x3 and refinement
                increments against fp64
  o=48, v=320   mixed (bf16x3, refinement polish) vs fp64 energies to 1e-9
                (|dE| <= 1e-7 after the polish), increments at the solution
                against the oracle, including i = o-1 and a = v-1 whose
                particle-ladder operand rows lie past element 2^31
  o=32, v=256   Lambda and response density: dE/dmu for a Fock and a rank-1
                integral perturbation against central differences (h = 1e-4)
                of energies converged to 1e-12, fp64 and mixed forms
  D2h, o=64, v=384 (8 irreps)  irrep-blocked engine: mixed vs fp64 energies,
                increments against the oracle, Lambda on the same lists

Each case frees its engine before the next one (the o=48 case holds ~220 GB).
The file sorts after the small-case GPU tests so that their failures show first.
"""
import gc

import numpy as np
import pytest

from oracle import slices

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu


def _free():
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def _c1_engine(o, v, seed):
    from aces2 import synth
    from aces2.ccsd import CCSD
    eps, B = synth.draw(o, v, seed, sigma=synth.scaled_sigma(o, v))
    blocks = synth.blocks_device(eps, B, o)
    cc = CCSD(o, v, blocks)
    del blocks
    _free()
    return cc, eps, B


def _pairs(o, v):
    return [(0, 0), (o - 1, v - 1), (o - 1, 0), (0, v - 1), (o // 2, v // 3)]


def _check_slices(S, t1, t2, i1, i2, pairs, rtol, what):
    """GPU increments (host copies) against the element-sampled oracle."""
    r1, sl = slices.increments(S, t1, t2, pairs)
    s1 = max(np.abs(r1).max(), 1e-300)
    e1 = np.abs(i1 - r1).max() / s1
    errs = []
    for (i, a) in pairs:
        ref = sl[(i, a)]
        errs.append(np.abs(i2[i, :, a, :] - ref).max() / max(np.abs(ref).max(), 1e-300))
    print('%s: inc1 rel err %.2e, inc2 slice rel errs %s' % (what, e1, ['%.1e' % x for x in errs]), flush=True)
    assert e1 < rtol
    assert max(errs) < rtol
    return r1, sl


def _check_triples(S, B, t1, t2, per, sample):
    """Per-triple (T) sums from the GPU against the oracle's for sampled occupied
    triples i >= j >= k (the (xp|qd) integrals built from the DF factors for the
    triple's occupied indices only)."""
    from oracle import triples
    o, v = S.o, S.v
    P = S.naux
    ovvv = {}
    for trip in sample:
        for x in trip:
            if x not in ovvv:
                ovvv[x] = (S.Bov[:, x, :].T @ S.Bvv.reshape(P, v * v)).reshape(v, v, v)
    g = {'ovvv': ovvv, 'ovoo': S.ovoo, 'ovov': S.ovov}
    eo = np.diag(S.f)[:o]
    ev = np.diag(S.f)[o:]
    refs = {trip: np.array(triples.triple_energy(t1, t2, g, eo, ev, *trip)) for trip in sample}
    # a triple i = j = k has no same-spin or mixed triples: its sums are rounding
    # noise of the six cancelling orderings, checked against the sample's scale
    scale = max(np.abs(r).max() for r in refs.values())
    for (i, j, k), ref in refs.items():
        t = i * (i + 1) * (i + 2) // 6 + j * (j + 1) // 2 + k
        err = np.abs(per[t] - ref).max() / (scale if i == j == k else np.abs(ref).max())
        print('(T) triple %s: rel err %.1e' % ((i, j, k), err), flush=True)
        assert err < 1e-11


def test_fullsize_o32_v256_fp64():
    o, v, seed = 32, 256, 20262
    print('case full size o=%d v=%d fp64' % (o, v), flush=True)
    cc, eps, B = _c1_engine(o, v, seed)
    S = slices.DFSystem(B, np.diag(eps), o)
    pairs = _pairs(o, v)
    # from T = 0, one fp64 cycle: T_new = MP1 and E(T_new) = the closed-form MP2 energy
    cc.set_diis_order(6)
    cc.reset_diis()
    cc.set_amplitudes(np.zeros((o, v)), np.zeros((o, o, v, v)))
    st = cc.cycle(1, 'fp64', 1e-30)
    e_mp2 = slices.mp2_energy(S)
    print('E(MP2) gpu %.15f closed form %.15f' % (st['energy'], e_mp2), flush=True)
    assert abs(st['energy'] - e_mp2) < 1e-12 * abs(e_mp2)
    t1, t2 = cc.get_amplitudes()
    assert float(t1.abs().max()) == 0.0
    t2h = t2.cpu().numpy()
    for (i, a) in pairs:
        ref = slices.mp1(S, i, a)
        assert np.abs(t2h[i, :, a, :] - ref).max() < 1e-13 * np.abs(ref).max()
    # the configuration's run: MP1 start, 20 cycles, 6 DIIS vectors
    h = cc.run(maxcyc=20, conv=14, order=6, start='mp1', raise_on_fail=False)
    print('20 cycles: E %.12f RESMITER %s' % (h['energy'][-1], ['%.1e' % x for x in h['resmiter'][::4]]), flush=True)
    assert len(h['resmiter']) == 20
    assert abs(h['resmiter'][-1]) < 1e-4 * abs(h['resmiter'][0])
    # increments at the 20-cycle amplitudes against the oracle
    t1, t2 = cc.get_amplitudes()
    i1, i2 = cc.increments('fp64')
    t1h, t2h = t1.cpu().numpy(), t2.cpu().numpy()
    i1h, i2h = i1.cpu().numpy(), i2.cpu().numpy()
    e_gpu = cc.energy(t1, t2)
    e_ref = slices.energy(S, t1h, t2h)
    assert abs(e_gpu - e_ref) < 1e-12 * abs(e_ref)
    _check_slices(S, t1h, t2h, i1h, i2h, pairs, 1e-11, 'fp64 o=32 v=256')
    # (T) at the 20-cycle amplitudes: sampled occupied triples against the oracle
    r = cc.triples(per_triple=True)
    per = r['per_triple'].cpu().numpy()
    print('(T): E[4]T %.12f E[5]ST %.12f, %d triples in %.0f ms (%.1f TF)' %
          (r['e4t'], r['e5st'], r['ntriples'], r['ms'], r['flops'] / r['ms'] / 1e9), flush=True)
    _check_triples(S, B, t1h, t2h, per, [(0, 0, 0), (o - 1, o - 1, o - 2), (o - 1, o // 2, 3), (o - 1, o - 2, o - 3)])
    rb = cc.triples(prec='bf16x3')
    eb = max(abs(rb[k] - r[k]) for k in ('e4t', 'e5st', 'e4t_aaa', 'e4t_baa')) / abs(r['e4t'])
    print('(T) bf16x3: %.0f ms (%.1f TF algorithmic), rel err vs fp64 %.1e' %
          (rb['ms'], rb['flops'] / rb['ms'] / 1e9, eb), flush=True)
    assert eb < 2e-5
    # bf16x3 increments: fp32-accumulated split products, relative error ~1e-6
    b1, b2 = cc.increments('bf16x3')
    eb = float((b2 - i2).abs().max()) / float(i2.abs().max())
    print('bf16x3 inc2 rel err %.2e' % eb, flush=True)
    assert 0 < eb < 5e-5
    # refinement: the first evaluation records the anchor (= fp64); one cycle later
    # the refined increments carry a small fraction of the bf16x3 error
    cc.reset_anchor()
    r1, r2 = cc.increments('refine')
    assert float((r2 - i2).abs().max()) < 1e-13 * float(i2.abs().max())
    cc.reset_diis()
    cc.cycle(1, 'refine', 1e-30)
    f1, f2 = cc.increments('fp64')
    r1, r2 = cc.increments('refine')
    b1, b2 = cc.increments('bf16x3')
    er = float((r2 - f2).abs().max())
    eb = float((b2 - f2).abs().max())
    print('after one refinement cycle: refine err %.2e bf16x3 err %.2e' % (er, eb), flush=True)
    assert er < 0.05 * eb
    cc.close()
    del t1, t2, i1, i2, b1, b2, r1, r2, f1, f2
    _free()


def test_fullsize_o48_v320_mixed_and_fp64():
    o, v, seed = 48, 320, 20260
    print('case full size o=%d v=%d mixed vs fp64' % (o, v), flush=True)
    cc, eps, B = _c1_engine(o, v, seed)
    h64 = cc.run(maxcyc=60, conv=9, order=6, start='mp1')
    hmx = cc.run(maxcyc=60, conv=7, order=6, start='mp1', prec='bf16x3', polish_conv=9, polish='refine')
    de = abs(hmx['final_energy'] - h64['final_energy'])
    print('fp64 E %.12f (%d cycles); bf16x3 + refinement E %.12f (%d + %d cycles); |dE| %.2e' %
          (h64['final_energy'], h64['cycles'], hmx['final_energy'], hmx['precision'].count('bf16x3'),
           hmx['precision'].count('refine'), de), flush=True)
    assert de <= 1e-7
    assert hmx['precision'][0] == 'bf16x3' and hmx['precision'][-1] == 'refine'
    # increments at the polished solution: fp64 against the oracle, the refined
    # products (anchored during the polish) and plain bf16x3 against fp64
    t1, t2 = cc.get_amplitudes()
    i1, i2 = cc.increments('fp64')
    r1, r2 = cc.increments('refine')
    b1, b2 = cc.increments('bf16x3')
    er = float((r2 - i2).abs().max()) / float(i2.abs().max())
    eb = float((b2 - i2).abs().max()) / float(i2.abs().max())
    print('refine rel err %.2e, bf16x3 rel err %.2e' % (er, eb), flush=True)
    assert eb < 5e-5 and er < 1e-3 * eb
    t1h, t2h = t1.cpu().numpy(), t2.cpu().numpy()
    i1h, i2h = i1.cpu().numpy(), i2.cpu().numpy()
    cc.close()
    del t1, t2, i1, i2, r1, r2, b1, b2
    _free()
    S = slices.DFSystem(B, np.diag(eps), o)
    assert abs(slices.energy(S, t1h, t2h) - hmx['final_energy']) < 1e-12 * abs(hmx['final_energy'])
    _check_slices(S, t1h, t2h, i1h, i2h, _pairs(o, v), 1e-11, 'fp64 o=48 v=320')


def test_fullsize_o32_v256_lambda_fd():
    from aces2 import synth
    from aces2.ccsd import CCSD
    o, v, seed = 32, 256, 20262
    n = o + v
    print('case full size o=%d v=%d Lambda / FD' % (o, v), flush=True)
    cc, eps, B = _c1_engine(o, v, seed)
    F0 = torch.tensor(np.diag(eps), dtype=torch.float64, device='cuda')
    rng = np.random.default_rng(seed + 1)
    P = rng.normal(0.0, 0.01, (n, n))
    P = torch.tensor(P + P.T, dtype=torch.float64, device='cuda')
    x = rng.normal(0.0, synth.scaled_sigma(o, v), (n, n))
    x = np.triu(x) + np.triu(x, 1).T
    h = 1e-4
    conv = 12

    def converge(c, start='mp1', prec='fp64'):
        if prec == 'fp64':
            r = c.run(maxcyc=100, conv=conv, order=6, start=start)
        else:
            r = c.run(maxcyc=100, conv=7, order=6, start=start, prec=prec, polish_conv=conv, polish='refine')
        return r['final_energy']

    def analytic(c, t1, t2):
        D = c.density()
        l1, l2 = c.get_lambda()
        an_f = float(torch.sum(D * P))
        # rank-1 (pq|rs) += mu x_pq x_rs: dE/dmu = E_x(T) + l . inc_x(T), both linear in the integrals
        bd = synth.blocks_device(np.zeros(n), x[None], o, fock_matrix=np.zeros((n, n)))
        cd = CCSD(o, v, bd)
        del bd
        cd.set_amplitudes(t1, t2)
        j1, j2 = cd.increments('fp64')
        an_g = cd.energy(t1, t2) + float(torch.sum(l1 * j1) + torch.sum(l2 * j2))
        cd.close()
        del j1, j2, l1, l2, D
        _free()
        return an_f, an_g

    E0 = converge(cc)
    T0 = cc.get_amplitudes()
    lh = cc.solve_lambda(maxcyc=100, conv=conv, order=6)
    an_f, an_g = analytic(cc, *T0)
    # mixed: bf16x3 + refinement polish, Lambda with mixed cycles
    Em = converge(cc, prec='bf16x3')
    Tm = cc.get_amplitudes()
    cc.solve_lambda(maxcyc=100, conv=conv, order=6, prec='refine')
    an_fm, an_gm = analytic(cc, *Tm)
    assert abs(Em - E0) < 1e-9
    e_f = []
    for mu in (h, -h):
        cc.set_fock(F0 + mu * P)
        cc.set_amplitudes(*T0)
        e_f.append(converge(cc, start='current'))
    cc.close()
    del cc, T0, Tm
    _free()
    fd_f = (e_f[0] - e_f[1]) / (2 * h)
    e_g = []
    Baug = np.concatenate([B, x[None]], axis=0)
    for mu in (h, -h):
        w = np.ones(Baug.shape[0])
        w[-1] = mu
        bp = synth.blocks_device(eps, Baug, o, weights=w)
        c2 = CCSD(o, v, bp)
        del bp
        e_g.append(converge(c2))
        c2.close()
        _free()
    fd_g = (e_g[0] - e_g[1]) / (2 * h)
    print('Lambda %d cycles; dE/dmu fock: analytic %.12e mixed %.12e FD %.12e; rank-1: analytic %.12e mixed %.12e '
          'FD %.12e' % (lh['cycles'], an_f, an_fm, fd_f, an_g, an_gm, fd_g), flush=True)
    for an in (an_f, an_fm):
        assert abs(an - fd_f) <= 1e-6 * abs(fd_f)
    for an in (an_g, an_gm):
        assert abs(an - fd_g) <= 1e-6 * abs(fd_g)


def test_fullsize_d2h_o64_v384():
    from aces2 import synth
    from aces2.ccsd import SymCCSD
    o, v, seed = 64, 384, 20260
    print('case full size D2h o=%d v=%d' % (o, v), flush=True)
    rng = np.random.default_rng(seed + 5)
    occ = synth.irrep_split(o, 8, 2, rng)
    vir = synth.irrep_split(v, 8, 16, rng)
    eps, B, irr = synth.sym_draw(occ, vir, seed)
    blocks = synth.blocks_device(eps, B, o)
    cc = SymCCSD(occ, vir, blocks)
    del blocks
    _free()
    h64 = cc.run(maxcyc=60, conv=9, order=6)
    hmx = cc.run(maxcyc=60, conv=7, order=6, prec='bf16x3', polish_conv=9, polish='refine')
    de = abs(hmx['final_energy'] - h64['final_energy'])
    print('D2h fp64 E %.12f (%d cycles); mixed E %.12f (%d cycles); |dE| %.2e' %
          (h64['final_energy'], h64['cycles'], hmx['final_energy'], hmx['cycles'], de), flush=True)
    assert de <= 1e-7
    t1, t2 = cc.get_amplitudes()
    i1, i2 = cc.increments('fp64')
    t1h, t2h = t1.cpu().numpy(), t2.cpu().numpy()
    i1h, i2h = i1.cpu().numpy(), i2.cpu().numpy()
    del t1, t2, i1, i2
    lh = cc.solve_lambda(maxcyc=60, conv=7, order=6)
    D = cc.density()
    asym = float((D - D.T).abs().max())
    print('D2h Lambda %d cycles, |D - D^T| %.1e' % (lh['cycles'], asym), flush=True)
    assert lh['converged'] and asym < 1e-12
    cc.close()
    del D
    _free()
    m1, m2 = synth.sym_masks(irr, o)
    assert np.abs(t2h[~m2]).max() == 0.0 and np.abs(i2h[~m2]).max() == 0.0
    S = slices.DFSystem(B, np.diag(eps), o)
    del B
    _check_slices(S, t1h, t2h, i1h, i2h, [(0, 0), (o - 1, v - 1), (o // 2, v // 2)], 1e-11, 'fp64 D2h o=64 v=384')


def test_fullsize_intprc_lists_o48_v320():
    """The GPU list producer at the o=48, v=320 size (n_aux = 1472): list 233
    (42 GB) and the other integral lists against (pq|rs) from the factors on
    the host for sampled columns, the denominators, and the time it takes."""
    import time
    from aces2 import lists, synth
    o, v, seed = 48, 320, 20260
    print('case full size intprc o=%d v=%d' % (o, v), flush=True)
    eps, B = synth.draw(o, v, seed, sigma=synth.scaled_sigma(o, v))
    lists.init()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lists.intprc_df(B, eps, [o], [v])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print('intprc lists 13/16/10/30/23/233 + denominators: %.2f s' % dt, flush=True)
    Bvv = B[:, o:, o:]
    Bov = B[:, :o, o:]
    nvp = v * (v + 1) // 2
    assert lists.shape(1, 233) == (nvp, v * v)
    pa, pb = synth.pair_le(v)
    for (c, d) in [(0, 0), (v - 1, v - 1), (v // 3, 7), (5, v - 2)]:
        col = lists.getlst(c + v * d + 1, 1, 1, 233).ravel()
        M = Bvv[:, :, c].T @ Bvv[:, :, d]            # (ac|bd) as M[a][b]
        ref = M[pa, pb]
        assert np.abs(col - ref).max() < 1e-12 * np.abs(ref).max(), (c, d)
    # list 16 <ab|ij> = (ai|bj): column (i, j) at i + o j
    for (i, j) in [(0, 0), (o - 1, 3), (o // 2, o - 1)]:
        col = lists.getlst(i + o * j + 1, 1, 1, 16).ravel().reshape(v, v)    # [b][a]
        ref = (Bov[:, i, :].T @ Bov[:, j, :]).T                             # [b][a] = (ai|bj)
        assert np.abs(col - ref).max() < 1e-12 * np.abs(ref).max(), (i, j)
    # list 30 <ab|ci> = (ib|ac): column (c, i) at c + v i
    for (c, i) in [(0, 0), (v - 1, o - 1), (v // 2, 5)]:
        col = lists.getlst(c + v * i + 1, 1, 1, 30).ravel().reshape(v, v)    # [b][a]
        ref = Bov[:, i, :].T @ Bvv[:, :, c]                               # [b][a] = (ib|ac)
        assert np.abs(col - ref).max() < 1e-12 * np.abs(ref).max(), (c, i)
    d1 = lists.getlst(1, 1, 9, 48).ravel()
    assert np.abs(d1 - 1.0 / (eps[:o][:, None] - eps[o:][None, :]).ravel()).max() < 1e-15 * np.abs(d1).max()
    lists.fin()
    _free()
