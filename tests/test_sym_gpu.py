"""Irrep-blocked engine (grouped MFMA launches over irrep blocks) against the C1
engine and the oracle on symmetric systems: increments, CC iterations,
mixed precision, reverse pass, Lambda and response density."""
import numpy as np
import pytest

from oracle import lambda_, rccsd, vcc_loop

pytestmark = pytest.mark.gpu


def _system(occ, vir, seed, nonhf=False):
    from aces2 import synth
    eps, B, irr = synth.sym_draw(occ, vir, seed)
    o = sum(occ)
    eri = synth.full_eri(B)
    f = np.diag(eps)
    if nonhf:
        rng = np.random.default_rng(seed + 7)
        P = rng.normal(0, 0.01, f.shape)
        P = (P + P.T) * ((irr[:, None] ^ irr[None, :]) == 0)
        f = f + P
    blk = dict(rccsd.blocks(eri, o))
    blk['fock'] = f
    blk['list233'] = synth.list233_from_eri(eri, o)
    return f, eri, blk, irr


def _rand_sym(irr, o, v, seed, s=0.05):
    from aces2 import synth
    m1, m2 = synth.sym_masks(irr, o)
    rng = np.random.default_rng(seed)
    t1 = rng.normal(0, s, (o, v)) * m1
    t2 = rng.normal(0, s, (o, o, v, v))
    t2 = 0.5 * (t2 + t2.transpose(1, 0, 3, 2)) * m2
    return t1, t2


CASES = [([3], [9], 1, True), ([2, 1], [4, 5], 2, True), ([1, 2, 1, 1], [3, 2, 4, 3], 3, True),
         ([1, 1, 0, 1, 1, 1, 0, 1], [2, 3, 2, 2, 3, 2, 2, 3], 4, True), ([2, 2, 1, 1], [5, 4, 3, 4], 5, False)]


@pytest.mark.parametrize('occ,vir,seed,nonhf', CASES)
def test_increments_match_c1_and_oracle(occ, vir, seed, nonhf):
    from aces2.ccsd import CCSD, SymCCSD
    print('case sym increments', occ, vir, seed, flush=True)
    o, v = sum(occ), sum(vir)
    f, eri, blk, irr = _system(occ, vir, seed, nonhf)
    t1, t2 = _rand_sym(irr, o, v, seed + 1)
    cs = SymCCSD(occ, vir, blk)
    cs.set_amplitudes(t1, t2)
    i1, i2 = cs.increments('fp64')
    e_sym = cs.energy()
    cs.close()
    r1, r2 = rccsd.increments(t1, t2, f, rccsd.blocks(eri, o), o)
    sc = max(np.abs(r1).max(), np.abs(r2).max())
    np.testing.assert_allclose(i1.cpu().numpy(), r1, rtol=0, atol=1e-12 * sc)
    np.testing.assert_allclose(i2.cpu().numpy(), r2, rtol=0, atol=1e-12 * sc)
    assert abs(e_sym - rccsd.energy(t1, t2, f, rccsd.blocks(eri, o), o)) < 1e-12
    cc = CCSD(o, v, blk)
    cc.set_amplitudes(t1, t2)
    c1, c2 = cc.increments('fp64')
    cc.close()
    np.testing.assert_allclose(i2.cpu().numpy(), c2.cpu().numpy(), rtol=0, atol=1e-12 * sc)


@pytest.mark.parametrize('occ,vir,seed,nonhf', CASES[1:3])
def test_cc_loop_matches_c1(occ, vir, seed, nonhf):
    from aces2.ccsd import CCSD, SymCCSD
    print('case sym loop', occ, vir, flush=True)
    o, v = sum(occ), sum(vir)
    f, eri, blk, irr = _system(occ, vir, seed, nonhf)
    cs = SymCCSD(occ, vir, blk)
    hs = cs.run(maxcyc=60, conv=10, order=5)
    cs.close()
    cc = CCSD(o, v, blk)
    hc = cc.run(maxcyc=60, conv=10, order=5)
    cc.close()
    assert hs['cycles'] == hc['cycles']
    np.testing.assert_allclose(hs['energy'], hc['energy'], rtol=0, atol=1e-11)
    np.testing.assert_allclose(hs['resmiter'], hc['resmiter'], rtol=1e-6, atol=1e-13)
    assert abs(hs['final_energy'] - hc['final_energy']) < 1e-11


@pytest.mark.parametrize('prec,tol', [('fp32', 2e-6), ('bf16x3', 5e-5)])
def test_lowp_increments(prec, tol):
    from aces2.ccsd import SymCCSD
    occ, vir = [2, 2, 1, 1], [5, 4, 3, 4]
    print('case sym lowp', prec, flush=True)
    o, v = sum(occ), sum(vir)
    f, eri, blk, irr = _system(occ, vir, 9)
    t1, t2 = _rand_sym(irr, o, v, 10)
    cs = SymCCSD(occ, vir, blk)
    cs.set_amplitudes(t1, t2)
    a1, a2 = cs.increments('fp64')
    b1, b2 = cs.increments(prec)
    cs.close()
    sc = float(a2.abs().max())
    assert float((a2 - b2).abs().max()) < tol * sc
    assert float((a1 - b1).abs().max()) < tol * sc


def test_lowp_increments_tiled():
    """bf16x3 increments at a size where the permutes run tiled and write the
    bf16x3 operand splits of the ring products as they go: against fp64, and
    the same twice in a row."""
    from aces2.ccsd import SymCCSD
    occ, vir = [4, 3, 4, 5], [20, 18, 22, 19]
    print('case sym lowp tiled', flush=True)
    o, v = sum(occ), sum(vir)
    f, eri, blk, irr = _system(occ, vir, 31)
    t1, t2 = _rand_sym(irr, o, v, 32)
    cs = SymCCSD(occ, vir, blk)
    cs.set_amplitudes(t1, t2)
    a1, a2 = cs.increments('fp64')
    b1, b2 = cs.increments('bf16x3')
    c1, c2 = cs.increments('bf16x3')
    cs.close()
    sc = float(a2.abs().max())
    assert float((a2 - b2).abs().max()) < 5e-5 * sc
    assert float((a1 - b1).abs().max()) < 5e-5 * sc
    assert float((b2 - c2).abs().max()) == 0.0 and float((b1 - c1).abs().max()) == 0.0


@pytest.mark.parametrize('occ,vir,seed,nonhf', [CASES[1], CASES[3]])
def test_adjoint_matches_oracle(occ, vir, seed, nonhf):
    from aces2.ccsd import SymCCSD
    print('case sym adjoint', occ, vir, flush=True)
    o, v = sum(occ), sum(vir)
    f, eri, blk, irr = _system(occ, vir, seed, nonhf)
    t1, t2 = _rand_sym(irr, o, v, seed + 1)
    l1, l2 = _rand_sym(irr, o, v, seed + 2, s=0.3)
    cs = SymCCSD(occ, vir, blk)
    cs.set_amplitudes(t1, t2)
    r1, r2, D = cs.adjoint(l1, l2, density=True)
    cs.close()
    g = {k: blk[k] for k in ('ovov', 'oovv', 'ovvo', 'ovoo', 'ovvv', 'oooo', 'vvvv')}
    d1, d2, df = lambda_.lagrangian_grads(t1, t2, l1, l2, f, g, o)
    d2 = 0.5 * (d2 + d2.transpose(1, 0, 3, 2))
    df = 0.5 * (df + df.T)
    from aces2 import synth
    m1, m2 = synth.sym_masks(irr, o)
    sc = max(np.abs(d1).max(), np.abs(d2).max())
    np.testing.assert_allclose(r1.cpu().numpy(), d1 * m1, rtol=0, atol=1e-12 * sc)
    np.testing.assert_allclose(r2.cpu().numpy(), d2 * m2, rtol=0, atol=1e-12 * sc)
    np.testing.assert_allclose(D.cpu().numpy(), df, rtol=0, atol=1e-12 * max(1.0, np.abs(df).max()))


def test_lambda_and_density_match_c1():
    from aces2.ccsd import CCSD, SymCCSD
    occ, vir = [2, 1, 1, 2], [4, 5, 3, 4]
    print('case sym lambda', flush=True)
    o, v = sum(occ), sum(vir)
    f, eri, blk, irr = _system(occ, vir, 12, nonhf=True)
    cs = SymCCSD(occ, vir, blk)
    cs.run(maxcyc=100, conv=11, order=6)
    ls = cs.solve_lambda(maxcyc=100, conv=11, order=6)
    s1, s2 = cs.get_lambda()
    Ds = cs.density().cpu().numpy()
    cs.close()
    cc = CCSD(o, v, blk)
    cc.run(maxcyc=100, conv=11, order=6)
    lc = cc.solve_lambda(maxcyc=100, conv=11, order=6)
    c1, c2 = cc.get_lambda()
    Dc = cc.density().cpu().numpy()
    cc.close()
    assert ls['cycles'] == lc['cycles']
    np.testing.assert_allclose(ls['resmiter'], lc['resmiter'], rtol=1e-6, atol=1e-13)
    np.testing.assert_allclose(s1.cpu().numpy(), c1.cpu().numpy(), atol=1e-10)
    np.testing.assert_allclose(s2.cpu().numpy(), c2.cpu().numpy(), atol=1e-10)
    np.testing.assert_allclose(Ds, Dc, atol=1e-10)


def test_refine_polish_sym():
    """Irrep-blocked engine: bf16x3 cycles to 1e-6, then refinement cycles (the
    large products as bf16x3 products of the differences to the fp64 anchor of
    the switch) to 1e-10: the fp64 loop's energy, and refined increments far
    inside the plain bf16x3 error at the final amplitudes."""
    from aces2.ccsd import SymCCSD
    occ, vir = [2, 2, 1, 1], [5, 4, 3, 4]
    print('case sym refine', flush=True)
    f, eri, blk, irr = _system(occ, vir, 9)
    cs = SymCCSD(occ, vir, blk)
    h64 = cs.run(maxcyc=80, conv=10, order=6, start='mp1')
    hr = cs.run(maxcyc=80, conv=6, order=6, start='mp1', prec='bf16x3', polish_conv=10, polish='refine')
    assert h64['converged'] and hr['converged']
    assert hr['precision'][0] == 'bf16x3' and hr['precision'][-1] == 'refine'
    assert abs(hr['final_energy'] - h64['final_energy']) < 1e-9
    i1r, i2r = cs.increments('refine')
    i1d, i2d = cs.increments('fp64')
    i1b, i2b = cs.increments('bf16x3')
    cs.close()
    err_r = max(float((i1r - i1d).abs().max()), float((i2r - i2d).abs().max()))
    err_b = max(float((i1b - i1d).abs().max()), float((i2b - i2d).abs().max()))
    print('sym refine err %.3e bf16x3 err %.3e' % (err_r, err_b), flush=True)
    assert err_b > 0 and err_r < 1e-3 * err_b


@pytest.mark.parametrize('occ,vir,seed', [([2, 1, 1, 2], [9, 8, 7, 8], 21), ([2, 1, 1, 2, 1, 2, 1, 1], [6, 5, 7, 4, 5, 6, 4, 5], 22)])
def test_lambda_refine_sym(occ, vir, seed):
    """Mixed Lambda cycles on the irrep-blocked engine (fp64 residual at an
    anchor plus the linear reverse pass on the difference, its ladder, W-build
    and ring adjoint products in bf16x3 over the irrep blocks) converge to the
    fp64 multipliers and density."""
    from aces2.ccsd import SymCCSD
    print('case sym lambda refine', occ, vir, flush=True)
    f, eri, blk, irr = _system(occ, vir, seed, nonhf=True)
    cs = SymCCSD(occ, vir, blk)
    cs.run(maxcyc=200, conv=11, order=6)
    h64 = cs.solve_lambda(maxcyc=200, conv=10, order=6)
    a1, a2 = cs.get_lambda()
    Da = cs.density()
    hmx = cs.solve_lambda(maxcyc=200, conv=10, order=6, prec='refine')
    b1, b2 = cs.get_lambda()
    Db = cs.density()
    cs.close()
    print('fp64 %d cycles, mixed %d cycles' % (h64['cycles'], hmx['cycles']), flush=True)
    assert h64['converged'] and hmx['converged']
    assert hmx['cycles'] <= h64['cycles'] + 4
    sc = max(float(a2.abs().max()), 1.0)
    assert float((a1 - b1).abs().max()) < 1e-8 * sc
    assert float((a2 - b2).abs().max()) < 1e-8 * sc
    assert float((Da - Db).abs().max()) < 1e-8
