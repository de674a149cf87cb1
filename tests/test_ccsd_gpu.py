"""GPU CCSD engine (through the C ABI) against the CPU oracle."""
import numpy as np
import pytest

from oracle import rccsd, vcc_loop

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu


def _system(o, v, seed, nonhf=False):
    from aces2 import synth
    eps, B = synth.draw(o, v, seed, sigma=synth.scaled_sigma(o, v))
    pert = None
    if nonhf:
        rng = np.random.default_rng(seed + 100)
        P = rng.normal(0, 0.01, (o + v, o + v))
        pert = P + P.T
    blk, eri = synth.blocks_numpy(eps, B, o, perturb=pert)
    return blk, eri


def _random_t(o, v, seed, scale=0.05):
    rng = np.random.default_rng(seed)
    t1 = rng.normal(0, scale, (o, v))
    t2 = rng.normal(0, scale, (o, o, v, v))
    t2 = 0.5 * (t2 + t2.transpose(1, 0, 3, 2))
    return t1, t2


# (9, 80): ragged LDS tiles of the permute, ladder-unpack and symmetrize kernels
@pytest.mark.parametrize('o,v,seed,nonhf', [(1, 3, 1, False), (2, 3, 2, True), (5, 19, 3, False),
                                            (5, 19, 4, True), (7, 40, 5, True), (12, 64, 6, False),
                                            (9, 80, 7, True)])
def test_increments_fp64(o, v, seed, nonhf):
    from aces2.ccsd import CCSD
    print('case increments', o, v, seed, nonhf, flush=True)
    blk, eri = _system(o, v, seed, nonhf)
    cc = CCSD(o, v, blk)
    t1, t2 = _random_t(o, v, seed + 7)
    cc.set_amplitudes(t1, t2)
    i1, i2 = cc.increments('fp64')
    r1, r2 = rccsd.increments(t1, t2, blk['fock'], rccsd.blocks(eri, o), o)
    s1 = max(np.abs(r1).max(), 1e-30)
    s2 = np.abs(r2).max()
    assert np.abs(i1.cpu().numpy() - r1).max() / s1 < 1e-11
    assert np.abs(i2.cpu().numpy() - r2).max() / s2 < 1e-11
    e_gpu = cc.energy(t1, t2)
    e_ref = rccsd.energy(t1, t2, blk['fock'], rccsd.blocks(eri, o), o)
    assert abs(e_gpu - e_ref) < 1e-12 * max(1.0, abs(e_ref))
    cc.close()


# (7, 45): no operand length is a multiple of 32 (v, v^2, ov, ov^2, o^2 v), so
# every split written straight into an operand slot has padding columns, and
# odd o gives the antisymmetric ladder pairs a ragged row count
@pytest.mark.parametrize('prec,tol,o,v', [('fp32', 2e-6, 8, 48), ('bf16x3', 5e-5, 8, 48), ('bf16x3', 5e-5, 7, 45)])
def test_increments_lowp(prec, tol, o, v):
    from aces2.ccsd import CCSD
    print('case lowp', prec, o, v, flush=True)
    blk, eri = _system(o, v, 11, True)
    cc = CCSD(o, v, blk)
    t1, t2 = _random_t(o, v, 12)
    cc.set_amplitudes(t1, t2)
    i1, i2 = cc.increments(prec)
    r1, r2 = rccsd.increments(t1, t2, blk['fock'], rccsd.blocks(eri, o), o)
    assert np.abs(i2.cpu().numpy() - r2).max() / np.abs(r2).max() < tol
    assert np.abs(i1.cpu().numpy() - r1).max() / np.abs(r1).max() < tol
    cc.close()


@pytest.mark.parametrize('o,v,seed,nonhf', [(5, 19, 21, False), (6, 30, 22, True), (4, 70, 23, False)])
def test_cc_loop_matches_oracle(o, v, seed, nonhf):
    """Per-cycle energies, RESMITER and RESMDIIS follow the oracle's reference loop."""
    from aces2.ccsd import CCSD
    print('case loop', o, v, seed, nonhf, flush=True)
    blk, eri = _system(o, v, seed, nonhf)
    g = rccsd.blocks(eri, o)
    ref = vcc_loop.run(blk['fock'], g, o, maxcyc=60, conv=10, order=6, start='mp1')
    cc = CCSD(o, v, blk)
    hist = cc.run(maxcyc=60, conv=10, order=6, start='mp1')
    assert hist['converged'] and ref['converged']
    assert hist['cycles'] == ref['cycles']
    np.testing.assert_allclose(hist['energy'], ref['energy'], rtol=0, atol=1e-10)
    np.testing.assert_allclose(hist['resmiter'], ref['resmiter'], rtol=1e-6, atol=1e-13)
    np.testing.assert_allclose(hist['resmdiis'], ref['resmdiis'], rtol=1e-5, atol=1e-12)
    assert abs(hist['final_energy'] - ref['final_energy']) < 1e-10
    cc.close()


@pytest.mark.parametrize('o,v,seed', [(6, 30, 24), (7, 45, 25)])
def test_mixed_then_fp64_polish(o, v, seed):
    """bf16x3 cycles to 1e-7, then fp64 cycles (DIIS restarted) to 1e-10: the
    same energy as the fp64 loop to 1e-10, and the polish takes fewer cycles
    than the fp64 loop needs from the start."""
    from aces2.ccsd import CCSD
    print('case polish', o, v, seed, flush=True)
    blk, eri = _system(o, v, seed, True)
    cc = CCSD(o, v, blk)
    h64 = cc.run(maxcyc=80, conv=10, order=6, start='mp1')
    hmx = cc.run(maxcyc=80, conv=7, order=6, start='mp1', prec='bf16x3', polish_conv=10, polish='fp64')
    assert h64['converged'] and hmx['converged']
    assert hmx['precision'][-1] == 'fp64' and hmx['precision'][0] == 'bf16x3'
    assert abs(hmx['final_energy'] - h64['final_energy']) < 1e-9
    assert hmx['precision'].count('fp64') < h64['cycles']
    cc.close()


@pytest.mark.parametrize('o,v,seed', [(6, 30, 24), (7, 45, 25)])
def test_refine_polish(o, v, seed):
    """bf16x3 cycles to 1e-6, then refinement cycles (bf16x3 products of the
    operands' differences to the fp64 anchor taken at the switch) to 1e-10: the
    fp64 loop's energy, and at the final amplitudes the refined increments
    agree with the fp64 ones far inside the plain bf16x3 error."""
    from aces2.ccsd import CCSD
    print('case refine', o, v, seed, flush=True)
    blk, eri = _system(o, v, seed, True)
    cc = CCSD(o, v, blk)
    h64 = cc.run(maxcyc=80, conv=10, order=6, start='mp1')
    hr = cc.run(maxcyc=80, conv=6, order=6, start='mp1', prec='bf16x3', polish_conv=10, polish='refine')
    assert h64['converged'] and hr['converged']
    assert hr['precision'][0] == 'bf16x3' and hr['precision'][-1] == 'refine'
    # both loops stop at |RESMITER| < 1e-10, which leaves the energy within ~1e-10 of its limit
    assert abs(hr['final_energy'] - h64['final_energy']) < 1e-9
    assert hr['precision'].count('refine') < h64['cycles']
    # increments at the converged amplitudes (the anchor is the switch point)
    i1r, i2r = cc.increments('refine')
    i1d, i2d = cc.increments('fp64')
    i1b, i2b = cc.increments('bf16x3')
    err_r = max(float((i1r - i1d).abs().max()), float((i2r - i2d).abs().max()))
    err_b = max(float((i1b - i1d).abs().max()), float((i2b - i2d).abs().max()))
    print('refine err %.3e bf16x3 err %.3e' % (err_r, err_b), flush=True)
    # the refined error scales with the distance to the anchor (switch at 1e-6):
    # bf16x3's relative error times that distance, far below the 1e-10 threshold
    assert err_b > 0 and err_r < 1e-3 * err_b and err_r < 1e-11
    cc.close()


def test_refine_anchor_cycle_is_fp64():
    """The first refinement evaluation after set-up (or reset_anchor) records
    the anchor and equals the fp64 evaluation; near convergence, a later one at
    other amplitudes (reached by refinement cycles) has a small fraction of
    plain bf16x3's error (the difference operands are small)."""
    from aces2.ccsd import CCSD
    o, v, seed = 5, 19, 26
    print('case refine anchor', o, v, seed, flush=True)
    blk, eri = _system(o, v, seed, True)
    cc = CCSD(o, v, blk)
    t1, t2 = _random_t(o, v, seed + 7, scale=0.02)
    cc.set_amplitudes(torch.tensor(t1), torch.tensor(t2))
    i1a, i2a = cc.increments('refine')
    i1d, i2d = cc.increments('fp64')
    assert float((i2a - i2d).abs().max()) < 1e-13 and float((i1a - i1d).abs().max()) < 1e-13
    # a new anchor near the solution, then three refinement cycles
    cc.run(maxcyc=60, conv=4, order=6, start='current')
    cc.reset_anchor()
    cc.reset_diis()
    for ic in range(1, 4):
        cc.cycle(ic, 'refine', 1e-30)
    i1r, i2r = cc.increments('refine')
    i1d, i2d = cc.increments('fp64')
    i1b, i2b = cc.increments('bf16x3')
    err_r = float((i2r - i2d).abs().max())
    err_b = float((i2b - i2d).abs().max())
    print('refine err %.3e bf16x3 err %.3e' % (err_r, err_b), flush=True)
    assert err_r < 0.05 * err_b
    cc.close()


def test_run_raises_not_converged():
    """Too few cycles: the host loop raises (the reference exits with status 1,
    vcc/vcc.F:1439-1441) after running exactly maxcyc cycles."""
    from aces2.ccsd import CCSD, NotConverged
    print('case not converged', flush=True)
    blk, eri = _system(4, 13, 27)
    cc = CCSD(4, 13, blk)
    with pytest.raises(NotConverged):
        cc.run(maxcyc=2, conv=10)
    h = cc.run(maxcyc=2, conv=10, raise_on_fail=False)
    assert not h['converged'] and len(h['resmiter']) == 2
    cc.close()
