"""(T) of CCSD(T) on the GPU (aces_ccsd_triples) against the CPU oracle."""
import numpy as np
import pytest

from oracle import rccsd, triples

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu


def _system(o, v, seed):
    from aces2 import synth
    eps, B = synth.draw(o, v, seed, sigma=synth.scaled_sigma(o, v))
    blk, eri = synth.blocks_numpy(eps, B, o)
    return blk, eri


# v = 10, 19, 30, 45: ragged 8x8x8 energy tiles; o = 1: the single triple (0, 0, 0)
@pytest.mark.parametrize('o,v,seed', [(1, 6, 1), (3, 10, 2), (5, 19, 3), (6, 30, 4), (4, 45, 5)])
def test_triples_match_oracle(o, v, seed):
    from aces2.ccsd import CCSD
    print('case triples', o, v, seed, flush=True)
    blk, eri = _system(o, v, seed)
    rng = np.random.default_rng(seed + 50)
    t1 = rng.normal(0, 0.05, (o, v))
    t2 = rng.normal(0, 0.05, (o, o, v, v))
    t2 = 0.5 * (t2 + t2.transpose(1, 0, 3, 2))
    cc = CCSD(o, v, blk)
    cc.set_amplitudes(t1, t2)
    r = cc.triples(per_triple=True)
    per = r['per_triple'].cpu().numpy()
    cc.close()
    ref = triples.rccsd_triples(t1, t2, blk['fock'], rccsd.blocks(eri, o), o)
    if o == 1:
        # two electrons have no triples: the sums cancel to rounding
        assert abs(r['e4t']) < 1e-14 and abs(r['e5st']) < 1e-14 and abs(r['e4t_aaa']) < 1e-14
        return
    for key in ('e4t', 'e5st', 'e4t_aaa', 'e4t_baa'):
        assert abs(r[key] - ref[key]) <= 1e-12 * abs(ref['e4t']), (key, r[key], ref[key])
    t = 0
    for i in range(o):
        for j in range(i + 1):
            for k in range(j + 1):
                np.testing.assert_allclose(per[t], ref['per_triple'][(i, j, k)], rtol=1e-11,
                                           atol=1e-13 * abs(ref['e4t']))
                t += 1
    assert r['ntriples'] == t


# bf16x3 W products (fp32 accumulation): the triples energy to ~1e-6 relative
@pytest.mark.parametrize('o,v,seed', [(3, 10, 2), (5, 19, 3), (4, 45, 5), (6, 64, 6)])
def test_triples_bf16x3_against_oracle(o, v, seed):
    from aces2.ccsd import CCSD
    print('case triples bf16x3', o, v, seed, flush=True)
    blk, eri = _system(o, v, seed)
    rng = np.random.default_rng(seed + 50)
    t1 = rng.normal(0, 0.05, (o, v))
    t2 = rng.normal(0, 0.05, (o, o, v, v))
    t2 = 0.5 * (t2 + t2.transpose(1, 0, 3, 2))
    cc = CCSD(o, v, blk)
    cc.set_amplitudes(t1, t2)
    r = cc.triples(prec='bf16x3')
    r64 = cc.triples(prec='fp64')
    cc.close()
    ref = triples.rccsd_triples(t1, t2, blk['fock'], rccsd.blocks(eri, o), o)
    for key in ('e4t', 'e5st', 'e4t_aaa', 'e4t_baa'):
        assert abs(r64[key] - ref[key]) <= 1e-12 * abs(ref['e4t']), (key, r64[key], ref[key])
    err = max(abs(r[key] - ref[key]) for key in ('e4t', 'e5st', 'e4t_aaa', 'e4t_baa')) / abs(ref['e4t'])
    print('bf16x3 (T) rel err %.2e' % err, flush=True)
    assert 0 < err < 2e-5
