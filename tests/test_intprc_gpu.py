"""The integral and denominator lists built on the GPU from factorised
integrals (aces_intprc_df, the role of xintprc) against the host writer's
lists of the same integrals, and the CC loop run from them against the oracle."""
import ctypes

import numpy as np
import pytest

from oracle import rccsd, vcc_loop

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu


def _compare_lists(blk, occ, vir, rtol=1e-12):
    from aces2 import lists
    o = sum(occ)
    if len(occ) == 1:
        want = {lst: (tl, tr, [data.reshape(-1, lists.c1_pairs(tl, o, sum(vir)))])
                for lst, (tl, tr, data) in lists.c1_integral_lists(blk, o).items()}
    else:
        want = lists.sym_integral_lists(blk, occ, vir)
    for lst, (tl, tr, groups) in want.items():
        assert lists.isytyp(lst) == (tl, tr), lst
        scale = max(np.abs(g).max() if g.size else 0.0 for g in groups)
        for g, a in enumerate(groups, start=1):
            nr, nc = lists.shape(g, lst)
            assert (nc, nr) == a.shape, (lst, g)
            if a.size:
                got = lists.getlst(1, nc, g, lst)
                err = np.abs(got - a).max() / scale
                assert err < rtol, (lst, g, err)


def _check_denominators(eps, occ, vir):
    from aces2 import lists
    o, v = sum(occ), sum(vir)
    eo, ev = eps[:o], eps[o:]
    D = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]
    blk = {'d50': (1.0 / D).transpose(1, 0, 3, 2)}   # [j][i][b][a] = C1 order a + v b + v^2 (i + o j)
    c1_50 = np.ascontiguousarray(blk['d50']).ravel()
    want50 = lists.dpd_groups_from_c1(c1_50, occ, vir, 13, 14)
    iu_o = [(i, j) for j in range(o) for i in range(j)]
    iu_v = [(a, b) for b in range(v) for a in range(b)]
    c1_48 = np.array([1.0 / (eo[i] + eo[j] - ev[a] - ev[b]) for (i, j) in iu_o for (a, b) in iu_v])
    want48 = lists.dpd_groups_from_c1(c1_48, occ, vir, 1, 3)
    for lst, want in ((50, want50), (48, want48)):
        for g, a in enumerate(want, start=1):
            nr, nc = lists.shape(g, lst)
            if a.size:
                np.testing.assert_allclose(lists.getlst(1, nc, g, lst), a, rtol=1e-14, atol=0)
    # T1 denominators: group 9 of list 48, the irrep-1 (a,i) pairs in DPD order
    pa, pi = lists.dpd_pairs(occ, vir, 9, 0)
    nr, nc = lists.shape(9, 48)
    got = lists.getlst(1, nc, 9, 48).ravel()
    np.testing.assert_allclose(got, 1.0 / (eo[pi] - ev[pa]), rtol=1e-14, atol=0)
    np.testing.assert_array_equal(lists.getrec('SCFEVALA', o + v), eps)


def _loop():
    from aces2 import _lib
    _lib.call('aces_vcc_mp1')
    _lib.call('aces_vcc_set_diis_order', 6)
    e0 = ctypes.c_double()
    _lib.call('aces_cmpeng', 43, 0, None, ctypes.byref(e0), None)
    energies, resm = [e0.value], []
    conv = ctypes.c_int(1)
    st = (ctypes.c_double * 8)()
    for icycle in range(1, 61):
        for name in ('aces_genint', 'aces_dre3en', 'aces_e4s', 'aces_newt2'):
            _lib.call(name)
        e = ctypes.c_double()
        _lib.call('aces_cmpeng', 60, 2, None, ctypes.byref(e), None)
        energies.append(e.value)
        _lib.call('aces_dodiis0', icycle, 1e-9, ctypes.byref(conv), st)
        resm.append(st[1])
        if conv.value == 0:
            break
        _lib.call('aces_drmove')
    return energies, resm, conv.value == 0


@pytest.mark.parametrize('occ,vir,seed', [([5], [19], 61), ([3], [10], 62),
                                          ([1, 1, 0, 1, 1, 1, 0, 1], [2, 3, 2, 2, 3, 2, 2, 3], 63),
                                          ([2, 1, 1, 1], [4, 3, 5, 3], 64)])
def test_intprc_lists_and_cc_loop(occ, vir, seed):
    from aces2 import lists, synth
    print('case intprc', occ, vir, seed, flush=True)
    o = sum(occ)
    if len(occ) == 1:
        eps, B = synth.draw(o, vir[0], seed, sigma=synth.scaled_sigma(o, vir[0]))
        blk, eri = synth.blocks_numpy(eps, B, o)
    else:
        eps, B, irr = synth.sym_draw(occ, vir, seed)
        eri = synth.full_eri(B)
        blk = dict(rccsd.blocks(eri, o))
        blk['fock'] = np.diag(eps)
        blk['list233'] = synth.list233_from_eri(eri, o)
    ref = vcc_loop.run(np.diag(eps), rccsd.blocks(eri, o), o, maxcyc=60, conv=9, order=6)
    lists.init()
    lists.intprc_df(B, eps, occ, vir)
    _compare_lists(blk, occ, vir)
    _check_denominators(eps, occ, vir)
    energies, resm, converged = _loop()
    lists.fin()
    assert converged and len(resm) == ref['cycles']
    np.testing.assert_allclose(energies, ref['energy'], rtol=0, atol=1e-10)
    np.testing.assert_allclose(resm, ref['resmiter'], rtol=1e-6, atol=1e-13)
