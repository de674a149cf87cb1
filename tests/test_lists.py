"""List/record store: DPD addressing (CPU) and device round trips + the xvcc driver (GPU)."""
import itertools

import numpy as np
import pytest

from oracle import rccsd, vcc_loop

# Pair types 1..25 as (left space, right space, left spin, right spin, packing) -- the
# convention of acescore/aces_com_symloc.F:11-34, enumerated independently below.
TYPES = {1: ('V', 'V', 0, 0, '<'), 2: ('V', 'V', 1, 1, '<'), 3: ('O', 'O', 0, 0, '<'), 4: ('O', 'O', 1, 1, '<'),
         5: ('V', 'V', 0, 0, '<='), 6: ('V', 'V', 1, 1, '<='), 7: ('O', 'O', 0, 0, '<='), 8: ('O', 'O', 1, 1, '<='),
         9: ('V', 'O', 0, 0, ''), 10: ('V', 'O', 1, 1, ''), 11: ('V', 'O', 0, 1, ''), 12: ('V', 'O', 1, 0, ''),
         13: ('V', 'V', 0, 1, ''), 14: ('O', 'O', 0, 1, ''), 15: ('V', 'V', 0, 1, ''), 16: ('O', 'V', 0, 0, ''),
         17: ('O', 'V', 1, 1, ''), 18: ('O', 'V', 0, 1, ''), 19: ('V', 'V', 0, 0, ''), 20: ('V', 'V', 1, 1, ''),
         21: ('O', 'O', 0, 0, ''), 22: ('O', 'O', 1, 1, ''), 23: ('V', 'V', 1, 0, ''), 24: ('O', 'O', 1, 0, ''),
         25: ('O', 'V', 1, 0, '')}


def _brute_irpdpd(pop, vrt, irrep, typ):
    """Count orbital pairs of total irrep `irrep` (0-based) by enumeration."""
    l, r, sl, sr, pk = TYPES[typ]
    orbs = lambda sp, spin: [(h, k) for h in range(len(pop[0])) for k in range((pop if sp == 'O' else vrt)[spin][h])]
    n = 0
    P, Q = orbs(l, sl), orbs(r, sr)
    same = (l == r and sl == sr)
    for (hp, kp), (hq, kq) in itertools.product(P, Q):
        if hp ^ hq != irrep:
            continue
        if pk:
            ip, iq = P.index((hp, kp)), Q.index((hq, kq))
            if not same:
                continue
            if (pk == '<' and not ip < iq) or (pk == '<=' and not ip <= iq):
                continue
        n += 1
    return n


def test_irpdpd_matches_enumeration():
    from aces2 import lists
    lists.init()
    popa = [3, 1, 2, 0, 2, 1, 1, 0]
    vrta = [5, 2, 3, 4, 2, 3, 1, 2]
    popb = [2, 1, 2, 1, 2, 1, 0, 0]
    vrtb = [6, 2, 3, 3, 2, 3, 2, 2]
    lists.set_symmetry(popa, vrta, popb, vrtb)
    pop, vrt = [popa, popb], [vrta, vrtb]
    for typ in range(1, 26):
        for irrep in range(8):
            assert lists.irpdpd(irrep + 1, typ) == _brute_irpdpd(pop, vrt, irrep, typ), (typ, irrep)
    lists.fin()


def _pop8(pop):
    return list(pop) + [0] * (8 - len(pop))


@pytest.mark.gpu
def test_list_round_trip_and_records():
    from aces2 import lists
    print('case list round trip', flush=True)
    lists.init()
    lists.set_symmetry([3, 2, 1, 2], [4, 3, 5, 2])
    lists.inipck(1, 13, 14, 46)
    rng = np.random.default_rng(3)
    total = 0
    for g in range(1, 5):
        nr, nc = lists.shape(g, 46)
        assert nr == lists.irpdpd(g, 13) and nc == lists.irpdpd(g, 14)
        x = rng.normal(size=(nc, nr))
        lists.putlst(x, 1, nc, g, 46)
        if nc >= 2:
            y = lists.getlst(2, nc - 1, g, 46)
            np.testing.assert_array_equal(y, x[1:])
        lists.sumsym2(x, g, 46)
        np.testing.assert_allclose(lists.getlst(1, nc, g, 46), 2 * x)
        total += nr * nc
    allv = lists.getall(1, 46, total)
    assert allv.size == total
    with pytest.raises(Exception):
        lists.getlst(1, 10 ** 6, 1, 46)  # GETLST range error (acescore/getlst.F:100-108)
    lists.putrec('SCFEVALA', np.arange(5.0))
    np.testing.assert_array_equal(lists.getrec('SCFEVALA', 5), np.arange(5.0))
    assert lists.reclen('NOSUCHRC') == -1
    np.testing.assert_array_equal(lists.getrec('NOSUCHRC', 3, iflag=-1), np.zeros(3))
    lists.fin()


@pytest.mark.gpu
@pytest.mark.parametrize('nonhf', [False, True])
def test_xvcc_driver_over_lists(nonhf):
    """The vcc.F call sequence over device lists reproduces the oracle loop."""
    import ctypes
    from aces2 import _lib, lists, synth
    print('case xvcc lists nonhf=%s' % nonhf, flush=True)
    o, v = 5, 19
    eps, B = synth.draw(o, v, 31, sigma=synth.scaled_sigma(o, v))
    pert = None
    if nonhf:
        rng = np.random.default_rng(5)
        P = rng.normal(0, 0.01, (o + v, o + v))
        pert = P + P.T
    blk, eri = synth.blocks_numpy(eps, B, o, perturb=pert)
    ref = vcc_loop.run(blk['fock'], rccsd.blocks(eri, o), o, maxcyc=60, conv=9, order=6)
    lists.init()
    lists.load_c1_system(blk, o, v, nonhf=nonhf)
    _lib.call('aces_vcc_mp1')
    _lib.call('aces_vcc_set_diis_order', 6)
    e0 = ctypes.c_double()
    _lib.call('aces_cmpeng', 43, 0, None, ctypes.byref(e0), None)
    energies, resm = [e0.value], []
    conv = ctypes.c_int(1)
    st = (ctypes.c_double * 8)()
    for icycle in range(1, 61):
        _lib.call('aces_genint')
        _lib.call('aces_dre3en')
        _lib.call('aces_e4s')
        _lib.call('aces_newt2')
        e = ctypes.c_double()
        _lib.call('aces_cmpeng', 60, 2, None, ctypes.byref(e), None)
        energies.append(e.value)
        _lib.call('aces_dodiis0', icycle, 1e-9, ctypes.byref(conv), st)
        resm.append(st[1])
        if conv.value == 0:
            break
        _lib.call('aces_drmove')
    ef = ctypes.c_double()
    _lib.call('aces_cmpeng', 43, 0, None, ctypes.byref(ef), None)
    lists.fin()
    assert conv.value == 0 and len(resm) == ref['cycles']
    np.testing.assert_allclose(energies, ref['energy'], rtol=0, atol=1e-10)
    np.testing.assert_allclose(resm, ref['resmiter'], rtol=1e-6, atol=1e-13)
    assert abs(ef.value - ref['final_energy']) < 1e-10


@pytest.mark.gpu
def test_xvcc_driver_mixed_then_refine():
    """The list driver at precision 2 (bf16x3) to 1e-6, then 3 (refinement of
    an fp64 anchor) to 1e-10 with DIIS restarted: the fp64 loop's energy."""
    import ctypes
    from aces2 import _lib, lists, synth
    print('case xvcc lists mixed+refine', flush=True)
    o, v = 6, 30
    eps, B = synth.draw(o, v, 32, sigma=synth.scaled_sigma(o, v))
    blk, eri = synth.blocks_numpy(eps, B, o)
    ref = vcc_loop.run(blk['fock'], rccsd.blocks(eri, o), o, maxcyc=80, conv=10, order=6)
    lists.init()
    lists.load_c1_system(blk, o, v)
    _lib.call('aces_vcc_mp1')
    _lib.call('aces_vcc_set_diis_order', 6)
    _lib.call('aces_vcc_set_precision', 2)
    conv = ctypes.c_int(1)
    st = (ctypes.c_double * 8)()
    tol, ic0, phase, n_refine = 1e-6, 0, 2, 0
    for icycle in range(1, 81):
        _lib.call('aces_genint')
        _lib.call('aces_dre3en')
        _lib.call('aces_e4s')
        _lib.call('aces_newt2')
        e = ctypes.c_double()
        _lib.call('aces_cmpeng', 60, 2, None, ctypes.byref(e), None)
        _lib.call('aces_dodiis0', icycle - ic0, tol, ctypes.byref(conv), st)
        n_refine += phase == 3
        if conv.value == 0:
            if phase == 2:
                phase, tol, ic0 = 3, 1e-10, icycle
                _lib.call('aces_vcc_set_precision', 3)
                _lib.call('aces_vcc_set_diis_order', 6)
                continue
            break
        _lib.call('aces_drmove')
    ef = ctypes.c_double()
    _lib.call('aces_cmpeng', 43, 0, None, ctypes.byref(ef), None)
    _lib.call('aces_vcc_set_precision', 0)
    lists.fin()
    assert conv.value == 0 and phase == 3 and n_refine < ref['cycles']
    assert abs(ef.value - ref['final_energy']) < 1e-9


def _lists_system(o, v, seed, nonhf):
    from aces2 import lists, synth
    eps, B = synth.draw(o, v, seed, sigma=synth.scaled_sigma(o, v))
    pert = None
    if nonhf:
        rng = np.random.default_rng(seed + 1)
        P = rng.normal(0, 0.01, (o + v, o + v))
        pert = P + P.T
    blk, eri = synth.blocks_numpy(eps, B, o, perturb=pert)
    lists.init()
    lists.load_c1_system(blk, o, v, nonhf=nonhf)
    return blk, eri


@pytest.mark.gpu
@pytest.mark.parametrize('nonhf', [False, True])
def test_cmpeng_spin_cases_over_lists(nonhf):
    """CMPENG over lists 46/90 (and 63/90 group 3): ECORR(1) = E(AA), ECORR(3) =
    E(AB), ETOT and ETOTT2 as libr/cmpeng.F returns them; ECORR(2) untouched."""
    import ctypes
    from aces2 import _lib, lists
    print('case cmpeng lists nonhf=%s' % nonhf, flush=True)
    o, v = 5, 19
    blk, eri = _lists_system(o, v, 33, nonhf)
    _lib.call('aces_vcc_mp1')
    rng = np.random.default_rng(9)
    g = rccsd.blocks(eri, o)
    for nlist2, nlist1 in ((43, 0), (60, 2)):
        t1 = rng.normal(0, 0.05, (o, v))
        t2 = rng.normal(0, 0.05, (o, o, v, v))
        t2 = 0.5 * (t2 + t2.transpose(1, 0, 3, 2))
        if nlist2 == 60:
            lists.inipck(1, 13, 14, 63)
        lists.putlst(t2.ravel(), 1, o * o, 1, nlist2 + 3)
        lists.putlst(t1.ravel(), 1, 1, nlist1 + 1, 90)
        ec = (ctypes.c_double * 3)(7.0, 7.0, 7.0)
        et, et2 = ctypes.c_double(), ctypes.c_double()
        _lib.call('aces_cmpeng', nlist2, nlist1, ec, ctypes.byref(et), ctypes.byref(et2))
        e_aa, e_ab, e_lin = rccsd.energy_parts(t1, t2, blk['fock'], g, o)
        assert abs(ec[0] - e_aa) < 1e-13 and abs(ec[2] - e_ab) < 1e-13 and ec[1] == 7.0
        assert abs(et.value - rccsd.energy(t1, t2, blk['fock'], g, o)) < 1e-13
        assert abs(et2.value - e_lin) < 1e-13
    lists.fin()


@pytest.mark.gpu
def test_drmove_fills_lambda_start_lists():
    """DRMOVE(...,100,.TRUE.) (vcc/vcc.F:1138): lists 61/63 -> 144/146 and list 90
    group 3 -> list 190 group 1; the T lists 44/46/90 are left unchanged."""
    from aces2 import _lib, lists
    print('case drmove 100', flush=True)
    o, v = 4, 11
    _lists_system(o, v, 34, False)
    _lib.call('aces_vcc_mp1')
    _lib.call('aces_vcc_set_diis_order', 6)
    for name in ('aces_genint', 'aces_dre3en', 'aces_e4s', 'aces_newt2'):
        _lib.call(name)
    before = {(lst, g): lists.getlst(1, lists.shape(g, lst)[1], g, lst).copy()
              for (lst, g) in ((44, 1), (46, 1), (90, 1), (61, 1), (63, 1), (90, 3))}
    _lib.call('aces_drmove_lists', 100, 1)
    for key, x in before.items():
        np.testing.assert_array_equal(lists.getlst(1, lists.shape(key[1], key[0])[1], key[1], key[0]), x)
    np.testing.assert_array_equal(lists.getlst(1, lists.shape(1, 144)[1], 1, 144), before[(61, 1)])
    np.testing.assert_array_equal(lists.getlst(1, lists.shape(1, 146)[1], 1, 146), before[(63, 1)])
    np.testing.assert_array_equal(lists.getlst(1, 1, 1, 190), before[(90, 3)])
    assert lists.isytyp(146) == lists.isytyp(63) == (13, 14)
    assert np.abs(before[(63, 1)] - before[(46, 1)]).max() > 0
    _lib.call('aces_drmove_lists', 0, 0)   # without SING1 the T1 list stays
    np.testing.assert_array_equal(lists.getlst(1, 1, 1, 90), before[(90, 1)])
    np.testing.assert_array_equal(lists.getlst(1, o * o, 1, 46), before[(63, 1)])
    lists.fin()


def test_dpd_writer_sizes_and_subblocks_match_library():
    """The host writer's pair enumeration (aces2.lists.dpd_pairs) against the
    library's IRPDPD / ISYMOFF: group sizes, and each sub-block (irrep of q)
    starting where ISYMOFF says, p fastest inside it."""
    from aces2 import lists
    lists.init()
    for pop, vrt in [([2, 1, 0, 1], [3, 2, 4, 1]), ([1, 1, 1, 1, 0, 2, 1, 1], [2, 3, 1, 2, 2, 1, 3, 2]), ([2, 3], [4, 1])]:
        lists.set_symmetry(pop, vrt)
        irr_o = np.repeat(np.arange(len(pop)), pop)
        irr_v = np.repeat(np.arange(len(vrt)), vrt)
        for typ in range(1, 26):
            sp, sq, pack = lists.PAIR_TYPES[typ]
            irr_p = irr_o if sp == 'O' else irr_v
            irr_q = irr_o if sq == 'O' else irr_v
            for g in range(len(pop)):
                p, q = lists.dpd_pairs(pop, vrt, typ, g)
                assert len(p) == lists.irpdpd(g + 1, typ), (typ, g)
                if len(p) == 0:
                    continue
                assert np.all(irr_p[p] ^ irr_q[q] == g)
                if pack == 1:
                    assert np.all(p < q)
                elif pack == 2:
                    assert np.all(p <= q)
                for h in sorted(set(irr_q[q].tolist())):
                    first = int(np.argmax(irr_q[q] == h))
                    assert first + 1 == lists.isymoff(h + 1, g + 1, typ), (typ, g, h)
                # inside a sub-block: column (q) major, p fastest
                key = q * 10 ** 6 + p
                assert np.all(np.diff(key) > 0)
    lists.fin()


def _sym_system(occ, vir, seed, nonhf=False):
    from aces2 import synth
    eps, B, irr = synth.sym_draw(occ, vir, seed)
    o = sum(occ)
    eri = synth.full_eri(B)
    f = np.diag(eps)
    if nonhf:
        rng = np.random.default_rng(seed + 7)
        P = rng.normal(0, 0.01, f.shape)
        f = f + (P + P.T) * ((irr[:, None] ^ irr[None, :]) == 0)
    blk = dict(rccsd.blocks(eri, o))
    blk['fock'] = f
    blk['list233'] = synth.list233_from_eri(eri, o)
    return blk, eri, irr


@pytest.mark.gpu
def test_symmetry_packed_lists_through_c1_order():
    """Device DPD <-> C1 conversion (aces_list_c1) against the host writer:
    every integral list of a D2h system back in C1 order equals the C1 list with
    the symmetry-forbidden elements zero, and C1 -> groups reproduces the groups."""
    import torch
    from aces2 import _lib, lists
    print('case dpd lists c1 round trip', flush=True)
    occ, vir = [1, 2, 0, 1, 1, 1, 1, 1], [2, 3, 2, 1, 3, 2, 2, 3]
    blk, eri, irr = _sym_system(occ, vir, 7)
    o, v = sum(occ), sum(vir)
    lists.init()
    lists.load_sym_system(blk, occ, vir)
    for lst, (tl, tr, data) in lists.c1_integral_lists(blk, o).items():
        rows, cols = lists.c1_pairs(tl, o, v), lists.c1_pairs(tr, o, v)
        got = lists.list_to_c1(lst, tl, tr, rows, cols)
        np.testing.assert_array_equal(got.ravel(), data.ravel())   # forbidden elements are zero in blk
        # C1 -> groups: write a scaled copy back and read the groups
        dev = torch.as_tensor(2.0 * data.ravel(), device='cuda')
        _lib.call('aces_list_c1', lst, tl, tr, 1, dev.data_ptr(), 0)
        want = lists.dpd_groups_from_c1(2.0 * data.ravel(), occ, vir, tl, tr)
        for g, blkg in enumerate(want, start=1):
            nr, nc = lists.shape(g, lst)
            if nr * nc:
                np.testing.assert_array_equal(lists.getlst(1, nc, g, lst), blkg)
    lists.fin()


@pytest.mark.gpu
@pytest.mark.parametrize('nonhf', [False, True])
def test_xvcc_driver_over_symmetry_packed_lists(nonhf):
    """vcc.F's call sequence over symmetry-packed lists (D2h counts): the
    irrep-blocked engine behind the list entry points reproduces the oracle
    loop, CMPENG's spin cases included."""
    import ctypes
    from aces2 import _lib, lists
    print('case xvcc symmetry-packed lists nonhf=%s' % nonhf, flush=True)
    occ, vir = [1, 1, 0, 1, 1, 1, 0, 1], [2, 3, 2, 2, 3, 2, 2, 3]
    blk, eri, irr = _sym_system(occ, vir, 21, nonhf)
    o = sum(occ)
    g = rccsd.blocks(eri, o)
    ref = vcc_loop.run(blk['fock'], g, o, maxcyc=60, conv=9, order=6)
    lists.init()
    lists.load_sym_system(blk, occ, vir, nonhf=nonhf)
    _lib.call('aces_vcc_mp1')
    _lib.call('aces_vcc_set_diis_order', 6)
    e0 = ctypes.c_double()
    _lib.call('aces_cmpeng', 43, 0, None, ctypes.byref(e0), None)
    energies, resm, spin = [e0.value], [], []
    conv = ctypes.c_int(1)
    st = (ctypes.c_double * 8)()
    for icycle in range(1, 61):
        for name in ('aces_genint', 'aces_dre3en', 'aces_e4s', 'aces_newt2'):
            _lib.call(name)
        e = ctypes.c_double()
        ec = (ctypes.c_double * 3)()
        _lib.call('aces_cmpeng', 60, 2, ec, ctypes.byref(e), None)
        energies.append(e.value)
        spin.append((ec[0], ec[2]))
        _lib.call('aces_dodiis0', icycle, 1e-9, ctypes.byref(conv), st)
        resm.append(st[1])
        if conv.value == 0:
            break
        _lib.call('aces_drmove')
    ef = ctypes.c_double()
    _lib.call('aces_cmpeng', 43, 0, None, ctypes.byref(ef), None)
    lists.fin()
    assert conv.value == 0 and len(resm) == ref['cycles']
    np.testing.assert_allclose(energies, ref['energy'], rtol=0, atol=1e-10)
    np.testing.assert_allclose(resm, ref['resmiter'], rtol=1e-6, atol=1e-13)
    assert abs(ef.value - ref['final_energy']) < 1e-10
    # the spin cases of the last cycle's T_new against the oracle's split
    assert all(abs(2 * aa + ab - e) < 1e-12 for (aa, ab), e in zip(spin, energies[1:]))
