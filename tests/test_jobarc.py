"""The job archive (JAINDX/JOBARC) and list files on disk.

The library's reader and writer are checked against the restatement of the
reference's Fortran in ``oracle/jobarc.py`` (byte for byte) and against the
Fortran runtime's own unformatted I/O (``tests/fortran/ja_check.f90``, built
with the image's flang and -fdefault-integer-8, reading the files the way
acescore/aces_ja_init.F and getrec.F do).  No archive or list file ships with
the reference, so these and the round trips below are the pins; the list
tests move real lists (the O2 and H2O integral lists of the golden fixtures
are covered by the driver tests that read them back) through the files.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import jobarc

HERE = os.path.dirname(os.path.abspath(__file__))


def _records():
    rng = np.random.default_rng(3)
    out = [('JADIRTY', np.array([1])), ('NREP', np.array([7])), ('SCFEVALA', rng.normal(size=40).view(np.int64)),
           ('IFLAGS', np.arange(100)), ('EMPTY', np.zeros(0, dtype=np.int64)),
           ('NEARLY', np.arange(120)), ('STRADDL', np.arange(50)),   # 50 ints from word 300: moved to record 4
           ('BIG', np.arange(700)), ('GROWS', np.arange(5)), ('TAIL', np.array([9, 8]))]
    return out


def _fill_both(tmp_path):
    from aces2 import lists
    arch = jobarc.Archive()
    lists.init()
    for name, data in _records():
        arch.putrec(name, data)
        lists.putrec(name, data)
    # rewrites: a shorter one keeps the slot and length, a longer one moves ('. LOST .')
    for a in (arch,):
        a.putrec('NREP', np.array([11]))
        a.putrec('GROWS', np.arange(30))
    lists.putrec('NREP', np.array([11]))
    lists.putrec('GROWS', np.arange(30))
    return arch


def test_archive_bytes_match_the_reference_algorithm(tmp_path):
    from aces2 import lists
    arch = _fill_both(tmp_path)
    ours, ref = tmp_path / 'ours', tmp_path / 'ref'
    ours.mkdir()
    ref.mkdir()
    lists.jobarc_write(str(ours))
    arch.write(str(ref))
    for f in ('JOBARC', 'JAINDX'):
        assert (ours / f).read_bytes() == (ref / f).read_bytes(), f
    # read back (both directions): every record, lost ones hidden
    lists.fin()
    lists.init()
    lists.jobarc_read(str(ref))
    for name, data in _records():
        want = arch.getrec(name)
        got = lists.getrec(name, want.size, dtype=np.int64)
        np.testing.assert_array_equal(got, want)
    assert lists.reclen('GROWS') == 30 and lists.reclen('NREP') == 1
    back = jobarc.Archive.read(str(ours))
    assert back.marker[:12] == arch.marker[:12] and back.rloc[:12] == arch.rloc[:12]
    assert '. LOST .' in back.marker
    lists.fin()


def _flang():
    for cand in ('/opt/rocm/lib/llvm/bin/flang', shutil.which('flang') or ''):
        if cand and os.path.exists(cand):
            return cand
    return None


def test_archive_readable_by_fortran_unformatted_io(tmp_path):
    from aces2 import lists
    fl = _flang()
    if fl is None:
        pytest.skip('no Fortran compiler')
    exe = tmp_path / 'ja_check'
    subprocess.run([fl, '-fdefault-integer-8', '-o', str(exe), os.path.join(HERE, 'fortran', 'ja_check.f90')],
                   check=True, cwd=str(tmp_path), capture_output=True)
    _fill_both(tmp_path)
    d = tmp_path / 'ours'
    d.mkdir()
    lists.jobarc_write(str(d))
    lists.fin()
    for name, data in _records():
        out = subprocess.run([str(exe), 'read', str(d), name], check=True, capture_output=True, text=True).stdout
        vals = [int(x) for x in out.split()]
        want = {'NREP': [11], 'GROWS': list(range(30))}.get(name, np.asarray(data).tolist())
        assert vals[0] >= len(want) and vals[1:1 + len(want)] == list(want), name
    # JAINDX written by the Fortran runtime equals ours
    c = tmp_path / 'copy'
    c.mkdir()
    subprocess.run([str(exe), 'copy', str(d), str(c)], check=True, capture_output=True)
    assert (c / 'JAINDX').read_bytes() == (d / 'JAINDX').read_bytes()


def _driver_loop(maxcyc=60, tol=1e-9):
    """vcc.F's call sequence over the lists in the store (MP1 start, DIIS order 6)."""
    import ctypes
    from aces2 import _lib
    _lib.call('aces_vcc_mp1')
    _lib.call('aces_vcc_set_diis_order', 6)
    e0 = ctypes.c_double()
    _lib.call('aces_cmpeng', 43, 0, None, ctypes.byref(e0), None)
    energies, resm = [e0.value], []
    conv = ctypes.c_int(1)
    st = (ctypes.c_double * 8)()
    for icycle in range(1, maxcyc + 1):
        for name in ('aces_genint', 'aces_dre3en', 'aces_e4s', 'aces_newt2'):
            _lib.call(name)
        e = ctypes.c_double()
        _lib.call('aces_cmpeng', 60, 2, None, ctypes.byref(e), None)
        energies.append(e.value)
        _lib.call('aces_dodiis0', icycle, tol, ctypes.byref(conv), st)
        resm.append(st[1])
        if conv.value == 0:
            break
        _lib.call('aces_drmove')
    return energies, resm, conv.value == 0


def _all_lists():
    from aces2 import lists
    out = {}
    for f in range(1, 501):
        for g in range(1, 11):
            try:
                nr, nc = lists.shape(g, f)
            except Exception:
                continue
            out[(g, f)] = lists.getlst(1, nc, g, f).T if nr * nc else np.zeros((nr, nc))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize('symmetric', [False, True])
def test_list_files_round_trip_and_cc_run(tmp_path, symmetric):
    """Integral lists -> MOINTS/MOABCD + archive -> a fresh process state ->
    the same lists (and the files as the reference's GETLST reads them), then
    the CC loop from the files equals the oracle's."""
    from aces2 import lists, synth
    from oracle import rccsd, vcc_loop
    print('case list files symmetric=%s' % symmetric, flush=True)
    if symmetric:
        occ, vir = [1, 1, 0, 1, 1, 1, 0, 1], [2, 3, 2, 2, 3, 2, 2, 3]
        eps, B, irr = synth.sym_draw(occ, vir, 12)
        o = sum(occ)
        eri = synth.full_eri(B)
        blk = dict(rccsd.blocks(eri, o))
        blk['fock'] = np.diag(eps)
        blk['list233'] = synth.list233_from_eri(eri, o)
    else:
        o, v = 5, 19
        eps, B = synth.draw(o, v, 13, sigma=synth.scaled_sigma(o, v))
        blk, eri = synth.blocks_numpy(eps, B, o)
    ref = vcc_loop.run(blk['fock'], rccsd.blocks(eri, o), o, maxcyc=60, conv=9, order=6)
    d = tmp_path / 'job'
    d.mkdir()
    lists.init()
    if symmetric:
        lists.load_sym_system(blk, occ, vir)
    else:
        lists.load_c1_system(blk, o, blk['fock'].shape[0] - o)
    before = _all_lists()
    lists.putrec('IFLAGS', np.array([0] * 36 + [1000] + [0] * 63))   # 1000-word physical records
    lists.lists_write(str(d))
    lists.jobarc_write(str(d))
    lists.fin()
    # the files as the reference reads them
    arch = jobarc.Archive.read(str(d))
    for (g, f), a in before.items():
        np.testing.assert_array_equal(jobarc.read_list(str(d), arch, g, f), a)
    assert (d / 'MOINTS').exists() and (d / 'MOABCD').exists()    # list 233 lives in the third file
    # a fresh state from the files
    lists.init()
    lists.jobarc_read(str(d))
    lists.lists_read(str(d))
    after = _all_lists()
    assert set(after) == set(before)
    for k in before:
        np.testing.assert_array_equal(after[k], before[k])
    for lst in (13, 16, 10, 30, 23, 233):
        assert lists.isytyp(lst) == {13: (14, 14), 16: (13, 14), 10: (14, 18), 30: (13, 11), 23: (9, 9),
                                     233: (5, 13)}[lst]
    energies, resm, converged = _driver_loop()
    lists.fin()
    assert converged and len(resm) == ref['cycles']
    np.testing.assert_allclose(energies, ref['energy'], rtol=0, atol=1e-10)
    np.testing.assert_allclose(resm, ref['resmiter'], rtol=1e-6, atol=1e-13)


@pytest.mark.gpu
def test_lists_written_by_the_reference_algorithm_are_read(tmp_path):
    """List files and archive written by the restatement of the reference's
    Fortran (placement of aces_list_touch.F, records of putrec.F; 100-word
    physical records, so lists start mid-record and span many) are read into
    the store unchanged."""
    from aces2 import lists
    print('case list files from the reference algorithm', flush=True)
    rng = np.random.default_rng(5)
    data = {(1, 16): rng.normal(size=(9, 13)), (1, 44): rng.normal(size=(3, 1)), (2, 16): rng.normal(size=(4, 7)),
            (1, 233): rng.normal(size=(31, 17)), (1, 150): np.zeros((0, 4)), (3, 490): rng.normal(size=(5, 5))}
    arch = jobarc.Archive()
    arch.putrec('JADIRTY', np.array([0]))
    arch.putrec('IFLAGS', np.array([0] * 36 + [100] + [0] * 63))
    jobarc.write_lists(str(tmp_path), arch, data, 100)
    arch.write(str(tmp_path))
    lists.init()
    lists.jobarc_read(str(tmp_path))
    lists.lists_read(str(tmp_path))
    for (g, f), a in data.items():
        nr, nc = lists.shape(g, f)
        assert (nr, nc) == a.shape
        if a.size:
            np.testing.assert_array_equal(lists.getlst(1, nc, g, f).T, a)
    lists.fin()
