"""Host mirror of the ACES II list/record interface over the device store.

Python spellings of GETLST/PUTLST/GETALL/PUTALL/SUMSYM2/INIPCK/UPDMOI and
GETREC/PUTREC (``acescore/getlst.F:19``, ``putlst.F:19``, ``acesutil/getall.F``,
``acesutil/sumsym2.F``, ``libr/inipck.F:3``, ``libr/updmoi.F:11``,
``acescore/getrec.F:29``) with the same argument meaning: lists are addressed
by (group = irrep 1..nirrep, list 1..500), columns are 1-based, blocks are
column major.  Failures raise ``AcesError`` (the reference ends the process
via aces_exit(1)).

Also: writers that lay out closed-shell C1 integral blocks as the reference's
MOINTS lists 13, 16, 10, 30, 23 and 233, and the records SCFEVALA/FOCKMOA.
"""
import ctypes

import numpy as np

from . import _lib


def init(stream=None):
    _lib.call('aces_init', stream)


def fin():
    _lib.call('aces_fin')


def set_symmetry(pop, vrt, popb=None, vrtb=None):
    nirrep = len(pop)
    arr = lambda x: (ctypes.c_int * 8)(*list(x) + [0] * (8 - len(x))) if x is not None else None
    _lib.call('aces_set_symmetry', nirrep, arr(pop), arr(popb), arr(vrt), arr(vrtb))


def irpdpd(irrep, typ):
    return _lib.lib().aces_irpdpd(irrep, typ)


def isymoff(irrep_q, irrep_pq, typ):
    return _lib.lib().aces_isymoff(irrep_q, irrep_pq, typ)


def inipck(irrepx, typel, typer, lst):
    _lib.call('aces_inipck', irrepx, typel, typer, lst)


def updmoi(ncols, nrows, group, lst):
    _lib.call('aces_updmoi', ncols, nrows, group, lst)


def shape(group, lst):
    r, c = ctypes.c_long(), ctypes.c_long()
    _lib.call('aces_list_shape', group, lst, ctypes.byref(r), ctypes.byref(c))
    return r.value, c.value


def isytyp(lst):
    a, b = ctypes.c_int(), ctypes.c_int()
    _lib.call('aces_list_isytyp', lst, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def getlst(ifirst, ncols, group, lst):
    """Columns ifirst..ifirst+ncols-1 (1-based) as a (ncols, nrows) array (column major block)."""
    nrows, _ = shape(group, lst)
    out = np.empty((ncols, nrows))
    _lib.call('aces_getlst', out.ctypes.data, ifirst, ncols, group, lst, 0)
    return out


def putlst(src, ifirst, ncols, group, lst):
    a = np.ascontiguousarray(src, dtype=np.float64)
    _lib.call('aces_putlst', a.ctypes.data, ifirst, ncols, group, lst, 0)


def getall(irrepx, lst, length):
    out = np.empty(length)
    _lib.call('aces_getall', out.ctypes.data, length, irrepx, lst, 0)
    return out


def putall(src, irrepx, lst):
    a = np.ascontiguousarray(src, dtype=np.float64).ravel()
    _lib.call('aces_putall', a.ctypes.data, a.size, irrepx, lst, 0)


def sumsym2(add, group, lst):
    a = np.ascontiguousarray(add, dtype=np.float64)
    _lib.call('aces_sumsym2', a.ctypes.data, group, lst, 0)


def putrec(name, data):
    a = np.ascontiguousarray(data)
    assert a.itemsize == 8
    _lib.call('aces_putrec', name.encode(), a.size, a.ctypes.data)


def getrec(name, length, dtype=np.float64, iflag=1):
    out = np.zeros(length, dtype=dtype)
    n = ctypes.c_long(length)
    _lib.call('aces_getrec', iflag, name.encode(), ctypes.byref(n), out.ctypes.data)
    return out


def reclen(name):
    n = ctypes.c_long(0)
    _lib.call('aces_getrec', 0, name.encode(), ctypes.byref(n), None)
    return n.value


# ---------------------------------------------------------------------------
# Closed-shell C1 integral blocks -> the reference's list layouts.
def c1_integral_lists(blk, o):
    """{list: (typel, typer, block)} with block in memory order (column major)."""
    ovov, oovv, ovoo, oooo, ovvv = blk['ovov'], blk['oovv'], blk['ovoo'], blk['oooo'], blk['ovvv']
    return {
        16: (13, 14, np.ascontiguousarray(ovov.transpose(2, 0, 3, 1))),   # <ab|ij> at a+v b+v^2(i+o j)
        13: (14, 14, np.ascontiguousarray(oooo.transpose(3, 1, 2, 0))),   # <mn|ij> at m+o n+o^2(i+o j)
        10: (14, 18, np.ascontiguousarray(ovoo.transpose(1, 3, 0, 2))),   # <ij|ka> at i+o j+o^2(k+o a)
        30: (13, 11, np.ascontiguousarray(ovvv.transpose(0, 3, 1, 2))),   # <ab|ci> at a+v b+v^2(c+v i)
        23: (9, 9, np.ascontiguousarray(oovv.transpose(1, 3, 0, 2) - ovov.transpose(2, 3, 0, 1))),
        233: (5, 13, np.ascontiguousarray(blk['list233'])),
    }


def load_c1_system(blk, o, v, nonhf=False):
    """Symmetry (C1), integral lists and orbital-energy records into the store."""
    set_symmetry([o], [v])
    for lst, (tl, tr, data) in c1_integral_lists(blk, o).items():
        inipck(1, tl, tr, lst)
        nrows, ncols = shape(1, lst)
        assert data.size == nrows * ncols, (lst, data.shape, nrows, ncols)
        putlst(data, 1, ncols, 1, lst)
    f = np.asarray(blk['fock'], dtype=np.float64)
    putrec('SCFEVALA', np.diag(f).copy())
    if nonhf:
        putrec('FOCKMOA', f.copy())
