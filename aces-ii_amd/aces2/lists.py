"""Host mirror of the ACES II list/record interface over the device store.

Python spellings of GETLST/PUTLST/GETALL/PUTALL/SUMSYM2/INIPCK/UPDMOI and
GETREC/PUTREC (``acescore/getlst.F:19``, ``putlst.F:19``, ``acesutil/getall.F``,
``acesutil/sumsym2.F``, ``libr/inipck.F:3``, ``libr/updmoi.F:11``,
``acescore/getrec.F:29``) with the same argument meaning: lists are addressed
by (group = irrep 1..nirrep, list 1..500), columns are 1-based, blocks are
column major.  Failures raise ``AcesError`` (the reference ends the process
via aces_exit(1)).

Also: writers that lay out closed-shell integral blocks as the reference's
MOINTS lists 13, 16, 10, 30, 23 and 233 (C1, or packed by irrep with the DPD
pair ordering of acescore/aces_com_symloc.F) and the records SCFEVALA/FOCKMOA.
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


def intprc_df(B, eps, occ, vir):
    """Symmetry (irrep counts occ, vir) and the integral/denominator lists from
    factors B (naux, n, n) and orbital energies eps, built on the GPU
    (aces_intprc_df)."""
    import torch
    set_symmetry(list(occ), list(vir))
    Bd = torch.as_tensor(np.ascontiguousarray(B, dtype=np.float64), device='cuda')
    e = np.ascontiguousarray(eps, dtype=np.float64)
    _lib.call('aces_intprc_df', Bd.data_ptr(), Bd.shape[0], e.ctypes.data)
    del Bd


def jobarc_read(path=''):
    """JAINDX + JOBARC in directory ``path`` -> the record store."""
    _lib.call('aces_jobarc_read', str(path).encode())


def jobarc_write(path=''):
    _lib.call('aces_jobarc_write', str(path).encode())


def lists_read(path=''):
    """The list files in ``path`` (placed by the archive's MOIO records) -> the list store."""
    _lib.call('aces_lists_read', str(path).encode())


def lists_write(path=''):
    _lib.call('aces_lists_write', str(path).encode())


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


# ---------------------------------------------------------------------------
# Symmetry-packed (DPD) lists.  Pair types (acescore/aces_com_symloc.F:11-34):
# spaces of the pair's members and packing (0 full, 1 p<q, 2 p<=q); for the
# closed-shell lists the alpha and beta counts are equal.
PAIR_TYPES = {1: ('V', 'V', 1), 2: ('V', 'V', 1), 3: ('O', 'O', 1), 4: ('O', 'O', 1), 5: ('V', 'V', 2),
              6: ('V', 'V', 2), 7: ('O', 'O', 2), 8: ('O', 'O', 2), 9: ('V', 'O', 0), 10: ('V', 'O', 0),
              11: ('V', 'O', 0), 12: ('V', 'O', 0), 13: ('V', 'V', 0), 14: ('O', 'O', 0), 15: ('V', 'V', 0),
              16: ('O', 'V', 0), 17: ('O', 'V', 0), 18: ('O', 'V', 0), 19: ('V', 'V', 0), 20: ('V', 'V', 0),
              21: ('O', 'O', 0), 22: ('O', 'O', 0), 23: ('V', 'V', 0), 24: ('O', 'O', 0), 25: ('O', 'V', 0)}


def dpd_pairs(pop, vrt, typ, irrep):
    """Global (p, q) orbital indices (each space irrep-sorted) of the pairs of
    pair type ``typ`` with total irrep ``irrep`` (0-based), in the list's row
    (column) order: sub-blocks by the irrep of q ascending; inside a sub-block p
    fastest, or the triangle q(q-1)/2 + p (p<q) / q(q+1)/2 + p (p<=q) when both
    lie in one irrep of a packed type."""
    sp, sq, pack = PAIR_TYPES[typ]
    cnt = {'O': list(pop), 'V': list(vrt)}
    off = {k: np.concatenate([[0], np.cumsum(c)[:-1]]).astype(int) for k, c in cnt.items()}
    ps, qs = [], []
    for hq in range(len(pop)):
        hp = hq ^ irrep
        if pack and hp > hq:
            continue
        npp, nq = cnt[sp][hp], cnt[sq][hq]
        for lq in range(nq):
            if pack and hp == hq:
                lps = range(lq) if pack == 1 else range(lq + 1)
            else:
                lps = range(npp)
            for lp in lps:
                ps.append(off[sp][hp] + lp)
                qs.append(off[sq][hq] + lq)
    return np.array(ps, dtype=np.int64), np.array(qs, dtype=np.int64)


def c1_pair_index(p, q, typ, n_p):
    """Index of pair (p, q) in the C1 order of pair type ``typ`` (n_p orbitals in p's space)."""
    pack = PAIR_TYPES[typ][2]
    if pack == 0:
        return p + n_p * q
    return q * (q - 1) // 2 + p if pack == 1 else q * (q + 1) // 2 + p


def dpd_groups_from_c1(c1, pop, vrt, tl, tr):
    """Column-major blocks of a totally symmetric list, one per irrep, from the
    list in C1 order (flat, column major, rows = C1 pairs of tl)."""
    o, v = sum(pop), sum(vrt)
    nsp = {'O': o, 'V': v}
    nrow = c1_pairs(tl, o, v)
    mat = np.asarray(c1, dtype=np.float64).reshape(-1, nrow)  # [col][row]
    out = []
    for g in range(len(pop)):
        pr, qr = dpd_pairs(pop, vrt, tl, g)
        pc, qc = dpd_pairs(pop, vrt, tr, g)
        rows = c1_pair_index(pr, qr, tl, nsp[PAIR_TYPES[tl][0]])
        cols = c1_pair_index(pc, qc, tr, nsp[PAIR_TYPES[tr][0]])
        out.append(np.ascontiguousarray(mat[np.ix_(cols, rows)]))  # [ncols][nrows] = column major block
    return out


def c1_pairs(typ, o, v):
    sp, sq, pack = PAIR_TYPES[typ]
    n = {'O': o, 'V': v}
    if pack == 0:
        return n[sp] * n[sq]
    return n[sp] * (n[sp] - 1) // 2 if pack == 1 else n[sp] * (n[sp] + 1) // 2


def sym_integral_lists(blk, occ, vir):
    """{list: (typel, typer, [group blocks])}: the C1 integral lists of
    ``c1_integral_lists`` (orbitals irrep-sorted, symmetry-forbidden elements
    zero) packed by irrep."""
    o = sum(occ)
    return {lst: (tl, tr, dpd_groups_from_c1(data.ravel(), occ, vir, tl, tr))
            for lst, (tl, tr, data) in c1_integral_lists(blk, o).items()}


def load_sym_system(blk, occ, vir, nonhf=False):
    """Symmetry (irrep counts), the symmetry-packed integral lists and the
    orbital-energy records (irrep-sorted order) into the store."""
    set_symmetry(list(occ), list(vir))
    for lst, (tl, tr, groups) in sym_integral_lists(blk, occ, vir).items():
        inipck(1, tl, tr, lst)
        for g, data in enumerate(groups, start=1):
            nrows, ncols = shape(g, lst)
            assert data.size == nrows * ncols, (lst, g, data.shape, nrows, ncols)
            if data.size:
                putlst(data, 1, ncols, g, lst)
    f = np.asarray(blk['fock'], dtype=np.float64)
    putrec('SCFEVALA', np.diag(f).copy())
    if nonhf:
        putrec('FOCKMOA', f.copy())


def list_to_c1(lst, typel, typer, rows, cols):
    """A totally symmetric list in C1 order (device round trip through
    aces_list_c1), as a (cols, rows) host array."""
    import torch
    out = torch.empty(rows * cols, dtype=torch.float64, device='cuda')
    _lib.call('aces_list_c1', lst, typel, typer, 1, out.data_ptr(), 1)
    return out.cpu().numpy().reshape(cols, rows)
