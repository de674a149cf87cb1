"""The ACES II job archive and list files, restated (ILP64: 8-byte integers).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): writes and reads the
files the way the reference's Fortran does, to check the library's reader and
writer byte for byte.

Archive (acescore/putrec.F, getrec.F, aces_ja_init.F, aces_ja_fin.F,
include/jobarc.h): up to 1000 logical records named by 8-character markers
('OPENSLOT' = free); a new record takes the first free slot and starts after
the previous slot's record (rloc(i) = rloc(i-1) + rsize(i-1)), moved to the
start of the next 128-integer physical record when a record of 2..128 integers
would straddle one; a record rewritten with a larger length leaves its slot as
'. LOST .' and takes a new one; a shorter rewrite keeps the slot and length.
JOBARC is a direct-access file of 1024-byte records; JAINDX one unformatted
sequential record (4-byte length markers) holding marker(1000), rloc(1000),
rsize(1000), nrecs.

Lists (acescore/aces_list_touch.F, getlst.F, include/lists.h): list (g, f)
lives in file 1 + (f-1)//100 of MOINTS, GAMLAM, MOABCD, DERINT, DERGAM; lists
are placed one after another (physical record pRec, word 1 + iIntOff), in
physical records of IFLAGS(37) integers, and their MOIO* arrays (10 x 500,
column major) are archive records.
"""
import os
import struct

import numpy as np

RECWD = 128
NSLOT = 1000
FILES = ('MOINTS', 'GAMLAM', 'MOABCD', 'DERINT', 'DERGAM')


class Archive:
    def __init__(self):
        self.marker = ['OPENSLOT'] * NSLOT
        self.rloc = [0] * NSLOT
        self.rsize = [0] * NSLOT
        self.rloc[0] = 1
        self.nrecs = 0
        self.words = {}   # physical record number -> np.int64[128]

    def _find(self, name):
        name = name.ljust(8)[:8]
        for i, m in enumerate(self.marker):
            if m == name:
                return i
        return -1

    def _rec(self, r):
        if r not in self.words:
            self.words[r] = np.zeros(RECWD, dtype=np.int64)
        return self.words[r]

    def putrec(self, name, data):
        data = np.asarray(data).view(np.int64).ravel() if np.asarray(data).dtype != np.int64 \
            else np.asarray(data).ravel()
        n = data.size
        i = self._find(name)
        if i >= 0 and n > self.rsize[i]:
            self.marker[i] = '. LOST .'
            i = -1
        if i < 0:
            i = self._find('OPENSLOT')
            self.marker[i] = name.ljust(8)[:8]
            if i != 0:
                self.rloc[i] = self.rloc[i - 1] + self.rsize[i - 1]
            if 2 <= n <= RECWD:
                r0 = 1 + (self.rloc[i] - 1) // RECWD
                r1 = 1 + (self.rloc[i] + n - 2) // RECWD
                if r0 != r1:
                    self.rloc[i] = 1 + (r1 - 1) * RECWD
            self.rsize[i] = n
        loc = self.rloc[i]
        for k in range(n):
            w = loc - 1 + k
            self._rec(1 + w // RECWD)[w % RECWD] = data[k]
            self.nrecs = max(self.nrecs, 1 + w // RECWD)

    def getrec(self, name):
        i = self._find(name)
        if i < 0:
            return None
        out = np.zeros(self.rsize[i], dtype=np.int64)
        for k in range(self.rsize[i]):
            w = self.rloc[i] - 1 + k
            out[k] = self._rec(1 + w // RECWD)[w % RECWD]
        return out

    def write(self, path):
        nrecs = max(self.nrecs, 1)   # the initial all-zero physical record
        with open(os.path.join(path, 'JOBARC'), 'wb') as f:
            for r in range(1, nrecs + 1):
                f.write(self._rec(r).tobytes())
        body = ''.join(self.marker).encode() + np.array(self.rloc, dtype=np.int64).tobytes() \
            + np.array(self.rsize, dtype=np.int64).tobytes() + struct.pack('<q', nrecs)
        with open(os.path.join(path, 'JAINDX'), 'wb') as f:
            f.write(struct.pack('<i', len(body)) + body + struct.pack('<i', len(body)))

    @classmethod
    def read(cls, path):
        a = cls()
        raw = open(os.path.join(path, 'JAINDX'), 'rb').read()
        (n,) = struct.unpack('<i', raw[:4])
        assert n == len(raw) - 8 and struct.unpack('<i', raw[-4:])[0] == n
        body = raw[4:-4]
        a.marker = [body[8 * i:8 * i + 8].decode() for i in range(NSLOT)]
        a.rloc = np.frombuffer(body[8 * NSLOT:16 * NSLOT], dtype=np.int64).tolist()
        a.rsize = np.frombuffer(body[16 * NSLOT:24 * NSLOT], dtype=np.int64).tolist()
        (a.nrecs,) = struct.unpack('<q', body[24 * NSLOT:24 * NSLOT + 8])
        data = np.fromfile(os.path.join(path, 'JOBARC'), dtype=np.int64)
        assert data.size == a.nrecs * RECWD
        for r in range(a.nrecs):
            a.words[r + 1] = data[r * RECWD:(r + 1) * RECWD].copy()
        return a


def write_lists(path, arch, lists, recwd, isytyp=None):
    """lists: {(group, family): 2-D array (nrows, ncols) column major data}
    placed in (family, group) order as aces_list_touch places them; the MOIO
    records and TOTRECMO/TOTWRDMO go into ``arch``."""
    moio = np.zeros((500, 10), dtype=np.int64)    # [family][group] = Fortran moio(g, f)
    wd, sz, ds, fl, mx = (np.zeros_like(moio) for _ in range(5))
    prec, ioff = [1] * 5, [0] * 5
    files = {}
    for (g, f) in sorted(lists, key=lambda k: (k[1], k[0])):
        a = np.asarray(lists[(g, f)], dtype=np.float64)
        nr, nc = a.shape
        h = (f - 1) // 100
        moio[f - 1, g - 1], wd[f - 1, g - 1] = prec[h], 1 + ioff[h]
        sz[f - 1, g - 1], ds[f - 1, g - 1], fl[f - 1, g - 1], mx[f - 1, g - 1] = nr, nc, 50 + h, nr * nc
        word0 = (prec[h] - 1) * recwd + ioff[h]
        buf = files.setdefault(h, {})
        for k, x in enumerate(a.ravel(order='F')):
            buf[word0 + k] = x
        words = nr * nc + ioff[h]
        prec[h] += words // recwd
        ioff[h] = words % recwd
    for h, buf in files.items():
        n = prec[h] * recwd
        arr = np.zeros(n, dtype=np.float64)
        for k, x in buf.items():
            arr[k] = x
        arr.tofile(os.path.join(path, FILES[h]))
    for name, v in (('MOIOVEC', moio), ('MOIOWRD', wd), ('MOIOSIZ', sz), ('MOIODIS', ds), ('MOIOFIL', fl),
                    ('MOIOMXSZ', mx)):
        arch.putrec(name, v.ravel())
    arch.putrec('ISYMTYP', np.zeros(1000, dtype=np.int64) if isytyp is None else isytyp)
    arch.putrec('TOTRECMO', np.array(prec, dtype=np.int64))
    arch.putrec('TOTWRDMO', np.array(ioff, dtype=np.int64))


def read_list(path, arch, g, f):
    """GETLST of the whole list (g, f) from the files (acescore/getlst.F)."""
    recwd = int(arch.getrec('IFLAGS')[36])
    col = lambda name: arch.getrec(name).reshape(500, 10)[f - 1, g - 1]
    nr, nc, unit = int(col('MOIOSIZ')), int(col('MOIODIS')), int(col('MOIOFIL'))
    word0 = (int(col('MOIOVEC')) - 1) * recwd + int(col('MOIOWRD')) - 1
    data = np.fromfile(os.path.join(path, FILES[unit - 50]), dtype=np.float64, count=nr * nc, offset=8 * word0)
    return data.reshape(nc, nr).T
