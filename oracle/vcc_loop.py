"""The reference's closed-shell CC iteration (Jacobi step + DIIS), restated.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Order of one cycle follows ``vcc/vcc.F:798-1437``:
  INITIN/GENINT/.../DRE3EN/E4S   increments            (``rccsd.increments``)
  E4S / NEWT2                    T_new = inc / D       (``vcc/e4s.F``, ``libr2/newt2.F:233-244``)
  ABTOAA                         AA from AB            (``libr2/abtoaa.F``)
  CMPENG(60,2)                   E(T_new), printed     (``vcc/vcc.F`` after NEWT2; ``libr/cmpeng.F``)
  DODIIS0                        symmetrise, error, B matrix, dgesv, extrapolate
                                 (``vcc/dodiis0.F``, ``libr/dodiis.F``, ``vcc/diislst.F``)
  DRMOVE                         T <- increment lists  (``libr/drmove.f``) unless converged
  on convergence: CMPENG(43,0) on the T of that cycle -> TOTENERG (``vcc/vcc.F:1100``,
  ``vcc/finish.F:140``).
``hist['energy_aa_ab']`` holds CMPENG's printed (AA, AB) pair of every energy line.

DIIS vector (RHF, C1) = [T1 (list 90, AI order), T2(Ab,Ij) (list 46, Ab rows,
Ij columns, column major), T2(A<B,I<J) (list 44, packed)] -- ``vcc/diislst.F:20-29``
with the loop order ISPIN = 3, 1 of ``vcc/dodiis0.F`` (AB before AA).
"""
import numpy as np
import scipy.linalg.lapack
from . import rccsd


def pack_lt(n):
    """Row/col indices of the ACES 'PACK' (p<q) ordering: q outer, p inner (``libr/assym.f``)."""
    p, q = [], []
    for j in range(1, n):
        for i in range(j):
            p.append(i)
            q.append(j)
    return np.array(p, int), np.array(q, int)


def diis_vector(t1, t2):
    """Concatenated DIIS vector in the reference's storage order."""
    o, v = t1.shape
    ab = t2.transpose(1, 0, 3, 2).ravel()  # list 46: element A + v*b + v^2*(I + o*j)
    pi, pj = pack_lt(o)
    pa, pb = pack_lt(v)
    # AA(A<B, I<J) = t2[I,J,A,B] - t2[J,I,A,B]; column (I<J) slowest, row (A<B) fastest
    x = t2[pi, pj] - t2[pj, pi]           # [pairIJ, a, b]
    aa = x[:, pa, pb]                      # [pairIJ, pairAB]
    return np.concatenate([t1.ravel(), ab, aa.ravel()])


def from_diis_vector(vec, o, v):
    """T1 and T2(AB) from a DIIS vector (the AA part is discarded, as DODIIS0 does)."""
    t1 = vec[:o * v].reshape(o, v)
    t2 = vec[o * v:o * v + o * o * v * v].reshape(o, o, v, v).transpose(1, 0, 3, 2).copy()
    return t1.copy(), t2


def symmetrize_t2(t2):
    """``SYMETRIZE_T2``: T(Ab,Ij) <- (T(Ab,Ij) + T(bA,jI)) / 2 (``libr2/symetrize_t2.F``)."""
    return 0.5 * (t2 + t2.transpose(1, 0, 3, 2))


def first_absmax(x):
    """ISAMAX semantics: signed value of the first element of largest magnitude."""
    return float(x[int(np.argmax(np.abs(x)))])


class DIIS:
    """DODIIS0 state: circular subspace of CC_EXPORDER vectors (``vcc/dodiis0.F:36-40``)."""

    def __init__(self, order):
        self.order = order
        self.sol = {}
        self.err = {}
        self.R = np.zeros((order, order))

    def step(self, icycle, told, tnew):
        n = self.order
        ndimr = min(icycle, n)
        idimr = 1 + (icycle - 1) % n
        e = tnew - told
        resmiter = first_absmax(e)
        self.sol[idimr] = told.copy()
        for i in range(1, ndimr + 1):
            if i == idimr:
                continue
            x = float(np.dot(self.err[i], e))
            self.R[i - 1, idimr - 1] = x
            self.R[idimr - 1, i - 1] = x
        self.err[idimr] = e.copy()
        self.R[idimr - 1, idimr - 1] = float(np.dot(e, e))
        c = solve_diis(self.R[:ndimr, :ndimr])
        # DIIS residual: c_idimr*e + sum_{others} c_i e_i (``vcc/dodiis0.F`` "EVALUATE DIIS RESIDUAL")
        res = c[idimr - 1] * e
        for i in range(ndimr, idimr, -1):
            res = res + c[i - 1] * self.err[i]
        for i in range(idimr - 1, 0, -1):
            res = res + c[i - 1] * self.err[i]
        resmdiis = first_absmax(res)
        newvec = None
        if ndimr >= 2:
            newvec = res.copy()
            for i in range(1, ndimr + 1):
                newvec = newvec + c[i - 1] * self.sol[i]
        return resmiter, resmdiis, c, newvec


def solve_diis(R):
    """``DODIIS``: augmented Pulay system solved with LU (dgesv), ``libr/dodiis.F:30-46``."""
    nd = R.shape[0]
    A = np.zeros((nd + 1, nd + 1))
    A[0, 1:] = -1.0
    A[1:, 0] = -1.0
    A[1:, 1:] = R
    rhs = np.zeros(nd + 1)
    rhs[0] = -1.0
    lu, piv, x, info = scipy.linalg.lapack.dgesv(A, rhs)
    if info != 0:
        raise FloatingPointError('DODIIS: linear solver failed, INFO = %d' % info)
    return x[1:]


def mp1_start(f, g, o):
    """MP1 amplitudes written by xintprc (``intprc/dfrmt2.F``): T2 = <ij|ab>/D, T1 = 0."""
    v = f.shape[0] - o
    d1, d2 = rccsd.denominators(f, o)
    t2 = g['ovov'].transpose(0, 2, 1, 3) / d2
    return np.zeros((o, v)), t2


def run(f, g, o, maxcyc=50, conv=7, order=6, start='mp1', increments=None, energy=None,
        t1=None, t2=None, log=None):
    """Run the reference's CC iteration; returns a dict with the per-cycle history.

    ``conv`` is CC_CONV (converged when |RESMITER| < 10**-conv, ``vcc/dodiis0.F:167``),
    ``order`` is CC_EXPORDER (DIIS subspace size), ``start`` 'mp1' (the reference's start)
    or 'zero' (T1 = T2 = 0).
    """
    inc_fn = increments or (lambda a, b: rccsd.increments(a, b, f, g, o))
    e_fn = energy or (lambda a, b: rccsd.energy(a, b, f, g, o))
    v = f.shape[0] - o
    d1, d2 = rccsd.denominators(f, o)
    if t1 is None:
        if start == 'mp1':
            t1, t2 = mp1_start(f, g, o)
        else:
            t1, t2 = np.zeros((o, v)), np.zeros((o, o, v, v))
    parts = lambda a, b: rccsd.energy_parts(a, b, f, g, o)[:2] if energy is None else (None, None)
    hist = {'energy': [e_fn(t1, t2)], 'resmiter': [], 'resmdiis': [], 'converged': False,
            'energy_aa_ab': [parts(t1, t2)]}
    diis = DIIS(order)
    tol = 10.0 ** (-conv)
    for icycle in range(1, maxcyc + 1):
        inc1, inc2 = inc_fn(t1, t2)
        t1n = inc1 / d1
        t2n = inc2 / d2
        hist['energy'].append(e_fn(t1n, t2n))
        hist['energy_aa_ab'].append(parts(t1n, t2n))
        t2 = symmetrize_t2(t2)
        t2n = symmetrize_t2(t2n)
        told = diis_vector(t1, t2)
        tnew = diis_vector(t1n, t2n)
        resmiter, resmdiis, c, newvec = diis.step(icycle, told, tnew)
        hist['resmiter'].append(resmiter)
        hist['resmdiis'].append(resmdiis)
        if log:
            log(icycle, hist['energy'][-1], resmiter, resmdiis)
        if abs(resmiter) < tol:
            hist['converged'] = True
            hist['final_energy'] = e_fn(t1, t2)
            hist['energy_aa_ab'].append(parts(t1, t2))
            hist['cycles'] = icycle
            # DODIIS0 extrapolates in the converged cycle too and writes the
            # extrapolant over the increment lists (``vcc/dodiis0.F:181-233``,
            # "SKIP ONLY IN THE FIRST ITERATION"); the amplitudes are not moved
            # (``vcc/vcc.F:1077-1087``).  That content of the increment lists is
            # what DRMOVE(...,100,SING1) hands to xlambda (``vcc/vcc.F:1130-1143``).
            hist['lambda_start'] = (from_diis_vector(newvec, o, v) if newvec is not None
                                    else (t1n.copy(), t2n.copy()))
            break
        if newvec is not None:
            t1, t2 = from_diis_vector(newvec, o, v)
        else:
            t1, t2 = t1n, t2n
    hist['t1'] = t1
    hist['t2'] = t2
    return hist
