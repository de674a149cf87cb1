"""Element-sampled closed-shell CCSD increments for large systems.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

``rccsd.increments`` forms every intermediate in full (W(abcd) alone is v^4),
which is out of reach at the sizes the GPU engine is measured at (o=48,
v=320).  This module evaluates the same increments (``rccsd.increments``; the
reference's terms ``vcc/genint.F``, ``vcc/drlad.f``, ``vcc/drrng.f``,
``vcc/t1int2a.f``, ``vcc/t1int2b.f``, ``vcc/t12int2.F``, ``vcc/feacont.f``,
``vcc/fmicont.f``, ``vcc/e4s.F``) for selected output slices only:

  inc2[i, :, a, :]  for the requested (i, a) pairs (all j and b),
  inc1              in full,

from the density-fitted factors ((pq|rs) = sum_P B[P,p,q] B[P,r,s]) and the
amplitudes.  Every term of ``rccsd.increments`` is re-associated so that the
cost per slice is O((ov)^2) or O(naux v^2): the ring terms whose W(mbej)
intermediate carries an output index are contracted with the fixed amplitude
vector first, and the v^3-sized integral slices are built from B per slice.
Only the O(o^2 v^2) blocks ovov, oovv, ovoo, oooo and the small intermediates
(F, L, W(mnij)) are held in full.  ``tests/test_oracle.py`` pins this module
against ``rccsd.increments`` at small sizes, every slice and inc1 in full.

Conventions as ``rccsd``: t1[i,a], t2[i,j,a,b] = t_{Ij}^{Ab}, chemists'
integrals ovov[i,a,j,b] = (ia|jb), oovv[i,j,a,b] = (ij|ab), ovoo[i,a,j,k] =
(ia|jk), oooo[i,j,k,l] = (ij|kl).
"""
import functools

import numpy as np

# every contraction through BLAS where the operand orders allow it
_e = functools.partial(np.einsum, optimize=True)


def _blas(x):
    # BLAS takes C- or Fortran-ordered operands; any other strided view is copied
    return x if (x.flags.c_contiguous or x.flags.f_contiguous) else np.ascontiguousarray(x)


def _mm(a, b):
    return np.matmul(_blas(a), _blas(b))


class DFSystem:
    """Integral blocks of a closed-shell system from DF factors B (naux, n, n)."""

    def __init__(self, B, f, o):
        B = np.asarray(B, dtype=np.float64)
        self.naux, n, _ = B.shape
        self.o, self.v = o, n - o
        o_, v = self.o, self.v
        self.f = np.asarray(f, dtype=np.float64)
        self.Boo = np.ascontiguousarray(B[:, :o_, :o_])
        self.Bov = np.ascontiguousarray(B[:, :o_, o_:])
        self.Bvv = np.ascontiguousarray(B[:, o_:, o_:])
        P = self.naux
        Bov2 = self.Bov.reshape(P, o_ * v)
        Boo2 = self.Boo.reshape(P, o_ * o_)
        Bvv2 = self.Bvv.reshape(P, v * v)
        self.ovov = _mm(Bov2.T, Bov2).reshape(o_, v, o_, v)
        self.oovv = _mm(Boo2.T, Bvv2).reshape(o_, o_, v, v)
        self.ovoo = _mm(Bov2.T, Boo2).reshape(o_, v, o_, o_)
        self.oooo = _mm(Boo2.T, Boo2).reshape(o_, o_, o_, o_)
        # ovov with its second and fourth index exchanged: ovovx[l,d,k,c] = (lc|kd)
        self.ovovx = np.ascontiguousarray(self.ovov.transpose(0, 3, 2, 1))

    def ovvv_a(self, a):
        """(kd|ac) for fixed a as [k, d, c]."""
        o, v = self.o, self.v
        return _mm(self.Bov.reshape(self.naux, o * v).T, self.Bvv[:, a, :]).reshape(o, v, v)

    def vvvv_a(self, a):
        """(ac|bd) for fixed a as [c, b, d]."""
        v = self.v
        return _mm(self.Bvv[:, a, :].T, self.Bvv.reshape(self.naux, v * v)).reshape(v, v, v)


def _dressed_factors(S, tau):
    """Half-transformed DF factors of tau (one occupied index at a time):
      M1[P,i,c] = sum_ld B[P,l,d] tau[i,l,c,d]
      M2[P,i,d] = sum_lc B[P,l,c] tau[i,l,c,d]
      M3[P,l,a] = sum_kd B[P,k,d] tau[k,l,a,d]"""
    o, v = S.o, S.v
    P = S.naux
    Bov2 = S.Bov.reshape(P, o * v)
    M1 = np.empty((P, o, v))
    M2 = np.empty((P, o, v))
    M3 = np.empty((P, o, v))
    for i in range(o):
        M1[:, i, :] = _mm(Bov2, tau[i].transpose(0, 2, 1).reshape(o * v, v))
        M2[:, i, :] = _mm(Bov2, tau[i].reshape(o * v, v))
        M3[:, i, :] = _mm(Bov2, tau[:, i].transpose(0, 2, 1).reshape(o * v, v))
    return M1, M2, M3


def _intermediates(S, t1, t2):
    """Foo, Fvv, Fov, Loo, Lvv (diagonal removed where rccsd removes it) and W(klij)."""
    o, v = S.o, S.v
    f = S.f
    fov, foo, fvv = f[:o, o:], f[:o, :o], f[o:, o:]
    ovov, ovoo = S.ovov, S.ovoo
    P = S.naux
    tau = t2 + _e('ia,jb->ijab', t1, t1)
    M1, M2, M3 = _dressed_factors(S, tau)
    Bov2k = S.Bov.transpose(1, 0, 2).reshape(o, P * v)                 # [k, (P,c)]
    # Foo[k,i] = f_ki + sum_lcd (2 (kc|ld) - (kd|lc)) tau[i,l,c,d]
    Foo = foo + 2 * _mm(Bov2k, M1.transpose(1, 0, 2).reshape(o, P * v).T) \
        - _mm(Bov2k, M2.transpose(1, 0, 2).reshape(o, P * v).T)
    # Fvv[a,c] = f_ac - sum_kld (2 (kc|ld) - (kd|lc)) tau[k,l,a,d]
    Bc = S.Bov.reshape(P * o, v)                                        # [(P,k), c]
    Fvv = fvv - 2 * _mm(M1.reshape(P * o, v).T, Bc) + _mm(M3.reshape(P * o, v).T, Bc)
    ovov2 = ovov.reshape(o * v, o * v)
    Fov = fov + (2 * (ovov2 @ t1.ravel()) - S.ovovx.reshape(o * v, o * v) @ t1.ravel()).reshape(o, v)
    Loo = (Foo + _e('kc,ic->ki', fov, t1)
           + 2 * _e('lcki,lc->ki', ovoo, t1) - _e('kcli,lc->ki', ovoo, t1))
    # Lvv t1 terms through the DF factors:
    #   sum_kd (kd|ac) t1[k,d] = sum_P x_P B[P,a,c],  x_P = sum_kd B[P,k,d] t1[k,d]
    #   sum_kd (kc|ad) t1[k,d] = sum_Pk B[P,k,c] Y[P,a,k],  Y[P,a,k] = sum_d B[P,a,d] t1[k,d]
    x = _e('pkd,kd->p', S.Bov, t1)
    term1 = _e('p,pac->ac', x, S.Bvv)
    Y = _mm(S.Bvv.reshape(P * v, v), t1.T).reshape(P, v, o)
    term2 = _mm(Y.transpose(1, 0, 2).reshape(v, P * o), S.Bov.reshape(P * o, v))
    Lvv = Fvv - _e('kc,ka->ac', fov, t1) + 2 * term1 - term2
    eo, ev = np.diag(foo), np.diag(fvv)
    # W(klij) = (ki|lj) + t1 parts + sum_cd (kc|ld) tau[i,j,c,d], one k at a time
    Woooo = (S.oooo.transpose(0, 2, 1, 3) + _e('lcki,jc->klij', ovoo, t1)
             + _e('kclj,ic->klij', ovoo, t1))
    taur = tau.reshape(o * o, v * v)
    for k in range(o):
        Woooo[k] += _mm(ovov[k].transpose(1, 0, 2).reshape(o, v * v), taur.T).reshape(o, o, o)
    return {'tau': tau, 'Foo_d': Foo - np.diag(eo), 'Fvv_d': Fvv - np.diag(ev), 'Fov': Fov,
            'Loo_d': Loo - np.diag(eo), 'Lvv_d': Lvv - np.diag(ev), 'Woooo': Woooo, 'M1': M1, 'M2': M2}


def _inc1(S, t1, t2, I):
    """inc1 in full (rccsd.increments' T1 equation), ovvv terms through the DF factors."""
    o, v = S.o, S.v
    f = S.f
    fov = f[:o, o:]
    ovov, oovv, ovoo = S.ovov, S.oovv, S.ovoo
    tau = I['tau']
    Fov = I['Fov']
    inc1 = fov.copy()
    inc1 -= 2 * _e('kc,ka,ic->ia', fov, t1, t1)
    inc1 += t1 @ I['Fvv_d'].T
    inc1 -= I['Foo_d'].T @ t1
    for k in range(o):
        inc1 += 2 * _e('c,ica->ia', Fov[k], t2[k])                     # sum_kc Fov[k,c] t2[k,i,c,a]
        inc1 -= _e('kiac,c->ia', oovv[k:k + 1], t1[k])                 # sum_kc (ki|ac) t1[k,c]
        # - 2 sum_lc (lc|ki) tau[k,l,a,c] + sum_lc (kc|li) tau[k,l,a,c]
        inc1 -= 2 * _e('lci,lac->ia', ovoo[:, :, k, :], tau[k])
    for i in range(o):
        inc1[i] -= _e('kc,kca->a', Fov, t2[i])                          # sum_kc Fov[k,c] t2[i,k,c,a]
    inc1 += _e('kc,ic,ka->ia', Fov, t1, t1)
    inc1 += 2 * (ovov.reshape(o * v, o * v) @ t1.ravel()).reshape(o, v)   # ovvo[k,c,a,i] = (kc|ia)
    # 2 sum_kcd (kd|ac) tau[i,k,c,d] - sum_kcd (kc|ad) tau[i,k,c,d]
    inc1 += 2 * _e('pac,pic->ia', S.Bvv, I['M1'])
    inc1 -= _e('pad,pid->ia', S.Bvv, I['M2'])
    # sum_klc (kc|li) tau[k,l,a,c]
    for k in range(o):
        inc1 += _e('cli,lac->ia', ovoo[k], tau[k])
    return inc1


def _wvoov_slice(S, t1, t2, a, i):
    """Wvoov[a,k,i,c] of rccsd for fixed (a, i) as [k, c]."""
    o, v = S.o, S.v
    ovov, ovovx, ovoo = S.ovov, S.ovovx, S.ovoo
    w = ovov[:, :, i, a].copy()                                        # ovvo[k,c,a,i]
    z = S.Bvv[:, a, :] @ t1[i]                                         # [P]: sum_d B[P,a,d] t1[i,d]
    w += _e('p,pkc->kc', z, S.Bov)                              # sum_d (kc|ad) t1[i,d]
    w -= _e('kcl,l->kc', ovoo[:, :, :, i], t1[:, a])            # sum_l (kc|li) t1[l,a]
    ovov2 = ovov.reshape(o * v, o * v)
    x = t2[i, :, :, a].reshape(o * v)                                  # [l,d]
    w -= 0.5 * (x @ ovov2).reshape(o, v)                               # (ld|kc) t2[i,l,d,a]
    y = t2[i, :, a, :].reshape(o * v)                                  # [l,d] = t2[i,l,a,d]
    w -= 0.5 * (y @ ovovx.reshape(o * v, o * v)).reshape(o, v)         # (lc|kd) t2[i,l,a,d]
    w -= (np.outer(t1[:, a], t1[i]).reshape(o * v) @ ovov2).reshape(o, v)
    w += (y @ ovov2).reshape(o, v)                                     # (ld|kc) t2[i,l,a,d]
    return w


def _wvovo_slice(S, t1, t2, a, i):
    """Wvovo[a,k,c,i] of rccsd for fixed (a, i) as [k, c]."""
    o, v = S.o, S.v
    ovovx, ovoo = S.ovovx, S.ovoo
    w = S.oovv[:, i, a, :].copy()                                      # (ki|ac)
    zk = _e('pkd,d->pk', S.Bov, t1[i])                          # sum_d B[P,k,d] t1[i,d]
    w += _e('pk,pc->kc', zk, S.Bvv[:, a, :])                    # sum_d (kd|ac) t1[i,d]
    w -= _e('lck,l->kc', ovoo[:, :, :, i], t1[:, a])            # sum_l (lc|ki) t1[l,a]
    x = t2[i, :, :, a].reshape(o * v)
    w -= 0.5 * (x @ ovovx.reshape(o * v, o * v)).reshape(o, v)         # (lc|kd) t2[i,l,d,a]
    w -= (np.outer(t1[:, a], t1[i]).reshape(o * v) @ ovovx.reshape(o * v, o * v)).reshape(o, v)
    return w


def _wvoov_contract(S, t1, t2, y):
    """sum_kc y[k,c] Wvoov[b,k,j,c] for all (j, b), as [j, b]."""
    o, v = S.o, S.v
    ovov, ovovx, ovoo = S.ovov, S.ovovx, S.ovoo
    ovov2 = ovov.reshape(o * v, o * v)
    yv = y.reshape(o * v)
    r = (yv @ ovov2).reshape(o, v)                                     # (kc|jb)
    wP = _e('pkc,kc->p', S.Bov, y)                              # sum_d (kc|bd) t1[j,d]
    r += _e('p,pbd,jd->jb', wP, S.Bvv, t1)
    q = _e('kclj,kc->lj', ovoo, y)                              # sum_l (kc|lj) t1[l,b]
    r -= q.T @ t1
    u = (ovov2 @ yv).reshape(o, v)                                     # u[l,d] = sum_kc (ld|kc) y[k,c]
    r -= 0.5 * _e('ld,jldb->jb', u, t2)
    u2 = (ovovx.reshape(o * v, o * v) @ yv).reshape(o, v)              # u2[l,d] = sum_kc (lc|kd) y[k,c]
    r -= 0.5 * _e('ld,jlbd->jb', u2, t2)
    r -= _e('ld,jd,lb->jb', u, t1, t1)
    r += _e('ld,jlbd->jb', u, t2)
    return r


def _wvovo_contract(S, t1, t2, y, Z1):
    """sum_kc y[k,c] Wvovo[b,k,c,j] for all (j, b), as [j, b]; Z1[P,k,j] = sum_d B[P,k,d] t1[j,d]."""
    o, v = S.o, S.v
    ovovx, ovoo = S.ovovx, S.ovoo
    r = _e('kc,kjbc->jb', y, S.oovv)                           # (kj|bc)
    Z2 = _e('pbc,kc->pbk', S.Bvv, y)                            # sum_d (kd|bc) t1[j,d]
    r += _e('pkj,pbk->jb', Z1, Z2)
    q = _e('lckj,kc->lj', ovoo, y)                              # sum_l (lc|kj) t1[l,b]
    r -= q.T @ t1
    u2 = (ovovx.reshape(o * v, o * v) @ y.reshape(o * v)).reshape(o, v)   # sum_kc (lc|kd) y[k,c]
    r -= 0.5 * _e('ld,jldb->jb', u2, t2)
    r -= _e('ld,jd,lb->jb', u2, t1, t1)
    return r


def _wvovo_i(S, t1, t2, i):
    """Wvovo[b,k,c,i] for fixed i as [b, k, c]."""
    o, v = S.o, S.v
    ovov, ovovx, ovoo = S.ovov, S.ovovx, S.ovoo
    W = S.oovv[:, i, :, :].transpose(1, 0, 2).copy()                  # (ki|bc) as [b,k,c]
    zk = _e('pkd,d->pk', S.Bov, t1[i])
    W += _e('pk,pbc->bkc', zk, S.Bvv)
    W -= _e('lck,lb->bkc', ovoo[:, :, :, i], t1)
    X = t2[i].transpose(2, 0, 1).reshape(v, o * v)                     # [b, (l,d)] = t2[i,l,d,b]
    W -= 0.5 * _mm(X, ovovx.reshape(o * v, o * v)).reshape(v, o, v)
    g = _e('lckd,d->lck', ovov, t1[i])                          # sum_d (lc|kd) t1[i,d] as [l,c,k]
    W -= _e('lck,lb->bkc', g, t1)
    return W


def _wvovo_a(S, t1, t2, a, Z1):
    """Wvovo[a,k,c,j] for fixed a as [j, k, c]."""
    o, v = S.o, S.v
    ovovx, ovoo = S.ovovx, S.ovoo
    W = S.oovv[:, :, a, :].transpose(1, 0, 2).copy()                  # (kj|ac) as [j,k,c]
    W += _e('pkj,pc->jkc', Z1, S.Bvv[:, a, :])
    W -= _e('lckj,l->jkc', ovoo, t1[:, a])
    X = t2[:, :, :, a].reshape(o, o * v)                               # [j, (l,d)] = t2[j,l,d,a]
    W -= 0.5 * _mm(X, ovovx.reshape(o * v, o * v)).reshape(o, o, v)
    g = _e('ldkc,l->dkc', ovovx, t1[:, a])                      # sum_l (lc|kd) t1[l,a] as [d,k,c]
    W -= _e('dkc,jd->jkc', g, t1)
    return W


def inc2_slice(S, t1, t2, I, i, a, Z1=None):
    """inc2[i, :, a, :] (rows j, columns b) of rccsd.increments."""
    o, v = S.o, S.v
    ovov, oovv, ovoo = S.ovov, S.oovv, S.ovoo
    tau = I['tau']
    P = S.naux
    if Z1 is None:
        Z1 = _e('pkd,jd->pkj', S.Bov, t1)
    r = np.zeros((o, v))
    # T1INT2B / T12INT2 (rccsd's first tmp2): tmp[i,j,a,b] + tmp[j,i,b,a]
    iacb = (S.Bvv.reshape(P, v * v).T @ S.Bov[:, i, a]).reshape(v, v)  # (ia|cb) as [c,b]
    m = _e('k,kbc->bc', t1[:, a], oovv[:, i, :, :])            # sum_k t1[k,a] (ki|bc)
    r += -(t1 @ m.T) + t1 @ iacb                                       # [j,b]
    X = _e('kjc,c->kj', oovv[:, :, a, :], t1[i])                # sum_c (kj|ac) t1[i,c]
    r += -(X.T @ t1)
    zc = S.Bvv[:, :, a] @ t1[i]                                        # [P]: sum_c B[P,c,a] t1[i,c]
    r += _e('pjb,p->jb', S.Bov, zc)
    # T1INT2A / T12INT2 (second tmp2): -(tmp[i,j,a,b] + tmp[j,i,b,a])
    M = ovov[:, :, i, a]                                               # (kc|ia) as [k,c]
    r -= (M @ t1.T).T @ t1                                             # sum_kc M[k,c] t1[j,c] t1[k,b]
    r -= ovoo[i, a] @ t1                                               # sum_k (ia|jk) t1[k,b] -> [j,b]
    u = np.outer(t1[:, a], t1[i]).reshape(o * v)
    r -= (u @ ovov.reshape(o * v, o * v)).reshape(o, v)                # sum_kc t1[k,a] t1[i,c] (kc|jb)
    r -= _e('jbk,k->jb', ovoo[:, :, i, :], t1[:, a])            # sum_k (jb|ik) t1[k,a]
    # <ij|ab>
    r += ovov[i, a]
    # hole ladder
    r += _e('klj,klb->jb', I['Woooo'][:, :, i, :], tau[:, :, a, :])
    # particle ladder: sum_cd Wvvvv[a,b,c,d] tau[i,j,c,d]
    G = S.vvvv_a(a)                                                     # (ac|bd) as [c,b,d]
    r += _e('cbd,jcd->jb', G, tau[i])
    H = S.ovvv_a(a)                                                     # (kd|ac) as [k,d,c]
    Xk = _e('kdc,jcd->kj', H, tau[i])
    r -= Xk.T @ t1
    yP = _e('k,pkc->pc', t1[:, a], S.Bov)                       # sum_k t1[k,a] B[P,k,c]
    Q = _mm(yP.T, S.Bvv.reshape(P, v * v)).reshape(v, v, v)            # [c,b,d]
    r -= _e('cbd,jcd->jb', Q, tau[i])
    # FEACONT
    r += _e('c,jcb->jb', I['Lvv_d'][a], t2[i])
    r += t2[:, i, :, a] @ I['Lvv_d'].T
    # FMICONT
    r -= _e('k,kjb->jb', I['Loo_d'][:, i], t2[:, :, a, :])
    r -= I['Loo_d'].T @ t2[:, i, :, a]
    # ring: tmp = 2 Wvoov.t2 - Wvovo.t2 (+ transpose), tmp2 = Wvoov.t2 (-), tmp3 = Wvovo.t2 (-)
    Wvoov_ai = _wvoov_slice(S, t1, t2, a, i)
    Wvovo_ai = _wvovo_slice(S, t1, t2, a, i)
    S1 = 2 * Wvoov_ai - Wvovo_ai
    r += _e('kc,kjcb->jb', S1, t2)
    y = t2[:, i, :, a]                                                 # [k,c] = t2[k,i,c,a]
    r += 2 * _wvoov_contract(S, t1, t2, y) - _wvovo_contract(S, t1, t2, y, Z1)
    r -= _e('kc,kjbc->jb', Wvoov_ai, t2)
    y2 = t2[:, i, a, :]                                                # [k,c] = t2[k,i,a,c]
    r -= _wvoov_contract(S, t1, t2, y2)
    Wi = _wvovo_i(S, t1, t2, i)                                        # [b,k,c]
    r -= _e('bkc,kjc->jb', Wi, t2[:, :, a, :])
    Wa = _wvovo_a(S, t1, t2, a, Z1)                                    # [j,k,c]
    r -= _e('jkc,kbc->jb', Wa, t2[:, i, :, :])
    return r


def increments(S, t1, t2, pairs):
    """(inc1, {(i, a): inc2[i, :, a, :]}) of the closed-shell CCSD increments of
    the system S (a DFSystem) at the amplitudes (t1, t2)."""
    t1 = np.asarray(t1, dtype=np.float64)
    t2 = np.asarray(t2, dtype=np.float64)
    I = _intermediates(S, t1, t2)
    inc1 = _inc1(S, t1, t2, I)
    Z1 = _e('pkd,jd->pkj', S.Bov, t1)
    out = {}
    for (i, a) in pairs:
        out[(i, a)] = inc2_slice(S, t1, t2, I, i, a, Z1)
    return inc1, out


def energy(S, t1, t2):
    """E = 2 f.t1 + sum (2(ia|jb) - (ib|ja)) (t_ijab + t_ia t_jb), as ``rccsd.energy``."""
    o = S.o
    e = 2.0 * float(np.sum(S.f[:o, o:] * t1))
    L = 2 * S.ovov - S.ovovx                                            # [i,a,j,b]
    for i in range(o):
        tau_i = t2[i] + np.outer(t1[i], t1.ravel()).reshape(S.v, o, S.v).transpose(1, 0, 2)  # [j,a,b]
        e += float(np.sum(L[i].transpose(1, 0, 2) * tau_i))
    return e


def mp1(S, i, a):
    """MP1 amplitudes t2[i, :, a, :] = (ia|jb) / D (``intprc/dfrmt2.F``)."""
    o = S.o
    e = np.diag(S.f)
    d = e[i] - e[o + a] + e[:o, None] - e[None, o:]
    return S.ovov[i, a] / d


def mp2_energy(S):
    """Closed-form MP2 energy sum (ia|jb)(2(ia|jb) - (ib|ja)) / D from the DF factors."""
    o = S.o
    e = np.diag(S.f)
    eo, ev = e[:o], e[o:]
    tot = 0.0
    for i in range(o):
        d = eo[i] - ev[:, None, None] + eo[None, :, None] - ev[None, None, :]     # [a,j,b]
        x = S.ovov[i]
        tot += float(np.sum(x * (2 * x - S.ovovx[i]) / d))
    return tot
