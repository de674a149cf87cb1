"""Closed-shell (RHF, spatial-orbital) CCSD increments.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference evaluates closed-shell CCSD on its alpha-beta lists only
(T(Ab,Ij), list 46; the same-spin amplitude is rebuilt by ``ABTOAA``,
``libr2/newt2.F:277-278``) with spin-adapted intermediates (e.g. list 56 =
2 W(MbEj) - W(MbeJ), ``vcc/dwmbej.F:500-505``).  This module is a
restatement of the same increments in spatial orbitals, written with the
intermediates Foo/Fvv/Fov, Loo/Lvv, Woooo, Wvvvv, Wvoov and Wvovo that
appear in the spin adaptation of the Stanton-Gauss equations (Hirata et al.,
J. Chem. Phys. 120, 2581 (2004) give the same closed-shell form).  Equality
with ``oracle.spinorb`` for arbitrary amplitudes is the test that pins it.

Conventions: t1[i,a]; t2[i,j,a,b] = t_{Ij}^{Ab} (i->a, j->b), so
t2[i,j,a,b] == t2[j,i,b,a].  ``eri`` is (pq|rs) over all n = o+v orbitals or
a dict of blocks (see ``blocks``).  Increments are the quantities that the
Jacobi step divides by D (``libr2/newt2.F:233-244``, ``vcc/e4s.F``):
inc1 = Omega1 + (f_ii - f_aa) t1, inc2 = Omega2 + (f_ii+f_jj-f_aa-f_bb) t2.
"""
import numpy as np


def blocks(eri, o):
    """Chemists' integral blocks (ia|jb) etc. from the full (pq|rs) tensor."""
    O = slice(0, o)
    V = slice(o, eri.shape[0])
    return {
        'oooo': np.ascontiguousarray(eri[O, O, O, O]),
        'ovoo': np.ascontiguousarray(eri[O, V, O, O]),
        'oovv': np.ascontiguousarray(eri[O, O, V, V]),
        'ovov': np.ascontiguousarray(eri[O, V, O, V]),
        'ovvo': np.ascontiguousarray(eri[O, V, V, O]),
        'ovvv': np.ascontiguousarray(eri[O, V, V, V]),
        'vvvv': np.ascontiguousarray(eri[V, V, V, V]),
    }


def _ein(s, *a):
    return np.einsum(s, *a, optimize=True)


def increments(t1, t2, f, g, o):
    """Closed-shell CCSD increments (inc1, inc2) for spatial f (n x n) and blocks g."""
    nocc, nvir = t1.shape
    fov = f[:o, o:]
    foo = f[:o, :o]
    fvv = f[o:, o:]
    eo = np.diag(foo)
    ev = np.diag(fvv)
    ovov, oovv, ovvo = g['ovov'], g['oovv'], g['ovvo']
    ovoo, ovvv, oooo, vvvv = g['ovoo'], g['ovvv'], g['oooo'], g['vvvv']

    tau = t2 + _ein('ia,jb->ijab', t1, t1)
    # F intermediates (QUAD2/QUAD3/MAKFME, T1RING pieces)
    Foo = foo + 2 * _ein('kcld,ilcd->ki', ovov, tau) - _ein('kdlc,ilcd->ki', ovov, tau)
    Fvv = fvv - 2 * _ein('kcld,klad->ac', ovov, tau) + _ein('kdlc,klad->ac', ovov, tau)
    Fov = fov + 2 * _ein('kcld,ld->kc', ovov, t1) - _ein('kdlc,ld->kc', ovov, t1)
    Loo = Foo + _ein('kc,ic->ki', fov, t1) + 2 * _ein('lcki,lc->ki', ovoo, t1) - _ein('kcli,lc->ki', ovoo, t1)
    Lvv = Fvv - _ein('kc,ka->ac', fov, t1) + 2 * _ein('kdac,kd->ac', ovvv, t1) - _ein('kcad,kd->ac', ovvv, t1)
    Foo_d = Foo - np.diag(eo)
    Fvv_d = Fvv - np.diag(ev)
    Loo_d = Loo - np.diag(eo)
    Lvv_d = Lvv - np.diag(ev)

    # ---- T1 increment ----
    inc1 = fov.copy()
    inc1 -= 2 * _ein('kc,ka,ic->ia', fov, t1, t1)
    inc1 += _ein('ac,ic->ia', Fvv_d, t1)
    inc1 -= _ein('ki,ka->ia', Foo_d, t1)
    inc1 += 2 * _ein('kc,kica->ia', Fov, t2)
    inc1 -= _ein('kc,ikca->ia', Fov, t2)
    inc1 += _ein('kc,ic,ka->ia', Fov, t1, t1)
    inc1 += 2 * _ein('kcai,kc->ia', ovvo, t1)
    inc1 -= _ein('kiac,kc->ia', oovv, t1)
    inc1 += 2 * _ein('kdac,ikcd->ia', ovvv, tau)
    inc1 -= _ein('kcad,ikcd->ia', ovvv, tau)
    inc1 -= 2 * _ein('lcki,klac->ia', ovoo, tau)
    inc1 += _ein('kcli,klac->ia', ovoo, tau)

    # ---- T2 increment ----
    # T1 x <ab||ej> / <mb||ij> pieces (T1INT2A/B, T12INT2)
    tmp2 = _ein('kibc,ka->abic', oovv, -t1) + ovvv.transpose(1, 3, 0, 2)
    tmp = _ein('abic,jc->ijab', tmp2, t1)
    inc2 = tmp + tmp.transpose(1, 0, 3, 2)
    tmp2 = _ein('kcai,jc->akij', ovvo, t1) + ovoo.transpose(1, 3, 0, 2)
    tmp = _ein('akij,kb->ijab', tmp2, t1)
    inc2 -= tmp + tmp.transpose(1, 0, 3, 2)
    inc2 += ovov.transpose(0, 2, 1, 3)  # <ij|ab> (INITIN)

    Woooo = (oooo.transpose(0, 2, 1, 3) + _ein('lcki,jc->klij', ovoo, t1)
             + _ein('kclj,ic->klij', ovoo, t1) + _ein('kcld,ijcd->klij', ovov, tau))
    Wvvvv = vvvv.transpose(0, 2, 1, 3) - _ein('kdac,kb->abcd', ovvv, t1) - _ein('kcbd,ka->abcd', ovvv, t1)
    Wvoov = (ovvo.transpose(2, 0, 3, 1) + _ein('kcad,id->akic', ovvv, t1) - _ein('kcli,la->akic', ovoo, t1)
             - 0.5 * _ein('ldkc,ilda->akic', ovov, t2) - 0.5 * _ein('lckd,ilad->akic', ovov, t2)
             - _ein('ldkc,id,la->akic', ovov, t1, t1) + _ein('ldkc,ilad->akic', ovov, t2))
    Wvovo = (oovv.transpose(2, 0, 3, 1) + _ein('kdac,id->akci', ovvv, t1) - _ein('lcki,la->akci', ovoo, t1)
             - 0.5 * _ein('lckd,ilda->akci', ovov, t2) - _ein('lckd,id,la->akci', ovov, t1, t1))

    # DRLAD: hole-hole and particle-particle ladders on tau
    inc2 += _ein('klij,klab->ijab', Woooo, tau)
    inc2 += _ein('abcd,ijcd->ijab', Wvvvv, tau)
    # FEACONT / FMICONT
    tmp = _ein('ac,ijcb->ijab', Lvv_d, t2)
    inc2 += tmp + tmp.transpose(1, 0, 3, 2)
    tmp = _ein('ki,kjab->ijab', Loo_d, t2)
    inc2 -= tmp + tmp.transpose(1, 0, 3, 2)
    # DRRNG: ring terms
    tmp = 2 * _ein('akic,kjcb->ijab', Wvoov, t2) - _ein('akci,kjcb->ijab', Wvovo, t2)
    inc2 += tmp + tmp.transpose(1, 0, 3, 2)
    tmp = _ein('akic,kjbc->ijab', Wvoov, t2)
    inc2 -= tmp + tmp.transpose(1, 0, 3, 2)
    tmp = _ein('bkci,kjac->ijab', Wvovo, t2)
    inc2 -= tmp + tmp.transpose(1, 0, 3, 2)
    # Diagonal Fock parts moved to the denominators: Loo_d/Lvv_d already
    # exclude f_ii/f_aa, and the D*T part is what NEWT2/E4S divide out.
    return inc1, inc2


def denominators(f, o):
    eo = np.diag(f)[:o]
    ev = np.diag(f)[o:]
    d1 = eo[:, None] - ev[None, :]
    d2 = d1[:, None, :, None] + d1[None, :, None, :]
    return d1, d2


def energy(t1, t2, f, g, o):
    """E = 2 f_ia t_ia + sum (2(ia|jb) - (ib|ja)) (t_ijab + t_ia t_jb).

    Equals the reference's 2 E(AA) + E(AB) (+ 2 f.t for non-HF), ``libr/cmpeng.F:102-180``.
    """
    ovov = g['ovov']
    tau = t2 + np.einsum('ia,jb->ijab', t1, t1)
    L = 2 * ovov - ovov.transpose(0, 3, 2, 1)
    e = 2 * np.einsum('ia,ia->', f[:o, o:], t1)
    e += np.einsum('iajb,ijab->', L, tau)
    return float(e)


def energy_parts(t1, t2, f, g, o):
    """CMPENG's spin-case pieces (``libr/cmpeng.F:99-180``, ``libr/tener.F``, ``libr/ftau.f``).

    Returns (E_AA, E_AB, E_lin) with
      E_AA  = sum_{A<B, I<J} tau(AB,IJ) <AB||IJ> + sum f(I,A) t(I,A)   (printed 'AA')
      E_AB  = sum tau(Ab,Ij) <Ab|Ij>                                   (printed 'AB')
      E_lin = ETOTT2, the same total 2 E_AA + E_AB with T2 in place of tau,
    where tau(AB,IJ) = T(AB,IJ) + t(I,A) t(J,B) - t(J,A) t(I,B), T(AB,IJ) =
    T(Ab,Ij) - T(Ab,Ji) (ABTOAA) and <AB||IJ> = (AI|BJ) - (AJ|BI).
    """
    ovov = g['ovov']
    iu_o = np.triu_indices(o, 1)
    v = t1.shape[1]
    iu_v = np.triu_indices(v, 1)
    W = ovov.transpose(0, 2, 1, 3)                       # <ab|ij> as [i,j,a,b] = (ia|jb)
    Waa = W - W.transpose(0, 1, 3, 2)
    Taa = t2 - t2.transpose(1, 0, 2, 3)
    ss = np.einsum('ia,jb->ijab', t1, t1)
    tau_aa = Taa + ss - ss.transpose(1, 0, 2, 3)
    pick = lambda x: x[iu_o[0], iu_o[1]][:, iu_v[0], iu_v[1]]
    e_aa_tau = float(np.sum(pick(tau_aa) * pick(Waa)))
    e_aa_lin = float(np.sum(pick(Taa) * pick(Waa)))
    e_ab_tau = float(np.sum((t2 + ss) * W))
    e_ab_lin = float(np.sum(t2 * W))
    esing = float(np.sum(f[:o, o:] * t1))
    return e_aa_tau + esing, e_ab_tau, 2.0 * (e_aa_lin + esing) + e_ab_lin
