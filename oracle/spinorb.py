"""Spin-orbital CCSD residual in the Stanton-Gauss intermediate form.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

This is the general (UHF / spin-orbital) form of the equations that ACES II
``xvcc`` evaluates per spin case on its DPD lists:

* intermediates F(ae), F(mi), F(me), W(mnij), W(abef), W(mbej) as built by
  ``GENINT`` (``vcc/genint.F:1``; coefficients ``:52-67``; definitions in the
  header of ``vcc/drrng.f:1-128`` and ``vcc/genint.F:3-9``);
* the T2 increment assembled by ``FEACONT`` (``vcc/feacont.f``), ``FMICONT``
  (``vcc/fmicont.f``), ``T1INT2A/B`` (``vcc/t1int2a.f``, ``vcc/t1int2b.f``),
  ``DRE3EN`` (ladders ``vcc/drlad.f``, ring ``vcc/drrng.f``, ``T12INT2``);
* the T1 increment assembled by ``FEACONT``/``FMICONT``/``FMECONT``/``T1INT1``
  and ``E4S`` (``vcc/e4s.F``).

Published source of the equations: Stanton, Gauss, Watts, Bartlett,
J. Chem. Phys. 94, 4334 (1991), Table III.

"Increment" means the right-hand side that the Jacobi step divides by the
orbital-energy denominator (``NEWT2``, ``libr2/newt2.F:233-244``; ``E4S``):
    inc = Omega(T) + D * T,   D_ia = f_ii - f_aa,  D_ijab = f_ii+f_jj-f_aa-f_bb
so that Omega(T) = <mu| exp(-T) H exp(T) |0> is recovered as inc - D*T.

Spin-orbital ordering used throughout: [occ alpha (o), occ beta (o),
vir alpha (v), vir beta (v)].
"""
import numpy as np


def spin_orbital_hamiltonian(f, eri, o):
    """Spin-orbital Fock matrix and antisymmetrised <pq||rs>.

    ``f`` is the spatial Fock matrix (n x n), ``eri`` the spatial chemists'
    integrals (pq|rs) (n^4), ``o`` the number of doubly occupied orbitals.
    Returns (F, G, no) with G[p,q,r,s] = <pq||rs> = <pq|rs> - <pq|sr>,
    <pq|rs> = (pr|qs) delta(spin p, spin r) delta(spin q, spin s).
    """
    n = f.shape[0]
    v = n - o
    idx = np.concatenate([np.arange(o), np.arange(o), np.arange(o, n), np.arange(o, n)])
    spin = np.concatenate([np.zeros(o, int), np.ones(o, int), np.zeros(v, int), np.ones(v, int)])
    F = f[np.ix_(idx, idx)] * (spin[:, None] == spin[None, :])
    g = eri[np.ix_(idx, idx, idx, idx)]  # (pq|rs) over spin-orbital labels
    same = spin[:, None] == spin[None, :]
    # <pq|rs> = (pr|qs) d(sp,sr) d(sq,ss)
    phys = g.transpose(0, 2, 1, 3) * same[:, None, :, None] * same[None, :, None, :]
    G = phys - phys.transpose(0, 1, 3, 2)
    return F, G, 2 * o


def closed_to_so_amplitudes(t1, t2):
    """Spin-orbital amplitudes from closed-shell spatial t1[i,a], t2[i,j,a,b].

    t2 is the alpha-beta amplitude t_{Ij}^{Ab} (i->a, j->b); the same-spin
    amplitude is t2 - t2.swap(i,j) (``ABTOAA`` relation, ``libr2/newt2.F:277``).
    """
    o, v = t1.shape
    no, nv = 2 * o, 2 * v
    T1 = np.zeros((no, nv))
    T1[:o, :v] = t1
    T1[o:, v:] = t1
    T2 = np.zeros((no, no, nv, nv))
    aa = t2 - t2.transpose(1, 0, 2, 3)
    A, B = slice(0, o), slice(o, no)
    a, b = slice(0, v), slice(v, nv)
    T2[A, A, a, a] = aa
    T2[B, B, b, b] = aa
    T2[A, B, a, b] = t2
    T2[B, A, b, a] = t2.transpose(1, 0, 3, 2)
    T2[A, B, b, a] = -t2.transpose(0, 1, 3, 2)
    T2[B, A, a, b] = -t2.transpose(1, 0, 2, 3)
    return T1, T2


def increments(T1, T2, F, G, no):
    """Stanton-Gauss CCSD increments (inc1[i,a], inc2[i,j,a,b]).

    Term-by-term the same quantities the reference accumulates into the
    T1/T2 increment lists (90 groups 3-4, lists 61-63).
    """
    o = slice(0, no)
    v = slice(no, F.shape[0])
    fov = F[o, v]
    foo = F[o, o]
    fvv = F[v, v]
    oooo = G[o, o, o, o]
    ooov = G[o, o, o, v]
    oovv = G[o, o, v, v]
    ovvo = G[o, v, v, o]
    ovov = G[o, v, o, v]
    ovvv = G[o, v, v, v]
    vvvv = G[v, v, v, v]
    vvvo = G[v, v, v, o]
    ovoo = G[o, v, o, o]
    vovv = G[v, o, v, v]
    oovo = G[o, o, v, o]

    tau_t = T2 + 0.5 * (np.einsum('ia,jb->ijab', T1, T1) - np.einsum('ib,ja->ijab', T1, T1))
    tau = T2 + np.einsum('ia,jb->ijab', T1, T1) - np.einsum('ib,ja->ijab', T1, T1)

    nv = fvv.shape[0]
    nO = foo.shape[0]
    # F(ae): vcc/genint.F (QUAD3 ``vcc/quad3.f``, T1RING, MAKFME pieces)
    Fae = fvv * (1 - np.eye(nv))
    Fae -= 0.5 * np.einsum('me,ma->ae', fov, T1)
    Fae += np.einsum('mf,mafe->ae', T1, G[o, v, v, v])
    Fae -= 0.5 * np.einsum('mnaf,mnef->ae', tau_t, oovv)
    # F(mi): QUAD2 (``vcc/quad2.f``), T1INW2 (``vcc/t1inw2.f``)
    Fmi = foo * (1 - np.eye(nO))
    Fmi += 0.5 * np.einsum('me,ie->mi', fov, T1)
    Fmi += np.einsum('ne,mnie->mi', T1, ooov)
    Fmi += 0.5 * np.einsum('inef,mnef->mi', tau_t, oovv)
    # F(me): MAKFME (``vcc/makfme.F``)
    Fme = fov + np.einsum('nf,mnef->me', T1, oovv)
    # W(mnij): QUAD1/Q1ALL (``vcc/q1all.F``) + T1INW2
    Wmnij = oooo.copy()
    tmp = np.einsum('je,mnie->mnij', T1, ooov)
    Wmnij += tmp - tmp.transpose(0, 1, 3, 2)
    Wmnij += 0.25 * np.einsum('ijef,mnef->mnij', tau, oovv)
    # W(abef): the ladder list <ab||cd> dressed by T1 (``vcc/ladab1.F``,
    # ``vcc/t1inw1.f``) plus the quarter tau<mn||ef> term
    Wabef = vvvv.copy()
    tmp = np.einsum('mb,amef->abef', T1, G[v, o, v, v])
    Wabef -= tmp - tmp.transpose(1, 0, 2, 3)
    Wabef += 0.25 * np.einsum('mnab,mnef->abef', tau, oovv)
    # W(mbej): DWMBEJ (``vcc/dwmbej.F``) + T1RING (``libr/t1ring.F``)
    Wmbej = G[o, v, v, o].copy()
    Wmbej += np.einsum('jf,mbef->mbej', T1, ovvv)
    Wmbej -= np.einsum('nb,mnej->mbej', T1, oovo)
    Wmbej -= np.einsum('jnfb,mnef->mbej', 0.5 * T2 + np.einsum('jf,nb->jnfb', T1, T1), oovv)

    # ---- T1 increment (FEACONT/FMICONT/FMECONT/T1INT1/E4S) ----
    inc1 = fov.copy()
    inc1 += np.einsum('ie,ae->ia', T1, Fae)
    inc1 -= np.einsum('ma,mi->ia', T1, Fmi)
    inc1 += np.einsum('imae,me->ia', T2, Fme)
    inc1 -= np.einsum('nf,naif->ia', T1, G[o, v, o, v])
    inc1 -= 0.5 * np.einsum('imef,maef->ia', T2, G[o, v, v, v])
    inc1 -= 0.5 * np.einsum('mnae,nmei->ia', T2, G[o, o, v, o])

    # ---- T2 increment ----
    inc2 = oovv.copy()  # <ij||ab>  (INITIN, ``vcc/initin.f``)
    # FEACONT: P(ab) t_ij^ae (F_be - 1/2 t_m^b F_me)
    Fbe = Fae - 0.5 * np.einsum('mb,me->be', T1, Fme)
    tmp = np.einsum('ijae,be->ijab', T2, Fbe)
    inc2 += tmp - tmp.transpose(0, 1, 3, 2)
    # FMICONT: -P(ij) t_im^ab (F_mj + 1/2 t_j^e F_me)
    Fmj = Fmi + 0.5 * np.einsum('je,me->mj', T1, Fme)
    tmp = np.einsum('imab,mj->ijab', T2, Fmj)
    inc2 -= tmp - tmp.transpose(1, 0, 2, 3)
    # DRLAD: 1/2 tau_mn^ab W_mnij + 1/2 tau_ij^ef W_abef
    inc2 += 0.5 * np.einsum('mnab,mnij->ijab', tau, Wmnij)
    inc2 += 0.5 * np.einsum('ijef,abef->ijab', tau, Wabef)
    # DRRNG + T12INT2: P(ij)P(ab) (t_im^ae W_mbej - t_i^e t_m^a <mb||ej>)
    tmp = np.einsum('imae,mbej->ijab', T2, Wmbej)
    tmp -= np.einsum('ie,ma,mbej->ijab', T1, T1, G[o, v, v, o])
    inc2 += tmp - tmp.transpose(1, 0, 2, 3) - tmp.transpose(0, 1, 3, 2) + tmp.transpose(1, 0, 3, 2)
    # T1INT2B: P(ij) t_i^e <ab||ej>
    tmp = np.einsum('ie,abej->ijab', T1, vvvo)
    inc2 += tmp - tmp.transpose(1, 0, 2, 3)
    # T1INT2A: -P(ab) t_m^a <mb||ij>
    tmp = np.einsum('ma,mbij->ijab', T1, ovoo)
    inc2 -= tmp - tmp.transpose(0, 1, 3, 2)
    return inc1, inc2


def denominators(F, no):
    eo = np.diag(F)[:no]
    ev = np.diag(F)[no:]
    d1 = eo[:, None] - ev[None, :]
    d2 = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]
    return d1, d2


def omega(T1, T2, F, G, no):
    """CC residual <mu|exp(-T)H exp(T)|0> = increment - D*T."""
    inc1, inc2 = increments(T1, T2, F, G, no)
    d1, d2 = denominators(F, no)
    return inc1 - d1 * T1, inc2 - d2 * T2


def energy(T1, T2, F, G, no):
    """E_corr = f_ia t_ia + 1/4 <ij||ab> t_ijab + 1/2 <ij||ab> t_ia t_jb (``libr/cmpeng.F``)."""
    o = slice(0, no)
    v = slice(no, F.shape[0])
    oovv = G[o, o, v, v]
    e = np.einsum('ia,ia->', F[o, v], T1)
    e += 0.25 * np.einsum('ijab,ijab->', oovv, T2)
    e += 0.5 * np.einsum('ijab,ia,jb->', oovv, T1, T1)
    return float(e)
