"""Non-iterative triples (T) of CCSD(T), restated.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference evaluates the triples correction after the CC iterations in
``TRPS`` (``trp/trps.F:4``, called from ``vcc/vcc.F:1199-1221``) and prints
E[4]T ("E4T (T(CCSD))"), its same-spin piece E4TAAA and mixed piece E4TBAA
(RHF: E[4]T = 2 (E4TAAA + E4TBAA)), the singles term E[5]ST and
CCSD(T) = CCSD + E[4]T + E[5]ST (``test/OUTPUT/out.zmat.014a``).

Two restatements of the published equations:

spinorb_triples  the spin-orbital form (Crawford & Schaefer, Rev. Comp. Chem. 14,
                 33 (2000); Raghavachari, Trucks, Pople, Head-Gordon, CPL 157, 479
                 (1989)): with P(i/jk) f = f(ijk) - f(jik) - f(kji),
                   D t_c = P(i/jk) P(a/bc) [sum_e t_jk^ae <ei||bc> - sum_m t_im^bc <ma||jk>]
                   D t_d = P(i/jk) P(a/bc) t_i^a <jk||bc>
                   E[4]T = 1/36 sum t_c D t_c,   E[5]ST = 1/36 sum t_c D t_d.
                 Brute force over o^3 v^3 spin-orbital elements (small systems).
rccsd_triples    the closed-shell spatial-orbital form (Rendell, Lee, Komornicki,
                 CPL 178, 462 (1991)): with P the sum over the six simultaneous
                 permutations of the pairs (ia), (jb), (kc),
                   W_ijk^abc = P [sum_d (ia|bd) t_kj^cd - sum_l (kc|jl) t_il^ab]
                   V_ijk^abc = W_ijk^abc + (jb|kc) t_i^a + (ia|kc) t_j^b + (ia|jb) t_k^c
                   E(T) = 1/3 sum (4 W_abc + W_bca + W_cab - 2 W_acb - 2 W_bac - 2 W_cba) V_abc / D_ijk^abc
                 (E[4]T with V = W), and the same-spin piece from the abc-
                 antisymmetrised W, E4TAAA = 1/36 sum (A W)^2 / D.  Evaluated one
                 occupied triple i >= j >= k at a time (weights 6, 3, 1), the
                 algorithm the GPU implementation follows.
Canonical HF references only (diagonal Fock matrix, as the reference's TRPS for
an HF reference).  ``tests/test_oracle.py`` pins rccsd_triples against
spinorb_triples; ``tests/test_golden_h2o.py`` pins both against the reference's
printed (T) pieces.
"""
import numpy as np


def _p_occ(x):
    """P(i/jk) on the first three axes of a 6-index array."""
    return x - x.transpose(1, 0, 2, 3, 4, 5) - x.transpose(2, 1, 0, 3, 4, 5)


def _p_vir(x):
    return x - x.transpose(0, 1, 2, 4, 3, 5) - x.transpose(0, 1, 2, 5, 4, 3)


def spinorb_triples(T1, T2, F, G, no):
    """(E[4]T, E[5]ST, E4TAAA) from spin-orbital amplitudes and <pq||rs> (``oracle.spinorb``
    ordering: occupied alpha, occupied beta, virtual alpha, virtual beta)."""
    eo = np.diag(F)[:no]
    ev = np.diag(F)[no:]
    D = (eo[:, None, None, None, None, None] + eo[None, :, None, None, None, None] + eo[None, None, :, None, None, None]
         - ev[None, None, None, :, None, None] - ev[None, None, None, None, :, None]
         - ev[None, None, None, None, None, :])
    X = (np.einsum('jkae,eibc->ijkabc', T2, G[no:, :no, no:, no:], optimize=True)
         - np.einsum('imbc,majk->ijkabc', T2, G[:no, no:, :no, :no], optimize=True))
    tc = _p_vir(_p_occ(X)) / D
    del X
    td = _p_vir(_p_occ(np.einsum('ia,jkbc->ijkabc', T1, G[:no, :no, no:, no:], optimize=True))) / D
    e4 = float(np.sum(tc * D * tc)) / 36.0
    e5 = float(np.sum(tc * D * td)) / 36.0
    o, v = no // 2, (F.shape[0] - no) // 2
    A, a = slice(0, o), slice(0, v)
    e4aaa = float(np.sum((tc * D * tc)[A, A, A, a, a, a])) / 36.0
    return e4, e5, e4aaa


_PERMS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


def triple_energy(t1, t2, g, eo, ev, i, j, k):
    """(E[4]T, E[4]T + E[5]ST, E4TAAA) contributions of the ordered occupied triple
    (i, j, k), summed over all a, b, c (closed-shell form)."""
    ovvv, ovoo, ovov = g['ovvv'], g['ovoo'], g['ovov']
    occ = (i, j, k)
    v = t1.shape[1]
    W = np.zeros((v, v, v))
    for p in _PERMS:
        x, y, z = occ[p[0]], occ[p[1]], occ[p[2]]
        # base(x y z; p q r) = sum_d (xp|qd) t_zy^rd - sum_l (zr|yl) t_xl^pq as [p][q][r]
        base = (np.einsum('pqd,rd->pqr', ovvv[x], t2[z, y], optimize=True)
                - np.einsum('rl,lpq->pqr', ovoo[z, :, y, :], t2[x], optimize=True))
        # W[abc] += base[p q r] with (p, q, r) the virtual indices in positions p of (a, b, c)
        inv = np.argsort(p)
        W += base.transpose(inv)
    Vd = (np.einsum('bc,a->abc', ovov[j, :, k, :], t1[i]) + np.einsum('ac,b->abc', ovov[i, :, k, :], t1[j])
          + np.einsum('ab,c->abc', ovov[i, :, j, :], t1[k]))
    V = W + Vd
    D = eo[i] + eo[j] + eo[k] - ev[:, None, None] - ev[None, :, None] - ev[None, None, :]
    # Z = 4 W_abc + W_bca + W_cab - 2 W_acb - 2 W_bac - 2 W_cba is unchanged by a
    # simultaneous permutation of (ijk) and (abc), so each ordered triple gives the
    # same sum over abc and the loop runs over i >= j >= k with weights
    Z = (4 * W + W.transpose(1, 2, 0) + W.transpose(2, 0, 1) - 2 * W.transpose(0, 2, 1) - 2 * W.transpose(1, 0, 2)
         - 2 * W.transpose(2, 1, 0))
    e4 = float(np.sum(Z * W / D)) / 3.0
    et = float(np.sum(Z * V / D)) / 3.0
    A = (W - W.transpose(0, 2, 1) - W.transpose(1, 0, 2) - W.transpose(2, 1, 0) + W.transpose(1, 2, 0)
         + W.transpose(2, 0, 1))
    eaaa = float(np.sum(A * A / D)) / 36.0
    return e4, et, eaaa


def rccsd_triples(t1, t2, f, g, o, triples=None):
    """Closed-shell (T): dict with e4t, e5st, e4t_aaa, e4t_baa, et = e4t + e5st and
    the per-triple contributions {(i, j, k): (e4, e4 + e5, e4aaa)} for i >= j >= k
    (``triples`` restricts the evaluation to those triples; the totals are then
    partial sums)."""
    eo = np.diag(f)[:o]
    ev = np.diag(f)[o:]
    if triples is None:
        triples = [(i, j, k) for i in range(o) for j in range(i + 1) for k in range(j + 1)]
    per = {}
    e4 = et = eaaa = 0.0
    for (i, j, k) in triples:
        w = 6.0 if (i != j and j != k) else (1.0 if i == j == k else 3.0)
        c = triple_energy(t1, t2, g, eo, ev, i, j, k)
        per[(i, j, k)] = c
        e4 += w * c[0]
        et += w * c[1]
        eaaa += w * c[2]
    return {'e4t': e4, 'e5st': et - e4, 'et': et, 'e4t_aaa': eaaa, 'e4t_baa': 0.5 * e4 - eaaa, 'per_triple': per}
