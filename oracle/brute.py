"""Second-quantised brute force for the CC residual on tiny systems.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Builds the N-electron determinant space of ``nso`` spin orbitals, the
operators E_pq = a_p^+ a_q in that space, the bare Hamiltonian
    H = sum_pq h_pq a_p^+ a_q + 1/4 sum_pqrs <pq||rs> a_p^+ a_q^+ a_s a_r
with h chosen so that the Fock matrix of the reference determinant equals the
given F, and evaluates  <mu| exp(-T) H exp(T) |0>  exactly (T is nilpotent, so
both exponential series terminate).  Conventions:
    T1 = sum t_ia a_a^+ a_i,   T2 = 1/4 sum t_ijab a_a^+ a_b^+ a_j a_i,
    |ia> = a_a^+ a_i |0>,     |ijab> = a_a^+ a_b^+ a_j a_i |0>.
Nothing here follows the reference's code: it is the definition of the CCSD
equations, used to pin ``oracle.spinorb`` (and through it everything else).
"""
import itertools
import numpy as np


class DetSpace:
    def __init__(self, nso, nel):
        self.nso = nso
        self.nel = nel
        self.dets = [sum(1 << p for p in c) for c in itertools.combinations(range(nso), nel)]
        self.index = {d: k for k, d in enumerate(self.dets)}
        self.dim = len(self.dets)
        # E[p][q] as dense matrices in the N-electron sector
        self.E = np.zeros((nso, nso, self.dim, self.dim))
        for k, d in enumerate(self.dets):
            for q in range(nso):
                if not (d >> q) & 1:
                    continue
                d1 = d & ~(1 << q)
                s1 = (-1) ** bin(d1 & ((1 << q) - 1)).count('1')
                for p in range(nso):
                    if (d1 >> p) & 1:
                        continue
                    d2 = d1 | (1 << p)
                    s2 = (-1) ** bin(d1 & ((1 << p) - 1)).count('1')
                    self.E[p, q, self.index[d2], k] += s1 * s2
        self.ref = self.index[(1 << nel) - 1]

    def vacuum(self):
        x = np.zeros(self.dim)
        x[self.ref] = 1.0
        return x


def bare_one_body(F, G, no):
    """h_pq = f_pq - sum_k <pk||qk> (k over the reference's occupied set)."""
    return F - np.einsum('pkqk->pq', G[:, :no, :, :no])


def hamiltonian(space, F, G, no):
    h = bare_one_body(F, G, no)
    E = space.E
    H = np.einsum('pq,pqxy->xy', h, E)
    # a_p^+ a_q^+ a_s a_r = E_pr E_qs - delta_qr E_ps
    M = np.einsum('pqrs,qsxy->prxy', G, E)
    H += 0.25 * np.einsum('prxy,pryz->xz', E, M)
    H -= 0.25 * np.einsum('pqqs,psxy->xy', G, E)
    return H


def excitation_operator(space, T1, T2, no):
    E = space.E
    nv = T1.shape[1]
    vi = np.arange(no, no + nv)
    T = np.einsum('ia,aixy->xy', T1, E[np.ix_(vi, range(no))])
    Eai = E[np.ix_(vi, range(no))]  # [a, i]
    # T2 = 1/4 t_ijab E_ai E_bj
    X = np.einsum('ijab,bjxy->aixy', T2, Eai)
    T += 0.25 * np.einsum('aixy,aiyz->xz', Eai, X)
    return T


def _expm_nilpotent(A, x, nmax):
    out = x.copy()
    term = x.copy()
    for k in range(1, nmax + 1):
        term = A @ term / k
        out = out + term
    return out


def projections(F, G, no, T1, T2):
    """Return (E_corr, Omega1[i,a], Omega2[i,j,a,b]) by brute force."""
    nso = F.shape[0]
    space = DetSpace(nso, no)
    H = hamiltonian(space, F, G, no)
    T = excitation_operator(space, T1, T2, no)
    phi = space.vacuum()
    x = _expm_nilpotent(T, phi, no + 1)
    y = H @ x
    z = _expm_nilpotent(-T, y, no + 1)
    e_ref = phi @ H @ phi
    ecorr = z[space.ref] - e_ref
    nv = nso - no
    E = space.E
    om1 = np.zeros((no, nv))
    om2 = np.zeros((no, no, nv, nv))
    for i in range(no):
        for a in range(nv):
            bra = E[no + a, i] @ phi
            om1[i, a] = bra @ z
    for i in range(no):
        for j in range(no):
            for a in range(nv):
                for b in range(nv):
                    bra = E[no + a, i] @ (E[no + b, j] @ phi)
                    om2[i, j, a, b] = bra @ z
    return float(ecorr), om1, om2
