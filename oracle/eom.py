"""EOM-CCSD singlet excitation energies of a closed-shell reference, from the
dense CC Jacobian.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference's xvee (vee/doeomee_david.F: HBARDIAG at :118 for the diagonal,
NEWGES at :257 for the guesses, HBARXC at :276 for the products Hbar C and
NEXTDAV at :296, vee/nextdav.F:336 printing the converged root) finds the
lowest eigenvalues of Hbar - E_CC on the singles and doubles space, one
symmetry block (ESTATE_SYM) at a time.  For converged amplitudes the Jacobian
of the CC residual Omega(T) = inc(T) - D T (``oracle.rccsd.increments``,
D = e_i + e_j - e_a - e_b) is that matrix, so its eigenvalues are the
excitation energies.

The Jacobian is built column by column by the complex step: Omega is a
polynomial with real coefficients, so Im Omega(t + i h e) / h is its exact
directional derivative for any h.  Coordinates: the t1 elements, then one per
pair-symmetric doubles orbit {t2[i,j,a,b], t2[j,i,b,a]} (the direction sets
both elements, the coordinate of a symmetric vector is its value at the orbit's
representative).  ``irrep`` (orbital irreps, occupied then virtual, XOR
products) and ``target`` keep the coordinates of one irrep.

Pinned by the two-electron case: there CCSD is exact and the S+D singlet space
is the whole singlet space apart from the reference, so the Jacobian spectrum
is the singlet FCI spectrum minus the ground state (``singlet_fci_levels``).
"""
import numpy as np

from . import rccsd


def coordinates(o, v, irrep=None, target=0):
    """(singles, doubles): index lists; singles (i, a); doubles orbit
    representatives (i, j, a, b) with (i, a) <= (j, b)."""
    ok = (lambda *orb: True) if irrep is None else \
        (lambda *orb: np.bitwise_xor.reduce([irrep[k] for k in orb]) == target)
    s = [(i, a) for i in range(o) for a in range(v) if ok(i, o + a)]
    d = []
    for i in range(o):
        for a in range(v):
            for j in range(o):
                for b in range(v):
                    if (i, a) <= (j, b) and ok(i, j, o + a, o + b):
                        d.append((i, j, a, b))
    return s, d


def residual(t1, t2, f, g, o):
    d1, d2 = rccsd.denominators(f, o)
    inc1, inc2 = rccsd.increments(t1, t2, f, g, o)
    return inc1 - d1 * t1, inc2 - d2 * t2


def jacobian(t1, t2, f, g, o, irrep=None, target=0, h=1e-20):
    """Dense Jacobian on the coordinates of ``coordinates``."""
    v = t1.shape[1]
    s, d = coordinates(o, v, irrep, target)
    n = len(s) + len(d)
    J = np.zeros((n, n))
    gc = {k: x.astype(np.complex128) for k, x in g.items()}
    fc = f.astype(np.complex128)
    for col in range(n):
        e1 = np.zeros_like(t1, dtype=np.complex128)
        e2 = np.zeros_like(t2, dtype=np.complex128)
        if col < len(s):
            i, a = s[col]
            e1[i, a] = 1.0
        else:
            i, j, a, b = d[col - len(s)]
            e2[i, j, a, b] = 1.0
            e2[j, i, b, a] = 1.0
        r1, r2 = residual(t1 + 1j * h * e1, t2 + 1j * h * e2, fc, gc, o)
        r1, r2 = r1.imag / h, r2.imag / h
        J[:len(s), col] = [r1[i, a] for (i, a) in s]
        J[len(s):, col] = [r2[i, j, a, b] for (i, j, a, b) in d]
    return J


def excitation_energies(t1, t2, f, g, o, irrep=None, target=0):
    """All eigenvalues of the Jacobian block, sorted by real part (complex
    ones kept as complex)."""
    w = np.linalg.eigvals(jacobian(t1, t2, f, g, o, irrep, target))
    return w[np.argsort(w.real)]


def singlet_fci_levels(f, eri):
    """Two-electron singlet FCI energies relative to the reference determinant
    (orbital 0 doubly occupied), ascending; same basis and Hamiltonian as
    ``oracle.fci.singlet_fci_corr``."""
    n = f.shape[0]
    h = f - 2.0 * eri[:, :, 0, 0] + eri[:, 0, 0, :]
    iu = np.triu_indices(n)
    m = len(iu[0])
    basis = np.zeros((m, n, n))
    for k, (p, q) in enumerate(zip(*iu)):
        if p == q:
            basis[k, p, p] = 1.0
        else:
            basis[k, p, q] = basis[k, q, p] = 1.0 / np.sqrt(2.0)
    HC = (np.einsum('pr,krq->kpq', h, basis) + np.einsum('qs,kps->kpq', h, basis)
          + np.einsum('prqs,krs->kpq', eri, basis))
    Hm = np.einsum('lpq,kpq->lk', basis, HC)
    e_hf = 2.0 * h[0, 0] + eri[0, 0, 0, 0]
    return np.linalg.eigvalsh(0.5 * (Hm + Hm.T)) - e_hf


def spin_orbital_excitation_energies(F, G, no, T1, T2, h=1e-20):
    """Eigenvalues of the spin-orbital CC Jacobian (``oracle.spinorb.omega``
    at amplitudes T1, T2) on the coordinates T1[i, a] and T2[i, j, a, b] with
    i < j, a < b: every EOM-CCSD state (singlets, triplets, all M_s) of the
    spin-orbital problem."""
    from . import spinorb
    nv = T1.shape[1]
    s = [(i, a) for i in range(no) for a in range(nv)]
    d = [(i, j, a, b) for i in range(no) for j in range(i + 1, no) for a in range(nv) for b in range(a + 1, nv)]
    n = len(s) + len(d)
    J = np.zeros((n, n))
    Fc, Gc = F.astype(np.complex128), G.astype(np.complex128)
    for col in range(n):
        e1 = np.zeros(T1.shape, dtype=np.complex128)
        e2 = np.zeros(T2.shape, dtype=np.complex128)
        if col < len(s):
            e1[s[col]] = 1.0
        else:
            i, j, a, b = d[col - len(s)]
            e2[i, j, a, b] = e2[j, i, b, a] = 1.0
            e2[j, i, a, b] = e2[i, j, b, a] = -1.0
        r1, r2 = spinorb.omega(T1 + 1j * h * e1, T2 + 1j * h * e2, Fc, Gc, no)
        J[:len(s), col] = [r1[p].imag / h for p in s]
        J[len(s):, col] = [r2[p].imag / h for p in d]
    w = np.linalg.eigvals(J)
    return w[np.argsort(w.real)]


def davidson(t1, t2, f, g, o, irrep=None, target=0, nroots=1, tol=1e-8, maxit=60, h=1e-20):
    """The nroots lowest roots of one symmetry block by a right-eigenvector
    Davidson iteration with matrix-free products J x = Im Omega(t + i h x) / h
    (one residual evaluation each), for systems whose dense Jacobian is too
    large to build; returns (omega, residual norms)."""
    v = t1.shape[1]
    d1, d2 = rccsd.denominators(f, o)
    if irrep is None:
        m1, m2 = np.ones_like(t1), np.ones_like(t2)
    else:
        ir = np.asarray(irrep)
        io, iv = ir[:o], ir[o:]
        m1 = ((io[:, None] ^ iv[None, :]) == target).astype(float)
        m2 = ((io[:, None, None, None] ^ io[None, :, None, None] ^ iv[None, None, :, None]
               ^ iv[None, None, None, :]) == target).astype(float)
    gc = {k: x.astype(np.complex128) for k, x in g.items()}
    fc = f.astype(np.complex128)
    n1 = t1.size

    def matvec(x):
        x1, x2 = x[:n1].reshape(t1.shape), x[n1:].reshape(t2.shape)
        r1, r2 = residual(t1 + 1j * h * x1, t2 + 1j * h * x2, fc, gc, o)
        return np.concatenate([(r1.imag / h * m1).ravel(), (r2.imag / h * m2).ravel()])

    delta = -np.concatenate([d1.ravel(), d2.ravel()])
    mask = np.concatenate([m1.ravel(), m2.ravel()])
    cand = np.argsort(np.where(mask[:n1] > 0, delta[:n1], np.inf))[:min(nroots + 3, int(mask[:n1].sum()))]
    V = []
    for c in cand:
        e = np.zeros(delta.size)
        e[c] = 1.0
        V.append(e)
    AV = [matvec(x) for x in V]
    for _ in range(maxit):
        Vm, AVm = np.array(V), np.array(AV)
        w, y = np.linalg.eig(Vm @ AVm.T)
        order = [q for q in np.argsort(w.real) if abs(w[q].imag) < 1e-10][:nroots]
        done, rn = True, []
        for q in order:
            yq = y[:, q].real
            x = yq @ Vm
            r = yq @ AVm - w[q].real * x
            rn.append(np.linalg.norm(r))
            if rn[-1] < tol:
                continue
            done = False
            den = w[q].real - delta
            den[np.abs(den) < 1e-3] = 1e-3
            c = mask * r / den
            c2 = c[n1:].reshape(t2.shape)
            c[n1:] = (0.5 * (c2 + c2.transpose(1, 0, 3, 2))).ravel()
            Vall = np.array(V)   # includes the corrections of this iteration
            for _ in range(2):
                c -= Vall.T @ (Vall @ c)
            nrm = np.linalg.norm(c)
            if nrm > 1e-12:
                V.append(c / nrm)
                AV.append(matvec(V[-1]))
        if done:
            return np.array([w[q].real for q in order]), np.array(rn)
    return np.array([w[q].real for q in order]), np.array(rn)
