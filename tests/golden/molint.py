"""Cartesian-Gaussian molecular integrals and RHF for the golden-vector checks.

TEST INFRASTRUCTURE ONLY.  Turns a molecule (geometry in bohr) and a GENBAS-
format basis into the MO-basis Fock matrix and (pq|rs), so that the
reference's own regression energies (test/zmat.015a TEST.DAT) can be compared
with the CCSD oracle and the GPU engine.

Integrals: McMurchie-Davidson (Hermite expansion coefficients E, Hermite
Coulomb integrals R from the Boys function), vectorised over primitives.
Contracted functions are normalised.  RHF: Roothaan-Hall with DIIS.
"""
import math

import numpy as np
from scipy.special import gammainc, gamma


def boys(n, T):
    T = np.asarray(T, dtype=float)
    out = np.empty_like(T)
    small = T < 1e-12
    out[small] = 1.0 / (2 * n + 1)
    Ts = T[~small]
    out[~small] = gamma(n + 0.5) * gammainc(n + 0.5, Ts) / (2.0 * Ts ** (n + 0.5))
    return out


def cart_components(l):
    return [(l - i - j, j, i) for i in range(l + 1) for j in range(l + 1 - i)][::-1] if l else [(0, 0, 0)]


def _cart(l):
    if l == 0:
        return [(0, 0, 0)]
    if l == 1:
        return [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
    if l == 2:
        return [(2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2)]
    raise ValueError('l > 2 not needed')


def dfact(n):
    return 1.0 if n <= 0 else float(np.prod(np.arange(n, 0, -2)))


class Shell:
    def __init__(self, center, l, exps, coefs):
        self.center = np.asarray(center, float)
        self.l = l
        self.exps = np.asarray(exps, float)
        self.comps = _cart(l)
        # primitive normalisation for each Cartesian component
        self.coefs = np.asarray(coefs, float)
        self.norm = np.zeros((len(self.comps), len(self.exps)))
        for ic, (lx, ly, lz) in enumerate(self.comps):
            N = ((2 * self.exps / np.pi) ** 0.75 * (4 * self.exps) ** (l / 2.0)
                 / math.sqrt(dfact(2 * lx - 1) * dfact(2 * ly - 1) * dfact(2 * lz - 1)))
            self.norm[ic] = N * self.coefs
        # contracted normalisation (per component)
        for ic, (lx, ly, lz) in enumerate(self.comps):
            s = 0.0
            for i, a in enumerate(self.exps):
                for j, b in enumerate(self.exps):
                    p = a + b
                    s += self.norm[ic, i] * self.norm[ic, j] * (np.pi / p) ** 1.5 * (
                        dfact(2 * lx - 1) * dfact(2 * ly - 1) * dfact(2 * lz - 1)) / (2 * p) ** l
            self.norm[ic] /= math.sqrt(s)


def hermite_E(la, lb, a, b, Ax, Bx):
    """E[i, j, t, prim_a, prim_b] Hermite coefficients for one Cartesian direction."""
    a = a[:, None]
    b = b[None, :]
    p = a + b
    mu = a * b / p
    Px = (a * Ax + b * Bx) / p
    XPA = Px - Ax
    XPB = Px - Bx
    E = np.zeros((la + 1, lb + 1, la + lb + 2) + p.shape)
    E[0, 0, 0] = np.exp(-mu * (Ax - Bx) ** 2)
    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            for t in range(i + j + 1):
                if i > 0:
                    v = XPA * E[i - 1, j, t]
                    if t > 0:
                        v = v + E[i - 1, j, t - 1] / (2 * p)
                    v = v + (t + 1) * E[i - 1, j, t + 1]
                else:
                    v = XPB * E[i, j - 1, t]
                    if t > 0:
                        v = v + E[i, j - 1, t - 1] / (2 * p)
                    v = v + (t + 1) * E[i, j - 1, t + 1]
                E[i, j, t] = v
    return E


def hermite_R(tmax, p, PC):
    """R[t, u, v, ...] Hermite Coulomb integrals (order 0) for exponent p and vector PC."""
    X, Y, Z = PC
    T = p * (X * X + Y * Y + Z * Z)
    nmax = tmax
    R = {}
    Fn = [boys(n, T) for n in range(nmax + 1)]
    # R^n_{000} = (-2p)^n F_n(T)
    base = {n: (-2 * p) ** n * Fn[n] for n in range(nmax + 1)}

    def rec(t, u, v, n):
        key = (t, u, v, n)
        if key in R:
            return R[key]
        if t < 0 or u < 0 or v < 0:
            return 0.0
        if t == u == v == 0:
            val = base[n]
        elif t > 0:
            val = (t - 1) * rec(t - 2, u, v, n + 1) + X * rec(t - 1, u, v, n + 1)
        elif u > 0:
            val = (u - 1) * rec(t, u - 2, v, n + 1) + Y * rec(t, u - 1, v, n + 1)
        else:
            val = (v - 1) * rec(t, u, v - 2, n + 1) + Z * rec(t, u, v - 1, n + 1)
        R[key] = val
        return val

    out = {}
    for t in range(tmax + 1):
        for u in range(tmax + 1 - t):
            for v in range(tmax + 1 - t - u):
                out[(t, u, v)] = rec(t, u, v, 0)
    return out


def _pair(sa, sb):
    Ex = hermite_E(sa.l, sb.l, sa.exps, sb.exps, sa.center[0], sb.center[0])
    Ey = hermite_E(sa.l, sb.l, sa.exps, sb.exps, sa.center[1], sb.center[1])
    Ez = hermite_E(sa.l, sb.l, sa.exps, sb.exps, sa.center[2], sb.center[2])
    p = sa.exps[:, None] + sb.exps[None, :]
    P = (sa.exps[:, None, None] * sa.center + sb.exps[None, :, None] * sb.center) / p[..., None]
    return Ex, Ey, Ez, p, P


def one_electron(shells, atoms):
    nbf = sum(len(s.comps) for s in shells)
    S = np.zeros((nbf, nbf))
    T = np.zeros((nbf, nbf))
    V = np.zeros((nbf, nbf))
    offs = np.cumsum([0] + [len(s.comps) for s in shells])
    for I, sa in enumerate(shells):
        for J, sb in enumerate(shells):
            Ex, Ey, Ez, p, P = _pair(sa, sb)
            # kinetic needs E with lb + 2
            Ex2 = hermite_E(sa.l, sb.l + 2, sa.exps, sb.exps, sa.center[0], sb.center[0])
            Ey2 = hermite_E(sa.l, sb.l + 2, sa.exps, sb.exps, sa.center[1], sb.center[1])
            Ez2 = hermite_E(sa.l, sb.l + 2, sa.exps, sb.exps, sa.center[2], sb.center[2])
            b = sb.exps[None, :]
            fac = (np.pi / p) ** 1.5
            Rs = [hermite_R(sa.l + sb.l, p, [P[..., k] - C[k] for k in range(3)]) for (Z, C) in atoms]
            for ia, (ax, ay, az) in enumerate(sa.comps):
                for ib, (bx, by, bz) in enumerate(sb.comps):
                    w = sa.norm[ia][:, None] * sb.norm[ib][None, :]
                    sx, sy, sz = Ex[ax, bx, 0], Ey[ay, by, 0], Ez[az, bz, 0]
                    S[offs[I] + ia, offs[J] + ib] = np.sum(w * sx * sy * sz * fac)

                    def kin1(E2, i, j):
                        return (j * (j - 1) * (E2[i, j - 2, 0] if j >= 2 else 0.0)
                                - 2 * b * (2 * j + 1) * E2[i, j, 0] + 4 * b * b * E2[i, j + 2, 0])
                    tx = kin1(Ex2, ax, bx) * sy * sz
                    ty = sx * kin1(Ey2, ay, by) * sz
                    tz = sx * sy * kin1(Ez2, az, bz)
                    T[offs[I] + ia, offs[J] + ib] = np.sum(w * (-0.5) * (tx + ty + tz) * fac)
                    v = 0.0
                    for (Z, C), R in zip(atoms, Rs):
                        acc = 0.0
                        for t in range(ax + bx + 1):
                            for u in range(ay + by + 1):
                                for vv in range(az + bz + 1):
                                    acc = acc + Ex[ax, bx, t] * Ey[ay, by, u] * Ez[az, bz, vv] * R[(t, u, vv)]
                        v = v - Z * acc
                    V[offs[I] + ia, offs[J] + ib] = np.sum(w * v * 2 * np.pi / p)
    return S, T, V


def two_electron(shells):
    nbf = sum(len(s.comps) for s in shells)
    offs = np.cumsum([0] + [len(s.comps) for s in shells])
    eri = np.zeros((nbf,) * 4)
    pairs = {}
    for I, sa in enumerate(shells):
        for J, sb in enumerate(shells):
            pairs[(I, J)] = _pair(sa, sb)
    ns = len(shells)
    for I in range(ns):
        for J in range(I + 1):
            ExA, EyA, EzA, p, P = pairs[(I, J)]
            sa, sb = shells[I], shells[J]
            for K in range(ns):
                for L in range(K + 1):
                    if (I * (I + 1) // 2 + J) < (K * (K + 1) // 2 + L):
                        continue
                    sc, sd = shells[K], shells[L]
                    ExC, EyC, EzC, q, Q = pairs[(K, L)]
                    pp = p[:, :, None, None]
                    qq = q[None, None, :, :]
                    alpha = pp * qq / (pp + qq)
                    PQ = [P[:, :, None, None, k] - Q[None, None, :, :, k] for k in range(3)]
                    ltot = sa.l + sb.l + sc.l + sd.l
                    R = hermite_R(ltot, alpha, PQ)
                    pref = 2 * np.pi ** 2.5 / (pp * qq * np.sqrt(pp + qq))
                    for ia, (ax, ay, az) in enumerate(sa.comps):
                        for ib, (bx, by, bz) in enumerate(sb.comps):
                            wab = sa.norm[ia][:, None] * sb.norm[ib][None, :]
                            for ic, (cx, cy, cz) in enumerate(sc.comps):
                                for id_, (dx, dy, dz) in enumerate(sd.comps):
                                    wcd = sc.norm[ic][:, None] * sd.norm[id_][None, :]
                                    acc = 0.0
                                    for t in range(ax + bx + 1):
                                        for u in range(ay + by + 1):
                                            for v in range(az + bz + 1):
                                                eab = (ExA[ax, bx, t] * EyA[ay, by, u] * EzA[az, bz, v])[:, :, None, None]
                                                for tau in range(cx + dx + 1):
                                                    for nu in range(cy + dy + 1):
                                                        for phi in range(cz + dz + 1):
                                                            ecd = (ExC[cx, dx, tau] * EyC[cy, dy, nu]
                                                                   * EzC[cz, dz, phi])[None, None, :, :]
                                                            sgn = -1.0 if (tau + nu + phi) % 2 else 1.0
                                                            acc = acc + sgn * eab * ecd * R[(t + tau, u + nu, v + phi)]
                                    val = np.sum(wab[:, :, None, None] * wcd[None, None, :, :] * pref * acc)
                                    a_, b_ = offs[I] + ia, offs[J] + ib
                                    c_, d_ = offs[K] + ic, offs[L] + id_
                                    for (w, x, y, z) in ((a_, b_, c_, d_), (b_, a_, c_, d_), (a_, b_, d_, c_),
                                                         (b_, a_, d_, c_), (c_, d_, a_, b_), (d_, c_, a_, b_),
                                                         (c_, d_, b_, a_), (d_, c_, b_, a_)):
                                        eri[w, x, y, z] = val
    return eri


def nuclear_repulsion(atoms):
    e = 0.0
    for i in range(len(atoms)):
        for j in range(i):
            e += atoms[i][0] * atoms[j][0] / np.linalg.norm(np.asarray(atoms[i][1]) - np.asarray(atoms[j][1]))
    return e


def reflection_labels(shells, C, S, axes=(0, 1, 2)):
    """Characters (+1/-1) of each MO under the reflections through the planes
    normal to the given axes (x -> -x etc.), for a molecule mapped onto itself
    by each of them: (sx, sy, sz) for D2h on the z axis, (sx, sy) for a C2v
    molecule in the yz plane.

    The reflection x -> -x multiplies a Cartesian function by (-1)^lx and maps
    its centre to the mirror image (which must be an atom of the same basis).
    """
    # a basis function is (centre, index of its shell among that centre's shells,
    # Cartesian component); mirror-image atoms carry the same shells in the same order
    funcs = []
    seen = {}
    for sh in shells:
        cen = tuple(np.round(sh.center, 10) + 0.0)
        ordinal = seen.get(cen, 0)
        seen[cen] = ordinal + 1
        for c in sh.comps:
            funcs.append((cen, ordinal, c))
    n = len(funcs)
    index = {f: k for k, f in enumerate(funcs)}
    assert len(index) == n
    ops = []
    for axis in axes:
        P = np.zeros((n, n))
        for i, (cen, ordinal, comp) in enumerate(funcs):
            img = list(cen)
            img[axis] = -img[axis]
            P[index[(tuple(np.asarray(img) + 0.0), ordinal, comp)], i] = (-1) ** comp[axis]
        ops.append(P)
    labels = []
    for k in range(C.shape[1]):
        labels.append(tuple(int(np.sign(np.round(C[:, k] @ S @ (P @ C[:, k]), 6))) for P in ops))
    return labels


def rhf(S, H, eri, nocc, conv=1e-12, maxiter=200, occ_rule=None):
    """Closed-shell Roothaan-Hall SCF with DIIS; returns (E_elec, C, eps, F_ao).

    occ_rule(C, eps) -> indices of the doubly occupied MOs (default: aufbau).
    """
    s, U = np.linalg.eigh(S)
    X = U @ np.diag(s ** -0.5) @ U.T
    F = H.copy()
    D = np.zeros_like(H)
    focks, errs = [], []
    e_old = 0.0
    for it in range(maxiter):
        Fp = X.T @ F @ X
        eps, Cp = np.linalg.eigh(Fp)
        C = X @ Cp
        occ = list(range(nocc)) if occ_rule is None else occ_rule(C, eps)
        Cocc = C[:, occ]
        D = Cocc @ Cocc.T
        J = np.einsum('pqrs,rs->pq', eri, D)
        K = np.einsum('prqs,rs->pq', eri, D)
        F = H + 2 * J - K
        e = np.sum(D * (H + F))
        err = F @ D @ S - S @ D @ F
        focks.append(F)
        errs.append(err)
        if len(focks) > 8:
            focks.pop(0)
            errs.pop(0)
        if len(focks) >= 2:
            n = len(focks)
            Bm = -np.ones((n + 1, n + 1))
            Bm[n, n] = 0
            for i in range(n):
                for j in range(n):
                    Bm[i, j] = np.sum(errs[i] * errs[j])
            rhs = np.zeros(n + 1)
            rhs[n] = -1
            c = np.linalg.solve(Bm, rhs)[:n]
            F = sum(ci * Fi for ci, Fi in zip(c, focks))
        if abs(e - e_old) < conv and np.abs(err).max() < 1e-9:
            break
        e_old = e
    Fp = X.T @ F @ X
    eps, Cp = np.linalg.eigh(Fp)
    C = X @ Cp
    occ = list(range(nocc)) if occ_rule is None else occ_rule(C, eps)
    virt = [k for k in range(C.shape[1]) if k not in occ]
    order = list(occ) + virt
    return e, C[:, order], eps[order], F


def read_genbas_entry(path, key):
    """Shells (l, exponents, coefficient matrix) of one GENBAS entry, e.g. 'O:STO-3G'."""
    with open(path) as f:
        lines = f.read().split('\n')
    i = [k for k, s in enumerate(lines) if s.strip() == key][0]
    toks = ' '.join(lines[i + 2:i + 400]).split()
    pos = 0
    nsh = int(toks[pos]); pos += 1
    ls = [int(x) for x in toks[pos:pos + nsh]]; pos += nsh
    ncon = [int(x) for x in toks[pos:pos + nsh]]; pos += nsh
    nprim = [int(x) for x in toks[pos:pos + nsh]]; pos += nsh
    shells = []
    for l, nc, npr in zip(ls, ncon, nprim):
        exps = [float(x) for x in toks[pos:pos + npr]]; pos += npr
        cm = np.array([float(x) for x in toks[pos:pos + npr * nc]]).reshape(npr, nc); pos += npr * nc
        shells.append((l, exps, cm))
    return shells


def build_shells(atoms_basis):
    """atoms_basis: list of (Z, center, [(l, exps, coef matrix), ...])."""
    shells = []
    for Z, C, blist in atoms_basis:
        for l, exps, cm in blist:
            for k in range(cm.shape[1]):
                nz = np.abs(cm[:, k]) > 0
                shells.append(Shell(C, l, np.asarray(exps)[nz], cm[nz, k]))
    return shells


def d2h_occupation(shells, S, counts):
    """occ_rule occupying the lowest MOs of each D2h irrep, counts = {(sx, sy, sz): n}."""
    def rule(C, eps):
        labels = reflection_labels(shells, C, S)
        occ = []
        for lab, nwant in counts.items():
            ks = [k for k in np.argsort(eps) if labels[k] == lab][:nwant]
            occ.extend(ks)
        return sorted(occ, key=lambda k: eps[k])
    return rule


def irrep_order(labels, o):
    """Irrep-sorted orbital order for the irrep-blocked engine.

    labels: per MO the characters under the symmetry generators; the irrep index
    is the bit vector (bit k set when the k-th character is -1), so that the
    irrep of a product is the XOR of the indices.  Returns (perm, occ, vir):
    occupied MOs (the first o) sorted by irrep, then the virtual ones, keeping
    the energy order inside each irrep, and the orbital counts per irrep.
    """
    irr = [sum(1 << k for k, c in enumerate(lab) if c < 0) for lab in labels]
    nirr = 1 << len(labels[0])
    occ_idx = sorted(range(o), key=lambda k: (irr[k], k))
    vir_idx = sorted(range(o, len(irr)), key=lambda k: (irr[k], k))
    occ = [sum(1 for k in range(o) if irr[k] == h) for h in range(nirr)]
    vir = [sum(1 for k in range(o, len(irr)) if irr[k] == h) for h in range(nirr)]
    return occ_idx + vir_idx, occ, vir


def mo_system(atoms_basis, nocc, occ_counts=None, axes=None):
    """(E_nuc, E_scf, MO Fock, MO (pq|rs), eps) for a closed-shell molecule; with
    ``axes`` also the MO symmetry labels under those reflections."""
    shells = build_shells(atoms_basis)
    atoms = [(Z, C) for Z, C, _ in atoms_basis]
    S, T, V = one_electron(shells, atoms)
    eri = two_electron(shells)
    H = T + V
    rule = d2h_occupation(shells, S, occ_counts) if occ_counts else None
    e_el, C, eps, F = rhf(S, H, eri, nocc, occ_rule=rule)
    enuc = nuclear_repulsion(atoms)
    Fmo = C.T @ F @ C
    eri_mo = np.einsum('pqrs,pi,qj,rk,sl->ijkl', eri, C, C, C, C, optimize=True)
    if axes is not None:
        return enuc, e_el + enuc, Fmo, eri_mo, eps, reflection_labels(shells, C, S, axes)
    return enuc, e_el + enuc, Fmo, eri_mo, eps


def water_atoms(basis, r_bohr, angle_rad):
    """Z-matrix H; O 1 R; H 2 R 1 A in a C2v frame: C2 axis along z (bisecting
    the HOH angle), molecular plane x = 0.  ``basis``: {'H': shells, 'O': shells}
    as returned by read_genbas_entry."""
    R, A = r_bohr, angle_rad
    atoms = [(1.0, [0.0, 0.0, R], basis['H']), (8.0, [0.0, 0.0, 0.0], basis['O']),
             (1.0, [0.0, R * np.sin(A), R * np.cos(A)], basis['H'])]
    c = np.array([a[1] for a in atoms])
    ez = (c[0] + c[2]) / 2 - c[1]
    ez /= np.linalg.norm(ez)
    ey = np.cross(ez, [1.0, 0.0, 0.0])
    ey /= np.linalg.norm(ey)
    ex = np.cross(ey, ez)
    Rm = np.array([ex, ey, ez])
    return [(Z, list(Rm @ np.asarray(C)), b) for (Z, C, b) in atoms]


def irrep_indices(labels):
    """Per MO the irrep index (bit k set when odd under generator k): the
    irrep of a product is the XOR of the indices."""
    return [sum(1 << k for k, c in enumerate(lab) if c < 0) for lab in labels]
