"""Pins of the CPU oracle (no GPU): brute force, spin adaptation, FCI, MP2, DIIS,
planted mistakes, the synthetic generator and the C-ABI symbol table."""
import os
import re
import types

import numpy as np
import pytest

from oracle import brute, fci, rccsd, spinorb, vcc_loop

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _ham(o, v, seed, nonhf=True, scale=0.3):
    from aces2 import synth
    rng = np.random.default_rng(seed)
    n = o + v
    B = rng.normal(0, scale, (3 * n, n, n))
    B = np.triu(B) + np.triu(B, 1).transpose(0, 2, 1)
    eri = synth.full_eri(B)
    f = np.diag(np.concatenate([np.sort(rng.uniform(-2, -0.4, o)), np.sort(rng.uniform(0.1, 4, v))]))
    if nonhf:
        P = rng.normal(0, 0.1, (n, n))
        f = f + P + P.T
    return f, eri


def _rand_closed_t(o, v, seed, s=0.1):
    rng = np.random.default_rng(seed)
    t1 = rng.normal(0, s, (o, v))
    t2 = rng.normal(0, s, (o, o, v, v))
    return t1, 0.5 * (t2 + t2.transpose(1, 0, 3, 2))


def _rand_so_t(no, nv, seed, s=0.1):
    """General (not spin-adapted) antisymmetric spin-orbital amplitudes."""
    rng = np.random.default_rng(seed)
    T1 = rng.normal(0, s, (no, nv))
    T2 = rng.normal(0, s, (no, no, nv, nv))
    T2 = T2 - T2.transpose(1, 0, 2, 3)
    T2 = T2 - T2.transpose(0, 1, 3, 2)
    return T1, T2


@pytest.mark.parametrize('o,v,seed', [(2, 2, 1), (3, 2, 2), (2, 3, 3)])
def test_spinorb_matches_brute_force(o, v, seed):
    """<mu| exp(-T) H exp(T) |0> by brute force vs the Stanton-Gauss residual, arbitrary T."""
    f, eri = _ham(o, v, seed)
    F, G, no = spinorb.spin_orbital_hamiltonian(f, eri, o)
    for T1, T2 in (spinorb.closed_to_so_amplitudes(*_rand_closed_t(o, v, seed + 10)),
                   _rand_so_t(2 * o, 2 * v, seed + 20)):
        e, om1, om2 = brute.projections(F, G, no, T1, T2)
        s1, s2 = spinorb.omega(T1, T2, F, G, no)
        assert abs(e - spinorb.energy(T1, T2, F, G, no)) < 1e-13
        np.testing.assert_allclose(s1, om1, atol=1e-13)
        np.testing.assert_allclose(s2, om2, atol=1e-13)


@pytest.mark.parametrize('o,v,seed,nonhf', [(2, 3, 4, True), (3, 4, 5, True), (4, 5, 6, False)])
def test_closed_shell_matches_spin_orbital(o, v, seed, nonhf):
    f, eri = _ham(o, v, seed, nonhf)
    F, G, no = spinorb.spin_orbital_hamiltonian(f, eri, o)
    t1, t2 = _rand_closed_t(o, v, seed + 1)
    T1, T2 = spinorb.closed_to_so_amplitudes(t1, t2)
    i1, i2 = spinorb.increments(T1, T2, F, G, no)
    g = rccsd.blocks(eri, o)
    r1, r2 = rccsd.increments(t1, t2, f, g, o)
    np.testing.assert_allclose(r1, i1[:o, :v], atol=1e-13)
    np.testing.assert_allclose(r2, i2[:o, o:, :v, v:], atol=1e-13)       # alpha-beta block
    np.testing.assert_allclose(r2 - r2.transpose(1, 0, 2, 3), i2[:o, :o, :v, :v], atol=1e-13)  # same spin
    assert abs(rccsd.energy(t1, t2, f, g, o) - spinorb.energy(T1, T2, F, G, no)) < 1e-13


@pytest.mark.parametrize('o,v,seed,nonhf', [(2, 3, 7, True), (3, 4, 8, False)])
def test_cmpeng_spin_cases_match_spin_orbital(o, v, seed, nonhf):
    """CMPENG's AA piece is the spin-orbital energy of the alpha orbitals alone;
    the AB piece is the rest (E = 2 E_AA + E_AB, E_BB = E_AA); ETOTT2 is that
    total with the t1 t1 part of tau removed."""
    f, eri = _ham(o, v, seed, nonhf)
    F, G, no = spinorb.spin_orbital_hamiltonian(f, eri, o)
    t1, t2 = _rand_closed_t(o, v, seed + 1)
    T1, T2 = spinorb.closed_to_so_amplitudes(t1, t2)
    g = rccsd.blocks(eri, o)
    e_aa, e_ab, e_lin = rccsd.energy_parts(t1, t2, f, g, o)
    n = o + v
    alpha = np.concatenate([np.arange(o), np.arange(2 * o, 2 * o + v)])
    e_alpha = spinorb.energy(T1[:o, :v], T2[:o, :o, :v, :v], F[np.ix_(alpha, alpha)],
                             G[np.ix_(alpha, alpha, alpha, alpha)], o)
    e_tot = spinorb.energy(T1, T2, F, G, no)
    assert abs(e_aa - e_alpha) < 1e-13
    assert abs(e_ab - (e_tot - 2 * e_alpha)) < 1e-13
    assert abs(e_lin - spinorb.energy(T1, T2, F, G, no) + 0.5 * np.einsum('ijab,ia,jb->', G[:no, :no, no:, no:], T1,
                                                                           T1)) < 1e-13
    assert n == f.shape[0]


@pytest.mark.parametrize('o,v,seed,nonhf', [(1, 3, 1, True), (3, 5, 2, True), (4, 7, 3, False), (5, 9, 4, True)])
def test_element_sampled_increments_match_rccsd(o, v, seed, nonhf):
    """The element-sampled oracle (re-associated terms, DF factors) equals
    rccsd.increments for every (i, a) slice and inc1, arbitrary amplitudes; its
    energy, MP1 and MP2 helpers equal rccsd's and the textbook closed form."""
    from aces2 import synth
    from oracle import slices
    eps, B = synth.draw(o, v, seed, sigma=0.1)
    rng = np.random.default_rng(seed)
    f0 = synth.fock(eps)
    f = f0
    if nonhf:
        P = rng.normal(0, 0.05, f.shape)
        f = f + P + P.T
    eri = synth.full_eri(B)
    g = rccsd.blocks(eri, o)
    t1, t2 = _rand_closed_t(o, v, seed + 3)
    r1, r2 = rccsd.increments(t1, t2, f, g, o)
    S = slices.DFSystem(B, f, o)
    pairs = [(i, a) for i in range(o) for a in range(v)]
    i1, sl = slices.increments(S, t1, t2, pairs)
    scale = np.abs(r2).max()
    np.testing.assert_allclose(i1, r1, atol=1e-14 * max(1.0, np.abs(r1).max()))
    for (i, a) in pairs:
        np.testing.assert_allclose(sl[(i, a)], r2[i, :, a, :], atol=1e-14 * scale)
    assert abs(slices.energy(S, t1, t2) - rccsd.energy(t1, t2, f, g, o)) < 1e-13
    S0 = slices.DFSystem(B, f0, o)
    m1, m2 = vcc_loop.mp1_start(f0, g, o)
    assert abs(slices.mp2_energy(S0) - rccsd.energy(m1, m2, f0, g, o)) < 1e-13
    for (i, a) in pairs:
        np.testing.assert_allclose(slices.mp1(S0, i, a), m2[i, :, a, :], atol=1e-15)


@pytest.mark.parametrize('o,v,seed', [(1, 4, 1), (2, 3, 2), (3, 4, 3), (2, 6, 4)])
def test_closed_shell_triples_match_spin_orbital(o, v, seed):
    """The closed-shell (T) (one occupied triple i >= j >= k at a time, weights
    6/3/1) equals the spin-orbital brute force: E[4]T, E[5]ST and the same-spin
    piece, arbitrary (pair-symmetric) amplitudes."""
    from oracle import triples
    f, eri = _ham(o, v, seed, nonhf=False)
    F, G, no = spinorb.spin_orbital_hamiltonian(f, eri, o)
    t1, t2 = _rand_closed_t(o, v, seed + 1)
    T1, T2 = spinorb.closed_to_so_amplitudes(t1, t2)
    e4, e5, eaaa = triples.spinorb_triples(T1, T2, F, G, no)
    r = triples.rccsd_triples(t1, t2, f, rccsd.blocks(eri, o), o)
    if o == 1:
        # two electrons have no triples: the closed-shell sums cancel to rounding
        assert abs(e4) == 0.0 and abs(r['e4t']) < 1e-14 and abs(r['e5st']) < 1e-14
        return
    scale = abs(e4)
    assert abs(r['e4t'] - e4) < 1e-13 * scale
    assert abs(r['e5st'] - e5) < 1e-13 * scale
    assert abs(r['e4t_aaa'] - eaaa) < 1e-13 * scale
    assert abs(2 * (r['e4t_aaa'] + r['e4t_baa']) - r['e4t']) < 1e-14 * scale


@pytest.mark.parametrize('v,seed', [(4, 1), (9, 2)])
def test_ccsd_equals_fci_for_two_electrons(v, seed):
    from aces2 import synth
    eps, B = synth.draw(1, v, seed, sigma=0.08)
    eri = synth.full_eri(B)
    f = synth.fock(eps)
    h = vcc_loop.run(f, rccsd.blocks(eri, 1), 1, maxcyc=100, conv=12, order=6)
    assert h['converged']
    assert abs(h['final_energy'] - fci.singlet_fci_corr(f, eri)) < 1e-11


def test_mp2_closed_form():
    from aces2 import synth
    o, v = 4, 12
    eps, B = synth.draw(o, v, 3, sigma=synth.scaled_sigma(o, v))
    eri = synth.full_eri(B)
    f = synth.fock(eps)
    g = rccsd.blocks(eri, o)
    ovov = g['ovov']
    d1, d2 = rccsd.denominators(f, o)
    e_mp2 = sum(ovov[i, a, j, b] * (2 * ovov[i, a, j, b] - ovov[i, b, j, a]) / d2[i, j, a, b]
                for i in range(o) for j in range(o) for a in range(v) for b in range(v))
    t1, t2 = vcc_loop.mp1_start(f, g, o)
    assert abs(rccsd.energy(t1, t2, f, g, o) - e_mp2) < 1e-13
    # from T = 0 the first Jacobi step reproduces MP1
    i1, i2 = rccsd.increments(np.zeros((o, v)), np.zeros((o, o, v, v)), f, g, o)
    np.testing.assert_allclose(i2 / d2, t2, atol=1e-15)
    np.testing.assert_allclose(i1, 0, atol=1e-15)


def test_diis_coefficients_textbook():
    """DODIIS solution = argmin |sum c_i e_i| s.t. sum c_i = 1 (Pulay 1980)."""
    rng = np.random.default_rng(0)
    E = rng.normal(size=(5, 40))
    R = E @ E.T
    c = vcc_loop.solve_diis(R)
    w = np.linalg.solve(R, np.ones(5))
    np.testing.assert_allclose(c, w / w.sum(), rtol=1e-10)
    assert abs(c.sum() - 1) < 1e-12


def test_diis_vector_layout():
    o, v = 3, 4
    t1, t2 = _rand_closed_t(o, v, 9)
    vec = vcc_loop.diis_vector(t1, t2)
    assert vec.size == o * v + o * o * v * v + (o * (o - 1) // 2) * (v * (v - 1) // 2)
    # list 46 position a + v b + v^2 (i + o j) holds T(Ab,Ij) = t2[i,j,a,b]
    i, j, a, b = 2, 1, 3, 0
    assert vec[o * v + a + v * b + v * v * (i + o * j)] == t2[i, j, a, b]
    # list 44 position pk(a<b) + nvp * pk(i<j) holds t2[i,j,a,b] - t2[j,i,a,b]
    i, j, a, b = 0, 2, 1, 3
    pk = lambda p, q: q * (q - 1) // 2 + p
    nvp = v * (v - 1) // 2
    assert abs(vec[o * v + o * o * v * v + pk(a, b) + nvp * pk(i, j)] - (t2[i, j, a, b] - t2[j, i, a, b])) < 1e-15
    t1b, t2b = vcc_loop.from_diis_vector(vec, o, v)
    np.testing.assert_array_equal(t1b, t1)
    np.testing.assert_array_equal(t2b, t2)


def _mutated_spinorb(old, new):
    src = open(spinorb.__file__).read()
    assert old in src
    mod = types.ModuleType('spinorb_mut')
    exec(compile(src.replace(old, new), 'spinorb_mut', 'exec'), mod.__dict__)
    return mod


@pytest.mark.parametrize('old,new', [
    ("Wmnij += 0.25 * np.einsum('ijef,mnef->mnij', tau, oovv)",
     "Wmnij += 0.5 * np.einsum('ijef,mnef->mnij', tau, oovv)"),
    ("inc1 -= 0.5 * np.einsum('mnae,nmei->ia', T2, G[o, o, v, o])",
     "inc1 += 0.5 * np.einsum('mnae,nmei->ia', T2, G[o, o, v, o])"),
    ("tmp = np.einsum('imae,mbej->ijab', T2, Wmbej)",
     "tmp = np.einsum('imae,mbej->ijab', T2, Wmbej.transpose(0, 2, 1, 3))"),
])
def test_planted_mistakes_are_caught(old, new):
    """A wrong factor, sign or index in the oracle fails the brute-force pin."""
    mut = _mutated_spinorb(old, new)
    f, eri = _ham(2, 3, 11)
    F, G, no = mut.spin_orbital_hamiltonian(f, eri, 2)
    T1, T2 = _rand_so_t(4, 6, 12)
    _, om1, om2 = brute.projections(F, G, no, T1, T2)
    s1, s2 = mut.omega(T1, T2, F, G, no)
    assert max(np.abs(s1 - om1).max(), np.abs(s2 - om2).max()) > 1e-6


def test_generator_symmetry_and_lists():
    from aces2 import synth
    o, v = 3, 7
    eps, B = synth.draw(o, v, 5, sigma=synth.scaled_sigma(o, v))
    eri = synth.full_eri(B)
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)]:
        np.testing.assert_allclose(eri, eri.transpose(perm), atol=1e-15)
    assert np.all(np.diff(eps[:o]) >= 0) and np.all(np.diff(eps[o:]) >= 0)
    assert eps[:o].max() < -0.4 + 1e-12 and eps[o:].min() > 0.1 - 1e-12
    L = synth.list233_from_eri(eri, o)
    p, q = synth.pair_le(v)
    for k in range(len(p)):
        a, b = p[k], q[k]
        for c in range(v):
            for d in range(v):
                assert L[d, c, k] == eri[o + a, o + c, o + b, o + d]


def test_c_abi_exports_every_declared_symbol():
    """The library loads without a GPU and exports every function in include/aces2_amd.h."""
    import ctypes
    from aces2 import _lib
    hdr = open(os.path.join(ROOT, 'include', 'aces2_amd.h')).read()
    names = set(re.findall(r'^\s*(?:const\s+)?\w+\**\s+\**(aces_\w+)\s*\(', hdr, re.M))
    assert len(names) > 40
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [nm for nm in sorted(names) if not hasattr(L, nm)]
    assert not missing, missing
    for nm in ('getlst_', 'putlst_', 'getall_', 'putall_', 'sumsym2_', 'inipck_', 'updmoi_', 'getrec_', 'putrec_',
               'genint_', 'dre3en_', 'e4s_', 'newt2_', 'cmpeng_', 'dodiis0_', 'drmove_'):
        assert hasattr(L, nm), nm


# ---------------------------------------------------------------------------
# Lambda / response density oracle (oracle/lambda_.py)

def _converged(o, v, seed, nonhf=False):
    from aces2 import synth
    eps, B = synth.draw(o, v, seed, sigma=synth.scaled_sigma(o, v))
    eri = synth.full_eri(B)
    f = synth.fock(eps)
    if nonhf:
        rng = np.random.default_rng(seed + 100)
        P = rng.normal(0, 0.01, f.shape)
        f = f + P + P.T
    g = rccsd.blocks(eri, o)
    h = vcc_loop.run(f, g, o, maxcyc=200, conv=12, order=6)
    assert h['converged']
    return f, eri, g, h


def test_lambda_torch_restatement_matches_numpy():
    import torch
    from oracle import lambda_
    f, eri = _ham(3, 5, 21)
    g = rccsd.blocks(eri, 3)
    t1, t2 = _rand_closed_t(3, 5, 22)
    G = {k: torch.tensor(x) for k, x in g.items()}
    i1, i2 = lambda_.increments_t(torch.tensor(t1), torch.tensor(t2), torch.tensor(f), G, 3)
    r1, r2 = rccsd.increments(t1, t2, f, g, 3)
    np.testing.assert_allclose(i1.numpy(), r1, atol=1e-14)
    np.testing.assert_allclose(i2.numpy(), r2, atol=1e-14)
    e = lambda_.energy_t(torch.tensor(t1), torch.tensor(t2), torch.tensor(f), G, 3)
    assert abs(float(e) - rccsd.energy(t1, t2, f, g, 3)) < 1e-14


def test_spatial_lambda_step_equals_spin_orbital_step():
    """The reference's Jacobi step on its spin-orbital lists equals l <- l + (dL/dT)/D.

    dL/dt(Ij,Ab) of the spin-orbital Lagrangian (AB amplitudes varied alone, by
    complex step on the brute-force-pinned spin-orbital residual) equals
    (2 + P(ij)) r2 / 3 of the spatial Lagrangian, and dL/dt(I,A) = r1 / 2.
    """
    from oracle import lambda_
    o, v = 2, 3
    f, eri = _ham(o, v, 3, nonhf=True)
    F, G, no = spinorb.spin_orbital_hamiltonian(f, eri, o)
    F, G = F.astype(complex), G.astype(complex)
    t1, t2 = _rand_closed_t(o, v, 4)
    L1, x = _rand_closed_t(o, v, 5)
    T1, T2 = spinorb.closed_to_so_amplitudes(t1, t2)
    LL1, LL2 = spinorb.closed_to_so_amplitudes(L1, x)

    def lag(T1c, T2c):
        om1, om2 = spinorb.omega(T1c, T2c, F, G, no)
        Fov, Goovv = F[:no, no:], G[:no, :no, no:, no:]
        E = (np.sum(Fov * T1c) + 0.25 * np.einsum('ijab,ijab->', Goovv, T2c)
             + 0.5 * np.einsum('ijab,ia,jb->', Goovv, T1c, T1c))
        return E + np.sum(LL1 * om1) + 0.25 * np.sum(LL2 * om2)

    h = 1e-20
    RAB = np.zeros((o, o, v, v))
    for i in range(o):
        for j in range(o):
            for a in range(v):
                for b in range(v):
                    P = np.zeros(T2.shape, complex)
                    P[i, o + j, a, v + b] += 1
                    P[o + j, i, v + b, a] += 1
                    P[i, o + j, v + b, a] -= 1
                    P[o + j, i, a, v + b] -= 1
                    RAB[i, j, a, b] = np.imag(lag(T1.astype(complex), T2 + 1j * h * P)) / h
    R1 = np.zeros((o, v))
    for i in range(o):
        for a in range(v):
            P = np.zeros(T1.shape, complex)
            P[i, a] = 1
            R1[i, a] = np.imag(lag(T1 + 1j * h * P, T2.astype(complex))) / h
    l1, l2 = lambda_.from_aces(L1, x)
    r1, r2 = lambda_.lagrangian_grads(t1, t2, l1, l2, f, rccsd.blocks(eri, o), o)[:2]
    r2 = 0.5 * (r2 + r2.transpose(1, 0, 3, 2))
    np.testing.assert_allclose(R1, r1 / 2, atol=1e-13)
    np.testing.assert_allclose(RAB, (2 * r2 + r2.transpose(1, 0, 2, 3)) / 3, atol=1e-13)
    # the two normalisations are inverse maps
    b1, b2 = lambda_.to_aces(*lambda_.from_aces(L1, x))
    np.testing.assert_allclose(b1, L1, atol=1e-15)
    np.testing.assert_allclose(b2, x, atol=1e-15)


@pytest.mark.parametrize('nonhf', [False, True])
def test_lambda_density_matches_central_fd(nonhf):
    """dE/dmu from (T, Lambda) vs central differences of the converged energy
    (h = 1e-4) for F + mu P and for a rank-1 integral perturbation."""
    from aces2 import synth
    from oracle import lambda_
    o, v = 3, 9
    f, eri, g, h = _converged(o, v, 5, nonhf)
    l1, l2, res = lambda_.solve_lambda(h['t1'], h['t2'], f, g, o)
    assert res < 1e-10
    D = lambda_.response_density(h['t1'], h['t2'], l1, l2, f, g, o)
    rng = np.random.default_rng(1)
    P = rng.normal(0, 0.01, f.shape)
    P = P + P.T
    n = o + v
    x = rng.normal(0, 0.1, (n, n))
    x = x + x.T
    dg = np.einsum('pq,rs->pqrs', x, x)

    def ecc(fm, em):
        return vcc_loop.run(fm, rccsd.blocks(em, o), o, maxcyc=200, conv=12, order=6)['final_energy']

    hh = 1e-4
    fd = (ecc(f + hh * P, eri) - ecc(f - hh * P, eri)) / (2 * hh)
    assert abs(np.sum(D * P) - fd) < 1e-8
    gd = rccsd.blocks(dg, o)
    z = np.zeros_like(f)
    i1, i2 = rccsd.increments(h['t1'], h['t2'], z, gd, o)
    an = rccsd.energy(h['t1'], h['t2'], z, gd, o) + np.sum(l1 * i1) + np.sum(l2 * i2)
    fd2 = (ecc(f, eri + hh * dg) - ecc(f, eri - hh * dg)) / (2 * hh)
    assert abs(an - fd2) < 1e-8
    # the trace of the correlation density vanishes (E_corr is invariant to f -> f + c I)
    assert abs(np.trace(D)) < 1e-12


def test_lambda_reference_iteration_reaches_the_linear_solution():
    from oracle import lambda_
    o, v = 3, 9
    f, eri, g, h = _converged(o, v, 7)
    l1, l2, _ = lambda_.solve_lambda(h['t1'], h['t2'], f, g, o)
    lh = lambda_.run_reference(f, g, o, h['t1'], h['t2'], maxcyc=100, conv=11, order=6)
    assert lh['converged']
    np.testing.assert_allclose(lh['l1'], l1, atol=1e-9)
    np.testing.assert_allclose(lh['l2'], l2, atol=1e-9)
