"""EOM-CCSD excitation energies (the reference's xvee EOMEE singlets, vee/).

Oracle pins (CPU): the dense Jacobian of the closed-shell CC residual at
converged amplitudes has the two-electron singlet FCI spectrum (CCSD is exact
there) and, for more electrons, every one of its eigenvalues is an eigenvalue of
the spin-orbital Jacobian (``oracle.spinorb``, itself pinned by brute force);
a residual with the orbital-energy term dropped fails the latter.  Golden
vectors: test/zmat.054a (H2O DZP at R = 0.957 A, A = 104.5 degrees, one root
in each C2v block, tests/golden/h2o_eom.json), matched by the oracle's
matrix-free Davidson and by the engine.  The reference's left-hand roots agree
with the exact eigenvalues to < 1e-8 a.u. (its right-hand ones, converged
more loosely, to 1.6e-7), so the checks are 3e-8 and 3e-7 respectively.

GPU: the engine's left-eigenvector Davidson (products J^T l by the reverse
pass) against the oracle's dense spectrum per irrep on synthetic systems whose
integrals carry a four-irrep symmetry, its left eigenvectors against the
oracle's Jacobian, and the zmat.054a roots.
"""
import json
import os
import sys

import numpy as np
import pytest

from oracle import eom, rccsd, spinorb, vcc_loop

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, 'golden'))
sys.path.insert(0, HERE)
import molint  # noqa: E402

EF = json.load(open(os.path.join(HERE, 'golden', 'h2o_eom.json')))
BASIS = json.load(open(os.path.join(HERE, 'golden', 'h2o_dzp.json')))['basis']
LEFT_TOL = 3e-8
RIGHT_TOL = 3e-7


def _ham(o, v, seed, irr=None, nonhf=True, scale=0.2):
    """Synthetic closed-shell Hamiltonian; with orbital irreps ``irr`` the
    integrals and Fock matrix are symmetric under the group (XOR products)."""
    from aces2 import synth
    rng = np.random.default_rng(seed)
    n = o + v
    B = rng.normal(0, scale, (4 * n, n, n))
    if irr is not None:
        ir = np.asarray(irr)
        aux = np.arange(4 * n) % 4
        B *= (aux[:, None, None] == (ir[None, :, None] ^ ir[None, None, :]))
    B = np.triu(B) + np.triu(B, 1).transpose(0, 2, 1)
    eri = synth.full_eri(B)
    f = np.diag(np.concatenate([np.sort(rng.uniform(-2, -0.4, o)), np.sort(rng.uniform(0.1, 4, v))]))
    if nonhf:
        P = rng.normal(0, 0.05, (n, n))
        if irr is not None:
            P *= (ir[:, None] == ir[None, :])
        f = f + P + P.T
    return f, eri


def _water():
    z = EF['zmatrix']
    bas = {el: [(b['l'], b['exps'], np.array(b['coefs'])) for b in BASIS[el]] for el in ('H', 'O')}
    atoms = molint.water_atoms(bas, z['R_angstrom'] / z['bohr_angstrom'], np.radians(z['A_degree']))
    enuc, escf, F, eri, eps, labels = molint.mo_system(atoms, EF['nocc'], axes=(0, 1))
    return enuc, escf, F, eri, molint.irrep_indices(labels)


@pytest.fixture(scope='module')
def water():
    return _water()


# ---------------------------------------------------------------------------- oracle pins
@pytest.mark.parametrize('v,seed', [(4, 1), (7, 3)])
def test_oracle_two_electron_spectrum_is_singlet_fci(v, seed):
    from aces2 import synth
    eps, B = synth.draw(1, v, seed, sigma=0.08)
    eri = synth.full_eri(B)
    f = synth.fock(eps)
    g = rccsd.blocks(eri, 1)
    h = vcc_loop.run(f, g, 1, maxcyc=100, conv=12, order=6)
    w = eom.excitation_energies(h['t1'], h['t2'], f, g, 1)
    lv = eom.singlet_fci_levels(f, eri)
    assert np.abs(w.imag).max() == 0.0
    np.testing.assert_allclose(np.sort(w.real), lv[1:] - lv[0], rtol=0, atol=1e-11)


def _so_distance(o, v, seed, nonhf, residual=None):
    f, eri = _ham(o, v, seed, nonhf=nonhf)
    g = rccsd.blocks(eri, o)
    h = vcc_loop.run(f, g, o, maxcyc=200, conv=12, order=6)
    assert h['converged']
    if residual is None:
        w = eom.excitation_energies(h['t1'], h['t2'], f, g, o)
    else:
        saved = eom.residual
        eom.residual = residual
        try:
            w = eom.excitation_energies(h['t1'], h['t2'], f, g, o)
        finally:
            eom.residual = saved
    F, G, no = spinorb.spin_orbital_hamiltonian(f, eri, o)
    T1, T2 = spinorb.closed_to_so_amplitudes(h['t1'], h['t2'])
    ws = eom.spin_orbital_excitation_energies(F, G, no, T1, T2)
    return max(np.min(np.abs(ws - x)) for x in w)


@pytest.mark.parametrize('o,v,seed,nonhf', [(2, 3, 4, True), (3, 3, 5, False)])
def test_oracle_singlets_are_spin_orbital_eigenvalues(o, v, seed, nonhf):
    assert _so_distance(o, v, seed, nonhf) < 1e-10


def test_oracle_check_catches_dropped_denominator_term():
    """Planted mistake: the Jacobian of the increments alone (no -D T term)."""
    bad = lambda t1, t2, f, g, o: rccsd.increments(t1, t2, f, g, o)  # noqa: E731
    assert _so_distance(2, 3, 4, True, residual=bad) > 1e-2


def test_oracle_davidson_matches_dense_spectrum():
    irr = [0, 1, 0, 2, 3, 1, 0, 2]
    o = 3
    f, eri = _ham(o, 5, 9, irr=irr)
    g = rccsd.blocks(eri, o)
    h = vcc_loop.run(f, g, o, maxcyc=200, conv=12, order=6)
    for tg in range(4):
        dense = eom.excitation_energies(h['t1'], h['t2'], f, g, o, irrep=irr, target=tg)
        real = np.sort(dense.real[np.abs(dense.imag) < 1e-10])
        w, rn = eom.davidson(h['t1'], h['t2'], f, g, o, irrep=irr, target=tg, nroots=2, tol=1e-9)
        np.testing.assert_allclose(w, real[:2], rtol=0, atol=1e-9)


def test_oracle_reproduces_reference_eom_roots(water):
    """zmat.054a: nuclear repulsion, SCF and CCSD energies, basis functions per
    irrep, then one root per symmetry block."""
    enuc, escf, F, eri, irr = water
    o = EF['nocc']
    assert abs(enuc - EF['nucrep']) < 1e-9
    assert abs(escf - EF['escf']) < 1e-9
    assert [irr.count(q) for q in range(4)] == EF['irrep_functions']
    assert [irr[:o].count(q) for q in range(4)] == EF['irrep_occupied']
    g = rccsd.blocks(eri, o)
    h = vcc_loop.run(F, g, o, maxcyc=60, conv=10, order=6)
    assert abs(h['final_energy'] + escf - EF['ccsd_total_energy']) < 1e-8
    for blk in EF['eom_blocks']:
        w, rn = eom.davidson(h['t1'], h['t2'], F, g, o, irrep=irr, target=blk['symmetry_block'] - 1, tol=1e-9)
        assert abs(w[0] - blk['left_from_total']) < LEFT_TOL
        assert abs(w[0] - blk['right_from_total']) < RIGHT_TOL


def test_small_eig_matches_numpy():
    """The projected-matrix eigensolver of the Davidson iteration (host code)."""
    import ctypes
    from aces2 import _lib
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 7, 24, 40):
        for kind in range(3):
            A = rng.normal(size=(n, n))
            if kind == 1:
                A = np.diag(np.arange(n, dtype=float)) + 0.05 * A   # near-diagonal, real spectrum
            elif kind == 2:
                A = A + A.T
            a = np.ascontiguousarray(A)
            wr = np.zeros(n)
            wi = np.zeros(n)
            _lib.call('aces_small_eig', n, a.ctypes.data, wr.ctypes.data, wi.ctypes.data)
            got = np.sort_complex(wr + 1j * wi)
            ref = np.sort_complex(np.linalg.eigvals(A))
            np.testing.assert_allclose(got, ref, rtol=0, atol=1e-10 * max(1.0, np.abs(ref).max()))
            for q in np.flatnonzero(np.abs(wi) == 0.0)[:3]:
                x = np.zeros(n)
                _lib.call('aces_small_eigvec', n, a.ctypes.data, ctypes.c_double(wr[q]),
                          x.ctypes.data)
                assert np.linalg.norm(A @ x - wr[q] * x) < 1e-8 * max(1.0, np.abs(A).max())


# ---------------------------------------------------------------------------- GPU
def _engine(o, v, f, eri):
    from aces2 import synth
    from aces2.ccsd import CCSD
    blk = dict(rccsd.blocks(eri, o))
    blk['fock'] = f
    blk['list233'] = synth.list233_from_eri(eri, o)
    cc = CCSD(o, v, blk)
    h = cc.run(maxcyc=200, conv=11, order=6)
    assert h['converged']
    return cc


def _rep_weights(o, v, s, d):
    """Left vectors in full (t1 | t2) storage -> the oracle's coordinates, with
    the orbit multiplicity (a left eigenvector u of the coordinate matrix J_c has
    u_k = (number of elements of orbit k) x element value)."""
    idx1 = [i * v + a for (i, a) in s]
    idx2 = [o * v + ((i * o + j) * v + a) * v + b for (i, j, a, b) in d]
    mult = np.array([1.0] * len(s) + [1.0 if (i, a) == (j, b) else 2.0 for (i, j, a, b) in d])
    return np.array(idx1 + idx2), mult


@pytest.mark.gpu
@pytest.mark.parametrize('o,v,seed,nroots', [(3, 5, 11, 3), (4, 7, 15, 2)])
def test_eom_matches_oracle_per_irrep(o, v, seed, nroots):
    print('case eom synthetic', o, v, seed, nroots, flush=True)
    rng = np.random.default_rng(seed)
    irr = [int(x) for x in rng.integers(0, 4, o + v)]
    f, eri = _ham(o, v, seed, irr=irr, scale=0.12)
    g = rccsd.blocks(eri, o)
    ref = vcc_loop.run(f, g, o, maxcyc=200, conv=12, order=6)
    cc = _engine(o, v, f, eri)
    for tg in range(4):
        s, d = eom.coordinates(o, v, irr, tg)
        J = eom.jacobian(ref['t1'], ref['t2'], f, g, o, irr, tg)
        wd = np.linalg.eigvals(J)
        real = np.sort(wd.real[np.abs(wd.imag) < 1e-10])
        if len(real) < nroots + 2:
            continue
        out = cc.eom(nroots=nroots, irrep=tg, orb_irrep=irr, tol=1e-8, maxit=200, vectors=True)
        assert max(out['resid']) < 1e-8
        np.testing.assert_allclose(out['omega'], real[:nroots], rtol=0, atol=1e-9)
        vecs = out['vectors'].cpu().numpy()
        idx, mult = _rep_weights(o, v, s, d)
        for r in range(nroots):
            u = vecs[r][idx] * mult
            assert np.linalg.norm(J.T @ u - out['omega'][r] * u) < 1e-7 * np.linalg.norm(u)
    cc.close()


@pytest.mark.gpu
def test_eom_c1_two_electrons_is_fci():
    from aces2 import synth
    print('case eom fci', flush=True)
    eps, B = synth.draw(1, 9, 21, sigma=0.08)
    eri = synth.full_eri(B)
    f = synth.fock(eps)
    cc = _engine(1, 9, f, eri)
    out = cc.eom(nroots=4, tol=1e-9, maxit=200)
    lv = eom.singlet_fci_levels(f, eri)
    np.testing.assert_allclose(out['omega'], (lv[1:] - lv[0])[:4], rtol=0, atol=1e-9)
    cc.close()


@pytest.mark.gpu
def test_eom_reproduces_reference_roots(water):
    print('case eom golden H2O zmat.054a', flush=True)
    enuc, escf, F, eri, irr = water
    o = EF['nocc']
    cc = _engine(o, F.shape[0] - o, F, eri)
    for blk in EF['eom_blocks']:
        out = cc.eom(nroots=1, irrep=blk['symmetry_block'] - 1, orb_irrep=irr, tol=1e-8)
        assert out['resid'][0] < 1e-8
        assert abs(out['omega'][0] - blk['left_from_total']) < LEFT_TOL
        assert abs(out['omega'][0] - blk['right_from_total']) < RIGHT_TOL
    cc.close()
