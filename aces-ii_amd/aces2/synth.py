"""Synthetic closed-shell inputs for the CCSD amplitude path.

Workload definition (orbital energies, density-fitted integrals with exact
8-fold symmetry, seeds):
  occupied eps ~ U(-2.0, -0.4) sorted, virtual eps ~ U(0.1, 4.0) sorted,
  Fock matrix diagonal (optionally plus a symmetric perturbation),
  (pq|rs) = sum_P B[P,p,q] B[P,r,s],  B[P] symmetric, entries N(0, 0.05^2),
  n_aux = 4 (o + v).
Draw order from ``numpy.random.default_rng(seed)``: eps_occ, eps_vir, B.
"""
import numpy as np


def draw(o, v, seed, naux=None, sigma=0.05):
    """Orbital energies (n,) and symmetric factors B (naux, n, n), float64."""
    n = o + v
    naux = 4 * n if naux is None else naux
    rng = np.random.default_rng(seed)
    eo = np.sort(rng.uniform(-2.0, -0.4, o))
    ev = np.sort(rng.uniform(0.1, 4.0, v))
    X = rng.normal(0.0, sigma, (naux, n, n))
    iu = np.triu_indices(n, 1)
    X[:, iu[1], iu[0]] = X[:, iu[0], iu[1]]
    return np.concatenate([eo, ev]), X


def full_eri(B):
    """(pq|rs) for small n (n^4 memory)."""
    naux, n, _ = B.shape
    Bm = B.reshape(naux, n * n)
    return (Bm.T @ Bm).reshape(n, n, n, n)


def fock(eps, perturb=None):
    f = np.diag(eps)
    if perturb is not None:
        f = f + perturb
    return f


def scaled_sigma(o, v):
    """Entry scale that keeps the ring coupling o*v*sigma^2 at its o=5, v=19 value.

    With a size-independent sigma = 0.05 the direct-ring series of the CCSD
    equations grows like o*v*sigma^2/|D|; beyond o*v of a few hundred the
    iterations diverge (o=10, v=60 reaches inf/NaN).  sigma(o, v) =
    0.05*sqrt(95/(o*v)) keeps every size in the convergent regime of the
    smallest configuration while leaving the workload's shapes, sparsity and
    draw order unchanged.
    """
    return 0.05 * float(np.sqrt(95.0 / (o * v)))


def pair_le(v):
    """(p, q) index arrays of the 'PCK2' p <= q packing: index q*(q+1)/2 + p."""
    q = np.repeat(np.arange(v), np.arange(1, v + 1))
    p = np.concatenate([np.arange(k + 1) for k in range(v)])
    return p, q


def list233_from_eri(eri, o):
    """Reference list 233 for C1: rows A<=B ('PCK2'), columns Cd (C fastest), column major.

    Memory order is [d][c][pair(a<=b)] holding <ab|cd> = (ac|bd).
    """
    v = eri.shape[0] - o
    V = eri[o:, o:, o:, o:]            # (ac|bd) as V[a,c,b,d]
    p, q = pair_le(v)
    # out[d, c, k] = V[p_k, c, q_k, d]
    return np.ascontiguousarray(V[p, :, q, :].transpose(2, 1, 0))


def blocks_numpy(eps, B, o, perturb=None):
    """Integral blocks (chemists' notation) and list 233 on the host, small sizes."""
    eri = full_eri(B)
    f = fock(eps, perturb)
    O = slice(0, o)
    V = slice(o, eri.shape[0])
    return {
        'fock': f,
        'ovov': np.ascontiguousarray(eri[O, V, O, V]),
        'oovv': np.ascontiguousarray(eri[O, O, V, V]),
        'ovoo': np.ascontiguousarray(eri[O, V, O, O]),
        'oooo': np.ascontiguousarray(eri[O, O, O, O]),
        'ovvv': np.ascontiguousarray(eri[O, V, V, V]),
        'list233': list233_from_eri(eri, o),
    }, eri


def blocks_device(eps, B, o, perturb=None, device='cuda', weights=None, fock_matrix=None):
    """Integral blocks and list 233 built on the GPU from the factors B (any size).

    (pq|rs) = sum_P w_P B[P,p,q] B[P,r,s] (w = 1 unless ``weights`` is given)
    evaluated with fp64 matrix products; the vvvv list is assembled one row
    index a at a time so that only the A<=B compressed list (v^3 (v+1)/2
    doubles) is ever resident.  ``fock_matrix`` replaces diag(eps) + perturb.
    """
    import torch
    naux, n, _ = B.shape
    v = n - o
    Bd = torch.as_tensor(B, device=device, dtype=torch.float64)
    if weights is None:
        Wd = Bd
    else:
        Wd = Bd * torch.as_tensor(np.asarray(weights, dtype=np.float64), device=device)[:, None, None]
    Boo = Bd[:, :o, :o].reshape(naux, o * o)
    Bov = Bd[:, :o, o:].reshape(naux, o * v)
    Bvv = Bd[:, o:, o:].reshape(naux, v * v)
    Woo = Wd[:, :o, :o].reshape(naux, o * o)
    Wov = Wd[:, :o, o:].reshape(naux, o * v)
    out = {}
    f = fock(eps, perturb) if fock_matrix is None else fock_matrix
    out['fock'] = torch.as_tensor(f, device=device, dtype=torch.float64).contiguous()
    out['ovov'] = (Wov.T @ Bov).reshape(o, v, o, v).contiguous()
    out['oovv'] = (Woo.T @ Bvv).reshape(o, o, v, v).contiguous()
    out['ovoo'] = (Wov.T @ Boo).reshape(o, v, o, o).contiguous()
    out['oooo'] = (Woo.T @ Boo).reshape(o, o, o, o).contiguous()
    out['ovvv'] = (Wov.T @ Bvv).reshape(o, v, v, v).contiguous()
    nvp = v * (v + 1) // 2
    L = torch.empty((v, v, nvp), device=device, dtype=torch.float64)   # [d][c][pair]
    B3 = Wd[:, o:, o:]                                                    # [P][a][c]
    for a in range(v):
        # T[c][b][d] = (ac|bd) for b >= a
        Ta = (B3[:, a, :].T @ Bvv).reshape(v, v, v)                      # [c][b][d]
        bs = torch.arange(a, v, device=device)
        cols = bs * (bs + 1) // 2 + a
        L[:, :, cols] = Ta[:, a:, :].permute(2, 0, 1)                     # [d][c][b]
        del Ta
    out['list233'] = L
    del Bd, Wd, Boo, Bov, Bvv, Woo, Wov, B3
    return out


# ---------------------------------------------------------------------------
# Symmetry-blocked inputs (abelian groups, irrep product = XOR of labels)

def irrep_split(total, nirr, minimum, rng):
    """Per-irrep counts: `minimum` each plus a multinomial share of the rest."""
    extra = total - minimum * nirr
    if extra < 0:
        raise ValueError('too few orbitals for %d irreps with at least %d each' % (nirr, minimum))
    return [minimum + int(x) for x in rng.multinomial(extra, [1.0 / nirr] * nirr)]


def sym_draw(occ, vir, seed, naux=None, sigma=None):
    """Orbital energies and DF factors of a symmetric system.

    The per-irrep generator: occupied eps ~ U(-2.0, -0.4) and virtual eps ~
    U(0.1, 4.0), sorted inside each irrep; B[P] symmetric with entries
    N(0, sigma^2), auxiliary function P of irrep P mod nirr, and B[P][p][q]
    zeroed unless irrep(p) ^ irrep(q) = irrep(P), so (pq|rs) vanishes unless
    irrep(p)^irrep(q)^irrep(r)^irrep(s) = 0.  Orbitals are ordered occupied
    (irrep 0..nirr-1) then virtual (irrep 0..nirr-1).  Returns (eps, B, irreps).
    """
    nirr = len(occ)
    o, v = sum(occ), sum(vir)
    n = o + v
    naux = 4 * n if naux is None else naux
    sigma = scaled_sigma(o, v) if sigma is None else sigma
    rng = np.random.default_rng(seed)
    eo = np.concatenate([np.sort(rng.uniform(-2.0, -0.4, c)) for c in occ])
    ev = np.concatenate([np.sort(rng.uniform(0.1, 4.0, c)) for c in vir])
    irr = np.concatenate([np.repeat(np.arange(nirr), occ), np.repeat(np.arange(nirr), vir)]).astype(np.int64)
    X = rng.normal(0.0, sigma, (naux, n, n))
    iu = np.triu_indices(n, 1)
    X[:, iu[1], iu[0]] = X[:, iu[0], iu[1]]
    pair_irr = irr[:, None] ^ irr[None, :]
    aux_irr = np.arange(naux) % nirr
    X *= (pair_irr[None, :, :] == aux_irr[:, None, None])
    return np.concatenate([eo, ev]), X, irr


def sym_masks(irr, o):
    """Boolean masks of the symmetry-allowed t1 (o, v) and t2 (o, o, v, v) elements."""
    io, iv = irr[:o], irr[o:]
    m1 = (io[:, None] ^ iv[None, :]) == 0
    m2 = (io[:, None, None, None] ^ io[None, :, None, None] ^ iv[None, None, :, None] ^ iv[None, None, None, :]) == 0
    return m1, m2
