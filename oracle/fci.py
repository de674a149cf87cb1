"""Two-electron singlet FCI (pins the full nonlinear CCSD at o = 1).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

For two electrons CCSD is exact, so the converged CCSD correlation energy of
a closed-shell system with one doubly occupied orbital must equal the singlet
FCI correlation energy.  The singlet space is spanned by symmetric spatial
coefficient matrices C[p,q]; (H C)_pq = sum_r h_pr C_rq + sum_s h_qs C_ps
+ sum_rs (pr|qs) C_rs with the bare one-electron h recovered from the Fock
matrix of the reference (orbital 0 doubly occupied).
"""
import numpy as np


def singlet_fci_corr(f, eri):
    n = f.shape[0]
    h = f - 2.0 * eri[:, :, 0, 0] + eri[:, 0, 0, :]
    iu = np.triu_indices(n)
    m = len(iu[0])
    # orthonormal symmetric basis: E_pq = (e_p e_q^T + e_q e_p^T)/norm
    basis = np.zeros((m, n, n))
    for k, (p, q) in enumerate(zip(*iu)):
        if p == q:
            basis[k, p, p] = 1.0
        else:
            basis[k, p, q] = basis[k, q, p] = 1.0 / np.sqrt(2.0)
    HC = (np.einsum('pr,krq->kpq', h, basis) + np.einsum('qs,kps->kpq', h, basis)
          + np.einsum('prqs,krs->kpq', eri, basis))
    Hm = np.einsum('lpq,kpq->lk', basis, HC)
    w = np.linalg.eigvalsh(0.5 * (Hm + Hm.T))
    e_hf = 2.0 * h[0, 0] + eri[0, 0, 0, 0]
    return float(w[0] - e_hf)
