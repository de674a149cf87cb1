"""Lambda (adjoint) equations and response density from the CCSD Lagrangian.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

L(T, lam; f, g) = E(T; f, g) + sum_mu lam_mu Omega_mu(T; f, g),   Omega = inc - D*T
(the CC residual whose roots the amplitude equations solve).  The reference's
xlambda solves the same stationarity conditions (lambda/vlamcc.F; equations
in lambda/README:1-40, Gauss, Stanton, Bartlett, J. Chem. Phys. 95, 2623 (1991)):
  dL/dT = 0  ->  lam,         D^(1)_pq = dL/df_pq  (response one-particle density).
Here the Lagrangian is differentiated by reverse-mode automatic differentiation
(torch.autograd, float64, CPU) of a restatement of ``oracle.rccsd.increments``;
it is pinned by central finite differences of the converged CCSD energy.

Normalisation: lam2 is indexed like t2 (all i,j,a,b with lam2[i,j,a,b] =
lam2[j,i,b,a]); the Lagrangian sums over every array element.
"""
import numpy as np
import torch

from . import rccsd
from . import vcc_loop as _vl


def _ein(s, *a):
    return torch.einsum(s, *a)


def increments_t(t1, t2, f, g, o):
    """Torch restatement of oracle.rccsd.increments (same terms, same grouping)."""
    fov = f[:o, o:]
    foo = f[:o, :o]
    fvv = f[o:, o:]
    eo = torch.diagonal(foo)
    ev = torch.diagonal(fvv)
    ovov, oovv, ovvo = g['ovov'], g['oovv'], g['ovvo']
    ovoo, ovvv, oooo, vvvv = g['ovoo'], g['ovvv'], g['oooo'], g['vvvv']
    tau = t2 + _ein('ia,jb->ijab', t1, t1)
    Foo = foo + 2 * _ein('kcld,ilcd->ki', ovov, tau) - _ein('kdlc,ilcd->ki', ovov, tau)
    Fvv = fvv - 2 * _ein('kcld,klad->ac', ovov, tau) + _ein('kdlc,klad->ac', ovov, tau)
    Fov = fov + 2 * _ein('kcld,ld->kc', ovov, t1) - _ein('kdlc,ld->kc', ovov, t1)
    Loo = Foo + _ein('kc,ic->ki', fov, t1) + 2 * _ein('lcki,lc->ki', ovoo, t1) - _ein('kcli,lc->ki', ovoo, t1)
    Lvv = Fvv - _ein('kc,ka->ac', fov, t1) + 2 * _ein('kdac,kd->ac', ovvv, t1) - _ein('kcad,kd->ac', ovvv, t1)
    Foo_d = Foo - torch.diag(eo)
    Fvv_d = Fvv - torch.diag(ev)
    Loo_d = Loo - torch.diag(eo)
    Lvv_d = Lvv - torch.diag(ev)
    inc1 = fov.clone()
    inc1 = inc1 - 2 * _ein('kc,ka,ic->ia', fov, t1, t1)
    inc1 = inc1 + _ein('ac,ic->ia', Fvv_d, t1) - _ein('ki,ka->ia', Foo_d, t1)
    inc1 = inc1 + 2 * _ein('kc,kica->ia', Fov, t2) - _ein('kc,ikca->ia', Fov, t2)
    inc1 = inc1 + _ein('kc,ic,ka->ia', Fov, t1, t1)
    inc1 = inc1 + 2 * _ein('kcai,kc->ia', ovvo, t1) - _ein('kiac,kc->ia', oovv, t1)
    inc1 = inc1 + 2 * _ein('kdac,ikcd->ia', ovvv, tau) - _ein('kcad,ikcd->ia', ovvv, tau)
    inc1 = inc1 - 2 * _ein('lcki,klac->ia', ovoo, tau) + _ein('kcli,klac->ia', ovoo, tau)
    tmp2 = _ein('kibc,ka->abic', oovv, -t1) + ovvv.permute(1, 3, 0, 2)
    tmp = _ein('abic,jc->ijab', tmp2, t1)
    inc2 = tmp + tmp.permute(1, 0, 3, 2)
    tmp2 = _ein('kcai,jc->akij', ovvo, t1) + ovoo.permute(1, 3, 0, 2)
    tmp = _ein('akij,kb->ijab', tmp2, t1)
    inc2 = inc2 - (tmp + tmp.permute(1, 0, 3, 2))
    inc2 = inc2 + ovov.permute(0, 2, 1, 3)
    Woooo = (oooo.permute(0, 2, 1, 3) + _ein('lcki,jc->klij', ovoo, t1) + _ein('kclj,ic->klij', ovoo, t1)
             + _ein('kcld,ijcd->klij', ovov, tau))
    Wvvvv = vvvv.permute(0, 2, 1, 3) - _ein('kdac,kb->abcd', ovvv, t1) - _ein('kcbd,ka->abcd', ovvv, t1)
    Wvoov = (ovvo.permute(2, 0, 3, 1) + _ein('kcad,id->akic', ovvv, t1) - _ein('kcli,la->akic', ovoo, t1)
             - 0.5 * _ein('ldkc,ilda->akic', ovov, t2) - 0.5 * _ein('lckd,ilad->akic', ovov, t2)
             - _ein('ldkc,id,la->akic', ovov, t1, t1) + _ein('ldkc,ilad->akic', ovov, t2))
    Wvovo = (oovv.permute(2, 0, 3, 1) + _ein('kdac,id->akci', ovvv, t1) - _ein('lcki,la->akci', ovoo, t1)
             - 0.5 * _ein('lckd,ilda->akci', ovov, t2) - _ein('lckd,id,la->akci', ovov, t1, t1))
    inc2 = inc2 + _ein('klij,klab->ijab', Woooo, tau) + _ein('abcd,ijcd->ijab', Wvvvv, tau)
    tmp = _ein('ac,ijcb->ijab', Lvv_d, t2)
    inc2 = inc2 + tmp + tmp.permute(1, 0, 3, 2)
    tmp = _ein('ki,kjab->ijab', Loo_d, t2)
    inc2 = inc2 - (tmp + tmp.permute(1, 0, 3, 2))
    tmp = 2 * _ein('akic,kjcb->ijab', Wvoov, t2) - _ein('akci,kjcb->ijab', Wvovo, t2)
    inc2 = inc2 + tmp + tmp.permute(1, 0, 3, 2)
    tmp = _ein('akic,kjbc->ijab', Wvoov, t2)
    inc2 = inc2 - (tmp + tmp.permute(1, 0, 3, 2))
    tmp = _ein('bkci,kjac->ijab', Wvovo, t2)
    inc2 = inc2 - (tmp + tmp.permute(1, 0, 3, 2))
    return inc1, inc2


def energy_t(t1, t2, f, g, o):
    ovov = g['ovov']
    tau = t2 + _ein('ia,jb->ijab', t1, t1)
    L = 2 * ovov - ovov.permute(0, 3, 2, 1)
    return 2 * torch.sum(f[:o, o:] * t1) + _ein('iajb,ijab->', L, tau)


def omega_t(t1, t2, f, g, o):
    inc1, inc2 = increments_t(t1, t2, f, g, o)
    eo = torch.diagonal(f)[:o]
    ev = torch.diagonal(f)[o:]
    d1 = eo[:, None] - ev[None, :]
    d2 = d1[:, None, :, None] + d1[None, :, None, :]
    return inc1 - d1 * t1, inc2 - d2 * t2


def _tt(x):
    return torch.as_tensor(np.asarray(x), dtype=torch.float64)


def lagrangian_grads(t1, t2, l1, l2, f, g, o):
    """(dL/dt1, dL/dt2, dL/df) at the given amplitudes and multipliers."""
    T1 = _tt(t1).requires_grad_(True)
    T2 = _tt(t2).requires_grad_(True)
    F = _tt(f).requires_grad_(True)
    G = {k: _tt(v) for k, v in g.items()}
    o1, o2 = omega_t(T1, T2, F, G, o)
    L = energy_t(T1, T2, F, G, o) + torch.sum(_tt(l1) * o1) + torch.sum(_tt(l2) * o2)
    d1, d2, df = torch.autograd.grad(L, (T1, T2, F))
    return d1.numpy(), d2.numpy(), df.numpy()


def solve_lambda(t1, t2, f, g, o, tol=1e-12):
    """Solve the linear Lambda equations dL/dT(lam) = 0 by GMRES.

    The map lam -> dL/dT(lam) - dL/dT(0) is linear; it is applied as a
    reverse-mode product (one autograd pass per GMRES iteration) and
    preconditioned by the orbital-energy denominators.  lam2 is kept in the
    symmetric subspace lam2[i,j,a,b] = lam2[j,i,b,a].
    """
    from scipy.sparse.linalg import LinearOperator, gmres
    d1, d2 = rccsd.denominators(f, o)
    n1 = t1.size
    shp2 = t2.shape

    def resid(lv):
        l1 = lv[:n1].reshape(t1.shape)
        l2 = lv[n1:].reshape(shp2)
        r1, r2 = lagrangian_grads(t1, t2, l1, l2, f, g, o)[:2]
        r2 = 0.5 * (r2 + r2.transpose(1, 0, 3, 2))
        return np.concatenate([r1.ravel(), r2.ravel()])

    n = n1 + t2.size
    r0 = resid(np.zeros(n))
    A = LinearOperator((n, n), matvec=lambda x: resid(x) - r0, dtype=np.float64)
    dinv = 1.0 / np.concatenate([d1.ravel(), d2.ravel()])
    M = LinearOperator((n, n), matvec=lambda x: dinv * x, dtype=np.float64)
    x, info = gmres(A, -r0, M=M, rtol=tol, atol=0.0, restart=200, maxiter=50)
    if info != 0:
        raise RuntimeError('Lambda equations did not converge (gmres info %d)' % info)
    l1 = x[:n1].reshape(t1.shape)
    l2 = x[n1:].reshape(shp2)
    return l1, 0.5 * (l2 + l2.transpose(1, 0, 3, 2)), np.abs(resid(x)).max()


def response_density(t1, t2, l1, l2, f, g, o):
    """D_pq = dL/df_pq, symmetrised (the perturbations are symmetric)."""
    df = lagrangian_grads(t1, t2, l1, l2, f, g, o)[2]
    return 0.5 * (df + df.T)


def to_aces(l1, l2):
    """Spatial multipliers -> the reference's Lambda lists (spin-orbital normalisation).

    L = E + sum_mu Lambda_mu Omega_mu over the unique spin-orbital excitations
    (L1 lists 190 alpha and beta, L2 lists 144-146) equals the spatial
    E + sum l1*Omega1 + sum l2*Omega2 when l1 = 2 Lambda1 and
    l2 = 2 Lambda(Ij,Ab) - Lambda(Ji,Ab); inverted here ((2 - P)^-1 = (2 + P)/3).
    """
    return 0.5 * l1, (2.0 * l2 + l2.transpose(1, 0, 2, 3)) / 3.0


def from_aces(L1, L2):
    return 2.0 * L1, 2.0 * L2 - L2.transpose(1, 0, 2, 3)


def lambda_start(f, g, o, t1, t2):
    """The Jacobi update at (t1, t2) in the list normalisation (L1, L2).

    After convergence xvcc creates the Lambda lists (INIT2, ``vcc/init2.F``) and
    moves the increment lists into them with DRMOVE(...,100,SING1)
    (``vcc/vcc.F:1130-1143``, ``libr/drmove.f``).  Those lists hold the
    converged cycle's DIIS extrapolant (``vcc/dodiis0.F:181-233``), which
    ``vcc_loop.run`` records as ``lambda_start``; they hold this Jacobi update,
    T1_new = inc1/D1 and SYMETRIZE_T2(inc2/D2), only when the run converged in
    its first cycle (no extrapolation yet).
    """
    i1, i2 = rccsd.increments(t1, t2, f, g, o)
    d1, d2 = rccsd.denominators(f, o)
    return i1 / d1, symmetrize_pair(i2 / d2)


def run_reference(f, g, o, t1, t2, maxcyc=50, conv=7, order=5, log=None, start=None):
    """xlambda's CCSD Lambda iteration (``lambda/vlamcc.F:599-830``), restated.

    Start: ``start`` = (L1, L2), what xvcc leaves on the increment lists and
    moves to the Lambda lists (the converged cycle's DIIS extrapolant,
    ``vcc_loop.run(...)['lambda_start']``); None takes the Jacobi update at
    (t1, t2) (``lambda_start``).  With the Jacobi update instead of the
    extrapolant the printed trace is off by up to 1e-4 relative, with a zero
    Lambda1 by up to 5 %.  Each cycle forms the increments,
    divides by the denominators (NEWT2) and calls the Lambda DODIIS0
    (``lambda/dodiis0.F``): same DIIS vector and subspace as the amplitude
    equations, but converged when |RESMITER| or |RESMDIIS| is below
    10**-conv.  DODIIS0 extrapolates in the converged cycle too and xlambda
    moves the result to the Lambda lists unconditionally (DRMOVE after
    DODIIS0, ``lambda/vlamcc.F:811-815``, ``lambda/dodiis0.F:131-170``), so the
    multipliers returned are that extrapolant.  The Jacobi step in the reference's normalisation equals
    l <- l + (dL/dT)/D in the spatial one (the map ``to_aces`` is linear and
    commutes with the denominators), so the increments are the Lagrangian
    gradient; DIIS runs on the reference's vectors.
    """
    d1, d2 = rccsd.denominators(f, o)
    L1, L2 = lambda_start(f, g, o, t1, t2) if start is None else (start[0].copy(), start[1].copy())
    diis = _vl.DIIS(order)
    tol = 10.0 ** (-conv)
    hist = {'resmiter': [], 'resmdiis': [], 'converged': False}
    for icycle in range(1, maxcyc + 1):
        l1, l2 = from_aces(L1, L2)
        r1, r2 = lagrangian_grads(t1, t2, l1, l2, f, g, o)[:2]
        r2 = 0.5 * (r2 + r2.transpose(1, 0, 3, 2))
        N1, N2 = to_aces(l1 + r1 / d1, l2 + r2 / d2)
        N2 = symmetrize_pair(N2)
        L2 = symmetrize_pair(L2)
        told = _vl.diis_vector(L1, L2)
        tnew = _vl.diis_vector(N1, N2)
        resmiter, resmdiis, c, newvec = diis.step(icycle, told, tnew)
        hist['resmiter'].append(resmiter)
        hist['resmdiis'].append(resmdiis)
        if log:
            log(icycle, resmiter, resmdiis)
        done = abs(resmiter) < tol or abs(resmdiis) < tol
        if newvec is not None:
            L1, L2 = _vl.from_diis_vector(newvec, o, t1.shape[1])
        else:
            L1, L2 = N1, N2
        if done:
            hist['converged'] = True
            hist['cycles'] = icycle
            break
    hist['L1'], hist['L2'] = L1, L2
    hist['l1'], hist['l2'] = from_aces(L1, L2)
    return hist


def symmetrize_pair(x):
    return 0.5 * (x + x.transpose(1, 0, 3, 2))
