"""Differentiation form on the GPU: the engine's reverse pass (Lambda residual,
response density), the xlambda loop, and dE/dmu through a torch.autograd.Function
checked against central finite differences (h = 1e-4)."""
import numpy as np
import pytest

from oracle import lambda_, rccsd, vcc_loop


def _blocks(o, v, seed, nonhf=False, eri=None, f=None):
    from aces2 import synth
    if eri is None:
        eps, B = synth.draw(o, v, seed, sigma=synth.scaled_sigma(o, v))
        eri = synth.full_eri(B)
        f = synth.fock(eps)
        if nonhf:
            rng = np.random.default_rng(seed + 100)
            P = rng.normal(0, 0.01, f.shape)
            f = f + P + P.T
    blk = dict(rccsd.blocks(eri, o))
    blk['fock'] = f
    blk['list233'] = synth.list233_from_eri(eri, o)
    return f, eri, blk


def _rand(o, v, seed, s=0.05, sym=True):
    rng = np.random.default_rng(seed)
    x1 = rng.normal(0, s, (o, v))
    x2 = rng.normal(0, s, (o, o, v, v))
    if sym:
        x2 = 0.5 * (x2 + x2.transpose(1, 0, 3, 2))
    return x1, x2


@pytest.mark.gpu
@pytest.mark.parametrize('o,v,seed,nonhf,sym', [(3, 9, 5, False, True), (4, 12, 6, True, True),
                                               (5, 16, 7, True, False), (1, 6, 8, True, True)])
def test_adjoint_matches_oracle(o, v, seed, nonhf, sym):
    """r = dL/dT and dL/dfock from the HIP reverse pass vs reverse-mode autodiff
    of the oracle's increments, at random amplitudes and multipliers."""
    from aces2.ccsd import CCSD
    print('case adjoint o=%d v=%d seed=%d nonhf=%s sym=%s' % (o, v, seed, nonhf, sym), flush=True)
    f, eri, blk = _blocks(o, v, seed, nonhf)
    t1, t2 = _rand(o, v, seed + 1)
    l1, l2 = _rand(o, v, seed + 2, s=0.3, sym=sym)
    cc = CCSD(o, v, blk)
    cc.set_amplitudes(t1, t2)
    r1, r2, D = cc.adjoint(l1, l2, density=True)
    cc.close()
    g = {k: blk[k] for k in ('ovov', 'oovv', 'ovvo', 'ovoo', 'ovvv', 'oooo', 'vvvv')}
    d1, d2, df = lambda_.lagrangian_grads(t1, t2, l1, l2, f, g, o)
    d2 = 0.5 * (d2 + d2.transpose(1, 0, 3, 2))
    df = 0.5 * (df + df.T)
    sc = max(np.abs(d1).max(), np.abs(d2).max())
    np.testing.assert_allclose(r1.cpu().numpy(), d1, rtol=0, atol=1e-12 * sc)
    np.testing.assert_allclose(r2.cpu().numpy(), d2, rtol=0, atol=1e-12 * sc)
    np.testing.assert_allclose(D.cpu().numpy(), df, rtol=0, atol=1e-12 * max(1.0, np.abs(df).max()))


@pytest.mark.gpu
@pytest.mark.parametrize('nonhf', [False, True])
def test_lambda_solve_and_density_vs_oracle(nonhf):
    from aces2.ccsd import CCSD
    o, v, seed = 4, 16, 11
    print('case lambda solve o=%d v=%d nonhf=%s' % (o, v, nonhf), flush=True)
    f, eri, blk = _blocks(o, v, seed, nonhf)
    g = {k: blk[k] for k in ('ovov', 'oovv', 'ovvo', 'ovoo', 'ovvv', 'oooo', 'vvvv')}
    h = vcc_loop.run(f, g, o, maxcyc=200, conv=12, order=6)
    l1, l2, _ = lambda_.solve_lambda(h['t1'], h['t2'], f, g, o)
    Dref = lambda_.response_density(h['t1'], h['t2'], l1, l2, f, g, o)
    cc = CCSD(o, v, blk)
    cc.run(maxcyc=200, conv=12, order=6)
    lh = cc.solve_lambda(maxcyc=200, conv=12, order=6)
    g1, g2 = cc.get_lambda()
    D = cc.density().cpu().numpy()
    cc.close()
    assert lh['converged']
    np.testing.assert_allclose(g1.cpu().numpy(), l1, atol=1e-9)
    np.testing.assert_allclose(g2.cpu().numpy(), l2, atol=1e-9)
    np.testing.assert_allclose(D, Dref, atol=1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize('o,v,seed,nonhf', [(6, 30, 12, True), (7, 45, 13, False)])
def test_lambda_refine_matches_fp64(o, v, seed, nonhf):
    """Mixed Lambda cycles (fp64 anchors, the linear part on the difference to
    the anchor with its large products in bf16x3) converge to the fp64
    multipliers and density."""
    from aces2.ccsd import CCSD
    print('case lambda refine o=%d v=%d seed=%d nonhf=%s' % (o, v, seed, nonhf), flush=True)
    f, eri, blk = _blocks(o, v, seed, nonhf)
    cc = CCSD(o, v, blk)
    cc.run(maxcyc=200, conv=11, order=6)
    h64 = cc.solve_lambda(maxcyc=200, conv=10, order=6)
    a1, a2 = cc.get_lambda()
    Da = cc.density()
    hmx = cc.solve_lambda(maxcyc=200, conv=10, order=6, prec='refine')
    b1, b2 = cc.get_lambda()
    Db = cc.density()
    cc.close()
    assert h64['converged'] and hmx['converged']
    assert hmx['cycles'] <= h64['cycles'] + 4
    sc = max(float(a2.abs().max()), 1.0)
    assert float((a1 - b1).abs().max()) < 1e-8 * sc
    assert float((a2 - b2).abs().max()) < 1e-8 * sc
    assert float((Da - Db).abs().max()) < 1e-8


class _EnergyOfFock:
    """E_CCSD(F) at fixed integrals as a torch.autograd.Function: forward runs the
    GPU amplitude equations, backward multiplies by the response density from
    the GPU Lambda equations."""

    @staticmethod
    def make(o, v, blk, conv=11):
        import torch
        from aces2.ccsd import CCSD

        class Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, F):
                b = dict(blk)
                b['fock'] = F.detach().cpu().numpy()
                cc = CCSD(o, v, b)
                h = cc.run(maxcyc=300, conv=conv, order=6)
                ctx.save_for_backward(F)
                if F.requires_grad:
                    cc.solve_lambda(maxcyc=300, conv=conv, order=6)
                    ctx.dens = cc.density().to(F.device)
                cc.close()
                return torch.tensor(h['final_energy'], dtype=torch.float64, device=F.device)

            @staticmethod
            def backward(ctx, gE):
                return gE * ctx.dens

        return Fn


@pytest.mark.gpu
def test_autograd_fock_perturbation_vs_central_fd():
    import torch
    o, v, seed = 6, 24, 21
    print('case autograd fock o=%d v=%d' % (o, v), flush=True)
    f, eri, blk = _blocks(o, v, seed, nonhf=True)
    Fn = _EnergyOfFock.make(o, v, blk)
    rng = np.random.default_rng(3)
    P = rng.normal(0, 0.01, f.shape)
    P = torch.tensor(P + P.T, dtype=torch.float64, device='cuda')
    F0 = torch.tensor(f, dtype=torch.float64, device='cuda')
    mu = torch.zeros((), dtype=torch.float64, device='cuda', requires_grad=True)
    E = Fn.apply(F0 + mu * P)
    (g,) = torch.autograd.grad(E, mu)
    hh = 1e-4
    with torch.no_grad():
        fd = (Fn.apply(F0 + hh * P) - Fn.apply(F0 - hh * P)) / (2 * hh)
    assert abs(float(g) - float(fd)) < 1e-6 * max(1.0, abs(float(fd)))


@pytest.mark.gpu
def test_integral_perturbation_vs_central_fd():
    """Rank-1 perturbation (pq|rs) + mu x_pq x_rs: dE/dmu = E(T; 0, dg) + l.inc(T; 0, dg)
    (the residual is linear in the Hamiltonian), with l from the GPU Lambda solve."""
    import torch
    from aces2.ccsd import CCSD
    o, v, seed = 6, 24, 22
    print('case integral perturbation o=%d v=%d' % (o, v), flush=True)
    f, eri, blk = _blocks(o, v, seed)
    n = o + v
    rng = np.random.default_rng(4)
    x = rng.normal(0, 0.1, (n, n))
    x = x + x.T
    dg = np.einsum('pq,rs->pqrs', x, x)
    cc = CCSD(o, v, blk)
    cc.run(maxcyc=300, conv=11, order=6)
    cc.solve_lambda(maxcyc=300, conv=11, order=6)
    t1, t2 = cc.get_amplitudes()
    l1, l2 = cc.get_lambda()
    cc.close()
    _, _, bd = _blocks(o, v, 0, eri=dg, f=np.zeros_like(f))
    cd = CCSD(o, v, bd)
    cd.set_amplitudes(t1, t2)
    i1, i2 = cd.increments('fp64')
    an = cd.energy(t1, t2) + float(torch.sum(l1 * i1) + torch.sum(l2 * i2))
    cd.close()

    def ecc(e):
        _, _, b = _blocks(o, v, 0, eri=e, f=f)
        c = CCSD(o, v, b)
        r = c.run(maxcyc=300, conv=11, order=6)['final_energy']
        c.close()
        return r

    hh = 1e-4
    fd = (ecc(eri + hh * dg) - ecc(eri - hh * dg)) / (2 * hh)
    assert abs(an - fd) < 1e-6 * max(1.0, abs(fd))
