"""aces_gemm_grouped_noalloc (split-K partials in the library's static device
array) against aces_gemm_grouped with a caller workspace of the same size, on
two calls in a row, the second smaller, through the C ABI."""
import ctypes

import pytest

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu

SCRATCH_ELEMS = 1 << 22   # the entry point's documented maximum (32 MiB of doubles)


def _args(prec, groups):
    ng = len(groups)
    I = ctypes.c_int * ng
    Lg = ctypes.c_long * ng
    P = ctypes.c_void_p * ng
    return (prec, ng, I(*[g['M'] for g in groups]), I(*[g['N'] for g in groups]), I(*[g['K'] for g in groups]),
            P(*[g['A'].data_ptr() for g in groups]),
            P(*[g['Alo'].data_ptr() for g in groups]) if prec == 2 else None,
            Lg(*[g['A'].shape[-1] for g in groups]), 1,
            P(*[g['B'].data_ptr() for g in groups]),
            P(*[g['Blo'].data_ptr() for g in groups]) if prec == 2 else None,
            Lg(*[g['B'].shape[-1] for g in groups]), 1,
            P(*[g['C'].data_ptr() for g in groups]), Lg(*[g['C'].shape[-1] for g in groups]))


def _problem(prec, shapes, seed):
    from aces2 import _lib
    g = torch.Generator(device='cuda').manual_seed(seed)
    groups = []
    for (M, N, K) in shapes:
        Am = torch.randn(M, K, dtype=torch.float64, device='cuda', generator=g)
        Bm = torch.randn(N, K, dtype=torch.float64, device='cuda', generator=g)
        C0 = torch.randn(M, N, dtype=torch.float64, device='cuda', generator=g)
        gr = {'M': M, 'N': N, 'K': K, 'C0': C0}
        if prec == 0:
            gr['A'], gr['B'] = Am, Bm
        else:
            for nm, X in (('A', Am), ('B', Bm)):
                hi = torch.empty(X.shape, dtype=torch.int16, device='cuda')
                lo = torch.empty(X.shape, dtype=torch.int16, device='cuda')
                _lib.call('aces_split_bf16', X.shape[0], X.shape[1], X.data_ptr(), X.shape[1], hi.data_ptr(),
                          lo.data_ptr(), X.shape[1], None)
                gr[nm], gr[nm + 'lo'] = hi, lo
        groups.append(gr)
    return groups


def _run(prec, groups, ws_elems, noalloc):
    from aces2 import _lib
    for gr in groups:
        gr['C'] = gr['C0'].clone()
    a = _args(prec, groups)
    if noalloc:
        _lib.call('aces_gemm_grouped_noalloc', *a, 0.5, 0.25, ws_elems, None)
    else:
        ws = torch.empty(ws_elems, dtype=torch.float64, device='cuda')
        _lib.call('aces_gemm_grouped', *a, 0.5, 0.25, ws.data_ptr(), ws_elems, None)
    torch.cuda.synchronize()
    return [gr['C'].clone() for gr in groups]


@pytest.mark.parametrize('prec', [0, 2])
def test_noalloc_matches_workspace_entry_point(prec):
    print('case noalloc grouped prec=%d' % prec, flush=True)
    # skinny long-K blocks take the split-K path (partials in the scratch array)
    calls = [([(96, 80, 6144), (40, 64, 4096), (128, 128, 2048)], SCRATCH_ELEMS),
             ([(32, 48, 3072), (16, 16, 1024)], 1 << 16)]
    for k, (shapes, ws_elems) in enumerate(calls):
        groups = _problem(prec, shapes, 100 + 10 * k + prec)
        ref = _run(prec, groups, ws_elems, noalloc=False)
        got = _run(prec, groups, ws_elems, noalloc=True)
        for r, x in zip(ref, got):
            assert torch.equal(r, x)     # same kernels, same split plan, same order of additions
        # and the product itself
        if prec == 0:
            for gr, x in zip(groups, got):
                ex = 0.5 * (gr['A'] @ gr['B'].t()) + 0.25 * gr['C0']
                assert torch.allclose(x, ex, rtol=1e-12, atol=1e-11 * gr['K'] ** 0.5)


def test_noalloc_rejects_workspace_beyond_maximum():
    from aces2 import _lib
    print('case noalloc limit', flush=True)
    groups = _problem(0, [(8, 8, 2048)], 7)
    for gr in groups:
        gr['C'] = gr['C0'].clone()
    with pytest.raises(_lib.AcesError, match='exceeds the library'):
        _lib.call('aces_gemm_grouped_noalloc', *_args(0, groups), 1.0, 0.0, SCRATCH_ELEMS + 1, None)
