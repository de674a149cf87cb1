"""MFMA GEMM engine vs a torch fp64 reference (GPU, through the C ABI)."""
import pytest

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu


def _gemm(prec, A, B, C, alpha, beta, a_kc, b_kc, A_lo=None, B_lo=None, batch=1, sA=0, sB=0, sC=0,
          M=None, N=None, K=None):
    from aces2 import _lib
    lda = A.shape[-1]
    ldb = B.shape[-1]
    _lib.call('aces_gemm', prec, M, N, K, A.data_ptr(), A_lo.data_ptr() if A_lo is not None else None, lda,
              a_kc, B.data_ptr(), B_lo.data_ptr() if B_lo is not None else None, ldb, b_kc, C.data_ptr(),
              C.shape[-1], alpha, beta, batch, sA, sB, sC, None)


@pytest.mark.parametrize('M,N,K', [(1, 1, 1), (17, 33, 5), (128, 128, 16), (300, 257, 1000), (129, 64, 47)])
@pytest.mark.parametrize('a_kc,b_kc', [(1, 1), (1, 0), (0, 1), (0, 0)])
def test_f64_layouts(M, N, K, a_kc, b_kc):
    print('case f64', M, N, K, a_kc, b_kc, flush=True)
    g = torch.Generator(device='cuda').manual_seed(M * 1000 + N + K)
    Am = torch.randn(M, K, dtype=torch.float64, device='cuda', generator=g)
    Bm = torch.randn(N, K, dtype=torch.float64, device='cuda', generator=g)
    C0 = torch.randn(M, N, dtype=torch.float64, device='cuda', generator=g)
    A = Am.contiguous() if a_kc else Am.t().contiguous()
    B = Bm.contiguous() if b_kc else Bm.t().contiguous()
    C = C0.clone()
    _gemm(0, A, B, C, 0.7, -0.3, a_kc, b_kc, M=M, N=N, K=K)
    torch.cuda.synchronize()
    ref = 0.7 * (Am @ Bm.t()) - 0.3 * C0
    assert torch.allclose(C, ref, rtol=1e-12, atol=1e-12 * (K ** 0.5))


def test_f64_batched():
    print('case f64 batched', flush=True)
    g = torch.Generator(device='cuda').manual_seed(5)
    bt, M, N, K = 7, 40, 50, 60
    A = torch.randn(bt, M, K, dtype=torch.float64, device='cuda', generator=g)
    B = torch.randn(bt, K, N, dtype=torch.float64, device='cuda', generator=g)
    C = torch.zeros(bt, M, N, dtype=torch.float64, device='cuda')
    _gemm(0, A, B, C, 1.0, 0.0, 1, 0, batch=bt, sA=M * K, sB=K * N, sC=M * N, M=M, N=N, K=K)
    torch.cuda.synchronize()
    assert torch.allclose(C, torch.bmm(A, B), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize('M,N,K', [(64, 64, 64), (300, 257, 1000), (1000, 777, 4096)])
def test_f32_and_bf16x3(M, N, K):
    from aces2 import _lib
    print('case lowp', M, N, K, flush=True)
    g = torch.Generator(device='cuda').manual_seed(K)
    A = torch.randn(M, K, dtype=torch.float64, device='cuda', generator=g)
    B = torch.randn(N, K, dtype=torch.float64, device='cuda', generator=g)
    ref = A @ B.t()
    scale = (A.abs() @ B.abs().t())
    # fp32 operands
    A32 = torch.empty(M, K, dtype=torch.float32, device='cuda')
    B32 = torch.empty(N, K, dtype=torch.float32, device='cuda')
    _lib.call('aces_convert_f32', M, K, A.data_ptr(), K, A32.data_ptr(), K, None)
    _lib.call('aces_convert_f32', N, K, B.data_ptr(), K, B32.data_ptr(), K, None)
    C = torch.zeros(M, N, dtype=torch.float64, device='cuda')
    _gemm(1, A32, B32, C, 1.0, 0.0, 1, 1, M=M, N=N, K=K)
    torch.cuda.synchronize()
    err32 = ((C - ref).abs() / scale).max().item()
    assert err32 < 4e-7, err32
    # bf16x3 operands (K padded to a multiple of 8 by construction here)
    Kp = (K + 7) // 8 * 8
    Ah = torch.empty(M, Kp, dtype=torch.int16, device='cuda')
    Al = torch.empty_like(Ah)
    Bh = torch.empty(N, Kp, dtype=torch.int16, device='cuda')
    Bl = torch.empty_like(Bh)
    _lib.call('aces_split_bf16', M, K, A.data_ptr(), K, Ah.data_ptr(), Al.data_ptr(), Kp, None)
    _lib.call('aces_split_bf16', N, K, B.data_ptr(), K, Bh.data_ptr(), Bl.data_ptr(), Kp, None)
    C.zero_()
    _gemm(2, Ah, Bh, C, 1.0, 0.0, 1, 1, A_lo=Al, B_lo=Bl, M=M, N=N, K=Kp)
    torch.cuda.synchronize()
    err16 = ((C - ref).abs() / scale).max().item()
    assert err16 < 3e-5, err16


def _grouped(prec, groups, a_kc, b_kc, alpha, beta, ws=None):
    import ctypes
    from aces2 import _lib
    ng = len(groups)
    I = ctypes.c_int * ng
    Lg = ctypes.c_long * ng
    P = ctypes.c_void_p * ng
    M = I(*[g['M'] for g in groups])
    N = I(*[g['N'] for g in groups])
    K = I(*[g['K'] for g in groups])
    A = P(*[g['A'].data_ptr() for g in groups])
    B = P(*[g['B'].data_ptr() for g in groups])
    Alo = P(*[g['Alo'].data_ptr() for g in groups]) if prec == 2 else None
    Blo = P(*[g['Blo'].data_ptr() for g in groups]) if prec == 2 else None
    C = P(*[g['C'].data_ptr() for g in groups])
    lda = Lg(*[g['A'].shape[-1] for g in groups])
    ldb = Lg(*[g['B'].shape[-1] for g in groups])
    ldc = Lg(*[g['C'].shape[-1] for g in groups])
    _lib.call('aces_gemm_grouped', prec, ng, M, N, K, A, Alo, lda, a_kc, B, Blo, ldb, b_kc, C, ldc, alpha, beta,
              ws.data_ptr() if ws is not None else None, ws.numel() if ws is not None else 0, None)


@pytest.mark.parametrize('a_kc,b_kc', [(1, 1), (1, 0), (0, 1), (0, 0)])
@pytest.mark.parametrize('split', [False, True])
def test_f64_grouped(a_kc, b_kc, split):
    """Uneven irrep-block shapes in one grouped launch (incl. an empty block and
    skinny long-K blocks that take the split-K path)."""
    print('case f64 grouped', a_kc, b_kc, split, flush=True)
    g = torch.Generator(device='cuda').manual_seed(11 + 2 * a_kc + b_kc)
    shapes = [(70, 50, 90), (0, 10, 10), (5, 7, 3000), (129, 3, 1), (16, 16, 5000), (200, 180, 64), (1, 1, 1)]
    groups, refs = [], []
    for (M, N, K) in shapes:
        Am = torch.randn(M, K, dtype=torch.float64, device='cuda', generator=g)
        Bm = torch.randn(N, K, dtype=torch.float64, device='cuda', generator=g)
        C0 = torch.randn(M, N, dtype=torch.float64, device='cuda', generator=g)
        A = Am.contiguous() if a_kc else Am.t().contiguous()
        B = Bm.contiguous() if b_kc else Bm.t().contiguous()
        if M == 0:
            A = torch.zeros(1, max(K, 1), dtype=torch.float64, device='cuda')
            C0 = torch.zeros(0, N, dtype=torch.float64, device='cuda')
        groups.append({'M': M, 'N': N, 'K': K, 'A': A, 'B': B, 'C': C0.clone()})
        refs.append(0.5 * (Am @ Bm.t()) + 0.25 * C0)
    ws = torch.empty(1 << 22, dtype=torch.float64, device='cuda') if split else None
    _grouped(0, groups, a_kc, b_kc, 0.5, 0.25, ws)
    torch.cuda.synchronize()
    for gr, ref in zip(groups, refs):
        assert torch.allclose(gr['C'], ref, rtol=1e-12, atol=1e-12 * (gr['K'] ** 0.5 + 1))


@pytest.mark.parametrize('prec', [1, 2])
def test_lowp_grouped(prec):
    from aces2 import _lib
    print('case lowp grouped', prec, flush=True)
    g = torch.Generator(device='cuda').manual_seed(31 + prec)
    shapes = [(130, 70, 64), (33, 200, 128), (256, 256, 256)]
    groups, refs = [], []
    for (M, N, K) in shapes:
        Am = torch.randn(M, K, dtype=torch.float64, device='cuda', generator=g)
        Bm = torch.randn(N, K, dtype=torch.float64, device='cuda', generator=g)
        C0 = torch.zeros(M, N, dtype=torch.float64, device='cuda')
        gr = {'M': M, 'N': N, 'K': K, 'C': C0}
        if prec == 1:
            gr['A'], gr['B'] = Am.float().contiguous(), Bm.float().contiguous()
        else:
            for nm, X in (('A', Am), ('B', Bm)):
                hi = torch.empty(X.shape, dtype=torch.int16, device='cuda')
                lo = torch.empty(X.shape, dtype=torch.int16, device='cuda')
                _lib.call('aces_split_bf16', X.shape[0], X.shape[1], X.data_ptr(), X.shape[1], hi.data_ptr(),
                          lo.data_ptr(), X.shape[1], None)
                gr[nm], gr[nm + 'lo'] = hi, lo
        groups.append(gr)
        refs.append((Am @ Bm.t(), Am.abs() @ Bm.abs().t()))
    _grouped(prec, groups, 1, 1, 1.0, 0.0)
    torch.cuda.synchronize()
    for gr, (ref, scale) in zip(groups, refs):
        err = ((gr['C'] - ref).abs() / scale).max().item()
        assert err < (4e-7 if prec == 1 else 3e-5), err


@pytest.mark.parametrize('M,N,K,ws', [(1500, 300, 512, False), (4100, 4200, 96, False), (1100, 2100, 1024, True),
                                      (1030, 1040, 2080, True), (1100, 1176, 1024, True), (1100, 1200, 1024, True),
                                      (1100, 1230, 1024, True), (16400, 1128, 96, False)])
def test_bf16x3_block_shapes(M, N, K, ws):
    """The bf16x3 block shapes (128x128, 256x128, 256x256; direct-to-LDS ring when
    K % 32 == 0) on ragged M/N, with and without split-K; last column tiles of
    256x256 blocks with 104, 152, 176 and 206 valid columns."""
    from aces2 import _lib
    print('case bf16x3 shape', M, N, K, ws, flush=True)
    g = torch.Generator(device='cuda').manual_seed(M + N + K)
    A = torch.randn(M, K, dtype=torch.float64, device='cuda', generator=g)
    B = torch.randn(N, K, dtype=torch.float64, device='cuda', generator=g)
    C0 = torch.randn(M, N, dtype=torch.float64, device='cuda', generator=g)
    ref = 0.75 * (A @ B.t()) + 0.5 * C0
    scale = 0.75 * (A.abs() @ B.abs().t()) + 0.5 * C0.abs()
    ops = []
    for X in (A, B):
        hi = torch.empty(X.shape, dtype=torch.int16, device='cuda')
        lo = torch.empty(X.shape, dtype=torch.int16, device='cuda')
        _lib.call('aces_split_bf16', X.shape[0], K, X.data_ptr(), K, hi.data_ptr(), lo.data_ptr(), K, None)
        ops.append((hi, lo))
    C = C0.clone()
    groups = [{'M': M, 'N': N, 'K': K, 'A': ops[0][0], 'Alo': ops[0][1], 'B': ops[1][0], 'Blo': ops[1][1], 'C': C}]
    wsb = torch.empty(1 << 24, dtype=torch.float64, device='cuda') if ws else None
    _grouped(2, groups, 1, 1, 0.75, 0.5, wsb)
    torch.cuda.synchronize()
    err = ((C - ref).abs() / scale).max().item()
    assert err < 3e-5, err


@pytest.mark.parametrize('M,N,K,ws', [(1300, 200, 512, False), (3000, 1176, 1024, False), (1100, 90, 4096, True),
                                      (2000, 320, 48, False)])
def test_f64_direct_lds(M, N, K, ws):
    """fp64 problems on the direct-to-LDS kernel (k-contiguous operands, K % 16 == 0,
    narrow outputs), ragged M/N, with and without split-K."""
    print('case f64 dl', M, N, K, ws, flush=True)
    g = torch.Generator(device='cuda').manual_seed(M + 7 * N + K)
    A = torch.randn(M, K, dtype=torch.float64, device='cuda', generator=g)
    B = torch.randn(N, K, dtype=torch.float64, device='cuda', generator=g)
    C0 = torch.randn(M, N, dtype=torch.float64, device='cuda', generator=g)
    C = C0.clone()
    wsb = torch.empty(1 << 22, dtype=torch.float64, device='cuda') if ws else None
    groups = [{'M': M, 'N': N, 'K': K, 'A': A, 'B': B, 'C': C}]
    _grouped(0, groups, 1, 1, -0.5, 1.5, wsb)
    torch.cuda.synchronize()
    ref = -0.5 * (A @ B.t()) + 1.5 * C0
    assert torch.allclose(C, ref, rtol=1e-12, atol=1e-12 * K ** 0.5)


@pytest.mark.parametrize('a_kc,b_kc', [(1, 0), (0, 1), (0, 0)])
@pytest.mark.parametrize('M,N,K,ws', [(1030, 500, 528, False), (300, 90, 4096, True), (1030, 501, 528, False),
                                      (2000, 320, 48, False), (1030, 300, 32, False)])
def test_f64_direct_lds_layouts(a_kc, b_kc, M, N, K, ws):
    """Row-contiguous operands on the direct-to-LDS fp64 kernel ([k][row] LDS
    tiles, row pairs clamped at the edge); an odd row count takes the
    register-staged kernel."""
    print('case f64 dl layout', a_kc, b_kc, M, N, K, ws, flush=True)
    g = torch.Generator(device='cuda').manual_seed(M * 3 + N + K + a_kc + 2 * b_kc)
    Am = torch.randn(M, K, dtype=torch.float64, device='cuda', generator=g)
    Bm = torch.randn(N, K, dtype=torch.float64, device='cuda', generator=g)
    C0 = torch.randn(M, N, dtype=torch.float64, device='cuda', generator=g)
    A = Am.contiguous() if a_kc else Am.t().contiguous()
    B = Bm.contiguous() if b_kc else Bm.t().contiguous()
    C = C0.clone()
    wsb = torch.empty(1 << 22, dtype=torch.float64, device='cuda') if ws else None
    _grouped(0, [{'M': M, 'N': N, 'K': K, 'A': A, 'B': B, 'C': C}], a_kc, b_kc, 0.25, -1.0, wsb)
    torch.cuda.synchronize()
    ref = 0.25 * (Am @ Bm.t()) - C0
    assert torch.allclose(C, ref, rtol=1e-12, atol=1e-12 * K ** 0.5)


@pytest.mark.parametrize('prec,a_kc,b_kc,M,N,K', [(2, 1, 1, 4352, 4096, 2048), (2, 1, 1, 4300, 4000, 1024),
                                                  (0, 0, 1, 2150, 1024, 2048), (0, 0, 0, 2150, 1000, 1024),
                                                  (0, 1, 1, 2200, 1024, 1024)])
def test_sparse_last_round(prec, a_kc, b_kc, M, N, K):
    """Products whose tile count leaves the last round of the chip's CUs under
    half full (about 272 tiles on 256 CUs) run as the rows that fill the full
    rounds plus the remaining rows with split-K: row offsets of k-contiguous and
    row-contiguous operands, alpha/beta on both parts."""
    from aces2 import _lib
    print('case sparse last round', prec, a_kc, b_kc, M, N, K, flush=True)
    g = torch.Generator(device='cuda').manual_seed(M + 3 * N + K + a_kc)
    Am = torch.randn(M, K, dtype=torch.float64, device='cuda', generator=g)
    Bm = torch.randn(N, K, dtype=torch.float64, device='cuda', generator=g)
    C0 = torch.randn(M, N, dtype=torch.float64, device='cuda', generator=g)
    ref = 0.75 * (Am @ Bm.t()) - 0.5 * C0
    C = C0.clone()
    wsb = torch.empty(1 << 24, dtype=torch.float64, device='cuda')
    if prec == 2:
        ops = []
        for X in (Am, Bm):
            hi = torch.empty(X.shape, dtype=torch.int16, device='cuda')
            lo = torch.empty(X.shape, dtype=torch.int16, device='cuda')
            _lib.call('aces_split_bf16', X.shape[0], K, X.data_ptr(), K, hi.data_ptr(), lo.data_ptr(), K, None)
            ops.append((hi, lo))
        groups = [{'M': M, 'N': N, 'K': K, 'A': ops[0][0], 'Alo': ops[0][1], 'B': ops[1][0], 'Blo': ops[1][1],
                   'C': C}]
        _grouped(2, groups, 1, 1, 0.75, -0.5, wsb)
        torch.cuda.synchronize()
        scale = 0.75 * (Am.abs() @ Bm.abs().t()) + 0.5 * C0.abs()
        err = ((C - ref).abs() / scale).max().item()
        assert err < 3e-5, err
    else:
        A = Am.contiguous() if a_kc else Am.t().contiguous()
        B = Bm.contiguous() if b_kc else Bm.t().contiguous()
        _grouped(0, [{'M': M, 'N': N, 'K': K, 'A': A, 'B': B, 'C': C}], a_kc, b_kc, 0.75, -0.5, wsb)
        torch.cuda.synchronize()
        assert torch.allclose(C, ref, rtol=1e-12, atol=1e-12 * K ** 0.5)


def _skinny_groups(shapes, a_kc, b_kc, seed):
    g = torch.Generator(device='cuda').manual_seed(seed)
    groups, refs = [], []
    for (M, N, K) in shapes:
        Am = torch.randn(M, K, dtype=torch.float64, device='cuda', generator=g)
        Bm = torch.randn(N, K, dtype=torch.float64, device='cuda', generator=g)
        C0 = torch.randn(M, N, dtype=torch.float64, device='cuda', generator=g)
        A = Am.contiguous() if a_kc else Am.t().contiguous()
        B = Bm.contiguous() if b_kc else Bm.t().contiguous()
        if M == 0:
            A = torch.zeros(1, max(K, 1), dtype=torch.float64, device='cuda')
        groups.append({'M': M, 'N': N, 'K': K, 'A': A, 'B': B, 'C': C0.clone()})
        refs.append((Am, Bm, C0))
    return groups, refs


@pytest.mark.parametrize('a_kc,b_kc', [(1, 1), (1, 0), (0, 1), (0, 0)])
@pytest.mark.parametrize('shapes', [
    [(20000, 6, 49)], [(6, 20000, 12)], [(5000, 49, 6)], [(49, 5000, 12)], [(4100, 33, 20)], [(4097, 17, 32)],
    [(20000, 6, 49), (0, 5, 40), (19000, 11, 40), (21000, 12, 46), (4096, 1, 1)],
    [(6, 6, 50000)], [(6, 49, 20000)], [(49, 49, 9000)], [(64, 64, 4096)], [(1, 33, 4099)], [(30, 64, 5000)],
    [(64, 20, 4100)],
    [(6, 6, 50000), (11, 11, 40000), (0, 3, 4096), (5, 57, 30001)]])
@pytest.mark.parametrize('beta', [0.0, 1.25])
def test_f64_skinny(a_kc, b_kc, shapes, beta):
    """Skinny fp64 products (one output side <= 64 with short K, or sides <= 32
    and <= 64 with long K) on the operand-streaming kernels, grouped with ragged
    and empty blocks, every operand order, against torch fp64 (49 x 49 and
    64 x 64 long-K outputs take the tiled split-K kernel)."""
    print('case f64 skinny', a_kc, b_kc, shapes, beta, flush=True)
    groups, refs = _skinny_groups(shapes, a_kc, b_kc, 17 + len(shapes) + 2 * a_kc + b_kc)
    ws = torch.empty(1 << 22, dtype=torch.float64, device='cuda')
    _grouped(0, groups, a_kc, b_kc, -0.75, beta, ws)
    torch.cuda.synchronize()
    for gr, (Am, Bm, C0) in zip(groups, refs):
        ref = -0.75 * (Am @ Bm.t()) + beta * C0
        assert torch.allclose(gr['C'], ref, rtol=1e-12, atol=1e-12 * (gr['K'] ** 0.5 + 1))


@pytest.mark.parametrize('ws_elems', [0, 40, 1 << 10])
def test_f64_skinny_small_workspace(ws_elems):
    """Long-K skinny products with no workspace (tiled kernels), one too small
    for the partial sums of every group, and one that holds a few splits."""
    print('case f64 skinny ws', ws_elems, flush=True)
    groups, refs = _skinny_groups([(6, 6, 50000), (5, 7, 40000)], 1, 1, 3)
    ws = torch.empty(ws_elems, dtype=torch.float64, device='cuda') if ws_elems else None
    _grouped(0, groups, 1, 1, 1.0, 0.5, ws)
    torch.cuda.synchronize()
    for gr, (Am, Bm, C0) in zip(groups, refs):
        ref = Am @ Bm.t() + 0.5 * C0
        assert torch.allclose(gr['C'], ref, rtol=1e-12, atol=1e-12 * (gr['K'] ** 0.5 + 1))
