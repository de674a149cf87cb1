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
