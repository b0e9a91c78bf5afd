"""BF16 MFMA GEMM against an fp32 torch reference of the same bf16-rounded operands."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _ref(A, B):
    return A.float() @ B.float()


@pytest.mark.parametrize("M,N,K,batch", [(64, 64, 64, 1), (128, 256, 96, 2), (33, 70, 17, 3), (256, 32, 512, 4),
                                         (1, 300, 6, 1), (200, 1, 128, 2)])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_gemm_layouts_match_torch(M, N, K, batch, ta, tb):
    from polarism_amd.device import init
    from polarism_amd.surrogate.kernels import gemm
    init()
    torch.manual_seed(M * 7 + N)
    dev = "cuda"
    A = torch.randn(batch, M, K, device=dev).to(torch.bfloat16)
    B = torch.randn(batch, K, N, device=dev).to(torch.bfloat16)
    As = A.transpose(1, 2).contiguous() if ta else A.contiguous()
    Bs = B.transpose(1, 2).contiguous() if tb else B.contiguous()
    sa = (1, M, M * K) if ta else (K, 1, M * K)
    sb = (1, K, K * N) if tb else (N, 1, K * N)
    C = torch.zeros(batch, M, N, device=dev)
    gemm(As, Bs, C, M, N, K, batch=batch, sa=sa, sb=sb, sc=(N, 1, M * N))
    torch.cuda.synchronize()
    ref = torch.bmm(A.float(), B.float())
    err = (C - ref).abs().max().item() / max(ref.abs().max().item(), 1e-6)
    assert err < 1e-4, err


@pytest.mark.parametrize("M,N,K,batch", [(256, 128, 64, 1), (512, 384, 192, 2), (768, 256, 128, 1)])
@pytest.mark.parametrize("tb", [False, True])
def test_large_tile_gemm_matches_torch(M, N, K, batch, tb):
    """Whole 256 x 128 tiles with K a multiple of 64 and a k-contiguous A take the large-tile
    kernel (double-buffered LDS; a row-major B transposed in registers, tb = False)."""
    from polarism_amd.device import init
    from polarism_amd.surrogate.kernels import gemm
    init()
    torch.manual_seed(M + N + K)
    A = torch.randn(batch, M, K, device="cuda").to(torch.bfloat16)
    B = torch.randn(batch, K, N, device="cuda").to(torch.bfloat16)
    Bs = B.transpose(1, 2).contiguous() if tb else B.contiguous()
    sb = (1, K, K * N) if tb else (N, 1, K * N)
    C = torch.zeros(batch, M, N, device="cuda")
    gemm(A.contiguous(), Bs, C, M, N, K, batch=batch, sa=(K, 1, M * K), sb=sb, sc=(N, 1, M * N))
    torch.cuda.synchronize()
    ref = torch.bmm(A.float(), B.float())
    err = (C - ref).abs().max().item() / ref.abs().max().item()
    assert err < 1e-4, err


@pytest.mark.parametrize("bias_mode", [1, 2])
def test_large_tile_gemm_epilogue(bias_mode):
    from polarism_amd.device import init
    from polarism_amd.surrogate.kernels import gemm
    init()
    torch.manual_seed(9 + bias_mode)
    M, N, K = 256, 256, 128
    A = torch.randn(M, K, device="cuda").to(torch.bfloat16)
    B = torch.randn(K, N, device="cuda").to(torch.bfloat16)
    bias = torch.randn(M if bias_mode == 1 else N, device="cuda")
    res = torch.randn(M, N, device="cuda")
    C = torch.randn(M, N, device="cuda")
    C0 = C.clone()
    gemm(A, B, C, M, N, K, sa=(K, 1, 0), sb=(N, 1, 0), sc=(N, 1, 0), alpha=0.5, beta=2.0, bias=bias,
         bias_mode=bias_mode, residual=res, gelu=True)
    torch.cuda.synchronize()
    b = bias[:, None] if bias_mode == 1 else bias[None, :]
    ref = torch.nn.functional.gelu(0.5 * (A.float() @ B.float()) + 2.0 * C0 + b + res)
    assert torch.allclose(C, ref, atol=1e-4, rtol=1e-4)
    Cb = torch.empty(M, N, device="cuda", dtype=torch.bfloat16)
    gemm(A, B, Cb, M, N, K, sa=(K, 1, 0), sb=(N, 1, 0), sc=(N, 1, 0))
    torch.cuda.synchronize()
    assert torch.allclose(Cb.float(), A.float() @ B.float(), atol=3e-2, rtol=1e-2)


def test_gemm_epilogue_bias_residual_gelu_and_beta():
    from polarism_amd.device import init
    from polarism_amd.surrogate.kernels import gemm
    init()
    torch.manual_seed(3)
    M, N, K = 96, 160, 40
    A = torch.randn(M, K, device="cuda").to(torch.bfloat16)
    B = torch.randn(K, N, device="cuda").to(torch.bfloat16)
    bias = torch.randn(M, device="cuda")
    res = torch.randn(M, N, device="cuda")
    C = torch.randn(M, N, device="cuda")
    C0 = C.clone()
    gemm(A, B, C, M, N, K, sa=(K, 1, 0), sb=(N, 1, 0), sc=(N, 1, 0), alpha=0.5, beta=2.0, bias=bias, bias_mode=1,
         residual=res, gelu=True)
    torch.cuda.synchronize()
    pre = 0.5 * (A.float() @ B.float()) + 2.0 * C0 + bias[:, None] + res
    ref = torch.nn.functional.gelu(pre)
    assert torch.allclose(C, ref, atol=1e-4, rtol=1e-4)
    Cb = torch.empty(M, N, device="cuda", dtype=torch.bfloat16)
    gemm(A, B, Cb, M, N, K, sa=(K, 1, 0), sb=(N, 1, 0), sc=(N, 1, 0), bias=bias[:N] if False else None)
    torch.cuda.synchronize()
    assert torch.allclose(Cb.float(), A.float() @ B.float(), atol=3e-2, rtol=1e-2)


def test_gemm_epilogue_in_place_residual_and_bf16_beta():
    """The epilogue reads a block's old C and residual values before its stores: a residual that is
    C itself (an in-place residual add) and a bf16 C read with beta stay exact per element."""
    from polarism_amd.device import init
    from polarism_amd.surrogate.kernels import gemm
    init()
    torch.manual_seed(4)
    M, N, K = 64, 1000, 64
    A = torch.randn(M, K, device="cuda").to(torch.bfloat16)
    B = torch.randn(K, N, device="cuda").to(torch.bfloat16)
    C = torch.randn(M, N, device="cuda")
    C0 = C.clone()
    gemm(A, B, C, M, N, K, sa=(K, 1, 0), sb=(N, 1, 0), sc=(N, 1, 0), residual=C)
    torch.cuda.synchronize()
    assert torch.allclose(C, A.float() @ B.float() + C0, atol=1e-4, rtol=1e-4)
    Cb = torch.randn(M, N, device="cuda").to(torch.bfloat16)
    Cb0 = Cb.float().clone()
    gemm(A, B, Cb, M, N, K, sa=(K, 1, 0), sb=(N, 1, 0), sc=(N, 1, 0), beta=0.5)
    torch.cuda.synchronize()
    assert torch.allclose(Cb.float(), A.float() @ B.float() + 0.5 * Cb0, atol=6e-2, rtol=1e-2)


@pytest.mark.parametrize("M,N", [(64, 64), (3, 128), (64, 6)])
def test_split_k_weight_gradient_matches_torch(M, N):
    """K = 2^17 pixels into a small fp32 output takes the split-K path (partials + fixed-order sum)."""
    from polarism_amd.device import init
    from polarism_amd.surrogate.kernels import _k_splits, gemm
    init()
    torch.manual_seed(11)
    K = 1 << 17
    dy = torch.randn(M, K, device="cuda").to(torch.bfloat16)
    x = torch.randn(N, K, device="cuda").to(torch.bfloat16)
    C = torch.empty(M, N, device="cuda")
    assert _k_splits(M, N, K, 1, (N, 1, 0), 0.0, None, None, False, C) > 1
    gemm(dy, x, C, M, N, K, sa=(K, 1, 0), sb=(1, K, 0), sc=(N, 1, 0))
    torch.cuda.synchronize()
    ref = dy.float() @ x.float().T
    err = (C - ref).abs().max().item() / ref.abs().max().item()
    assert err < 1e-4, err


def test_gemm_unaligned_views_use_element_loads():
    """Operands offset by one element cannot use 16-byte vectors; results must not change."""
    from polarism_amd.device import init
    from polarism_amd.surrogate.kernels import gemm
    init()
    torch.manual_seed(5)
    M, N, K = 72, 90, 48
    Abuf = torch.randn(M * K + 1, device="cuda").to(torch.bfloat16)
    Bbuf = torch.randn(K * N + 1, device="cuda").to(torch.bfloat16)
    A, B = Abuf[1:], Bbuf[1:]
    C = torch.empty(M, N, device="cuda")
    gemm(A, B, C, M, N, K, sa=(K, 1, 0), sb=(N, 1, 0), sc=(N, 1, 0))
    torch.cuda.synchronize()
    ref = A.float().reshape(M, K) @ B.float().reshape(K, N)
    assert (C - ref).abs().max().item() < 1e-4 * ref.abs().max().item()


@pytest.mark.parametrize("rows,cols", [(1 << 20, 64), (1000, 3), (7, 5)])
def test_colsum_contiguous_columns(rows, cols):
    from polarism_amd.device import init
    from polarism_amd.surrogate.kernels import colsum
    init()
    torch.manual_seed(2)
    x = torch.randn(cols, rows, device="cuda")
    out = torch.empty(cols, device="cuda")
    colsum(x, out, rows, cols, 1, rows)
    torch.cuda.synchronize()
    ref = x.double().sum(dim=1).float()
    assert torch.allclose(out, ref, rtol=1e-4, atol=1e-3 * rows ** 0.5)
