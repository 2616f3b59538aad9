"""MFMA BF16 GEMM kernels of the surrogate against a plain PyTorch fp32
reference of the same product (bf16 inputs, fp32 accumulation)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(128, 192, 96), (256, 384, 192), (384, 128, 512)],
                         ids=["64x64-tiles", "128x128-tiles", "128x128-tiles-long-k"])
@pytest.mark.parametrize("ta,tb", [(0, 1), (0, 0), (1, 0), (1, 1)])
def test_gemm_bf16_matches_torch(ta, tb, shape):
    """Both tile shapes (64x64 on the K=16 MFMA; 128x128 on the gfx950 K=32
    MFMA when M, N are multiples of 128 and K of 64) in every operand
    orientation."""
    import torch
    from jdsurfg_amd import swd
    L = swd.lib()
    L.jdsg_gemm_bf16.argtypes = [C.c_int] * 5 + [C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                                 C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                                 C.c_void_p, C.c_int, C.c_void_p]
    M, N, K = shape
    g = torch.Generator(device="cpu").manual_seed(3)
    a = torch.randn(M, K, generator=g).to(torch.bfloat16).cuda()
    b = torch.randn(K, N, generator=g).to(torch.bfloat16).cuda()
    A = a.t().contiguous() if ta else a
    B = b.t().contiguous() if tb else b
    lda = A.shape[1]
    ldb = B.shape[1]
    Cm = torch.zeros(M, N, dtype=torch.float32, device="cuda")
    print(f"case gemm ta={ta} tb={tb} M,N,K={shape}", flush=True)
    rc = L.jdsg_gemm_bf16(ta, tb, M, N, K, A.data_ptr(), lda, B.data_ptr(), ldb, Cm.data_ptr(), N,
                          0, None, None, 0, None)
    assert rc == 0
    torch.cuda.synchronize()
    ref = a.float() @ b.float()
    err = (Cm - ref).abs().max().item() / ref.abs().max().item()
    assert err < 1e-5, err


@pytest.mark.parametrize("M,N,K", [(64, 64, 32), (256, 128, 192)])
def test_gemm_bias_gelu_epilogue(M, N, K):
    import torch
    from jdsurfg_amd import swd
    L = swd.lib()
    g = torch.Generator(device="cpu").manual_seed(5)
    x = torch.randn(M, K, generator=g).to(torch.bfloat16).cuda()
    w = torch.randn(N, K, generator=g).to(torch.bfloat16).cuda()
    bias = torch.randn(N, generator=g).cuda()
    Z = torch.zeros(M, N, device="cuda")
    Y = torch.zeros(M, N, dtype=torch.bfloat16, device="cuda")
    rc = L.jdsg_gemm_bf16(0, 1, M, N, K, x.data_ptr(), K, w.data_ptr(), K, Z.data_ptr(), N, 1,
                          bias.data_ptr(), Y.data_ptr(), N, None)
    assert rc == 0
    torch.cuda.synchronize()
    zr = x.float() @ w.float().t() + bias
    yr = torch.nn.functional.gelu(zr, approximate="tanh")
    assert (Z - zr).abs().max().item() < 1e-4 * zr.abs().max().item()
    assert (Y.float() - yr).abs().max().item() < 1e-2 * yr.abs().max().item()
