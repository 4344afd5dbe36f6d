"""Kernel-level checks of the FP64 grouped GEMM (MFMA) against torch FP64 on
the GPU: edge tiles, K = 0, split-K, both transposed-operand paths."""
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1), (16, 16, 4), (64, 64, 16), (65, 63, 17), (5, 130, 3), (200, 7, 1000), (37, 41, 0),
          (64, 128, 32), (70, 129, 20), (300, 65, 7),
          (128, 96, 5000), (966, 101, 9000),
          # one shape per column tile width (16 .. 128 in steps of 16)
          (90, 17, 33), (150, 47, 30), (130, 81, 50), (129, 112, 31), (100, 200, 20), (64, 987, 98)]


@pytest.mark.parametrize("transA,transB", [(False, False), (True, False), (False, True)])
@pytest.mark.parametrize("M,N,K", SHAPES)
def test_gemm_f64(gpu_engine, M, N, K, transA, transB):
    import torch
    g = torch.Generator(device="cpu").manual_seed(M * 1000003 + N * 1009 + K)
    a = torch.randn(M, K, generator=g, dtype=torch.float64).to(gpu_engine.device)
    b = torch.randn(K, N, generator=g, dtype=torch.float64).to(gpu_engine.device)
    c = gpu_engine.gemm(a, b, transA, transB, kchunk=512)
    ref = (a.cpu() @ b.cpu()) if K else torch.zeros(M, N, dtype=torch.float64)
    err = (c.cpu() - ref).abs().max().item()
    scale = max(1.0, ref.abs().max().item()) * max(K, 1) ** 0.5
    assert err <= 1e-14 * scale, (M, N, K, transA, transB, err)
