"""The BF16x3 form (MQ_PREC_BF16X3) on the GPU: the FP32-operand GEMM kernel in
FP32 and BF16x3 products against the numpy model of the three bf16 products,
sigma vectors against the FP64 oracle, and eigensolvers that reach the FP64
eigenvalue through the FP64 continuation.

Bars: GEMM within FP32 accumulation error of the model (1e-6 of sum |a||b|,
far below the 2^-16 of the bf16 split itself); sigma within 1e-4 of |y|
(relative norm); eigenvalues within 1e-10 of the dense lowest ones."""
import numpy as np
import pytest

from bf16 import bf16x3_matmul
from test_gpu0_kernels import SHAPES

pytestmark = pytest.mark.gpu

TILES = [(64, 64), (128, 128), (64, 16), (128, 48), (64, 112)]


@pytest.mark.parametrize("tm,tn", TILES)
@pytest.mark.parametrize("transA,transB", [(False, False), (True, False), (False, True)])
def test_gemm32_fp32_and_bf16x3(gpu_engine, tm, tn, transA, transB):
    import torch
    for (M, N, K) in SHAPES:
        if M * N * K > 3e8:
            continue
        rng = np.random.default_rng(M * 7 + N * 13 + K)
        a = rng.standard_normal((M, K)).astype(np.float32)
        b = rng.standard_normal((K, N)).astype(np.float32)
        ad = torch.as_tensor(a).to(gpu_engine.device)
        bd = torch.as_tensor(b).to(gpu_engine.device)
        exact = a.astype(np.float64) @ b.astype(np.float64)
        scale = np.abs(a).astype(np.float64) @ np.abs(b).astype(np.float64) + 1e-30
        print(f"gemm32 {M}x{N}x{K} tile {tm}x{tn} trans {transA}{transB}", flush=True)
        c32 = gpu_engine.gemm32(ad, bd, transA, transB, kchunk=512, tm=tm, tn=tn).cpu().numpy()
        c3 = gpu_engine.gemm32(ad, bd, transA, transB, kchunk=512, tm=tm, tn=tn, bf16x3=True).cpu().numpy()
        assert (np.abs(c32 - exact) / scale).max() <= 1e-6, (M, N, K)
        model = bf16x3_matmul(a, b)
        assert (np.abs(c3 - model) / scale).max() <= 1e-6, (M, N, K, (np.abs(c3 - model) / scale).max())
        assert (np.abs(c3 - exact) / scale).max() <= 3 * 2.0 ** -16


def _site(fixtures, oracle_engine, gpu_engine, name, site=2):
    from qcmaquis_amd.sweep import Hamiltonian, TwoSiteDMRG
    fx = fixtures[name]
    ham = Hamiltonian(fx["L"], fx["nelec"], fx["spin"], [(tuple(k), v) for k, v in fx["integrals"]])
    d = TwoSiteDMRG(ham, oracle_engine, 40)
    for i in range(2):
        d.microiteration(0, i)
    psi, L, R, W = d.site_problem(site)
    Lg = gpu_engine.upload_boundary(L.layout, L.data)
    Rg = gpu_engine.upload_boundary(R.layout, R.data)
    return ham, psi, L, R, W, Lg, Rg


@pytest.mark.parametrize("name", ["lih", "benzene"])
def test_sigma_bf16x3_vs_oracle(fixtures, oracle_engine, gpu_engine, name):
    from qcmaquis_amd.desc import MQ_PREC_BF16X3, MQ_PREC_FP32
    ham, psi, L, R, W, Lg, Rg = _site(fixtures, oracle_engine, gpu_engine, name)
    x = np.random.default_rng(5).standard_normal(psi.to_flat().size)
    y_ref, _ = oracle_engine.sigma(ham.symm, psi, L, R, W, x=x)
    sp = gpu_engine.sigma_plan(ham.symm, psi, Lg, Rg, W)
    try:
        xd = gpu_engine.to_dev(x)
        y3 = sp.apply(xd, prec=MQ_PREC_BF16X3).cpu().numpy()
        y32 = sp.apply(xd, prec=MQ_PREC_FP32).cpu().numpy()
        X = gpu_engine.torch.as_tensor(np.stack([x, -2 * x, x[::-1].copy()])).to(gpu_engine.device)
        Yb = sp.apply_batch(X, prec=MQ_PREC_BF16X3).cpu().numpy()
    finally:
        sp.close()
    e3 = np.linalg.norm(y3 - y_ref) / np.linalg.norm(y_ref)
    e32 = np.linalg.norm(y32 - y_ref) / np.linalg.norm(y_ref)
    print(f"{name}: relative sigma error bf16x3 {e3:.2e}, fp32 {e32:.2e}", flush=True)
    assert e3 <= 1e-4 and e32 <= 1e-5
    assert np.array_equal(Yb[0], y3)
    assert np.allclose(Yb[1], -2 * y3, rtol=0, atol=1e-4 * np.abs(y3).max())


def test_solvers_bf16x3_reach_fp64_eigenvalue(fixtures, oracle_engine, gpu_engine):
    from qcmaquis_amd.desc import MQ_PREC_BF16X3
    ham, psi, L, R, W, Lg, Rg = _site(fixtures, oracle_engine, gpu_engine, "benzene")
    n = psi.to_flat().size
    H = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        H[:, j], _ = oracle_engine.sigma(ham.symm, psi, L, R, W, x=e)
    w = np.linalg.eigvalsh(0.5 * (H + H.T))
    sp = gpu_engine.sigma_plan(ham.symm, psi, Lg, Rg, W)
    try:
        x = gpu_engine.to_dev(psi.to_flat())
        tj, _, ij = sp.jd(x, 60, 1e-11, prec=MQ_PREC_BF16X3)
        td, _, _ = sp.davidson(x, 200, 1e-11, prec=MQ_PREC_BF16X3, diag=sp.diagonal())
        X0 = gpu_engine.torch.as_tensor(np.random.default_rng(4).standard_normal((4, n))).to(gpu_engine.device)
        tb, U, ib = sp.block_davidson(X0, 200, 1e-11, prec=MQ_PREC_BF16X3, diag=sp.diagonal())
    finally:
        sp.close()
    print(f"BF16x3 JD iterations {ij}, block Davidson iterations {ib}", flush=True)
    assert abs(tj - w[0]) <= 1e-10 and abs(td - w[0]) <= 1e-10
    assert np.abs(np.array(tb) - w[:4]).max() <= 1e-10, (tb, w[:4])
