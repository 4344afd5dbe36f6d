"""The allocation-free GEMM entry point (mq_gemm_f64_noalloc) against
mq_gemm_f64 on the GPU, through the C ABI.

mq_gemm_f64_noalloc runs the same kernels as mq_gemm_f64, with its task lists
and split-K partial tiles in the library's fixed per-device scratch array
instead of allocated device memory, so the two must agree bit for bit.  Two
calls in a row (the second smaller, so stale partials of the first would
show), a split-K product (K long, output small: the partials are used) and a
product beyond the scratch size (rejected with an error, no launch)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(eng, fn, A, B, M, N, K, kchunk):
    from qcmaquis_amd.gpu import lib
    t = eng.torch
    Cm = t.full((N, M), float("nan"), dtype=t.float64, device=eng.device)
    t.cuda.synchronize()
    rc = fn(eng.ctx, 0, 0, M, N, K, C.c_void_p(A.data_ptr()), max(M, 1), C.c_void_p(B.data_ptr()), max(K, 1),
            C.c_void_p(Cm.data_ptr()), max(M, 1), kchunk)
    return rc, Cm


@pytest.mark.parametrize("shapes", [[(300, 70, 5000), (130, 40, 3000)], [(1000, 200, 96), (77, 13, 40)]])
def test_noalloc_gemm_equals_gemm(gpu_engine, shapes):
    from qcmaquis_amd.gpu import lib
    L = lib()
    t = gpu_engine.torch
    vp = C.c_void_p
    L.mq_gemm_f64_noalloc.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int,
                                      vp, C.c_int, C.c_int]
    for M, N, K in shapes:
        print(f"noalloc gemm {M}x{N}x{K}", flush=True)
        rng = np.random.default_rng(M + N + K)
        A = t.as_tensor(rng.standard_normal(M * K)).to(gpu_engine.device)  # column-major M x K
        B = t.as_tensor(rng.standard_normal(K * N)).to(gpu_engine.device)  # column-major K x N
        rc0, C0 = _call(gpu_engine, L.mq_gemm_f64, A, B, M, N, K, 512)
        rc1, C1 = _call(gpu_engine, L.mq_gemm_f64_noalloc, A, B, M, N, K, 512)
        assert rc0 == 0 and rc1 == 0, gpu_engine.err()
        c0, c1 = C0.cpu().numpy(), C1.cpu().numpy()
        assert np.array_equal(c0, c1)
        ref = A.cpu().numpy().reshape(K, M).T @ B.cpu().numpy().reshape(N, K).T
        assert np.abs(c1.T - ref).max() <= 1e-12 * np.abs(ref).max()


def test_noalloc_gemm_rejects_beyond_scratch(gpu_engine):
    from qcmaquis_amd.gpu import lib
    L = lib()
    t = gpu_engine.torch
    vp = C.c_void_p
    L.mq_gemm_f64_noalloc.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int,
                                      vp, C.c_int, C.c_int]
    # 896 x 512 output (28 tiles), K = 131072 in 64 chunks of 2048: 64 partial
    # tiles of 896 x 512 doubles are 235 MB, beyond the 32 MiB scratch array
    M, N, K = 896, 512, 131072
    A = t.zeros(M * K, dtype=t.float64, device=gpu_engine.device)
    B = t.zeros(K * N, dtype=t.float64, device=gpu_engine.device)
    Cm = t.zeros(M * N, dtype=t.float64, device=gpu_engine.device)
    t.cuda.synchronize()
    rc = L.mq_gemm_f64_noalloc(gpu_engine.ctx, 0, 0, M, N, K, vp(A.data_ptr()), M, vp(B.data_ptr()), K,
                               vp(Cm.data_ptr()), M, 2048)
    assert rc == -1
    assert "scratch" in gpu_engine.err()
