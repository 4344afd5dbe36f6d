"""Boundary binding of a sigma plan (mq_plan_bind) on the GPU.

A plan arranges the boundaries into its GEMM panels once and reuses them
while the (L, R) pointer pair stays the same.  A boundary freed and replaced
by a new one at the same address (which torch's caching allocator does
routinely) must be re-arranged with mq_plan_bind before the next sigma; the
sigma then equals the oracle's for the new values."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_bind_after_boundary_reuse(fixtures, oracle_engine, gpu_engine):
    from conftest import ham_from_fixture
    from pyoracle import HostBoundary
    from qcmaquis_amd.gpu import GpuBoundary
    from qcmaquis_amd.sweep import TwoSiteDMRG
    ham = ham_from_fixture(fixtures["benzene"])
    d = TwoSiteDMRG(ham, oracle_engine, 40)
    for i in range(2):
        d.microiteration(0, i)
    psi, L, R, W = d.site_problem(2)
    x = np.random.default_rng(1).standard_normal(psi.to_flat().size)
    t = gpu_engine.torch
    Lg = gpu_engine.upload_boundary(L.layout, L.data)
    Rg = gpu_engine.upload_boundary(R.layout, R.data)
    sp = gpu_engine.sigma_plan(ham.symm, psi, Lg, Rg, W)
    try:
        xd = gpu_engine.to_dev(x)
        y0 = sp.apply(xd).cpu().numpy()
        ref0, _ = oracle_engine.sigma(ham.symm, psi, L, R, W, x=x)
        assert _rel(y0, ref0) <= 1e-12
        # a different left boundary of the same layout, written where the old one was
        newL = np.random.default_rng(2).standard_normal(L.data.size)
        ptr = Lg.data.data_ptr()
        n_el = Lg.data.numel()
        del Lg
        sp.L = None
        t.cuda.synchronize()
        Lg2 = GpuBoundary(L.layout, t.empty(n_el, dtype=t.float64, device=gpu_engine.device))
        Lg2.data.copy_(t.as_tensor(newL))
        t.cuda.synchronize()
        print(f"old boundary at {ptr:#x}, new one at {Lg2.data.data_ptr():#x}", flush=True)
        sp.bind(L=Lg2)
        y1 = sp.apply(xd).cpu().numpy()
        ref1, _ = oracle_engine.sigma(ham.symm, psi, HostBoundary(L.layout, newL), R, W, x=x)
        assert _rel(y1, ref1) <= 1e-12
        assert _rel(ref1, ref0) > 1e-3  # the two boundaries give different sigma vectors
        # new values written in place: the pointer pair is unchanged, so only an
        # explicit bind picks them up
        Lg2.data.mul_(2.0)
        sp.bind()
        y2 = sp.apply(xd).cpu().numpy()
        assert _rel(y2, 2.0 * ref1) <= 1e-12
    finally:
        sp.close()
