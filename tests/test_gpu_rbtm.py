"""The rbtm sigma (MQ_OPT_SCHEDULE = MQ_SCHED_REFERENCE) on the GPU against the
oracle: at every site of a sweep where the reference picks rbtm
(site_hamil.hpp:35), sigma in FP64 within 1e-12 of the oracle (mixed and BF16x3
within 1e-5 / 1e-4), the batched sigma equal to single sigma vectors, and the
Jacobi-Davidson eigenvalue equal to the lbtm plan's; at full size (L=24 m=512,
the two-site problem right of the centre) sigma within 1e-12 and executed FLOPs
within 5% of the reference's F."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_rbtm_sites_vs_oracle(oracle_engine, gpu_engine):
    from autotune_helpers import schedule
    from qcmaquis_amd.desc import MQ_PREC_BF16X3, MQ_PREC_FP32, MQ_PREC_FP64
    from qcmaquis_amd.integrals import synthetic
    from qcmaquis_amd.sweep import Hamiltonian, TwoSiteDMRG
    ham = Hamiltonian(8, 8, 0, synthetic(8, seed=1008))
    d = TwoSiteDMRG(ham, oracle_engine, 48)
    d.sweep(0)
    n = 0
    with schedule(gpu_engine, sched=1):
        for site in range(ham.L - 1):
            psi, L, R, W = d.site_problem(site)
            Lg = gpu_engine.upload_boundary(L.layout, L.data)
            Rg = gpu_engine.upload_boundary(R.layout, R.data)
            sp = gpu_engine.sigma_plan(ham.symm, psi, Lg, Rg, W)
            try:
                if sp.rbtm:
                    n += 1
                    print(f"rbtm site {site}, n {sp.n}", flush=True)
                    x = np.random.default_rng(site).standard_normal(sp.n)
                    y0, fl = oracle_engine.sigma(ham.symm, psi, L, R, W, x=x)
                    assert sp.ref_flops == fl
                    xd = gpu_engine.to_dev(x)
                    y1 = sp.apply(xd, prec=MQ_PREC_FP64).cpu().numpy()
                    y32 = sp.apply(xd, prec=MQ_PREC_FP32).cpu().numpy()
                    y3 = sp.apply(xd, prec=MQ_PREC_BF16X3).cpu().numpy()
                    assert _rel(y1, y0) <= 1e-12
                    assert _rel(y32, y0) <= 1e-5 and _rel(y3, y0) <= 1e-4
                    X = gpu_engine.torch.as_tensor(np.stack([x, -x, x[::-1].copy()])).to(gpu_engine.device)
                    Y = sp.apply_batch(X, prec=MQ_PREC_FP64).cpu().numpy()
                    assert np.array_equal(Y[0], y1)
                    assert np.array_equal(Y[1], -y1)
                    t1, _, _ = sp.jd(gpu_engine.to_dev(psi.to_flat()), 30, 1e-10, prec=MQ_PREC_FP64)
                    t0, _, _ = oracle_engine.jd(ham.symm, psi, L, R, W, 30, 1e-10)
                    assert abs(t1 - t0) <= 1e-9 * max(1.0, abs(t0)), (site, t1, t0)
            finally:
                sp.close()
            d.microiteration(1, site)
    assert n >= 3


def test_rbtm_fullsize(gpu_engine, oracle_engine):
    from autotune_helpers import schedule
    from pyoracle import HostBoundary
    from qcmaquis_amd.desc import MQ_PREC_BF16X3, MQ_PREC_FP64
    from qcmaquis_amd.workload import centre_problem, load_sectors
    _, sec = load_sectors(os.path.join(HERE, "golden", "sectors_L24_m512.json"))
    print("building the L=24 m=512 problem at sites 12, 13", flush=True)
    p = centre_problem(gpu_engine, 24, sectors=sec, site=12)
    x = np.random.default_rng(3).standard_normal(p.psi.to_flat().size)
    with schedule(gpu_engine, sched=1):
        sp = gpu_engine.sigma_plan(p.ham.symm, p.psi, p.L, p.R, p.W)
    try:
        assert sp.rbtm
        y1 = sp.apply(gpu_engine.to_dev(x), prec=MQ_PREC_FP64).cpu().numpy()
        y3 = sp.apply(gpu_engine.to_dev(x), prec=MQ_PREC_BF16X3).cpu().numpy()
        ex, F = sp.exec_flops, sp.ref_flops
    finally:
        sp.close()
    y0, f0 = oracle_engine.sigma(p.ham.symm, p.psi, HostBoundary(p.L.layout, p.L.data.cpu().numpy()),
                                 HostBoundary(p.R.layout, p.R.data.cpu().numpy()), p.W, x=x)
    print(f"rbtm L=24 site 12: F {F / 1e9:.2f} GF, executed {ex / 1e9:.2f} GF, rel err {_rel(y1, y0):.2e}, "
          f"bf16x3 {_rel(y3, y0):.2e}", flush=True)
    assert F == f0
    assert _rel(y1, y0) <= 1e-12
    assert _rel(y3, y0) <= 2e-5
    assert abs(ex - F) <= 0.05 * F
