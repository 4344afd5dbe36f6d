"""GPU hot path (libmqgpu on the MI355X) against the CPU oracle.

Bars: sigma vectors and boundary steps <= 1e-12 relative (FP64); Davidson
eigenvalues <= 1e-10; sweep energies <= 1e-10 Eh against the oracle sweep
and within the golden tolerance of the reference's test energies."""
import numpy as np
import pytest

from conftest import ham_from_fixture

pytestmark = pytest.mark.gpu


def _sweep_pair(ham, oracle_engine, gpu_engine, m, nmicro, check):
    from qcmaquis_amd.sweep import TwoSiteDMRG, micro_site
    d = TwoSiteDMRG(ham, oracle_engine, m)
    for i in range(nmicro):
        check(d, micro_site(ham.L, i))
        d.microiteration(0, i)
    return d


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name", ["lih", "benzene"])
def test_sigma_parity_fixtures(fixtures, oracle_engine, gpu_engine, name):
    ham = ham_from_fixture(fixtures[name])

    def check(d, site):
        psi, L, R, W = d.site_problem(site)
        Lg = gpu_engine.upload_boundary(L.layout, L.data)
        Rg = gpu_engine.upload_boundary(R.layout, R.data)
        x = np.random.default_rng(site).standard_normal(psi.to_flat().size)
        y0, f0 = oracle_engine.sigma(ham.symm, psi, L, R, W, x=x)
        y1, f1 = gpu_engine.sigma(ham.symm, psi, Lg, Rg, W, x=x)
        assert _rel(y1, y0) <= 1e-12, (site, _rel(y1, y0))
        assert f0 == f1

    _sweep_pair(ham, oracle_engine, gpu_engine, 60, 2 * (ham.L - 2), check)


def test_sigma_and_steps_parity_config0(oracle_engine, gpu_engine):
    from qcmaquis_amd.integrals import synthetic
    from qcmaquis_amd.sweep import Hamiltonian
    ham = Hamiltonian(8, 8, 0, synthetic(8, seed=1008))

    def check(d, site):
        psi, L, R, W = d.site_problem(site)
        Lg = gpu_engine.upload_boundary(L.layout, L.data)
        Rg = gpu_engine.upload_boundary(R.layout, R.data)
        x = psi.to_flat()
        y0, _ = oracle_engine.sigma(ham.symm, psi, L, R, W, x=x)
        y1, _ = gpu_engine.sigma(ham.symm, psi, Lg, Rg, W, x=x)
        assert _rel(y1, y0) <= 1e-12
        A = d.mps[site]
        l0 = oracle_engine.left_step(ham.symm, A, L, ham.mpo[site]).dicts()
        l1 = gpu_engine.left_step(ham.symm, A, Lg, ham.mpo[site]).dicts()
        B = d.mps[site + 1]
        r0 = oracle_engine.right_step(ham.symm, B, R, ham.mpo[site + 1]).dicts()
        r1 = gpu_engine.right_step(ham.symm, B, Rg, ham.mpo[site + 1]).dicts()
        for got, exp in ((l1, l0), (r1, r0)):
            assert len(got) == len(exp)
            for b in range(len(exp)):
                assert set(got[b]) == set(exp[b])
                for k in exp[b]:
                    assert np.abs(got[b][k] - exp[b][k]).max() <= 1e-12 * max(1.0, np.abs(exp[b][k]).max())

    _sweep_pair(ham, oracle_engine, gpu_engine, 64, 8, check)


def test_jacobi_davidson_parity(fixtures, oracle_engine, gpu_engine):
    ham = ham_from_fixture(fixtures["benzene"])

    def check(d, site):
        psi, L, R, W = d.site_problem(site)
        Lg = gpu_engine.upload_boundary(L.layout, L.data)
        Rg = gpu_engine.upload_boundary(R.layout, R.data)
        t0, u0, i0 = oracle_engine.jd(ham.symm, psi, L, R, W, 10, 1e-8)
        t1, u1, i1 = gpu_engine.jd(ham.symm, psi, Lg, Rg, W, 10, 1e-8)
        assert abs(t1 - t0) <= 1e-10 * abs(t0)
        assert i0 == i1
        assert min(np.abs(u1 - u0).max(), np.abs(u1 + u0).max()) <= 1e-8

    _sweep_pair(ham, oracle_engine, gpu_engine, 60, 4, check)


def test_jacobi_davidson_long_solves(fixtures, oracle_engine, gpu_engine):
    """Long solves (up to 40 iterations, no restart) from a random start: the GPU
    eigenvalue equals the oracle's and the lowest eigenvalue of the explicit
    H_eff, in FP64 and -- after the FP64 defect correction -- in the mixed form."""
    from qcmaquis_amd.desc import MQ_PREC_FP32, MQ_PREC_FP64
    from qcmaquis_amd.integrals import synthetic
    from qcmaquis_amd.sweep import Hamiltonian
    ham = Hamiltonian(8, 8, 0, synthetic(8, seed=1008))

    def check(d, site):
        psi, L, R, W = d.site_problem(site)
        n = psi.to_flat().size
        if n < 200:
            return
        H = np.zeros((n, n))
        for k in range(n):
            e = np.zeros(n)
            e[k] = 1.0
            H[:, k], _ = oracle_engine.sigma(ham.symm, psi, L, R, W, x=e)
        lam = np.linalg.eigvalsh(0.5 * (H + H.T))[0]
        Lg = gpu_engine.upload_boundary(L.layout, L.data)
        Rg = gpu_engine.upload_boundary(R.layout, R.data)
        rng = np.random.default_rng(site)
        x = rng.standard_normal(n)
        sp = gpu_engine.sigma_plan(ham.symm, psi, Lg, Rg, W)
        try:
            xd = gpu_engine.to_dev(x)
            t64, _, i64 = sp.jd(xd, 40, 1e-11, prec=MQ_PREC_FP64)
            t32, _, _ = sp.jd(xd, 40, 1e-11, prec=MQ_PREC_FP32)
        finally:
            sp.close()
        assert abs(t64 - lam) <= 1e-9 * max(1.0, abs(lam)), (site, t64, lam, i64)
        assert abs(t32 - lam) <= 1e-8, (site, t32, lam)

    _sweep_pair(ham, oracle_engine, gpu_engine, 48, 6, check)


@pytest.mark.parametrize("name", ["h2", "lih", "benzene"])
def test_golden_energies_gpu(fixtures, gpu_engine, name):
    from qcmaquis_amd.sweep import TwoSiteDMRG
    fx = fixtures[name]
    E = TwoSiteDMRG(ham_from_fixture(fx), gpu_engine, 100).run(2)
    assert abs(E - fx["energy"]) <= fx["rtol"] * abs(fx["energy"]), (E, fx["energy"])


def test_config0_sweep_energies_match_oracle(oracle_engine, gpu_engine):
    """L=8, N=8, S=0, m=64 two-site sweeps: GPU engine vs oracle, per microiteration."""
    from qcmaquis_amd.integrals import synthetic
    from qcmaquis_amd.sweep import Hamiltonian, TwoSiteDMRG
    ints = synthetic(8, seed=1008)
    e = []
    for eng in (oracle_engine, gpu_engine):
        ham = Hamiltonian(8, 8, 0, ints)
        d = TwoSiteDMRG(ham, eng, 64)
        d.run(2)
        e.append(np.array(d.energies))
    assert np.abs(e[0] - e[1]).max() <= 1e-10, np.abs(e[0] - e[1]).max()


@pytest.mark.parametrize("fuse", [0, 2])
def test_sigma_and_left_step_fuse_modes(oracle_engine, fuse):
    """The same sigma and left step with MPO columns applied term by term
    (fuse 0) and with every column summed on the left boundary first (fuse 2)."""
    from qcmaquis_amd.gpu import GpuEngine
    from qcmaquis_amd.integrals import synthetic
    from qcmaquis_amd.sweep import Hamiltonian
    ham = Hamiltonian(8, 8, 0, synthetic(8, seed=1008))
    eng = GpuEngine(0, fuse=fuse)
    try:
        def check(d, site):
            psi, L, R, W = d.site_problem(site)
            Lg = eng.upload_boundary(L.layout, L.data)
            Rg = eng.upload_boundary(R.layout, R.data)
            x = np.random.default_rng(site).standard_normal(psi.to_flat().size)
            y0, _ = oracle_engine.sigma(ham.symm, psi, L, R, W, x=x)
            y1, _ = eng.sigma(ham.symm, psi, Lg, Rg, W, x=x)
            assert _rel(y1, y0) <= 1e-12
            A = d.mps[site]
            l0 = oracle_engine.left_step(ham.symm, A, L, ham.mpo[site]).dicts()
            l1 = eng.left_step(ham.symm, A, Lg, ham.mpo[site]).dicts()
            assert len(l1) == len(l0)
            for b in range(len(l0)):
                assert set(l1[b]) == set(l0[b])
                for k in l0[b]:
                    assert np.abs(l1[b][k] - l0[b][k]).max() <= 1e-12 * max(1.0, np.abs(l0[b][k]).max())

        _sweep_pair(ham, oracle_engine, None, 48, 6, check)
    finally:
        eng.close()


@pytest.mark.parametrize("symm", ["su2u1", "2u1"])
def test_expval_gpu(fixtures, oracle_engine, gpu_engine, symm):
    """expval through the GPU boundary steps equals the GPU sweep energy and the
    oracle's expval of the same MPS."""
    from qcmaquis_amd.sweep import Hamiltonian, TwoSiteDMRG, expval
    fx = fixtures["benzene"]
    ham = Hamiltonian(fx["L"], fx["nelec"], fx["spin"], [(tuple(k), v) for k, v in fx["integrals"]], symm=symm)
    d = TwoSiteDMRG(ham, gpu_engine, 400)
    E = d.run(2)
    e_gpu = expval(gpu_engine, ham, d.mps)
    e_ora = expval(oracle_engine, ham, d.mps, from_right=True)
    assert abs(e_gpu - E) <= 1e-10 * abs(E), (e_gpu, E)
    assert abs(e_gpu - e_ora) <= 1e-12 * abs(E), (e_gpu, e_ora)
