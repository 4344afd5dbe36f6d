"""The spin-orbital (U(1) x U(1)) path on the GPU against the CPU oracle:
abelian sigma, boundary steps and Jacobi-Davidson (<= 1e-12), and the
reference's golden energies reached by GPU sweeps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ham2u1(fx):
    from qcmaquis_amd.sweep import Hamiltonian
    return Hamiltonian(fx["L"], fx["nelec"], fx["spin"], [(tuple(k), v) for k, v in fx["integrals"]], symm="2u1")


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_twou1_sigma_steps_jd_parity(fixtures, oracle_engine, gpu_engine):
    from qcmaquis_amd.sweep import TwoSiteDMRG, micro_site, n_microiterations
    ham = _ham2u1(fixtures["benzene"])
    d = TwoSiteDMRG(ham, oracle_engine, 60)
    for i in range(n_microiterations(ham.L)):
        site = micro_site(ham.L, i)
        psi, L, R, W = d.site_problem(site)
        Lg = gpu_engine.upload_boundary(L.layout, L.data)
        Rg = gpu_engine.upload_boundary(R.layout, R.data)
        x = np.random.default_rng(i).standard_normal(psi.to_flat().size)
        y0, f0 = oracle_engine.sigma(ham.symm, psi, L, R, W, x=x)
        y1, f1 = gpu_engine.sigma(ham.symm, psi, Lg, Rg, W, x=x)
        assert _rel(y1, y0) <= 1e-12
        assert f1 == f0
        A, B = d.mps[site], d.mps[site + 1]
        for got, exp in ((gpu_engine.left_step(ham.symm, A, Lg, ham.mpo[site]).dicts(),
                          oracle_engine.left_step(ham.symm, A, L, ham.mpo[site]).dicts()),
                         (gpu_engine.right_step(ham.symm, B, Rg, ham.mpo[site + 1]).dicts(),
                          oracle_engine.right_step(ham.symm, B, R, ham.mpo[site + 1]).dicts())):
            assert len(got) == len(exp)
            for b in range(len(exp)):
                assert set(got[b]) == set(exp[b])
                for k in exp[b]:
                    assert np.abs(got[b][k] - exp[b][k]).max() <= 1e-12 * max(1.0, np.abs(exp[b][k]).max())
        t1, _, i1 = gpu_engine.jd(ham.symm, psi, Lg, Rg, W, 6, 1e-10)
        t0, _, i0 = oracle_engine.jd(ham.symm, psi, L, R, W, 6, 1e-10)
        assert i1 == i0 and abs(t1 - t0) <= 1e-10 * max(1.0, abs(t0))
        d.microiteration(0, i)


@pytest.mark.parametrize("name", ["h2", "lih", "benzene"])
def test_twou1_golden_energies_gpu(fixtures, gpu_engine, name):
    from qcmaquis_amd.sweep import TwoSiteDMRG
    fx = fixtures[name]
    E = TwoSiteDMRG(_ham2u1(fx), gpu_engine, 100).run(2)
    assert abs(E - fx["energy"]) <= fx["rtol"] * abs(fx["energy"]), (E, fx["energy"])
