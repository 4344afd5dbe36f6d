"""Davidson with the diagonal preconditioner (optimize/davidson.h:64-161,
optimize/ietl_davidson.h:78-160) and the SU2 diagonal of H_eff
(non-abelian/h_diag.hpp:20-152), on the CPU: the oracle against the dense
H_eff built column by column, and the diagonal plan's task lists run by the
numpy emulator against the oracle."""
import numpy as np
import pytest

from conftest import ham_from_fixture


def _problem(fixtures, oracle_engine, name, herm=True, m=40, nmicro=2, site=2):
    from qcmaquis_amd.sweep import Hamiltonian, TwoSiteDMRG
    fx = fixtures[name]
    ham = Hamiltonian(fx["L"], fx["nelec"], fx["spin"], [(tuple(k), v) for k, v in fx["integrals"]], herm=herm)
    d = TwoSiteDMRG(ham, oracle_engine, m)
    for i in range(nmicro):
        d.microiteration(0, i)
    return ham, d.site_problem(min(site, ham.L - 2))


def _dense(ham, oracle_engine, psi, L, R, W):
    n = psi.to_flat().size
    H = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        H[:, j], _ = oracle_engine.sigma(ham.symm, psi, L, R, W, x=e)
    return H


@pytest.mark.parametrize("name", ["lih", "benzene"])
def test_diagonal_equals_dense_diagonal(fixtures, oracle_engine, name):
    """With every boundary slice stored (an MPO without hermitian-partner bonds)
    the h_diag.hpp construction is the exact diagonal of H_eff: the couplings,
    the spin case of each entry and the row offsets all enter it."""
    ham, (psi, L, R, W) = _problem(fixtures, oracle_engine, name, herm=False)
    H = _dense(ham, oracle_engine, psi, L, R, W)
    D = oracle_engine.diagonal(ham.symm, psi, L, R, W)
    assert np.abs(D - np.diag(H)).max() <= 1e-13 * np.abs(np.diag(H)).max()


def test_diagonal_skips_unstored_partner_slices(fixtures, oracle_engine):
    """With hermitian-partner bonds the partner slices of L and R are not stored,
    so the diagonal misses their (small) contributions, as in the reference."""
    ham, (psi, L, R, W) = _problem(fixtures, oracle_engine, "benzene", herm=True)
    assert any(W.left_herm[b] < b for b in range(W.row_dim)) or any(W.right_herm[b] < b for b in range(W.col_dim))
    H = _dense(ham, oracle_engine, psi, L, R, W)
    D = oracle_engine.diagonal(ham.symm, psi, L, R, W)
    diff = np.abs(D - np.diag(H)).max()
    assert 0 < diff < 1e-6 * np.abs(np.diag(H)).max()


def test_davidson_matches_dense_eigensolver(fixtures, oracle_engine):
    """Lowest eigenpair from the MPS guess and from random guesses (which run
    through the restart to two vectors at 20, davidson.h:152-156)."""
    ham, (psi, L, R, W) = _problem(fixtures, oracle_engine, "benzene")
    H = _dense(ham, oracle_engine, psi, L, R, W)
    w = np.linalg.eigvalsh(H)
    th, u, its = oracle_engine.davidson(ham.symm, psi, L, R, W, max_iter=100, tol=1e-10)
    assert abs(th - w[0]) < 1e-10
    restarted = 0
    for seed in range(3):
        x0 = np.random.default_rng(seed).standard_normal(w.size)
        th, u, its = oracle_engine.davidson(ham.symm, psi, L, R, W, max_iter=200, tol=1e-10, x0=x0)
        restarted += its > 20
        assert abs(th - w[0]) < 1e-10, (seed, th - w[0])
        assert abs(u @ u - 1) < 1e-12
        assert np.linalg.norm(H @ u - th * u) <= 1e-10 * abs(th) * 1.01
    assert restarted >= 2


def test_davidson_projects_out_ortho_vectors(fixtures, oracle_engine):
    """With the exact ground state as constraint the solver returns the second
    eigenvalue (SingleSiteVS::project, ietl_lanczos_solver.h:90-95)."""
    ham, (psi, L, R, W) = _problem(fixtures, oracle_engine, "benzene")
    H = _dense(ham, oracle_engine, psi, L, R, W)
    w, Z = np.linalg.eigh(H)
    th, u, its = oracle_engine.davidson(ham.symm, psi, L, R, W, max_iter=200, tol=1e-10, ortho=[2.0 * Z[:, 0]])
    assert abs(th - w[1]) < 1e-9
    assert abs(u @ Z[:, 0]) < 1e-9


@pytest.mark.parametrize("name,herm", [("lih", True), ("benzene", True), ("benzene", False)])
def test_diagonal_plan_emulation_matches_oracle(fixtures, oracle_engine, name, herm):
    import emulate
    from qcmaquis_amd.desc import SiteDesc
    from qcmaquis_amd.gpu import HostPlan
    from qcmaquis_amd.sweep import Hamiltonian, TwoSiteDMRG, micro_site, n_microiterations
    fx = fixtures[name]
    ham = Hamiltonian(fx["L"], fx["nelec"], fx["spin"], [(tuple(k), v) for k, v in fx["integrals"]], herm=herm)
    d = TwoSiteDMRG(ham, oracle_engine, 60)
    for i in range(n_microiterations(ham.L)):
        psi, L, R, W = d.site_problem(micro_site(ham.L, i))
        D = oracle_engine.diagonal(ham.symm, psi, L, R, W)
        hp = HostPlan(SiteDesc(ham.symm, psi.phys, psi.left, psi.right, W, L.layout, R.layout), "diag")
        De = emulate.run_diag(hp, L.data, R.data)
        assert np.abs(D - De).max() <= 1e-13 * max(1.0, np.abs(D).max()), i
        d.microiteration(0, i)


def test_davidson_is_su2_only(fixtures, oracle_engine):
    """MultDiagonal throws for non-SU2 symmetry groups (ietl_davidson.h:60-76)."""
    from qcmaquis_amd.desc import SiteDesc
    from qcmaquis_amd.gpu import HostPlan, MQError
    from qcmaquis_amd.sweep import Hamiltonian, TwoSiteDMRG
    fx = fixtures["lih"]
    ham = Hamiltonian(fx["L"], fx["nelec"], fx["spin"], [(tuple(k), v) for k, v in fx["integrals"]], symm="2u1")
    d = TwoSiteDMRG(ham, oracle_engine, 20)
    psi, L, R, W = d.site_problem(0)
    with pytest.raises(RuntimeError, match="spin-adapted"):
        oracle_engine.diagonal(ham.symm, psi, L, R, W)
    with pytest.raises(MQError, match="spin-adapted"):
        HostPlan(SiteDesc(ham.symm, psi.phys, psi.left, psi.right, W, L.layout, R.layout), "diag")


@pytest.mark.parametrize("name", ["h2", "lih"])
def test_sweeps_with_davidson_reach_golden_energies(fixtures, oracle_engine, name):
    """Two-site sweeps driven by the Davidson eigensolver (IETL_DAVIDSON) reach
    the reference's golden energies (tests/test1.cpp:93, tests/test2.cpp:57)."""
    from qcmaquis_amd.sweep import TwoSiteDMRG
    fx = fixtures[name]
    E = TwoSiteDMRG(ham_from_fixture(fx), oracle_engine, 100, eigensolver="davidson").run(2)
    assert abs(E - fx["energy"]) <= fx["rtol"] * abs(fx["energy"]), (E, fx["energy"])
