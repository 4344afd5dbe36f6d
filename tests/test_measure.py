"""RDM measurements (qcmaquis_amd.measure) on the CPU oracle engine against
brute-force FCI: the spin-summed one-particle RDM and the orbit-averaged
two-particle RDM of the DMRG ground state at full bond dimension equal the FCI
ones, their traces are fixed by the electron number, and the energy they give
is the DMRG energy."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def l4_ground(oracle_engine):
    from fci import fci_ground_rdms
    from qcmaquis_amd.integrals import dense_tensors, synthetic
    from qcmaquis_amd.sweep import Hamiltonian, TwoSiteDMRG
    L = 4
    ints = synthetic(L, seed=404)
    h, g, ec = dense_tensors(L, ints)
    ham = Hamiltonian(L, L, 0, ints)
    d = TwoSiteDMRG(ham, oracle_engine, 64, jd_tol=1e-12, jd_maxiter=40)
    E = d.run(4)
    return ham, d, E, (h, g, ec), fci_ground_rdms(h, g, ec, L, L // 2, L // 2)


def test_rdm1_matches_fci(oracle_engine, l4_ground):
    from qcmaquis_amd.measure import rdm1
    ham, d, E, _, (e_fci, gamma_fci, _) = l4_ground
    assert abs(E - e_fci) < 1e-9
    gamma = rdm1(oracle_engine, ham, d.mps)
    assert np.abs(gamma - gamma.T).max() < 1e-12
    assert abs(np.trace(gamma) - ham.nelec) < 1e-9
    assert np.abs(gamma - gamma_fci).max() < 1e-6, np.abs(gamma - gamma_fci).max()


def test_rdm2_matches_fci_and_energy(oracle_engine, l4_ground):
    from qcmaquis_amd.measure import rdm1, rdm2
    ham, d, E, (h, g, ec), (e_fci, gamma_fci, Gamma_fci) = l4_ground
    Gamma = rdm2(oracle_engine, ham, d.mps)
    # the FCI Gamma averaged over the 8-fold orbits of (pq|rs)
    sym = (Gamma_fci + Gamma_fci.transpose(1, 0, 2, 3) + Gamma_fci.transpose(0, 1, 3, 2)
           + Gamma_fci.transpose(1, 0, 3, 2)) / 4
    sym = (sym + sym.transpose(2, 3, 0, 1)) / 2
    assert np.abs(Gamma - sym).max() < 1e-6, np.abs(Gamma - sym).max()
    assert np.abs(Gamma_fci - Gamma_fci.transpose(2, 3, 0, 1)).max() < 1e-12  # the 4-fold symmetry
    assert np.abs(Gamma_fci - Gamma_fci.transpose(1, 0, 3, 2)).max() < 1e-12
    # sum_pq Gamma_ppqq = N (N - 1); the energy from the RDMs is the DMRG energy
    N = ham.nelec
    assert abs(np.einsum("ppqq->", Gamma) - N * (N - 1)) < 1e-8
    gamma = rdm1(oracle_engine, ham, d.mps)
    e_rdm = np.einsum("pq,pq->", h, gamma) + 0.5 * np.einsum("pqrs,pqrs->", g, Gamma) + ec
    assert abs(e_rdm - E) < 1e-9, (e_rdm, E)
