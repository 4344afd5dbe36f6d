"""RDM measurements through the HIP boundary steps: the one-particle RDM and the
orbit-averaged two-particle RDM of a DMRG state measured with the GPU engine
equal the ones the CPU oracle engine measures on the same MPS (FP64 boundary
kernels, so to round-off), and the energy they give is the state's energy."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_rdms_gpu_vs_oracle(gpu_engine, oracle_engine):
    from qcmaquis_amd.integrals import dense_tensors, synthetic
    from qcmaquis_amd.measure import rdm1, rdm2
    from qcmaquis_amd.sweep import Hamiltonian, TwoSiteDMRG, expval
    L = 6
    ints = synthetic(L, seed=606)
    h, g, ec = dense_tensors(L, ints)
    ham = Hamiltonian(L, L, 0, ints)
    d = TwoSiteDMRG(ham, gpu_engine, 32, jd_tol=1e-10, jd_maxiter=30)
    d.run(2)
    print(f"rdm L={L} m=32", flush=True)
    g1 = rdm1(gpu_engine, ham, d.mps)
    o1 = rdm1(oracle_engine, ham, d.mps)
    assert np.abs(g1 - o1).max() < 1e-11, np.abs(g1 - o1).max()
    assert abs(np.trace(g1) - ham.nelec) < 1e-9
    g2 = rdm2(gpu_engine, ham, d.mps)
    o2 = rdm2(oracle_engine, ham, d.mps)
    assert np.abs(g2 - o2).max() < 1e-11, np.abs(g2 - o2).max()
    E = expval(gpu_engine, ham, d.mps)
    e_rdm = np.einsum("pq,pq->", h, g1) + 0.5 * np.einsum("pqrs,pqrs->", g, g2) + ec
    assert abs(e_rdm - E) < 1e-9, (e_rdm, E)
