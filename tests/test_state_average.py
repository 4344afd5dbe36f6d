"""State-averaged two-site DMRG on the CPU (the oracle engine): the shared-basis
split against its definition (the leading eigenvectors of the roots' averaged
reduced density matrix) and whole sweeps against brute-force FCI."""
import numpy as np
import pytest


def _fci_roots(L, seed, k):
    """Lowest Sz=0 energies, and the singlets among them (those without an Sz=1
    partner), of the synthetic L-orbital, L-electron system."""
    from fci import fci_energies
    from qcmaquis_amd.integrals import dense_tensors, synthetic
    ints = synthetic(L, seed=seed)
    h, g, ec = dense_tensors(L, ints)
    e0 = np.sort(fci_energies(h, g, ec, L, L // 2, L // 2, k=3 * k))
    e1 = np.sort(fci_energies(h, g, ec, L, L // 2 + 1, L // 2 - 1, k=3 * k))
    singlets = [e for e in e0 if np.abs(e1 - e).min() > 1e-7]
    return ints, e0, np.array(singlets)


def _site_tensors(oracle_engine, k, seed=3):
    """k random two-site tensors of one LiH-like site problem."""
    from qcmaquis_amd.integrals import synthetic
    from qcmaquis_amd.mps import MPSTensor
    from qcmaquis_amd.sweep import Hamiltonian, TwoSiteDMRG
    ham = Hamiltonian(6, 6, 0, synthetic(6, seed=11))
    d = TwoSiteDMRG(ham, oracle_engine, 30)
    d.microiteration(0, 0)
    psi, L, R, W = d.site_problem(1)
    rng = np.random.default_rng(seed)
    us = [MPSTensor.from_flat(psi.phys, psi.left, psi.right, rng.standard_normal(psi.to_flat().size))
          for _ in range(k)]
    return ham, d, psi, us


def test_sa_split_of_one_root_is_the_plain_split(oracle_engine):
    from qcmaquis_amd.mps import split_twosite, split_twosite_sa
    ham, d, psi, us = _site_tensors(oracle_engine, 1)
    left, right = d.mps[1].left, d.mps[2].right
    for l2r in (True, False):
        A, B, info = split_twosite(us[0].copy(), ham.phys, ham.phys, left, right, 12, 1e-16, l2r=l2r)
        if l2r:
            A2, Bs, info2 = split_twosite_sa([us[0].copy()], ham.phys, ham.phys, left, right, 12, 1e-16, l2r=True)
            B2 = Bs[0]
        else:
            As, B2, info2 = split_twosite_sa([us[0].copy()], ham.phys, ham.phys, left, right, 12, 1e-16, l2r=False)
            A2 = As[0]
        assert info["bond_dimension"] == info2["bond_dimension"]
        for X, Y in ((A, A2), (B, B2)):
            X.make_left_paired(), Y.make_left_paired()
            assert set(X.data.blocks) == set(Y.data.blocks)
            for key in X.data.blocks:
                assert np.abs(X.data.blocks[key] - Y.data.blocks[key]).max() < 1e-12


@pytest.mark.parametrize("l2r", [True, False])
def test_sa_split_keeps_the_averaged_density_matrix_eigenvectors(oracle_engine, l2r):
    """l2r: the shared left tensor spans the leading eigenvectors of
    sum_r psi_r psi_r^T / k in every symmetry block (r2l: of sum_r psi_r^T psi_r / k),
    and every root keeps exactly its projection onto that space."""
    from qcmaquis_amd.mps import split_twosite_sa, ts_left_to_both, unreduce_left
    k, m = 3, 6
    ham, d, psi, us = _site_tensors(oracle_engine, k)
    left, right = d.mps[1].left, d.mps[2].right
    boths = []
    for u in us:
        v = u.copy().make_left_paired()
        boths.append(ts_left_to_both(ham.phys, ham.phys, left, right,
                                     unreduce_left(ham.phys, ham.phys, left, right, v.data)))
    rho_ev = {}
    for key in boths[0].blocks:
        mats = [b.blocks[key] for b in boths]
        rho = sum(x @ x.T for x in mats) / k if l2r else sum(x.T @ x for x in mats) / k
        rho_ev[key] = np.linalg.eigh(rho)
    allw = np.sort(np.concatenate([w for w, _ in rho_ev.values()]))[::-1]
    res = split_twosite_sa([u.copy() for u in us], ham.phys, ham.phys, left, right, m, 1e-16, l2r=l2r)
    if l2r:
        A, Bs, info = res
        shared = {key: A.data.blocks[key] for key in A.data.blocks}
    else:
        As, B, info = res
        shared = {key: B.data.blocks[key].T for key in B.data.blocks}
    # the reference's rule keeps every value >= the (m+1)-th largest (estimate_truncation)
    nkeep = int(np.sum(allw >= allw[m]))
    assert info["bond_dimension"] == nkeep == m + 1
    assert abs(info["truncated_weight"] - allw[nkeep:].sum() / allw.sum()) < 1e-12
    kept = 0
    for key, Q in shared.items():
        w, Z = rho_ev[key]
        n = Q.shape[1]
        kept += n
        assert np.abs(Q.T @ Q - np.eye(n)).max() < 1e-12
        top = Z[:, -n:]
        # same subspace: the projectors agree
        assert np.abs(Q @ Q.T - top @ top.T).max() < 1e-10
        assert np.all(w[-n:] >= allw[m] - 1e-12)
    assert kept == nkeep
    for r, b in enumerate(boths):
        for key, Q in shared.items():
            if l2r:
                proj = Q @ (Bs[r].data.blocks[key] if Bs[r].pairing == "L" else Bs[r].make_right_paired().data.blocks[key])
                ref = Q @ (Q.T @ b.blocks[key])
            else:
                Ar = As[r].make_left_paired()
                proj = Ar.data.blocks[key] @ Q.T
                ref = b.blocks[key] @ Q @ Q.T
            assert np.abs(proj - ref).max() < 1e-12


def test_sa_sweeps_reach_fci_singlets(oracle_engine):
    """At full bond dimension the three state-averaged roots of the synthetic
    L=8, N=8, S=0 system are the three lowest FCI singlets."""
    from qcmaquis_amd.sweep import Hamiltonian, StateAveragedDMRG
    ints, e0, singlets = _fci_roots(8, 1008, 3)
    d = StateAveragedDMRG(Hamiltonian(8, 8, 0, ints), oracle_engine, 400, 3, jd_maxiter=40, jd_tol=1e-10)
    E = np.array(d.run(3))
    assert np.abs(E - singlets[:3]).max() < 1e-9, (E, singlets[:3])


def test_sa_sweeps_twou1_reach_lowest_sz0_states(oracle_engine):
    """Under TwoU1 the roots are the lowest Sz=0 states, triplet components included."""
    from qcmaquis_amd.sweep import Hamiltonian, StateAveragedDMRG
    ints, e0, singlets = _fci_roots(6, 606, 4)
    d = StateAveragedDMRG(Hamiltonian(6, 6, 0, ints, symm="2u1"), oracle_engine, 400, 4, jd_maxiter=40,
                          jd_tol=1e-10)
    E = np.array(d.run(3))
    assert np.abs(E - e0[:4]).max() < 1e-9, (E, e0[:4])


def test_sa_truncated_roots_are_upper_bounds(oracle_engine):
    """With a truncated shared basis every root stays above its exact singlet
    (min-max principle for the Ritz values of a subspace), and a larger basis
    lowers every root."""
    from qcmaquis_amd.sweep import Hamiltonian, StateAveragedDMRG
    ints, e0, singlets = _fci_roots(8, 1008, 3)
    Es = []
    for m in (12, 24):
        d = StateAveragedDMRG(Hamiltonian(8, 8, 0, ints), oracle_engine, m, 3, jd_maxiter=40, jd_tol=1e-10)
        E = np.array(d.run(2))
        assert np.all(np.diff(E) > 0)
        assert np.all(E > singlets[:3]), (m, E, singlets[:3])
        assert max(r["truncated_weight"] for r in d.records[-4:]) > 0
        Es.append(E)
    assert np.all(Es[1] < Es[0])


def test_independent_starts_replaces_dependent_vectors():
    from qcmaquis_amd.gpu import MQError, independent_starts
    rng = np.random.default_rng(0)
    a, b = rng.standard_normal(30), rng.standard_normal(30)
    X = independent_starts([a, b, 2 * a - b, np.zeros(30)])
    assert np.abs(X @ X.T - np.eye(4)).max() < 1e-12
    assert abs(abs(X[0] @ a) - np.linalg.norm(a)) < 1e-12
    with pytest.raises(MQError):
        independent_starts([rng.standard_normal(2) for _ in range(3)])
