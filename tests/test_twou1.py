"""The spin-orbital (U(1) x U(1)) model: Hamiltonian MPO, oracle sweeps and the
abelian task lists of the GPU plans, all on the CPU.

Pinned by: brute-force FCI in every (N_alpha, N_beta) sector of the dense
MPO (H2, LiH, L=3 synthetic); the reference's golden energies of H2
(dmrg/tests/test1.cpp:93, all symmetries share one value), LiH
(dmrg/tests/test2.cpp:57) and benzene (dmrg/tests/FEAST/ElectronicFEAST.cpp:192);
lbtm vs rbtm-free emulation of the abelian plans against the oracle
(contractions/abelian/site_hamil.hpp:15-87)."""
import itertools

import numpy as np
import pytest

from conftest import ham_from_fixture


def _ham2u1(fx):
    from qcmaquis_amd.sweep import Hamiltonian
    return Hamiltonian(fx["L"], fx["nelec"], fx["spin"], [(tuple(k), v) for k, v in fx["integrals"]], symm="2u1")


def _dense_mpo(ham):
    """Full operator of the MPO on the product basis (test infrastructure)."""
    chg = [c for c, _ in ham.phys]
    pos = {c: i for i, c in enumerate(chg)}
    d = len(chg)
    acc = None
    for W in ham.mpo:
        T = np.zeros((W.row_dim, W.col_dim, d, d))
        for (b1, b2), terms in W.entries.items():
            for tag, sc in terms:
                for (ci, co), m in W.table[tag].blocks.items():
                    T[b1, b2, pos[co], pos[ci]] += sc * m[0, 0]
        if acc is None:
            acc = T[0]
        else:
            n = acc.shape[1]
            acc = np.einsum("lij,lrab->riajb", acc, T).reshape(T.shape[1], n * d, n * d)
    H = acc[0] + ham.ecore * np.eye(acc.shape[1])
    occ = [(sum(chg[s][0] for s in st), sum(chg[s][1] for s in st))
           for st in itertools.product(range(d), repeat=len(ham.mpo))]
    return H, occ


@pytest.mark.parametrize("name", ["h2", "lih", "synthetic3"])
def test_twou1_mpo_spectrum_equals_fci(fixtures, name):
    from fci import fci_energies
    from qcmaquis_amd.integrals import dense_tensors, synthetic
    from qcmaquis_amd.sweep import Hamiltonian
    if name == "synthetic3":
        L, ints = 3, synthetic(3, seed=1003)
        ham = Hamiltonian(3, 3, 1, ints, symm="2u1")
    else:
        fx = fixtures[name]
        L, ints = fx["L"], [(tuple(k), v) for k, v in fx["integrals"]]
        ham = _ham2u1(fx)
    H, occ = _dense_mpo(ham)
    assert np.abs(H - H.T).max() < 1e-12
    h, g, ec = dense_tensors(L, ints)
    checked = 0
    for na in range(L + 1):
        for nb in range(L + 1):
            idx = [i for i, o in enumerate(occ) if o == (na, nb)]
            if len(idx) < 2:
                continue
            e = np.linalg.eigvalsh(H[np.ix_(idx, idx)])
            ef = np.sort(fci_energies(h, g, ec, L, na, nb, k=len(idx)))
            assert np.abs(e[:len(ef)] - ef).max() < 1e-10, (na, nb)
            checked += 1
    assert checked >= 3


@pytest.mark.parametrize("name", ["h2", "lih", "benzene"])
def test_twou1_golden_energy_oracle(fixtures, oracle_engine, name):
    from qcmaquis_amd.sweep import TwoSiteDMRG
    fx = fixtures[name]
    E = TwoSiteDMRG(_ham2u1(fx), oracle_engine, 100).run(2)
    assert abs(E - fx["energy"]) <= fx["rtol"] * abs(fx["energy"]), (E, fx["energy"])


def test_twou1_plan_emulation_matches_oracle(fixtures, oracle_engine):
    import emulate
    from qcmaquis_amd.desc import SiteDesc
    from qcmaquis_amd.gpu import HostPlan
    from qcmaquis_amd.sweep import TwoSiteDMRG, micro_site, n_microiterations
    fx = fixtures["benzene"]
    ham = _ham2u1(fx)
    d = TwoSiteDMRG(ham, oracle_engine, 40)
    for i in range(n_microiterations(ham.L)):
        site = micro_site(ham.L, i)
        psi, Lb, Rb, W = d.site_problem(site)
        x = np.random.default_rng(i).standard_normal(psi.to_flat().size)
        y, fl = oracle_engine.sigma(ham.symm, psi, Lb, Rb, W, x=x)
        for fuse in (0, 1, 2):
            hp = HostPlan(SiteDesc(ham.symm, psi.phys, psi.left, psi.right, W, Lb.layout, Rb.layout), "sigma", fuse)
            ye = emulate.run_sigma(hp, Lb.data, Rb.data, x)
            assert np.abs(y - ye).max() <= 1e-13 * max(1.0, np.abs(y).max())
            assert hp.ref_flops == fl
        for kind, A, B, Wm in (("lstep", d.mps[site], Lb, ham.mpo[site]),
                               ("rstep", d.mps[site + 1], Rb, ham.mpo[site + 1])):
            if kind == "lstep":
                ref = oracle_engine.left_step(ham.symm, A, B, Wm)
                sd = SiteDesc(ham.symm, A.phys, A.left, A.right, Wm, B.layout, None)
            else:
                ref = oracle_engine.right_step(ham.symm, A, B, Wm)
                sd = SiteDesc(ham.symm, A.phys, A.left, A.right, Wm, None, B.layout)
            hps = HostPlan(sd, kind)
            got = hps.out_layout.to_dicts(emulate.run_step(hps, kind, B.data, A.to_flat()))
            exp = ref.dicts()
            for b in range(len(exp)):
                assert set(got[b]) == set(exp[b])
                for k in exp[b]:
                    assert np.abs(got[b][k] - exp[b][k]).max() <= 1e-13 * max(1.0, np.abs(exp[b][k]).max())
        d.microiteration(0, i)


def test_twou1_and_su2_energies_agree(oracle_engine):
    """One Hamiltonian, two symmetries: the singlet ground state of the L=6
    synthetic system is the lowest state of the N_alpha = N_beta sector."""
    from qcmaquis_amd.integrals import synthetic
    from qcmaquis_amd.sweep import Hamiltonian, TwoSiteDMRG
    ints = synthetic(6, seed=1006)
    e_su2 = TwoSiteDMRG(Hamiltonian(6, 6, 0, ints), oracle_engine, 200).run(3)
    e_2u1 = TwoSiteDMRG(Hamiltonian(6, 6, 0, ints, symm="2u1"), oracle_engine, 400).run(3)
    assert abs(e_su2 - e_2u1) < 1e-9, (e_su2, e_2u1)
