"""The CPU oracle pinned against the reference's golden values, closed forms
and invariants (mirrors dmrg/tests/test_wigner.cpp, test1.cpp, test2.cpp,
FEAST/ElectronicFEAST.cpp, test_mps_mpo_ops/test_siteproblem.cpp and
SweepOptimizationTools/BoundaryPropagatorElectronic.cpp)."""
import math

import numpy as np
import pytest

from conftest import ham_from_fixture


def test_wigner9j_golden(shapes):
    import pyoracle
    from qcmaquis_amd import wigner
    tab = shapes["wigner9j"]
    assert len(tab["cases"]) == 24
    for args, val in tab["cases"]:
        for got in (pyoracle.wigner9j(*args), wigner.ninej(*args)):
            if val == 0:
                assert abs(got) < 1e-14
            else:
                assert abs(got - val) <= tab["rtol"] * abs(val), (args, got, val)


def test_wigner6j_closed_forms():
    """{a b c; 0 c b} = (-1)^{(a+b+c)/2} / sqrt((b+1)(c+1)) and 6j orthogonality."""
    import pyoracle
    for a in range(0, 7):
        for b in range(0, 7):
            for c in range(abs(a - b), a + b + 1, 2):
                ref = (-1) ** ((a + b + c) // 2) / math.sqrt((b + 1) * (c + 1))
                assert abs(pyoracle.wigner6j(a, b, c, 0, c, b) - ref) < 1e-13
    # sum_x (x+1)(f+1) {a b x; c d f}{c d x; a b f'} = delta_ff'
    a, b, c, d = 3, 2, 3, 4
    fs = [f for f in range(0, 12) if pyoracle.wigner6j(a, b, 1, c, d, f) != 0 or True]
    for f in range(0, 10):
        for fp in range(0, 10):
            s = 0.0
            for x in range(0, 12):
                s += (x + 1) * (f + 1) * pyoracle.wigner6j(a, b, x, c, d, f) * pyoracle.wigner6j(c, d, x, a, b, fp)
            valid = ((a + d + f) % 2 == 0 and abs(a - d) <= f <= a + d and (b + c + f) % 2 == 0
                     and abs(b - c) <= f <= b + c)
            exp = 1.0 if (f == fp and valid) else 0.0
            assert abs(s - exp) < 1e-12, (f, fp, s)
    assert fs


def test_wigner9j_zero_reduces_to_6j():
    """{a b c; d e f; g h 0} = delta_cf delta_gh (-1)^{(b+c+d+g)/2} / sqrt((c+1)(g+1)) {a b c; e d g}."""
    import pyoracle
    for a, b, c, d, e, g in [(1, 1, 2, 1, 1, 2), (2, 1, 3, 1, 2, 3), (3, 2, 1, 1, 2, 2), (2, 2, 2, 2, 2, 2)]:
        v = pyoracle.wigner9j(a, b, c, d, e, c, g, g, 0)
        ref = (-1) ** ((b + c + d + g) // 2) / math.sqrt((c + 1) * (g + 1)) * pyoracle.wigner6j(a, b, c, e, d, g)
        assert abs(v - ref) < 1e-13


def test_mpo_shape_benzene(shapes):
    """Term count, MPO bond dimensions/hermitian pairs and two-site MPO shapes of the
    reference's benzene su2u1 run (examples/TD-DMRG/Benzene_6e6o.TD.TS.out:15-28)."""
    import os
    from conftest import GOLDEN
    from qcmaquis_amd.integrals import read_fcidump
    from qcmaquis_amd.sweep import Hamiltonian
    d = read_fcidump(os.path.join(GOLDEN, "FCIDUMP_benzene_cas66"))
    ham = Hamiltonian(d["norb"], d["nelec"], 0, d["integrals"])
    ref = shapes["benzene_su2_ts"]
    assert len(ham.terms) == ref["n_terms"]
    got = [[w.col_dim, sum(1 for b in range(w.col_dim) if w.right_herm[b] > b)] for w in ham.mpo]
    assert got == ref["bonds"]
    ts = []
    for p in range(ham.L - 1):
        w = ham.ts_mpo(p)
        ts.append([w.row_dim, w.col_dim, sum(len(v) for v in w.entries.values())])
    assert ts == ref["ts_mpo"]


@pytest.mark.parametrize("name,m,sweeps", [("h2", 100, 2), ("lih", 100, 2), ("benzene", 100, 2)])
def test_golden_energy_oracle(fixtures, oracle_engine, name, m, sweeps):
    from qcmaquis_amd.sweep import TwoSiteDMRG
    fx = fixtures[name]
    ham = ham_from_fixture(fx)
    E = TwoSiteDMRG(ham, oracle_engine, m).run(sweeps)
    assert abs(E - fx["energy"]) <= fx["rtol"] * abs(fx["energy"]), (E, fx["energy"])


def _dmrg_state(ham, engine, m, nmicro):
    from qcmaquis_amd.sweep import TwoSiteDMRG
    d = TwoSiteDMRG(ham, engine, m)
    for i in range(nmicro):
        d.microiteration(0, i)
    return d


def test_sigma_schedules_and_hermiticity(fixtures, oracle_engine):
    """lbtm and rbtm schedules agree; <a|H b> = <H a|b>."""
    fx = fixtures["benzene"]
    ham = ham_from_fixture(fx)
    d = _dmrg_state(ham, oracle_engine, 40, 2)
    psi, L, R, W = d.site_problem(2)
    x = psi.to_flat()
    rng = np.random.default_rng(5)
    a = rng.standard_normal(x.size)
    b = rng.standard_normal(x.size)
    ya, _ = oracle_engine.sigma(ham.symm, psi, L, R, W, x=a, schedule=0)
    ya1, _ = oracle_engine.sigma(ham.symm, psi, L, R, W, x=a, schedule=1)
    yb, _ = oracle_engine.sigma(ham.symm, psi, L, R, W, x=b, schedule=0)
    assert np.abs(ya - ya1).max() <= 1e-13 * np.abs(ya).max()
    assert abs(b @ ya - a @ yb) <= 1e-12 * abs(b @ ya)


def test_siteproblem_energy_equals_boundary_energy(fixtures, oracle_engine):
    """<psi|H_eff|psi> + E_core equals the energy from the fully propagated left
    boundary (test_siteproblem.cpp:84-94, BoundaryPropagatorElectronic.cpp:56-65)."""
    from qcmaquis_amd.mps import left_normalize_and_return, multiply_from_left
    fx = fixtures["lih"]
    ham = ham_from_fixture(fx)
    d = _dmrg_state(ham, oracle_engine, 100, 1)
    psi, L, R, W = d.site_problem(1)
    x = psi.to_flat()
    y, _ = oracle_engine.sigma(ham.symm, psi, L, R, W, x=x)
    e_site = (x @ y) / (x @ x) + ham.ecore
    # propagate left boundaries through the whole (left-normalised) MPS
    mps = [t.copy() for t in d.mps]
    Lb = oracle_engine.seed_boundary((0, 0))
    for i in range(ham.L):
        if i < ham.L - 1:
            Rf = left_normalize_and_return(mps[i])
            multiply_from_left(mps[i + 1], Rf)
        Lb = oracle_engine.left_step(ham.symm, mps[i], Lb, ham.mpo[i])
    assert Lb.layout.aux == 1 and Lb.layout.nblocks == 1
    norm = mps[-1].norm() ** 2
    e_bnd = float(Lb.data[0]) / norm + ham.ecore
    assert abs(e_site - e_bnd) <= 1e-10 * abs(e_bnd), (e_site, e_bnd)


def test_jd_matches_dense_eigensolver(fixtures, oracle_engine):
    """Jacobi-Davidson lowest eigenvalue vs the dense H_eff built column by column."""
    fx = fixtures["lih"]
    ham = ham_from_fixture(fx)
    d = _dmrg_state(ham, oracle_engine, 100, 1)
    psi, L, R, W = d.site_problem(1)
    n = psi.to_flat().size
    H = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        H[:, j], _ = oracle_engine.sigma(ham.symm, psi, L, R, W, x=e)
    assert np.abs(H - H.T).max() < 1e-12
    w = np.linalg.eigvalsh(H)
    theta, u, its = oracle_engine.jd(ham.symm, psi, L, R, W, max_iter=min(n, 40), tol=1e-12)
    assert abs(theta - w[0]) < 1e-10


def test_config0_dmrg_reaches_fci(oracle_engine):
    """Synthetic L=8, N=8, S=0 system: DMRG at full bond dimension equals brute-force FCI."""
    from fci import fci_energies
    from qcmaquis_amd.integrals import dense_tensors, synthetic
    from qcmaquis_amd.sweep import Hamiltonian, TwoSiteDMRG
    L = 8
    ints = synthetic(L, seed=1008)
    h, g, ec = dense_tensors(L, ints)
    e0 = fci_energies(h, g, ec, L, 4, 4, k=2)[0]
    e1 = fci_energies(h, g, ec, L, 5, 3, k=1)[0]
    assert e0 < e1 - 1e-6  # the Sz=0 ground state is a singlet
    ham = Hamiltonian(L, L, 0, ints)
    E = TwoSiteDMRG(ham, oracle_engine, 400).run(3)
    assert abs(E - e0) < 1e-9, (E, e0)


@pytest.mark.parametrize("symm", ["su2u1", "2u1"])
def test_expval_equals_sweep_energy(fixtures, oracle_engine, symm):
    """expval (mps_mpo_ops.h:42-103) of the converged, untruncated LiH MPS equals
    the final sweep energy, from the left and from the right."""
    from qcmaquis_amd.sweep import Hamiltonian, TwoSiteDMRG, expval
    fx = fixtures["lih"]
    ham = Hamiltonian(fx["L"], fx["nelec"], fx["spin"], [(tuple(k), v) for k, v in fx["integrals"]], symm=symm)
    d = TwoSiteDMRG(ham, oracle_engine, 200)
    E = d.run(2)
    e_l = expval(oracle_engine, ham, d.mps)
    e_r = expval(oracle_engine, ham, d.mps, from_right=True)
    assert abs(e_l - E) <= 1e-10 * abs(E), (e_l, E)
    assert abs(e_r - E) <= 1e-10 * abs(E), (e_r, E)
