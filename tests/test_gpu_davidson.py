"""Diagonal of H_eff and the diagonal-preconditioned Davidson on the GPU
(mq_diag_plan_create / mq_diagonal / mq_davidson) against the CPU oracle and
the dense H_eff.

Bars: the diagonal <= 1e-13 relative to the oracle (FP64); Davidson
eigenvalues <= 1e-10 of the dense lowest eigenvalue in FP64 and <= 1e-8 in
the mixed form (FP64 defect correction)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(fixtures, oracle_engine, name, herm=True, m=40, nmicro=2, site=2):
    from qcmaquis_amd.sweep import Hamiltonian, TwoSiteDMRG
    fx = fixtures[name]
    ham = Hamiltonian(fx["L"], fx["nelec"], fx["spin"], [(tuple(k), v) for k, v in fx["integrals"]], herm=herm)
    d = TwoSiteDMRG(ham, oracle_engine, m)
    for i in range(nmicro):
        d.microiteration(0, i)
    return ham, d


def _dense(ham, oracle_engine, psi, L, R, W):
    n = psi.to_flat().size
    H = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        H[:, j], _ = oracle_engine.sigma(ham.symm, psi, L, R, W, x=e)
    return H


@pytest.mark.parametrize("name,herm", [("lih", True), ("benzene", True), ("benzene", False)])
def test_diagonal_gpu_vs_oracle(fixtures, oracle_engine, gpu_engine, name, herm):
    from qcmaquis_amd.sweep import micro_site, n_microiterations
    ham, d = _problem(fixtures, oracle_engine, name, herm=herm, nmicro=0, m=60)
    for i in range(n_microiterations(ham.L)):
        psi, L, R, W = d.site_problem(micro_site(ham.L, i))
        D0 = oracle_engine.diagonal(ham.symm, psi, L, R, W)
        Lg = gpu_engine.upload_boundary(L.layout, L.data)
        Rg = gpu_engine.upload_boundary(R.layout, R.data)
        print(f"diagonal {name} herm={herm} micro {i} n={D0.size}", flush=True)
        D1 = gpu_engine.diagonal(ham.symm, psi, Lg, Rg, W)
        assert np.abs(D1 - D0).max() <= 1e-13 * max(1.0, np.abs(D0).max()), i
        d.microiteration(0, i)


def test_diagonal_gpu_config0(oracle_engine, gpu_engine):
    from qcmaquis_amd.integrals import synthetic
    from qcmaquis_amd.sweep import Hamiltonian, TwoSiteDMRG, micro_site
    ham = Hamiltonian(8, 8, 0, synthetic(8, seed=1008))
    d = TwoSiteDMRG(ham, oracle_engine, 64)
    for i in range(6):
        psi, L, R, W = d.site_problem(micro_site(ham.L, i))
        D0 = oracle_engine.diagonal(ham.symm, psi, L, R, W)
        Lg = gpu_engine.upload_boundary(L.layout, L.data)
        Rg = gpu_engine.upload_boundary(R.layout, R.data)
        print(f"diagonal config0 micro {i} n={D0.size}", flush=True)
        D1 = gpu_engine.diagonal(ham.symm, psi, Lg, Rg, W)
        assert np.abs(D1 - D0).max() <= 1e-13 * max(1.0, np.abs(D0).max()), i
        d.microiteration(0, i)


def test_davidson_gpu_vs_dense_and_oracle(fixtures, oracle_engine, gpu_engine):
    """MPS guess and random guesses (through the restart at 20 vectors), FP64 and
    mixed; the ortho constraint selects the second eigenvalue."""
    from qcmaquis_amd.desc import MQ_PREC_FP32, MQ_PREC_FP64
    ham, d = _problem(fixtures, oracle_engine, "benzene")
    psi, L, R, W = d.site_problem(2)
    H = _dense(ham, oracle_engine, psi, L, R, W)
    w, Z = np.linalg.eigh(H)
    Lg = gpu_engine.upload_boundary(L.layout, L.data)
    Rg = gpu_engine.upload_boundary(R.layout, R.data)
    sp = gpu_engine.sigma_plan(ham.symm, psi, Lg, Rg, W)
    try:
        diag = sp.diagonal()
        restarted = 0
        for seed in (None, 0, 1, 2):
            x = psi.to_flat() if seed is None else np.random.default_rng(seed).standard_normal(w.size)
            t0, u0, i0 = oracle_engine.davidson(ham.symm, psi, L, R, W, max_iter=200, tol=1e-10, x0=x)
            print(f"davidson benzene seed {seed} n={w.size}", flush=True)
            xd = gpu_engine.to_dev(x)
            t1, u1, i1 = sp.davidson(xd, 200, 1e-10, prec=MQ_PREC_FP64, diag=diag)
            t2, _, _ = sp.davidson(xd, 200, 1e-10, prec=MQ_PREC_FP32, diag=diag)
            u1 = u1.cpu().numpy()
            restarted += i1 > 20
            assert abs(t1 - w[0]) <= 1e-10, (seed, t1 - w[0], i1)
            assert abs(t1 - t0) <= 1e-10, (seed, t1, t0)
            if seed is None:
                # from a random start the restarted iteration count follows the rounding
                # (the oracle's threaded sigma sums in varying order); from the MPS guess
                # it is stable
                assert abs(i1 - i0) <= 2, (i1, i0)
            assert i1 < 200
            assert np.linalg.norm(H @ u1 - t1 * u1) <= 1.01e-10 * abs(t1)
            assert abs(abs(u1 @ u0) - 1) <= 1e-9
            assert abs(t2 - w[0]) <= 1e-8, (seed, t2 - w[0])
        assert restarted >= 2
        t3, u3, _ = sp.davidson(gpu_engine.to_dev(psi.to_flat()), 200, 1e-10, ortho=[2.0 * Z[:, 0]],
                                prec=MQ_PREC_FP64, diag=diag)
        assert abs(t3 - w[1]) <= 1e-9
        assert abs(u3.cpu().numpy() @ Z[:, 0]) <= 1e-9
    finally:
        sp.close()


def test_davidson_gpu_rejects_twou1(fixtures, oracle_engine, gpu_engine):
    from qcmaquis_amd.gpu import MQError
    from qcmaquis_amd.sweep import Hamiltonian, TwoSiteDMRG
    fx = fixtures["lih"]
    ham = Hamiltonian(fx["L"], fx["nelec"], fx["spin"], [(tuple(k), v) for k, v in fx["integrals"]], symm="2u1")
    d = TwoSiteDMRG(ham, oracle_engine, 20)
    psi, L, R, W = d.site_problem(0)
    Lg = gpu_engine.upload_boundary(L.layout, L.data)
    Rg = gpu_engine.upload_boundary(R.layout, R.data)
    sp = gpu_engine.sigma_plan(ham.symm, psi, Lg, Rg, W)
    try:
        with pytest.raises(MQError, match="spin-adapted"):
            sp.diagonal()
        with pytest.raises(MQError, match="spin-adapted"):
            sp.davidson(gpu_engine.to_dev(psi.to_flat()), 10, 1e-8, diag=gpu_engine.to_dev(np.ones(psi.to_flat().size)))
    finally:
        sp.close()


@pytest.mark.parametrize("name", ["h2", "lih", "benzene"])
def test_golden_energies_gpu_davidson(fixtures, gpu_engine, name):
    from conftest import ham_from_fixture
    from qcmaquis_amd.sweep import TwoSiteDMRG
    fx = fixtures[name]
    E = TwoSiteDMRG(ham_from_fixture(fx), gpu_engine, 100, eigensolver="davidson").run(2)
    assert abs(E - fx["energy"]) <= fx["rtol"] * abs(fx["energy"]), (E, fx["energy"])


@pytest.mark.parametrize("nroots", [1, 3, 4])
def test_block_davidson_lowest_roots(fixtures, oracle_engine, gpu_engine, nroots):
    """The nroots lowest eigenvalues of the dense H_eff from random start vectors
    (through restarts), in FP64 and in the mixed form with the FP64 Rayleigh
    quotient per root; Ritz vectors orthonormal with small residuals."""
    from qcmaquis_amd.desc import MQ_PREC_FP32, MQ_PREC_FP64
    ham, d = _problem(fixtures, oracle_engine, "benzene")
    psi, L, R, W = d.site_problem(2)
    H = _dense(ham, oracle_engine, psi, L, R, W)
    w = np.linalg.eigvalsh(H)
    Lg = gpu_engine.upload_boundary(L.layout, L.data)
    Rg = gpu_engine.upload_boundary(R.layout, R.data)
    sp = gpu_engine.sigma_plan(ham.symm, psi, Lg, Rg, W)
    try:
        X0 = gpu_engine.torch.as_tensor(np.random.default_rng(nroots).standard_normal((nroots, w.size)))
        X0 = X0.to(gpu_engine.device)
        print(f"block davidson nroots {nroots} n={w.size}", flush=True)
        th, U, its = sp.block_davidson(X0, 200, 1e-10, prec=MQ_PREC_FP64)
        th32, _, _ = sp.block_davidson(X0, 200, 1e-10, prec=MQ_PREC_FP32)
        thd, _, itsd = sp.block_davidson(X0, 200, 1e-10, prec=MQ_PREC_FP64, diag=sp.diagonal())
    finally:
        sp.close()
    U = U.cpu().numpy()
    print(f"block iterations: {its} (JD corrections), {itsd} (diagonal preconditioner)", flush=True)
    assert its < 200 and itsd < 200
    assert np.abs(np.array(thd) - w[:nroots]).max() <= 1e-9, (thd, w[:nroots])
    assert np.abs(np.array(th) - w[:nroots]).max() <= 1e-9, (th, w[:nroots])
    assert np.abs(np.array(th32) - w[:nroots]).max() <= 1e-8, (th32, w[:nroots])
    assert np.abs(U @ U.T - np.eye(nroots)).max() <= 1e-10
    for r in range(nroots):
        assert np.linalg.norm(H @ U[r] - th[r] * U[r]) <= 1.01e-10 * abs(th[r])


def test_mixed_solvers_refine_in_fp64(fixtures, oracle_engine, gpu_engine):
    """A tolerance below the mixed form's rounding floor: after the mixed
    iterations the FP64 residual misses it, the solve continues in FP64, and
    the eigenvalue equals the FP64 solve's (JD, Davidson, block Davidson)."""
    from qcmaquis_amd.desc import MQ_PREC_FP32, MQ_PREC_FP64
    ham, d = _problem(fixtures, oracle_engine, "benzene")
    psi, L, R, W = d.site_problem(2)
    w = np.linalg.eigvalsh(_dense(ham, oracle_engine, psi, L, R, W))
    Lg = gpu_engine.upload_boundary(L.layout, L.data)
    Rg = gpu_engine.upload_boundary(R.layout, R.data)
    sp = gpu_engine.sigma_plan(ham.symm, psi, Lg, Rg, W)
    try:
        x = gpu_engine.to_dev(psi.to_flat())
        t64, _, i64 = sp.jd(x, 60, 1e-13, prec=MQ_PREC_FP64)
        t32, _, i32 = sp.jd(x, 60, 1e-13, prec=MQ_PREC_FP32)
        d64, _, _ = sp.davidson(x, 200, 1e-13, prec=MQ_PREC_FP64)
        d32, _, _ = sp.davidson(x, 200, 1e-13, prec=MQ_PREC_FP32)
        X0 = gpu_engine.torch.as_tensor(np.random.default_rng(2).standard_normal((2, w.size))).to(gpu_engine.device)
        b32, _, _ = sp.block_davidson(X0, 200, 1e-13, prec=MQ_PREC_FP32)
    finally:
        sp.close()
    print(f"JD iterations fp64 {i64}, mixed + fp64 refinement {i32}", flush=True)
    assert abs(t32 - w[0]) <= 1e-11 and abs(t64 - w[0]) <= 1e-11
    assert abs(d32 - w[0]) <= 1e-11 and abs(d64 - w[0]) <= 1e-11
    assert np.abs(np.array(b32) - w[:2]).max() <= 1e-11
