"""The live-coordinate mask of a sigma plan (mq_plan_live_mask) on the CPU.

The reduced two-site tensor has slots that no operator term reaches: the
spin-1 pair of two singly occupied sites next to a spin-0 bond state, which the
reference's two-site reduction never fills (ts_reduction.h:103-105) and whose
contributions its lbtm kernel filters (apply_op.hpp:75).  Their rows and
columns of H_eff are zero, so their eigenvalue 0 lies below every physical root
of a Hamiltonian whose spectrum is positive, and a block solver started from
vectors with weight there converges to them.  The mask must be exactly the
zero columns of the dense H_eff built with the oracle.
"""
import numpy as np
import pytest


def _dense_heff(ora, sp):
    n = sp.psi.to_flat().size
    H = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        H[:, i], _ = ora.sigma(sp.ham.symm, sp.psi, sp.L, sp.R, sp.W, x=e)
    return H


@pytest.mark.parametrize("L,m", [(6, 16), (8, 24)])
def test_live_mask_is_the_nonzero_columns_of_heff(oracle_engine, L, m):
    from qcmaquis_amd.desc import SiteDesc
    from qcmaquis_amd.gpu import HostPlan
    from qcmaquis_amd.workload import centre_problem
    sp = centre_problem(oracle_engine, L, m=m)
    psi = sp.psi
    H = _dense_heff(oracle_engine, sp)
    zero_cols = np.all(H == 0.0, axis=0)
    assert np.array_equal(zero_cols, np.all(H == 0.0, axis=1))
    assert zero_cols.sum() > 0  # the workload has such slots (11 at L=6, 34 at L=8)
    for fuse in (0, 1, 2):
        hp = HostPlan(SiteDesc(sp.ham.symm, psi.phys, psi.left, psi.right, sp.W, sp.L.layout, sp.R.layout),
                      "sigma", fuse)
        live = hp.parts["live"].astype(bool)
        assert live.size == psi.to_flat().size
        assert np.array_equal(~live, zero_cols), fuse
    # the MPS guess has no weight there; the dense spectrum is the spectrum on the
    # live coordinates plus one zero per dead slot
    assert np.abs(psi.to_flat()[~live]).max() == 0.0
    ev = np.linalg.eigvalsh(H)
    evl = np.linalg.eigvalsh(H[np.ix_(live, live)])
    assert np.allclose(np.sort(np.concatenate([evl, np.zeros(zero_cols.sum())])), ev, atol=1e-10)


def test_independent_starts_respect_the_mask():
    from qcmaquis_amd.gpu import independent_starts
    rng = np.random.default_rng(3)
    n = 200
    live = rng.random(n) > 0.1
    v0 = np.where(live, rng.standard_normal(n), 0.0)
    # the same vector three times: the repeats are replaced by random vectors, which
    # must vanish on the dead coordinates as the given one does
    X = independent_starts([v0, v0, v0, rng.standard_normal(n)], seed=1, live=live)
    assert X.shape == (4, n)
    assert np.all(X[:, ~live] == 0.0)
    assert np.allclose(X @ X.T, np.eye(4), atol=1e-12)
