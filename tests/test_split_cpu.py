"""The two-site split's gather map (mq_split, built by the library's host code)
on the CPU: the both-paired blocks it produces from a random two-site tensor
equal those of the host restatement (mps.unreduce_left for SU2U1, then
mps.ts_left_to_both), block for block, in descending charge order; and the
truncation rule shared by the host and device splits keeps what
estimate_truncation keeps."""
import numpy as np
import pytest


def _problem(symm, L, m, site, seed):
    from qcmaquis_amd.mps import bounded_sectors, make_twosite_mps, random_mps_on
    if symm in ("su2u1", "stretched"):
        from qcmaquis_amd.operators import su2_phys_index
        phys, end, su2 = su2_phys_index(), (L, 0), True
    else:
        from qcmaquis_amd.operators import twou1_phys_index
        phys, end, su2 = twou1_phys_index(), (L // 2, L // 2), False
    if symm == "stretched":  # the long-chain ground-state structure (SU2U1)
        import os
        from qcmaquis_amd.workload import stretched_sectors
        here = os.path.dirname(os.path.abspath(__file__))
        sec = stretched_sectors(os.path.join(here, "golden", "sectors_L24_m512.json"), L, m, phys, end)
    else:
        sec = bounded_sectors(L, phys, end, m, su2=su2)
    mps = random_mps_on(sec, phys, seed=seed)
    psi = make_twosite_mps(mps[site], mps[site + 1], su2=su2)
    return phys, mps[site].left, mps[site + 1].right, psi, su2


@pytest.mark.parametrize("symm,L,m,site", [("su2u1", 6, 16, 2), ("su2u1", 8, 40, 3), ("su2u1", 12, 128, 5),
                                           ("2u1", 6, 16, 2), ("2u1", 8, 40, 3)])
def test_split_map_matches_host_reshape(symm, L, m, site):
    from qcmaquis_amd.gpu import host_split_both
    from qcmaquis_amd.mps import MPSTensor, ts_left_to_both, unreduce_left
    phys, left, right, psi, su2 = _problem(symm, L, m, site, seed=1)
    x = np.random.default_rng(7).standard_normal(psi.to_flat().size)
    u = MPSTensor.from_flat(psi.phys, psi.left, psi.right, x)
    lp = unreduce_left(phys, phys, left, right, u.data) if su2 else u.data
    both = ts_left_to_both(phys, phys, left, right, lp)
    keys = sorted(both.blocks, reverse=True)
    assert all(lc == rc for lc, rc in keys)
    expect = np.concatenate([both.blocks[k].ravel() for k in keys])
    got = host_split_both(phys, phys, left, right, x, su2=su2)
    assert got.shape == expect.shape
    assert np.abs(got - expect).max() <= 1e-14 * np.abs(expect).max()


def test_truncation_keep_rule():
    from qcmaquis_amd.mps import truncation_keep
    s = [np.array([5.0, 3.0, 1.0]), np.array([4.0, 3.0]), np.array([2.0, 1e-20])]
    # Mmax 3: the threshold is the 4th largest value (3.0), kept with its equal
    keep, info = truncation_keep(s, 3, 1e-16)
    assert keep == [2, 2, 0] and info["bond_dimension"] == 4
    # cutoff only: 1e-20 < 1e-16 * 5
    keep, info = truncation_keep(s, 100, 1e-16)
    assert keep == [3, 2, 1]
    allv = np.array([5, 3, 1, 4, 3, 2, 1e-20])
    assert info["truncated_weight"] == pytest.approx(1e-40 / np.sum(allv ** 2))
    keep, _ = truncation_keep(s, 100, 0.5)  # threshold 2.5
    assert keep == [2, 2, 0]


@pytest.mark.parametrize("shape", [(40, 25), (25, 40), (33, 33), (1, 7), (7, 1), (64, 3)])
def test_jacobi_svd_executor(shape):
    """The device block SVD's algorithm (tests/emulate.jacobi_svd, the kernel's
    rotation order, test and ordering) against LAPACK: singular values to
    round-off, orthonormal factors, and the product back to the matrix; rank
    deficiency included."""
    from emulate import jacobi_svd
    rng = np.random.default_rng(sum(shape))
    M = rng.standard_normal(shape)
    if min(shape) > 3:  # a rank-deficient block and a wide spread of singular values
        M[:, 1] = M[:, 0]
        M = M @ np.diag(np.logspace(0, -9, shape[1]))
    U, s, Vt, sweeps = jacobi_svd(M)
    assert 0 < sweeps < 30 or min(shape) < 2
    ref = np.linalg.svd(M, compute_uv=False)
    assert np.abs(s - ref).max() <= 1e-14 * ref[0]
    assert np.all(np.diff(s) <= 0)
    assert np.abs(U @ np.diag(s) @ Vt - M).max() <= 1e-14 * ref[0]
    big = s > 1e-12 * s[0]
    assert np.abs(U[:, big].T @ U[:, big] - np.eye(big.sum())).max() < 1e-13
    assert np.abs(Vt @ Vt.T - np.eye(len(s))).max() < 1e-13
