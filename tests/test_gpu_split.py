"""The two-site split on the device (GpuEngine.split_twosite: mq_split's gather
map, block Jacobi SVD and assembly) against the host split (mps.split_twosite,
LAPACK SVD): the same kept bond and truncated weight, the same truncated
two-site tensor A.B block by block, and orthonormal site tensors, up to the
L=50 m=1024 centre tensor (blocks of several hundred columns); and whole sweeps
whose truncation runs on the device reproduce the host-truncated sweeps'
energies."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("l2r", [True, False])
@pytest.mark.parametrize("symm,L,m,site,Mmax", [("su2u1", 12, 128, 5, 100), ("su2u1", 24, 512, 11, 400),
                                                ("2u1", 12, 128, 5, 100), ("stretched", 50, 1024, 24, 1024)])
def test_device_split_matches_host(gpu_engine, symm, L, m, site, Mmax, l2r):
    from test_split_cpu import _problem
    from qcmaquis_amd.mps import MPSTensor, split_twosite
    phys, left, right, psi, su2 = _problem(symm, L, m, site, seed=3)
    x = np.random.default_rng(11).standard_normal(psi.to_flat().size)
    print(f"split {symm} L={L} m={m} site={site} l2r={l2r} n={x.size}", flush=True)
    A, B, info = gpu_engine.split_twosite(x, phys, phys, left, right, Mmax, 1e-12, l2r=l2r, su2=su2)
    Ah, Bh, infoh = split_twosite(MPSTensor.from_flat(psi.phys, psi.left, psi.right, x), phys, phys, left, right,
                                  Mmax, 1e-12, l2r=l2r, su2=su2)
    assert info["bond_dimension"] == infoh["bond_dimension"]
    assert abs(info["truncated_weight"] - infoh["truncated_weight"]) <= 1e-12 * max(1e-300, infoh["truncated_weight"]) \
        + 1e-15
    assert list(A.right) == list(Ah.right)
    A.make_left_paired()
    Ah.make_left_paired()
    B.make_right_paired()
    Bh.make_right_paired()
    scale = max(np.abs(x).max(), 1.0)
    for k in Ah.data.blocks:
        P = A.data.blocks[k] @ B.data.blocks[k]
        Ph = Ah.data.blocks[k] @ Bh.data.blocks[k]
        assert np.abs(P - Ph).max() <= 1e-11 * scale, (k, np.abs(P - Ph).max())
        Q = A.data.blocks[k] if l2r else B.data.blocks[k].T
        assert np.abs(Q.T @ Q - np.eye(Q.shape[1])).max() < 1e-12


def test_device_split_sweeps_match_host_split(gpu_engine):
    """L=8, N=8, S=0 synthetic system, m=24 (truncating), two sweeps with the
    truncation on the device and on the host: equal energies and bonds."""
    from qcmaquis_amd.integrals import synthetic
    from qcmaquis_amd.sweep import Hamiltonian, TwoSiteDMRG
    ints = synthetic(8, seed=1008)
    runs = []
    for split in ("host", "engine"):
        d = TwoSiteDMRG(Hamiltonian(8, 8, 0, ints), gpu_engine, 24, split=split, cutoff=1e-12)
        d.run(2)
        runs.append(d)
    e = [np.array(d.energies) for d in runs]
    assert np.abs(e[0] - e[1]).max() <= 1e-10, np.abs(e[0] - e[1]).max()
    assert [r["bond"] for r in runs[0].records] == [r["bond"] for r in runs[1].records]
