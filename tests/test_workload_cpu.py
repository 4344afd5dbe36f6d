"""The stretched sector model (workload.stretched_sectors) on the CPU: every
bond holds only sectors the long chain allows, each within its number of
states, with the chain's end charges; the central bonds carry about m states
in a few tens of sectors around the scaled mean filling; and consecutive bonds
can be joined by one site (every sector of bond i+1 is reachable from bond i)."""
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("L,m", [(24, 512), (36, 768), (50, 1024)])
def test_stretched_sectors(L, m):
    from qcmaquis_amd.blocks import fuse
    from qcmaquis_amd.mps import allowed_sectors
    from qcmaquis_amd.operators import su2_phys_index
    from qcmaquis_amd.workload import stretched_sectors
    phys = su2_phys_index()
    sec = stretched_sectors(os.path.join(HERE, "golden", "sectors_L24_m512.json"), L, m, phys, (L, 0))
    allowed = allowed_sectors(L, phys, (L, 0), 1 << 30, True)
    assert len(sec) == L + 1
    assert list(sec[0]) == [((0, 0), 1)] and list(sec[L]) == [((L, 0), 1)]
    for i, b in enumerate(sec):
        for c, n in b:
            assert allowed[i].has(c) and 1 <= n <= allowed[i].size_of_block(c), (i, c, n)
    centre = sec[L // 2]
    assert 0.9 * m <= centre.sum_of_sizes() <= 1.1 * m
    assert 8 <= len(centre) <= 40
    nbar = sum(c[0] * n for c, n in centre) / centre.sum_of_sizes()
    assert abs(nbar - 18.03 * L / 24) < 3  # the recorded centre filling scaled to this chain
    for i in range(L):
        reach = {fuse(c, s) for c, _ in sec[i] for s, _ in phys}
        assert all(c in reach for c, _ in sec[i + 1]), i
    if L == 24:  # the recorded structure itself, sizes scaled by m / 512
        from qcmaquis_amd.workload import load_sectors
        _, rec = load_sectors(os.path.join(HERE, "golden", "sectors_L24_m512.json"))
        assert all(set(dict(a).keys()) == set(dict(b).keys()) for a, b in zip(sec, rec))


def test_prepared_twosite_mpos(oracle_engine):
    """Hamiltonian.prepare_twosite builds every site pair's flattened two-site MPO
    in forked workers: the arrays equal those of the tensor built in process, the
    full tensor is built on demand, and an oracle sweep gives the same energies."""
    import numpy as np
    from qcmaquis_amd.desc import MPOArrays
    from qcmaquis_amd.integrals import synthetic
    from qcmaquis_amd.sweep import FlatTwoSiteMPO, Hamiltonian, TwoSiteDMRG
    L = 8
    ints = synthetic(L, seed=1000 + L)
    ham, ref = Hamiltonian(L, L, 0, ints), Hamiltonian(L, L, 0, ints)
    ham.prepare_twosite(workers=3)
    for p in range(L - 1):
        w = ham._ts[p]
        assert isinstance(w, FlatTwoSiteMPO) and w._full is None
        a, b = w._mq_arrays, MPOArrays(ref.ts_mpo(p))
        for k, v in vars(b).items():
            assert np.array_equal(np.asarray(getattr(a, k)), np.asarray(v)), (p, k)
    assert ham.ts_mpo(3).table is not None and ham._ts[3]._full is not None
    e = [TwoSiteDMRG(h, oracle_engine, 16, seed=5).sweep(0) for h in (ham, ref)]
    assert abs(e[0] - e[1]) <= 1e-12  # the oracle's threaded sums differ in the last bits between runs
