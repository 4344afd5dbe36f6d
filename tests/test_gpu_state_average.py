"""State-averaged two-site DMRG through the GPU engine: every local problem
solved for all roots by the block Davidson with one batched sigma per block
iteration, in FP64, FP32 and BF16x3 (each continued in FP64 to the tolerance),
against brute-force FCI."""
import numpy as np
import pytest

from test_state_average import _fci_roots

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_gpu_sa_sweeps_reach_fci_singlets(oracle_engine, gpu_engine, prec):
    from qcmaquis_amd import desc
    from qcmaquis_amd.sweep import Hamiltonian, StateAveragedDMRG
    ints, e0, singlets = _fci_roots(8, 1008, 4)
    old = gpu_engine.prec
    gpu_engine.prec = {"fp64": desc.MQ_PREC_FP64, "fp32": desc.MQ_PREC_FP32, "bf16x3": desc.MQ_PREC_BF16X3}[prec]
    try:
        print(f"state-averaged L=8 m=400 4 roots {prec}", flush=True)
        d = StateAveragedDMRG(Hamiltonian(8, 8, 0, ints), gpu_engine, 400, 4, jd_maxiter=60, jd_tol=1e-10)
        E = np.array(d.run(3))
    finally:
        gpu_engine.prec = old
    print("block iterations per problem", [r["iters"] for r in d.records[-12:]], flush=True)
    assert np.abs(E - singlets[:4]).max() < 1e-9, (E, singlets[:4])


def test_gpu_sa_twou1_matches_oracle_sweep(oracle_engine, gpu_engine):
    """TwoU1 (no diagonal preconditioner): the GPU block solve and the oracle's
    sequential constrained JD give the same truncated roots."""
    from qcmaquis_amd.sweep import Hamiltonian, StateAveragedDMRG
    ints, e0, singlets = _fci_roots(6, 606, 3)
    Es = []
    for eng in (oracle_engine, gpu_engine):
        d = StateAveragedDMRG(Hamiltonian(6, 6, 0, ints, symm="2u1"), eng, 400, 3, jd_maxiter=60, jd_tol=1e-10)
        Es.append(np.array(d.run(3)))
    assert np.abs(Es[0] - e0[:3]).max() < 1e-9
    assert np.abs(Es[1] - e0[:3]).max() < 1e-9, (Es[1], e0[:3])


def test_gpu_sa_compare_sweep_matches_fci(oracle_engine):
    """The sweep CLI's compare mode with 3 roots: FP64 state-averaged sweeps reach the
    FCI singlets, and every block solve repeated in FP32 and BF16x3 agrees with FP64."""
    from qcmaquis_amd.run_dmrg import run_compare
    from qcmaquis_amd.sweep import Hamiltonian
    ints, e0, singlets = _fci_roots(8, 1008, 3)
    print("compare sweeps L=8, 3 roots", flush=True)
    res = run_compare(Hamiltonian(8, 8, 0, ints), [400, 400, 400], 60, 1e-10, 42, roots=3)
    last = res[-1]
    assert np.abs(np.array(last["energies"]) - singlets[:3]).max() < 1e-9, (last["energies"], singlets[:3])
    assert last["max_abs_dtheta_mixed"] < 1e-9 and last["max_abs_dtheta_bf16x3"] < 1e-9, last
