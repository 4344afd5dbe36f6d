"""Write the oracle's two-site sweep energies of the L=24 workload Hamiltonian
(tests/golden/sweep_L24_m64.json), the CPU side of the GPU sweep comparison.

    python tests/golden/make_sweep_fixture.py

The L=24, N=24, S=0 SU(2)xU(1) Hamiltonian of the synthetic integrals (seed
1024, the bench workload's), a random initial MPS (seed 42, initial bond
dimension 5) and two two-site sweeps at m=64 with the reference's local-solver
defaults (Jacobi-Davidson, at most 10 iterations, tolerance 1e-8,
DmrgParameters.h:53-55), run by the FP64 oracle engine.  The record holds the
energy of every microiteration, its Jacobi-Davidson iteration count and the
bond dimension kept.
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

L, M, SWEEPS, SEED_INT = 24, 64, 2, 1024


def main():
    from pyoracle import OracleEngine
    from qcmaquis_amd.integrals import synthetic
    from qcmaquis_amd.sweep import Hamiltonian, TwoSiteDMRG
    t0 = time.time()
    ham = Hamiltonian(L, L, 0, synthetic(L, seed=SEED_INT))
    d = TwoSiteDMRG(ham, OracleEngine(), M, jd_tol=1e-8, jd_maxiter=10)
    for s in range(SWEEPS):
        E = d.sweep(s)
        print(f"sweep {s}: E {E:.12f} ({time.time() - t0:.0f} s)", flush=True)
    rec = dict(L=L, nelec=L, spin=0, integral_seed=SEED_INT, mps_seed=42, m=M, sweeps=SWEEPS, jd_tol=1e-8,
               jd_maxiter=10, energies=[r["energy"] for r in d.records], jd_iters=[r["jd_iters"] for r in d.records],
               bonds=[r["bond"] for r in d.records])
    with open(os.path.join(HERE, "sweep_L24_m64.json"), "w") as f:
        json.dump(rec, f, indent=0)


if __name__ == "__main__":
    main()
