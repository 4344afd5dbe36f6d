"""Record the per-bond symmetry-sector structure of a DMRG ground state.

The sigma-vector and boundary workloads are shaped by the sector sizes of
the MPS bonds, not by their values.  This script runs two-site DMRG sweeps on
the synthetic L-orbital half-filled Hamiltonian (integrals from
qcmaquis_amd.integrals.synthetic with seed 1000 + L, optionally decaying
with |p - q|) in the SU(2) (charges (N, 2S)) or U(1)xU(1) (charges
(N_alpha, N_beta)) model, with a bond-dimension ramp, and writes the final
MPS's bonds as lists of (charge, charge, size) together with the sweep
energies.  The benchmark and the full-size GPU tests build random
MPS tensors on exactly these sectors.

    python tests/golden/make_sectors.py --L 24 --ramp 64,128,256,512 \
        --out tests/golden/sectors_L24_m512.json [--engine gpu|oracle] [--symm 2u1 --decay 8]

With --engine gpu it needs the built HIP library and a GPU.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from qcmaquis_amd.integrals import synthetic  # noqa: E402
from qcmaquis_amd.sweep import Hamiltonian, TwoSiteDMRG  # noqa: E402


def bonds_of(mps, right_end):
    out = [[[int(q[0]), int(q[1]), int(n)] for q, n in site.left] for site in mps]
    out.append([[int(right_end[0]), int(right_end[1]), 1]])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--L", type=int, default=24)
    ap.add_argument("--ramp", default="64,128,256,512,512")
    ap.add_argument("--jd-maxiter", type=int, default=8)
    ap.add_argument("--engine", default="gpu")
    ap.add_argument("--symm", default="su2u1", choices=["su2u1", "2u1"])
    ap.add_argument("--decay", type=float, default=None, help="integral decay length (exp(-|p-q|/decay))")
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    L = a.L
    ham = Hamiltonian(L, L, 0, synthetic(L, seed=1000 + L, decay=a.decay), symm=a.symm)
    if a.engine == "gpu":
        from qcmaquis_amd.gpu import GpuEngine
        eng = GpuEngine(0)
    else:
        from pyoracle import OracleEngine
        eng = OracleEngine()
    ramp = [int(x) for x in a.ramp.split(",")]
    t0 = time.time()

    def log(r):
        if r["micro"] % 4 == 0:
            print(f"  sweep {r['sweep']} micro {r['micro']} E {r['energy']:.10f} m {r['bond']} "
                  f"jd {r['jd_iters']} t_jd {r['t_jd']:.2f}s t_rest {r['t_rest']:.2f}s", flush=True)

    d = TwoSiteDMRG(ham, eng, ramp[0], jd_maxiter=a.jd_maxiter, log=log)
    sweeps = []
    for s, m in enumerate(ramp):
        d.m = m
        E = d.sweep(s)
        sweeps.append(dict(m=m, energy=E, max_bond=max(r["bond"] for r in d.records if r["sweep"] == s)))
        print(f"sweep {s} m {m} E {E:.10f} elapsed {time.time() - t0:.1f}s", flush=True)
        rec = dict(source="tests/golden/make_sectors.py", L=L, nelec=L, spin=0, symm=a.symm, decay=a.decay,
                   integrals=f"synthetic(L={L}, seed={1000 + L}, decay={a.decay})", ramp=ramp[:s + 1],
                   jd_maxiter=a.jd_maxiter,
                   engine=a.engine, sweeps=sweeps, bonds=bonds_of(d.mps, ham.right_end))
        with open(a.out, "w") as f:
            json.dump(rec, f)
    if hasattr(eng, "close"):
        eng.close()


if __name__ == "__main__":
    main()
