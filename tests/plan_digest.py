"""Field-wise digests of the host plans of small synthetic chains, printed as one
JSON object: run in a subprocess by tests/test_plans_cpu.py with different host
thread counts (MQ_HOST_THREADS), whose plans must be identical."""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.dirname(HERE)]

import numpy as np  # noqa: E402


def digest(hp, names):
    d = hashlib.sha1()
    for name in names:
        arr = hp.parts[name]
        for f in (arr.dtype.names or [None]):
            if f == "pad":
                continue  # padding fields are not written by the builder
            d.update(np.ascontiguousarray(arr if f is None else arr[f]).tobytes())
    d.update(json.dumps(hp.sizes, sort_keys=True).encode())
    d.update(repr((hp.ref_flops, hp.exec_flops, hp.min_bytes, hp.rbtm)).encode())
    return d.hexdigest()


def main():
    from qcmaquis_amd import gpu
    from qcmaquis_amd.desc import BoundaryLayout, SiteDesc
    from qcmaquis_amd.integrals import synthetic
    from qcmaquis_amd.mps import make_twosite_mps, random_mps_on
    from qcmaquis_amd.sweep import Hamiltonian
    from qcmaquis_amd.workload import bounded_sectors
    out = {}
    L = 8
    for symm in ("su2u1", "2u1"):
        ham = Hamiltonian(L, L, 0, synthetic(L, seed=7), symm=symm)
        sec = bounded_sectors(L, ham.phys, ham.right_end, 40, su2=ham.su2)
        mps = random_mps_on(sec, ham.phys, seed=42)
        Rl = [None] * (L + 1)
        Rl[L] = BoundaryLayout.seed(ham.right_end)
        for i in range(L - 1, 0, -1):
            hp = gpu.HostPlan(SiteDesc(ham.symm, mps[i].phys, mps[i].left, mps[i].right, ham.mpo[i], None, Rl[i + 1]),
                              "rstep")
            Rl[i] = hp.out_layout
            out[f"{symm} rstep {i}"] = digest(hp, gpu.PART_NAMES)
        Ll = [None] * (L + 1)
        Ll[0] = BoundaryLayout.seed((0, 0))
        for i in range(L - 1):
            sd = SiteDesc(ham.symm, mps[i].phys, mps[i].left, mps[i].right, ham.mpo[i], Ll[i], None)
            for fuse in (0, 1, 2):
                hp = gpu.HostPlan(sd, "lstep", fuse)
                out[f"{symm} lstep {i} f{fuse}"] = digest(hp, gpu.PART_NAMES)
            Ll[i + 1] = hp.out_layout
        for c in range(L - 1):
            W = ham.ts_mpo(c)
            psi = make_twosite_mps(mps[c], mps[c + 1], su2=ham.su2)
            sd = SiteDesc(ham.symm, psi.phys, psi.left, psi.right, W, Ll[c], Rl[c + 2])
            kinds = [("sigma", 0, 0), ("sigma", 1, 0), ("sigma", 2, 0), ("sigma", 1, 1)]
            if ham.su2:
                kinds.append(("diag", 1, 0))  # the diagonal is built for SU2 only
            for kind, fuse, sched in kinds:
                hp = gpu.HostPlan(sd, kind, fuse, schedule=sched)
                out[f"{symm} {kind} {c} f{fuse} s{sched}"] = digest(hp, gpu.PART_NAMES)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
