"""Two-site DMRG sweeps on one GPU.

    python -m qcmaquis_amd.run_dmrg --L 24 --ramp 64,128,256,512,512,512 \
        [--fcidump FILE] [--prec fp64|mixed|bf16x3|both|all|compare|fp64,bf16x3] [--jd-maxiter 8] \
        [--converge 1e-9 --max-sweeps 30] [--out result.json]

Without --fcidump the synthetic L-orbital half-filled singlet Hamiltonian
(qcmaquis_amd.integrals.synthetic, seed 1000 + L) is used.  The sweep order,
truncation and Davidson semantics follow the reference's two-site optimizer;
each local problem is solved by the GPU Jacobi-Davidson (mixed: FP32 sigma
vectors; bf16x3: the same with BF16x3 products; both with FP64 Rayleigh
quotients and an FP64 continuation when the FP64 residual misses the
tolerance).  With --prec both (fp64 and mixed) or all (fp64, mixed and bf16x3)
the forms run from the same initial MPS and the energy differences per sweep
are reported.  With --prec compare the sweep runs in FP64 and every local
problem is solved again in the mixed and BF16x3 forms, on the same plan from the
same start vector: the largest |theta_form - theta_fp64| per sweep measures
each form's energy error on the sweep's own trajectory.
"""
import argparse
import json
import sys
import time

from .desc import MQ_PREC_BF16X3, MQ_PREC_FP32, MQ_PREC_FP64
from .integrals import read_fcidump, synthetic
from .sweep import Hamiltonian, TwoSiteDMRG


LOW_FORMS = (("mixed", MQ_PREC_FP32), ("bf16x3", MQ_PREC_BF16X3))


class _CompareEngine:
    """GPU engine whose local solves run in FP64 (driving the sweep) and, on the
    same plan and start vector, in the mixed and BF16x3 forms; records
    |theta_form - theta_fp64| per solve."""

    def __init__(self, eng):
        self.eng = eng
        self.diffs = {n: [] for n, _ in LOW_FORMS}

    def __getattr__(self, k):
        return getattr(self.eng, k)

    def jd(self, symm, psi, L, R, W, max_iter=10, tol=1e-8, ortho=()):
        sp = self.eng.sigma_plan(symm, psi, L, R, W)
        try:
            x = self.eng.to_dev(psi.to_flat())
            t64, u, its = sp.jd(x, max_iter, tol, ortho, prec=MQ_PREC_FP64)
            for name, prec in LOW_FORMS:
                t, _, _ = sp.jd(x, max_iter, tol, ortho, prec=prec)
                self.diffs[name].append(abs(t - t64))
            return t64, u.cpu().numpy(), its
        finally:
            sp.close()


    def block_solve(self, symm, psis, L, R, W, max_iter=10, tol=1e-8):
        """The roots in FP64 (driving the sweep) and, on the same plan from the same
        start vectors, in the mixed and BF16x3 forms; records the largest
        |theta_form - theta_fp64| over the roots."""
        import numpy as np
        from .gpu import independent_starts
        sp = self.eng.sigma_plan(symm, psis[0], L, R, W)
        try:
            X0 = self.eng.torch.as_tensor(independent_starts([p.to_flat() for p in psis],
                                                             live=sp.live_mask())).to(self.eng.device)
            diag = sp.diagonal() if symm == "su2u1" else None
            t64, U, its = sp.block_davidson(X0, max_iter, tol, prec=MQ_PREC_FP64, diag=diag)
            for name, prec in LOW_FORMS:
                t, _, _ = sp.block_davidson(X0, max_iter, tol, prec=prec, diag=diag)
                self.diffs[name].append(float(np.max(np.abs(np.array(t) - np.array(t64)))))
            Uh = U.cpu().numpy()
            return t64, [Uh[r] for r in range(len(psis))], its
        finally:
            sp.close()


def _driver(ham, eng, m, roots, jd_maxiter, jd_tol, seed, eigensolver="jd", split="engine"):
    from .sweep import StateAveragedDMRG
    if roots > 1:
        return StateAveragedDMRG(ham, eng, m, roots, jd_maxiter=jd_maxiter, jd_tol=jd_tol, seed=seed)
    return TwoSiteDMRG(ham, eng, m, jd_maxiter=jd_maxiter, jd_tol=jd_tol, seed=seed, eigensolver=eigensolver,
                       split=split)


def _energies(d, E):
    return list(d.root_energies[-1]) if hasattr(d, "root_energies") else [E]


def run_compare(ham, ramp, jd_maxiter, jd_tol, seed, log=print, roots=1, mpo_workers=None):
    """FP64 sweeps; every local problem is also solved in the mixed and BF16x3 forms."""
    from .gpu import GpuEngine
    ham.prepare_twosite(workers=mpo_workers, log=log)  # forks: before the GPU is initialised
    eng = _CompareEngine(GpuEngine(0, prec=MQ_PREC_FP64))
    eng.set_ref_flops(False)
    try:
        d = _driver(ham, eng, ramp[0], roots, jd_maxiter, jd_tol, seed)
        out = []
        t0 = time.perf_counter()
        for s, m in enumerate(ramp):
            d.m = m
            eng.diffs = {n: [] for n, _ in LOW_FORMS}
            E = d.sweep(s)
            rec = dict(sweep=s, m=m, energy=E, energies=_energies(d, E), solves=len(eng.diffs["mixed"]),
                       seconds=time.perf_counter() - t0)
            for n, _ in LOW_FORMS:
                rec[f"max_abs_dtheta_{n}"] = max(eng.diffs[n])
                rec[f"mean_abs_dtheta_{n}"] = sum(eng.diffs[n]) / len(eng.diffs[n])
            out.append(rec)
            log(f"  sweep {s} m {m} E(fp64) {E:.12f} max |theta - theta_fp64|: "
                + ", ".join(f"{n} {rec[f'max_abs_dtheta_{n}']:.3e}" for n, _ in LOW_FORMS)
                + f" Eh over {rec['solves']} solves, t {time.perf_counter() - t0:.1f}s")
        return out
    finally:
        eng.eng.close()


NAMES = {MQ_PREC_FP64: "fp64", MQ_PREC_FP32: "mixed", MQ_PREC_BF16X3: "bf16x3"}


TIMERS = ("t_site", "t_plan", "t_iter", "t_down", "t_split", "t_step")


def run(ham, prec, ramp, jd_maxiter, jd_tol, seed, log=print, eigensolver="jd", roots=1, converge=0.0,
        max_sweeps=0, split="engine", mpo_workers=None):
    """Sweeps with the bond dimensions of `ramp`; with converge > 0, further sweeps
    at the last bond dimension until the sweep energy changes by less than
    `converge` (at most max_sweeps sweeps in all).  Each sweep's record sums the
    microiterations' time per part: two-site tensor and MPO (t_site), plan build
    with boundary binding (t_plan), the solve (t_iter), result download (t_down),
    SVD truncation (t_split: on the device, or with split="host" on the host) and
    boundary step (t_step).  The two-site MPOs are built first, on mpo_workers
    processes (default min(16, CPUs); 1 builds them in the sweep as needed)."""
    from .gpu import GpuEngine
    ham.prepare_twosite(workers=mpo_workers, log=log)  # forks: before the GPU is initialised
    eng = GpuEngine(0, prec=prec)
    eng.set_ref_flops(False)  # the sweep needs no reference FLOP counts
    try:
        d = _driver(ham, eng, ramp[0], roots, jd_maxiter, jd_tol, seed, eigensolver, split)
        out = []
        t0 = time.perf_counter()
        ms = list(ramp)
        s = 0
        while s < len(ms):
            m = ms[s]
            d.m = m
            ts = time.perf_counter()
            E = d.sweep(s)
            recs = [r for r in d.records if r["sweep"] == s]
            rec = dict(sweep=s, m=m, energy=E, energies=_energies(d, E), max_bond=max(r["bond"] for r in recs),
                       jd_seconds=sum(r.get("t_jd", r.get("t_solve", 0.0)) for r in recs),
                       sweep_seconds=time.perf_counter() - ts, seconds=time.perf_counter() - t0,
                       jd_iterations=sum(r.get("jd_iters", 0) for r in recs),
                       fp64_iterations=sum(r.get("fp64_iterations", 0) for r in recs))
            for k in TIMERS:
                rec[k] = sum(r.get(k, 0.0) for r in recs)
            out.append(rec)
            log(f"  {NAMES[prec]} sweep {s} m {m} E {E:.12f} t {time.perf_counter() - t0:.1f}s ("
                + ", ".join(f"{k[2:]} {rec[k]:.2f}" for k in TIMERS) + ")")
            if converge > 0 and s == len(ms) - 1 and len(ms) < max_sweeps and (
                    len(out) < 2 or abs(out[-1]["energy"] - out[-2]["energy"]) >= converge):
                ms.append(ms[-1])
            s += 1
        return out
    finally:
        eng.close()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--L", type=int, default=8)
    ap.add_argument("--fcidump")
    ap.add_argument("--ramp", default="64,64,64,64")
    ap.add_argument("--prec", default="fp64",
                    help="fp64, mixed, bf16x3, both (fp64,mixed), all, compare, or a comma list of forms")
    ap.add_argument("--jd-maxiter", type=int, default=8)
    ap.add_argument("--jd-tol", type=float, default=1e-8)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--eigensolver", default="jd", choices=["jd", "davidson"],
                    help="IETL_JCD (default) or IETL_DAVIDSON (SU2 only)")
    ap.add_argument("--roots", type=int, default=1,
                    help="states optimised together (state-averaged sweeps with the block Davidson)")
    ap.add_argument("--converge", type=float, default=0.0,
                    help="after the ramp, sweep at its last bond dimension until the energy changes by less")
    ap.add_argument("--max-sweeps", type=int, default=30)
    ap.add_argument("--split", default="engine", choices=["engine", "host"],
                    help="SVD truncation on the device (engine, default) or on the host")
    ap.add_argument("--mpo-workers", type=int, default=None,
                    help="processes that build the two-site MPOs before the sweeps (default min(16, CPUs); 1: none)")
    ap.add_argument("--out")
    a = ap.parse_args(argv)
    if a.fcidump:
        fd = read_fcidump(a.fcidump)
        ham = Hamiltonian(fd["norb"], fd["nelec"], fd["ms2"], fd["integrals"])
    else:
        ham = Hamiltonian(a.L, a.L, 0, synthetic(a.L, seed=1000 + a.L))
    ramp = [int(v) for v in a.ramp.split(",")]
    if a.prec == "compare":
        res = run_compare(ham, ramp, a.jd_maxiter, a.jd_tol, a.seed, roots=a.roots, mpo_workers=a.mpo_workers)
        rec = dict(L=ham.L, ramp=ramp, roots=a.roots, jd_maxiter=a.jd_maxiter, jd_tol=a.jd_tol, mode="compare",
                   results=res)
        if a.out:
            with open(a.out, "w") as f:
                json.dump(rec, f, indent=1)
        return rec
    forms = {"fp64": MQ_PREC_FP64, "mixed": MQ_PREC_FP32, "bf16x3": MQ_PREC_BF16X3}
    names = {"both": "fp64,mixed", "all": "fp64,mixed,bf16x3"}.get(a.prec, a.prec).split(",")
    unknown = [n for n in names if n not in forms]
    if unknown:
        ap.error(f"unknown form(s) {unknown} in --prec {a.prec}")
    precs = [(n, forms[n]) for n in names]
    res = {name: run(ham, p, ramp, a.jd_maxiter, a.jd_tol, a.seed, eigensolver=a.eigensolver, roots=a.roots,
                     converge=a.converge, max_sweeps=a.max_sweeps, split=a.split, mpo_workers=a.mpo_workers)
           for name, p in precs}
    rec = dict(L=ham.L, ramp=ramp, roots=a.roots, jd_maxiter=a.jd_maxiter, jd_tol=a.jd_tol,
               eigensolver=a.eigensolver, split=a.split, results=res)
    for name, _ in LOW_FORMS:
        if "fp64" in res and name in res:
            rec[f"delta_E_{name}"] = [m["energy"] - f["energy"] for f, m in zip(res["fp64"], res[name])]
            for s, dE in enumerate(rec[f"delta_E_{name}"]):
                print(f"sweep {s}: E({name}) - E(fp64) = {dE:.3e} Eh")
            rec[f"final_delta_E_{name}"] = res[name][-1]["energy"] - res["fp64"][-1]["energy"]
            print(f"final: E({name}) - E(fp64) = {rec[f'final_delta_E_{name}']:.3e} Eh after "
                  f"{len(res[name])} / {len(res['fp64'])} sweeps")
    if a.out:
        with open(a.out, "w") as f:
            json.dump(rec, f, indent=1)
    return rec


if __name__ == "__main__":
    main(sys.argv[1:])
