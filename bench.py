"""Benchmark of the DMRG sigma-vector hot path on one MI355X.

Workload: the centre two-site problem (sites 11, 12) of the L=24-orbital,
24-electron singlet SU(2)xU(1) Hamiltonian (synthetic integrals, seed 1024),
m=512, on the bond sector structure recorded from a DMRG ramp to m=512
(tests/golden/sectors_L24_m512.json).  MPS values are random (seed 42),
right-normalised; the boundaries are built on the GPU with the library's own
boundary steps.

A step is one sigma-vector H.psi with the boundaries and psi resident in HBM.
value = algorithmic FLOPs of the reference schedule / wall time per step.
Three forms are measured: the plain FP64 sigma, the mixed form (FP32 panels,
intermediates and MFMA, FP64 output; Davidson in FP64 with an FP64 defect
correction per solve) and the BF16x3 form (the mixed form with every GEMM
product as bf16 hi/lo MFMA products accumulated in FP32).  Every form solves
the site problem from the same start vector (the MPS guess) to a residual of
1e-8 |theta|; the value is the fastest form whose eigenvalue agrees with the
FP64 one to 1e-8 Eh.  The record of each solve gives the iterations in the
form, the FP64 continuation iterations and the form's own Ritz value.
The 4-root block Davidson (one batched sigma of the block per iteration) is
timed in each form from the MPS guess plus random vectors restricted to the
coordinates H_eff couples, with its lowest root compared with the single-root
eigenvalue.
The JSON line also carries ms per Jacobi-Davidson iteration, the cost of one
site solve (boundary binding plus the solve), per-stage HIP event timings on
the library stream, the roofline of the sigma (F and the minimum bytes B_min
of the reference schedule, the binding roof and the fraction of the form's
MFMA peak reached), the measured HBM traffic of the run's schedule when a
counter profile of it is committed, and the CPU oracle timed on a bounded
sample of the same sigma.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "H·ψ TFLOP/s + ms/Davidson iter, L=24 m=512; mixed vs FP64 at ΔE≤1e-8 Eh"
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
# BF16x3: three bf16 MFMA products (dense bf16 peak 2.5 PFLOP/s) per FP32-equivalent product
PEAK_BF16X3_TFLOPS = 2500.0 / 3
PEAK_HBM_GBS = 8000.0
STAGES = ["arrange", "reshape", "gemm_T", "w_apply", "gemm_out"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(ham, psi, Lb, Rb, W, y_gpu, seconds):
    """The FP64 oracle on the host cores, over a bounded number of sigmas."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    from pyoracle import HostBoundary, OracleEngine, blas_name, lib as orclib

    ora = OracleEngine()
    Lo = HostBoundary(Lb.layout, Lb.data.cpu().numpy())
    Ro = HostBoundary(Rb.layout, Rb.data.cpu().numpy())
    x = psi.to_flat()
    n, t0 = 0, time.perf_counter()
    while True:
        y0, f0 = ora.sigma(ham.symm, psi, Lo, Ro, W, x=x)
        n += 1
        dt = time.perf_counter() - t0
        if dt >= seconds or n >= 50:
            break
    per = dt / n
    err = float(np.abs(y_gpu - y0).max() / np.abs(y0).max())
    return dict(value=f0 / per / 1e12, unit="TFLOP/s", cores=int(orclib().orc_num_threads()), kind="port",
                sample=f"{n} FP64 oracle sigma(s) of the same centre-bond problem, {per * 1e3:.1f} ms each, "
                       f"host BLAS {blas_name()}"), err


def sigma_kernels(form):
    """Kernel names (as rocprofv3 reports them) of the stages of one sigma in a form."""
    t = "double" if form == "fp64" else "float"
    x3 = "true" if form == "bf16x3" else "false"
    return {"reshape": f"k_copy2d<0, double, {t}>", "gemm_T": f"k_gemm<0, {t}, {t}, {x3}>",
            "w_apply": f"k_wapply<0, {t}, false>", "gemm_out": f"k_gemm<1, {t}, double, {x3}>",
            "split_reduce": f"k_split_reduce<1, {t}, double>"}


def plan_key(form, plan, sched=None):
    """Identifies a plan's task lists for a form: schedules that build the same
    lists (e.g. fused-column ratios that fuse the same columns) share it; a
    fixed GEMM tile shape is part of the key."""
    key = f"{form} | executed {plan.exec_flops:.0f} | workspace {plan.workspace_bytes}"
    if sched is not None and (sched.tile_m or sched.tile_n):
        key += f" | tiles {sched.tile_m}x{sched.tile_n}"
    return key


def traffic_profile(a, sec_src, key):
    """Counter-measured HBM bytes per launch of the sigma kernels for this
    workload and plan (key: plan_key of the form's plan), from the committed
    profile (None when the run builds a plan that has not been profiled)."""
    if not (sec_src == "tests/golden/sectors_L24_m512.json" and a.L == 24 and a.symm == "su2u1"):
        return None
    path = os.path.join(ROOT, "profiles", "sigma_L24_m512_traffic.json")
    if not os.path.exists(path):
        return None
    rec = json.load(open(path)).get(key)
    if rec is None:
        return None
    return dict(rec, file=os.path.relpath(path, ROOT))


def config_records():
    """Per-form summaries of the other workloads (TwoU1 L=40 m=1024; SU2U1 L=50
    m=1024, 4 roots), read from the JSON lines their own runs of this script
    wrote under profiles/ (python bench.py --config 3 / --config 4): recorded
    by those runs, not measured by this one."""
    out = {}
    for cfg in (3, 4):
        path = os.path.join(ROOT, "profiles", f"bench_config{cfg}.json")
        if not os.path.exists(path):
            continue
        rec = json.loads(open(path).read().strip().splitlines()[-1])
        forms = [k for k in ("fp64", "mixed", "bf16x3") if isinstance(rec.get(k), dict)]
        blk = rec.get("block_davidson") or {}
        out[str(cfg)] = dict(
            recorded_by=f"python bench.py --config {cfg}", file=os.path.relpath(path, ROOT),
            workload=rec["config"]["workload"], sectors=rec["config"].get("sectors"),
            psi_len=rec["config"].get("psi_len"), mpo_bond=rec["config"].get("mpo_bond"),
            flops_per_sigma=rec.get("flops_per_sigma"),
            forms={k: dict(ms_per_sigma=rec[k]["ms_per_sigma"], tflops=rec[k]["tflops"],
                           ms_per_davidson_iter=rec[k].get("ms_per_davidson_iter"),
                           ms_per_block_iter=(blk.get(k) or {}).get("ms_per_block_iter"),
                           delta_E_vs_fp64_Eh=rec.get(f"delta_E_{k}_vs_fp64_Eh"),
                           block_max_delta_E_vs_fp64=(blk.get(k) or {}).get("max_delta_E_vs_fp64"))
                   for k in forms},
            cpu_baseline=rec.get("cpu_baseline"), parity_max_rel_err_vs_oracle=rec.get("parity_max_rel_err_vs_oracle"))
    return out or None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--jd-iters", type=int, default=10)
    ap.add_argument("--L", type=int, default=24)
    ap.add_argument("--m", type=int, default=512)
    ap.add_argument("--sectors", default=None,
                    help="recorded bond sectors (default: tests/golden/sectors_L24_m512.json for SU2U1 L=24)")
    ap.add_argument("--symm", default="su2u1", choices=["su2u1", "2u1"])
    ap.add_argument("--decay", type=float, default=None, help="integral decay length (exp(-|p-q|/decay))")
    ap.add_argument("--roots", type=int, default=4, help="block Davidson roots to time (0: skip)")
    ap.add_argument("--forms", default="fp64,mixed,bf16x3",
                    help="comma-separated forms to measure (fp64 is always included)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-delta-e", action="store_true", help="skip the converged-solve energy comparison")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the default schedule instead of tuning it per form on the workload")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="first run a plain device copy of this many GiB (reference for HBM counters)")
    ap.add_argument("--lean", action="store_true",
                    help="free each form's plan before building the next (workloads near the HBM size)")
    ap.add_argument("--config", type=int, choices=[1, 3, 4], default=None,
                    help="preset workload: 1 the SU2U1 L=24 m=512 centre (default), 3 TwoU1 L=40 m=1024 decay 8 "
                         "(lean, default schedule), 4 SU2U1 L=50 m=1024 on the stretched sector model, FP64 and "
                         "BF16x3, 4 roots (lean, default schedule)")
    # a preset sets the defaults of its workload; flags given explicitly still win
    pre = argparse.ArgumentParser(add_help=False)
    pre.add_argument("--config", type=int, default=None)
    cfg = pre.parse_known_args()[0].config
    if cfg == 3:
        ap.set_defaults(L=40, m=1024, symm="2u1", decay=8.0, lean=True, no_autotune=True)
    elif cfg == 4:
        ap.set_defaults(L=50, m=1024, symm="su2u1", decay=None, sectors="stretched", roots=4, forms="fp64,bf16x3",
                        lean=True, no_autotune=True)
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the sigma-vector path is built and timed on a single MI355X; "
              "run it with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from qcmaquis_amd.desc import MQ_PREC_BF16X3, MQ_PREC_FP32, MQ_PREC_FP64
    from dataclasses import asdict

    from qcmaquis_amd.autotune import DEFAULT as DEFAULT_SCHEDULE, table_json, tune
    from qcmaquis_amd.gpu import GpuEngine
    from qcmaquis_amd.workload import centre_problem, load_sectors, stretched_sectors

    eng = GpuEngine(0)
    if a.copy_calibration > 0:
        n = int(a.copy_calibration * (1 << 30)) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dst.copy_(src)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        log(f"copy calibration: {n * 8 / 2**30:.2f} GiB read + written in {dt * 1e3:.2f} ms "
            f"({2 * n * 8 / dt / 1e9:.0f} GB/s)")
        del src, dst
    t0 = time.perf_counter()
    sectors, sec_src = None, f"degeneracy-proportional sectors capped at m={a.m}"
    if a.sectors is None and a.symm == "su2u1" and a.L == 24 and a.decay is None:
        a.sectors = os.path.join(ROOT, "tests", "golden", "sectors_L24_m512.json")
    if a.sectors == "stretched":
        # the recorded L=24 m=512 ground-state structure stretched to this chain
        from qcmaquis_amd.operators import su2_phys_index, twou1_phys_index
        if a.symm != "su2u1":
            raise SystemExit("--sectors stretched models SU2U1 chains")
        rec_path = os.path.join(ROOT, "tests", "golden", "sectors_L24_m512.json")
        sectors = stretched_sectors(rec_path, a.L, a.m, su2_phys_index(), (a.L, 0))
        sec_src = f"{os.path.relpath(rec_path, ROOT)} stretched to L={a.L}, m={a.m} (workload.stretched_sectors)"
    elif a.sectors and os.path.exists(a.sectors):
        rec, sectors = load_sectors(a.sectors)
        if rec["L"] != a.L or rec.get("symm", "su2u1") != a.symm or rec.get("decay") != a.decay:
            raise SystemExit(f"{a.sectors} does not describe L={a.L} {a.symm} decay={a.decay}")
        sec_src = os.path.relpath(a.sectors, ROOT)
    sp_ = centre_problem(eng, a.L, m=a.m, sectors=sectors, symm=a.symm, decay=a.decay,
                         log=(lambda msg: log(f"  {msg} ({time.perf_counter() - t0:.0f}s)")))
    ham, psi, Lb, Rb, W, site = sp_.ham, sp_.psi, sp_.L, sp_.R, sp_.W, sp_.site
    log(f"setup {time.perf_counter() - t0:.1f}s: site {site}, psi {psi.to_flat().size}, W {W.row_dim}x{W.col_dim}")
    # schedule per form (fused-column rule, split-K chunk): tuned on this problem
    want = set(a.forms.split(",")) | {"fp64"}
    FORMS = tuple((n, p) for n, p in (("fp64", MQ_PREC_FP64), ("mixed", MQ_PREC_FP32), ("bf16x3", MQ_PREC_BF16X3))
                  if n in want)
    tuned = {}
    for name, prec in FORMS:
        if a.no_autotune:
            tuned[name] = (DEFAULT_SCHEDULE, [])
        else:
            t0 = time.perf_counter()
            tuned[name] = tune(eng, ham.symm, psi, Lb, Rb, W, prec=prec)
            log(f"autotune {name}: {tuned[name][0].label()} ({tuned[name][1][0][1]:.3f} ms; default "
                f"{[r[1] for r in tuned[name][1] if r[0] == DEFAULT_SCHEDULE][0]:.3f} ms) "
                f"in {time.perf_counter() - t0:.1f}s")
    # one form at a time: its plan is built, measured (sigma timing, the site solve
    # and the block solve) and, with --lean, freed before the next form's
    plans, pinfo = {}, {}
    t_plan = 0.0
    x = eng.to_dev(psi.to_flat())
    y = torch.empty_like(x)

    def measure(sp, prec):
        torch.cuda.synchronize()
        for _ in range(a.warmup):
            sp.apply(x, y, prec)
        eng.sync()
        torch.cuda.synchronize()
        # per-site cost of binding the boundaries: arrangement into panels, the fused
        # rows and (mixed forms) the FP32 copies, measured as bind + sigma - sigma
        t0 = time.perf_counter()
        for _ in range(3):
            sp.apply(x, y, prec)
        eng.sync()
        t_sig = (time.perf_counter() - t0) / 3
        t0 = time.perf_counter()
        for _ in range(3):
            sp.bind()
            sp.apply(x, y, prec)
        eng.sync()
        bind_ms = max((time.perf_counter() - t0) / 3 - t_sig, 0.0) * 1e3
        tot = [0.0] * 5
        work = None
        t0 = time.perf_counter()
        for _ in range(a.steps):
            ms, work = sp.timed(x, y, prec)
            for i in range(5):
                tot[i] += ms[i]
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / a.steps
        avg = [t / a.steps for t in tot]
        ycopy = y.cpu().numpy()
        # ms per Jacobi-Davidson iteration (tolerance 0: every iteration runs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        theta, u, its = sp.jd(x, a.jd_iters, 0.0, prec=prec)
        eng.sync()
        torch.cuda.synchronize()
        ms_jd = (time.perf_counter() - t0) / max(its, 1) * 1e3
        return dict(dt=dt, avg=avg, work=work, theta=theta, its=its, ms_jd=ms_jd, y=ycopy, bind_ms=bind_ms)

    # energy agreement: every form solves the site problem from the same start
    # vector (the MPS guess) to |r| <= 1e-8 |theta| (the reference's default
    # ietl_jcd_tol, DmrgParameters.h:53; the eigenvalue error is then about
    # |r|^2 / gap, far below 1e-8 Eh), with restarted solves of JD_SOLVE
    # iterations; each record keeps the iterations in the form, the FP64
    # continuation iterations, the form's own last Ritz value and the wall time
    JD_SOLVE, RESTARTS, TOL = 50, 8, 1e-8

    def solve(p_, prec):
        u, rec = x, dict(form_iterations=0, fp64_iterations=0, restarts=0, converged=False)
        eng.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for r_ in range(RESTARTS):
            th, u, its = p_.jd(u, JD_SOLVE, TOL, prec=prec)
            info = p_.solve_info()
            if r_ == 0:
                rec["theta_form_first_solve"] = info["theta_form"] + ham.ecore
                rec["theta_rayleigh_fp64_first_solve"] = info["theta_rayleigh_fp64"] + ham.ecore
            rec["form_iterations"] += info["form_iterations"]
            rec["fp64_iterations"] += info["fp64_iterations"]
            rec["restarts"] = r_
            done = info["fp64_iterations"] < JD_SOLVE if prec != MQ_PREC_FP64 and info["fp64_iterations"] > 0 \
                else info["form_iterations"] < JD_SOLVE
            if done:
                rec["converged"] = True
                break
        eng.sync()
        rec["solve_s"] = time.perf_counter() - t0
        rec["theta"] = th + ham.ecore
        return rec

    res, conv = {}, ({} if not a.no_delta_e else None)
    block = {"roots": a.roots} if a.roots > 0 else None
    X0, ref = None, None
    for name, prec in FORMS:
        eng.set_schedule(tuned[name][0])
        t0 = time.perf_counter()
        bp = eng.sigma_plan(ham.symm, psi, Lb, Rb, W)
        t_plan = max(t_plan, time.perf_counter() - t0)
        pinfo[name] = dict(ref_flops=bp.ref_flops, exec_flops=bp.exec_flops, workspace_bytes=bp.workspace_bytes,
                           min_bytes=bp.min_bytes, n=bp.n, key=plan_key(name, bp, tuned[name][0]))
        log(f"{name} plan {time.perf_counter() - t0:.2f}s: F {bp.ref_flops / 1e9:.3f} GF, executed "
            f"{bp.exec_flops / 1e9:.3f} GF, FP64 workspace {bp.workspace_bytes / 1e9:.2f} GB")
        res[name] = measure(bp, prec)
        log(f"{name} stages ms: " + ", ".join(f"{n} {v:.3f}" for n, v in zip(STAGES, res[name]["avg"])))
        if conv is not None:
            conv[name] = solve(bp, prec)
            c_ = conv[name]
            log(f"solve {name}: theta {c_['theta']:.12f}, {c_['form_iterations']} iterations in the form + "
                f"{c_['fp64_iterations']} FP64, {c_['solve_s']:.2f} s, converged {c_['converged']}"
                + (f", dE vs fp64 {abs(c_['theta'] - conv['fp64']['theta']):.2e}" if name != "fp64" else ""))
        # block Davidson over `roots` roots: ms per block iteration (one batched sigma
        # of the block) in each form, and the converged roots against the FP64 ones
        if block is not None:
            from qcmaquis_amd.gpu import independent_starts
            k = a.roots
            bp.release()  # the single-root solve's arenas, before the batched ones
            if X0 is None:
                rng = np.random.default_rng(11)
                # the MPS guess and random vectors on the coordinates H_eff couples (the
                # structurally zero columns of H_eff have eigenvalue 0, below every
                # physical root of this Hamiltonian)
                live = bp.live_mask()
                X0 = independent_starts([psi.to_flat()] + [rng.standard_normal(bp.n) for _ in range(k - 1)],
                                        live=live)
                X0 = torch.as_tensor(X0).to(eng.device)
                block["dead_coordinates"] = int((~live).sum())
            diag = bp.diagonal() if ham.symm == "su2u1" else None
            bp.block_davidson(X0, 1, 0.0, prec=prec, diag=diag)  # builds the plan's batched task lists
            eng.sync()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, _, its = bp.block_davidson(X0, a.jd_iters, 0.0, prec=prec, diag=diag)
            eng.sync()
            torch.cuda.synchronize()
            ms_it = (time.perf_counter() - t0) / max(its, 1) * 1e3
            t0 = time.perf_counter()
            if a.lean:
                # restarted from the current Ritz vectors every 100 iterations, with a
                # progress line per restart (a sigma takes tenths of a second here)
                Xc, itc, conv_ = X0, 0, False
                for r_ in range(8):
                    th, Xc, k_ = bp.block_davidson(Xc, 100, 1e-9, prec=prec, diag=diag)
                    itc += k_
                    log(f"  block Davidson {name}: restart {r_}, {itc} iterations, lowest root {th[0]:.9f} "
                        f"({time.perf_counter() - t0:.0f}s)")
                    if k_ < 100:
                        conv_ = True
                        break
                if not conv_:
                    log(f"  block Davidson {name}: not converged in {itc} iterations")
            else:
                th, _, itc = bp.block_davidson(X0, 400, 1e-9, prec=prec, diag=diag)
                conv_ = itc < 400
            eng.sync()
            t_conv = time.perf_counter() - t0
            th = [t + ham.ecore for t in th]
            if ref is None:
                ref = th
            info = bp.solve_info()
            block[name] = dict(ms_per_block_iter=ms_it, converged=conv_, converged_iterations=itc, converge_s=t_conv,
                               energies=th,
                               form_iterations=info["form_iterations"], fp64_iterations=info["fp64_iterations"],
                               max_delta_E_vs_fp64=max(abs(x_ - y_) for x_, y_ in zip(th, ref)))
            if conv is not None:
                block[name]["root0_minus_single_root_theta"] = th[0] - conv["fp64"]["theta"]
            bp.release()  # the batched arenas of one form at a time
            log(f"block Davidson {k} roots {name}: {ms_it:.2f} ms/iteration; converged in {itc} iterations "
                f"({t_conv:.2f} s), roots {', '.join(f'{e:.9f}' for e in th)}, max dE vs fp64 "
                f"{block[name]['max_delta_E_vs_fp64']:.2e}")
        if a.lean:
            bp.close()
        else:
            plans[name] = bp
    fp = pinfo["fp64"]
    r64 = res["fp64"]
    LOW = [n for n, _ in FORMS if n != "fp64"]
    sig_err = {k: float(np.abs(res[k]["y"] - r64["y"]).max() / np.abs(r64["y"]).max()) for k in LOW}
    dE = None if conv is None else {k: abs(conv[k]["theta"] - conv["fp64"]["theta"]) for k in LOW}
    for k in LOW:
        log(f"{k}: sigma rel err vs fp64 {sig_err[k]:.2e}")
    ok = ["fp64"] + [k for k in LOW if dE is not None and dE[k] <= 1e-8]
    prec_name = min(ok, key=lambda k: res[k]["dt"])
    ch = res[prec_name]
    dt, avg, work = ch["dt"], ch["avg"], ch["work"]
    if block is not None:
        for name in LOW:
            block[name]["speedup_vs_fp64"] = block["fp64"]["ms_per_block_iter"] / block[name]["ms_per_block_iter"]

    # roofline of the sigma on the reference schedule's terms: F (algorithmic
    # flops) and B_min = |L| + |R| + |psi| + |sigma| in FP64 (the bytes any
    # schedule must move); the binding roof is the one with the larger lower bound
    # on the time, F / P_MFMA(form) or B_min / BW_HBM, and achieved = F / t with t
    # the kernel time of one sigma from HIP events on the library stream
    peak_f = {"fp64": PEAK_FP64_TFLOPS, "mixed": PEAK_FP32_TFLOPS, "bf16x3": PEAK_BF16X3_TFLOPS}[prec_name]
    F, Bmin = fp["ref_flops"], fp["min_bytes"]
    t_sig = sum(avg[1:]) * 1e-3
    kern = sigma_kernels(prec_name)
    prof = traffic_profile(a, sec_src, pinfo[prec_name]["key"])
    traffic = None
    if prof is not None and all(k in prof["kernels"] for k in kern.values()):
        traffic = sum(prof["kernels"][k]["traffic_bytes"] * prof["kernels"][k].get("per_sigma", 1.0)
                      for k in kern.values())
    mfma_bound = F / (peak_f * 1e12) >= Bmin / (PEAK_HBM_GBS * 1e9)
    if mfma_bound:
        ach = F / t_sig / 1e12
        roof = dict(bound="mfma", achieved=ach, peak=peak_f, unit="TFLOP/s", frac=ach / peak_f, traffic=traffic)
    else:
        ach = Bmin / t_sig / 1e9
        roof = dict(bound="hbm", achieved=ach, peak=PEAK_HBM_GBS, unit="GB/s", frac=ach / PEAK_HBM_GBS,
                    traffic=traffic)
    roof.update(kernel="sigma: " + ", ".join(kern.values()), flops=F, algorithmic_bytes=Bmin,
                intensity_flop_per_byte=F / Bmin, avg_launch_ms=t_sig * 1e3,
                traffic_over_algorithmic=None if traffic is None else traffic / Bmin,
                traffic_profile=None if prof is None else prof["file"])
    # the slowest stage kernel on its own: executed flops and the implementation's
    # own bytes (its operands and intermediates) per launch
    dom = max(range(2, 5), key=lambda i: avg[i])
    kname = {2: kern["gemm_T"], 3: kern["w_apply"], 4: kern["gemm_out"]}[dom]
    t_s = avg[dom] * 1e-3
    ex_f = {2: work[0], 3: work[1], 4: work[2]}[dom]
    ex_b = {2: work[3], 3: work[4], 4: work[5]}[dom]
    roof["dominant_kernel"] = dict(
        kernel=kname, avg_launch_ms=avg[dom], executed_flops=ex_f, tflops=ex_f / t_s / 1e12,
        frac_of_mfma_peak=ex_f / t_s / 1e12 / peak_f, implementation_bytes=ex_b, gbs=ex_b / t_s / 1e9,
        traffic=None if prof is None or kname not in prof["kernels"] else prof["kernels"][kname]["traffic_bytes"])

    cpu, err = None, None
    if not a.no_cpu:
        cpu, err = cpu_baseline(ham, psi, Lb, Rb, W, r64["y"], a.cpu_seconds)
        log(f"cpu baseline {cpu['value'] * 1e3:.2f} GFLOP/s; GPU FP64 vs oracle max rel err {err:.2e}")

    def summary(name):
        r = res[name]
        pk = {"fp64": PEAK_FP64_TFLOPS, "mixed": PEAK_FP32_TFLOPS, "bf16x3": PEAK_BF16X3_TFLOPS}[name]
        out_ = dict(tflops=F / r["dt"] / 1e12, ms_per_sigma=r["dt"] * 1e3,
                    frac_of_mfma_peak=F / (sum(r["avg"][1:]) * 1e-3) / 1e12 / pk,
                    ms_per_davidson_iter=r["ms_jd"], stages_ms=dict(zip(STAGES, r["avg"])),
                    stage_implementation_bytes=dict(zip(STAGES[2:], r["work"][3:6])),
                    theta_after_jd_iters=r["theta"] + ham.ecore, bind_ms=r["bind_ms"])
        if conv is not None:
            c_ = conv[name]
            out_["solve"] = c_
            # one site solve: binding the boundaries, then the solve from the MPS guess
            out_["site_solve_ms"] = r["bind_ms"] + c_["solve_s"] * 1e3
        return out_

    out = {
        "metric": METRIC,
        "value": F / dt / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "mixed": "f32", "bf16x3": "bf16x3 (f32 accumulation)"}[prec_name],
        "data": "synthetic",
        "config": {
            "workload": f"{'SU2U1' if a.symm == 'su2u1' else 'TwoU1'} two-site sigma, L={a.L} N={a.L} S=0 "
                        f"m={a.m}{'' if a.decay is None else f' decay={a.decay:g}'}, centre bond "
                        f"(sites {site},{site + 1})",
            "sectors": sec_src,
            "psi_len": int(fp["n"]),
            "mpo_bond": [int(W.row_dim), int(W.col_dim)],
            "form": prec_name,
            "parallelism": "single GPU",
        },
        "ms_per_davidson_iter": ch["ms_jd"],
        "davidson_iters": ch["its"],
        "flops_per_sigma": F,
        "executed_flops_per_sigma": pinfo[prec_name]["exec_flops"],
        "plan_build_s": t_plan,
        "schedule": {k: asdict(v[0]) for k, v in tuned.items()},
        "plan_keys": {k: v["key"] for k, v in pinfo.items()},
        "autotune": None if a.no_autotune else {k: table_json(v[1][:5]) for k, v in tuned.items()},
        "fp64": summary("fp64"),
        **{k: summary(k) for k in LOW},
        **{f"{k}_speedup_vs_fp64": r64["dt"] / res[k]["dt"] for k in LOW},
        **{f"delta_E_{k}_vs_fp64_Eh": None if dE is None else dE[k] for k in LOW},
        "min_bytes_per_sigma": Bmin,
        **{f"{k}_sigma_max_rel_err": sig_err[k] for k in LOW},
        "block_davidson": block,
        "parity_max_rel_err_vs_oracle": err,
        "roofline": roof,
        "cpu_baseline": cpu,
        "other_configs": config_records() if a.config in (None, 1) else None,
    }
    print(json.dumps(out), flush=True)
    for p_ in plans.values():
        p_.close()
    eng.close()


if __name__ == "__main__":
    main()
