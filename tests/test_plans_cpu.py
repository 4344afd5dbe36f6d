"""Host-side checks of libmqgpu without a GPU: the C ABI exports, and the
device task lists executed by a numpy emulator against the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT, ham_from_fixture


def test_library_exports_every_declared_symbol():
    from qcmaquis_amd import gpu
    hdr = open(os.path.join(ROOT, "include", "mqgpu.h")).read()
    names = sorted(set(re.findall(r"\b(mq_[a-z0-9_]+)\s*\(", hdr)))
    assert len(names) >= 20
    lib = C.CDLL(gpu.LIBPATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def _check_site(ham, d, site, oracle_engine):
    import emulate
    from qcmaquis_amd.desc import SiteDesc
    from qcmaquis_amd.gpu import HostPlan
    psi, Lb, Rb, W = d.site_problem(site)
    x = psi.to_flat()
    y, fl = oracle_engine.sigma(ham.symm, psi, Lb, Rb, W, x=x)
    # fuse 0: every MPO term applied on its own; 1: high fan-in columns summed on the
    # left boundary first; 2: every column summed that way
    for fuse in (0, 1, 2):
        hp = HostPlan(SiteDesc(ham.symm, psi.phys, psi.left, psi.right, W, Lb.layout, Rb.layout), "sigma", fuse)
        if fuse == 2:
            assert len(hp.parts["comb_strips"]) > 0
        ye = emulate.run_sigma(hp, Lb.data, Rb.data, x)
        assert np.abs(y - ye).max() <= 1e-13 * max(1.0, np.abs(y).max()), fuse
        assert hp.ref_flops == fl  # algorithmic FLOPs of the reference schedule
    for kind, A, B, Wm in (("lstep", d.mps[site], Lb, ham.mpo[site]),
                           ("rstep", d.mps[site + 1], Rb, ham.mpo[site + 1])):
        if kind == "lstep":
            ref = oracle_engine.left_step(ham.symm, A, B, Wm)
            sd = SiteDesc(ham.symm, A.phys, A.left, A.right, Wm, B.layout, None)
        else:
            ref = oracle_engine.right_step(ham.symm, A, B, Wm)
            sd = SiteDesc(ham.symm, A.phys, A.left, A.right, Wm, None, B.layout)
        for fuse in ((0, 1, 2) if kind == "lstep" else (1,)):
            hps = HostPlan(sd, kind, fuse)
            out = emulate.run_step(hps, kind, B.data, A.to_flat())
            got = hps.out_layout.to_dicts(out)
            exp = ref.dicts()
            assert len(got) == len(exp)
            for b in range(len(exp)):
                assert set(got[b]) == set(exp[b]), (kind, b)
                for k in exp[b]:
                    assert np.abs(got[b][k] - exp[b][k]).max() <= 1e-13 * max(1.0, np.abs(exp[b][k]).max()), fuse


@pytest.mark.parametrize("name", ["lih", "benzene"])
def test_plan_emulation_matches_oracle(fixtures, oracle_engine, name):
    from qcmaquis_amd.sweep import TwoSiteDMRG, micro_site, n_microiterations
    ham = ham_from_fixture(fixtures[name])
    d = TwoSiteDMRG(ham, oracle_engine, 60)
    for i in range(n_microiterations(ham.L)):
        _check_site(ham, d, micro_site(ham.L, i), oracle_engine)
        d.microiteration(0, i)


def test_plan_emulation_synthetic(oracle_engine):
    from qcmaquis_amd.integrals import synthetic
    from qcmaquis_amd.sweep import Hamiltonian, TwoSiteDMRG, micro_site
    ham = Hamiltonian(8, 8, 0, synthetic(8, seed=1008))
    d = TwoSiteDMRG(ham, oracle_engine, 48)
    for i in range(5):
        if i in (2, 3, 4):
            _check_site(ham, d, micro_site(ham.L, i), oracle_engine)
        d.microiteration(0, i)


def test_plan_rejects_mismatched_boundary(fixtures, oracle_engine):
    from qcmaquis_amd.desc import BoundaryLayout, SiteDesc
    from qcmaquis_amd.gpu import HostPlan, MQError
    from qcmaquis_amd.sweep import TwoSiteDMRG
    ham = ham_from_fixture(fixtures["lih"])
    d = TwoSiteDMRG(ham, oracle_engine, 20)
    psi, Lb, Rb, W = d.site_problem(0)
    bad = BoundaryLayout(W.row_dim + 1, [0] * (W.row_dim + 2), [], [], [], [], [], 0)
    with pytest.raises(MQError):
        HostPlan(SiteDesc(ham.symm, psi.phys, psi.left, psi.right, W, bad, Rb.layout), "sigma")


def test_rbtm_plans_match_oracle(fixtures, oracle_engine):
    """The reference's per-site schedule (MQ_HOST_SCHED_REFERENCE): at the sites
    where site_hamil.hpp:35 picks rbtm, the rbtm plan (t = psi R, rbtm W
    application, S = L^T BUF, right-to-left reshape) executed by the emulator
    equals the oracle's sigma, executes the reference's FLOPs to within 5%, and
    has the same live mask as the lbtm plan."""
    import emulate
    from qcmaquis_amd.desc import SiteDesc
    from qcmaquis_amd.gpu import HostPlan
    from qcmaquis_amd.integrals import synthetic
    from qcmaquis_amd.sweep import Hamiltonian, TwoSiteDMRG
    n_rbtm = 0
    for ham in (ham_from_fixture(fixtures["benzene"]), Hamiltonian(8, 8, 0, synthetic(8, seed=1008))):
        d = TwoSiteDMRG(ham, oracle_engine, 48)
        d.sweep(0)
        for site in range(ham.L - 1):
            psi, Lb, Rb, W = d.site_problem(site)
            sd = SiteDesc(ham.symm, psi.phys, psi.left, psi.right, W, Lb.layout, Rb.layout)
            h1 = HostPlan(sd, "sigma", schedule=1)
            h0 = HostPlan(sd, "sigma")
            assert not h0.rbtm
            x = np.random.default_rng(site).standard_normal(psi.to_flat().size)
            y, fl = oracle_engine.sigma(ham.symm, psi, Lb, Rb, W, x=x)
            ye = emulate.run_sigma(h1, Lb.data, Rb.data, x)
            assert np.abs(y - ye).max() <= 1e-13 * max(1.0, np.abs(y).max()), site
            assert h1.ref_flops == fl
            assert np.array_equal(h1.parts["live"], h0.parts["live"])
            if h1.rbtm:
                n_rbtm += 1
                assert abs(h1.exec_flops - fl) <= 0.05 * fl, (site, h1.exec_flops, fl)
            d.microiteration(1, site)
    assert n_rbtm >= 4


def test_plans_do_not_depend_on_the_host_thread_count():
    """The builder gathers columns, rows and fused groups on several host threads
    and emits them in a fixed order: plans built on 1 and on 8 threads are equal
    field by field (every task list, arena size and FLOP count)."""
    import json
    import subprocess
    import sys
    res = {}
    for nt in ("1", "8"):
        env = dict(os.environ, MQ_HOST_THREADS=nt)
        p = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "plan_digest.py")], env=env,
                           capture_output=True, text=True, timeout=600)
        assert p.returncode == 0, p.stderr[-2000:]
        res[nt] = json.loads(p.stdout.strip().splitlines()[-1])
    assert len(res["1"]) > 100
    diff = [k for k in res["1"] if res["1"][k] != res["8"].get(k)]
    assert not diff, diff[:10]
