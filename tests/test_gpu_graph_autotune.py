"""Captured HIP graphs of the sigma stages and the schedule autotuner, on the
GPU: both change how the work is issued, never the result."""
import numpy as np
import pytest

from conftest import ham_from_fixture

pytestmark = pytest.mark.gpu


def _benzene(oracle_engine, gpu_engine, fixtures):
    from qcmaquis_amd.sweep import TwoSiteDMRG
    ham = ham_from_fixture(fixtures["benzene"])
    d = TwoSiteDMRG(ham, oracle_engine, 60)
    for i in range(3):
        d.microiteration(0, i)
    psi, L, R, W = d.site_problem(3)
    return ham, psi, gpu_engine.upload_boundary(L.layout, L.data), gpu_engine.upload_boundary(R.layout, R.data), W


def test_sigma_graph_replays_match_eager(fixtures, oracle_engine, gpu_engine):
    """Several replays with new inputs, eager sigma vectors and host syncs in
    between, then a precision change (new capture): every replay equals the
    eager result bit for bit."""
    from qcmaquis_amd.desc import MQ_PREC_FP32, MQ_PREC_FP64
    ham, psi, Lg, Rg, W = _benzene(oracle_engine, gpu_engine, fixtures)
    sp = gpu_engine.sigma_plan(ham.symm, psi, Lg, Rg, W)
    try:
        rng = np.random.default_rng(11)
        for prec in (MQ_PREC_FP64, MQ_PREC_FP32, MQ_PREC_FP64):
            for rep in range(4):
                x = gpu_engine.to_dev(rng.standard_normal(sp.n))
                print(f"graph replay prec {prec} rep {rep}", flush=True)
                yg = sp.apply_graph(x, prec=prec).cpu().numpy()
                other = gpu_engine.to_dev(rng.standard_normal(sp.n))
                sp.apply(other, prec=prec)
                gpu_engine.sync()
                ye = sp.apply(x, prec=prec).cpu().numpy()
                assert np.array_equal(yg, ye), (prec, rep, np.abs(yg - ye).max())
    finally:
        sp.close()


def test_davidson_drivers_with_graphs(fixtures, oracle_engine, gpu_engine):
    ham, psi, Lg, Rg, W = _benzene(oracle_engine, gpu_engine, fixtures)
    x = psi.to_flat()
    res = {}
    try:
        for on in (False, True):
            gpu_engine.set_graph(on)
            sp = gpu_engine.sigma_plan(ham.symm, psi, Lg, Rg, W)
            try:
                xd = gpu_engine.to_dev(x)
                res[on] = (sp.jd(xd, 12, 1e-12)[::2], sp.davidson(xd, 40, 1e-12)[::2])
            finally:
                sp.close()
    finally:
        gpu_engine.set_graph(False)
    (j0, d0), (j1, d1) = res[False], res[True]
    assert j0[1] == j1[1] and d0[1] == d1[1]
    assert abs(j0[0] - j1[0]) <= 1e-13 * abs(j0[0])
    assert abs(d0[0] - d1[0]) <= 1e-13 * abs(d0[0])


def test_autotune_picks_a_schedule_with_the_same_sigma(fixtures, oracle_engine, gpu_engine):
    from qcmaquis_amd.autotune import DEFAULT, candidates, tune
    from qcmaquis_amd.desc import MQ_PREC_FP64
    ham, psi, Lg, Rg, W = _benzene(oracle_engine, gpu_engine, fixtures)
    best, table = tune(gpu_engine, ham.symm, psi, Lg, Rg, W, prec=MQ_PREC_FP64, reps=2)
    assert len(table) == len(candidates())
    # the fastest candidate, or the default when that is within 5% of the fastest
    dms = [r[1] for r in table if r[0] == DEFAULT][0]
    assert best == table[0][0] or (best == DEFAULT and table[0][1] > 0.95 * dms)
    assert all(table[i][1] <= table[i + 1][1] for i in range(len(table) - 1))
    x = gpu_engine.to_dev(np.random.default_rng(3).standard_normal(psi.to_flat().size))
    ys = []
    for sched in (DEFAULT, best):
        gpu_engine.set_schedule(sched)
        sp = gpu_engine.sigma_plan(ham.symm, psi, Lg, Rg, W)
        try:
            ys.append(sp.apply(x).cpu().numpy())
        finally:
            sp.close()
    gpu_engine.set_schedule(DEFAULT)
    assert np.abs(ys[0] - ys[1]).max() <= 1e-13 * np.abs(ys[0]).max()


def test_every_candidate_schedule_gives_the_same_sigma(fixtures, oracle_engine, gpu_engine):
    """Each candidate schedule (fused-column rules, split-K chunks, the 64-row
    tile) in FP64, FP32 and BF16x3 against the default schedule's FP64 sigma."""
    from qcmaquis_amd.autotune import DEFAULT, candidates
    from qcmaquis_amd.desc import MQ_PREC_BF16X3, MQ_PREC_FP32, MQ_PREC_FP64
    ham, psi, Lg, Rg, W = _benzene(oracle_engine, gpu_engine, fixtures)
    x = gpu_engine.to_dev(np.random.default_rng(4).standard_normal(psi.to_flat().size))
    ref = None
    try:
        for sched in [DEFAULT] + candidates():
            gpu_engine.set_schedule(sched)
            sp = gpu_engine.sigma_plan(ham.symm, psi, Lg, Rg, W)
            try:
                for prec, bar in ((MQ_PREC_FP64, 1e-13), (MQ_PREC_FP32, 1e-5), (MQ_PREC_BF16X3, 1e-4)):
                    y = sp.apply(x, prec=prec).cpu().numpy()
                    if ref is None:
                        ref = y
                    err = np.abs(y - ref).max() / np.abs(ref).max()
                    assert err <= bar, (sched.label(), prec, err)
            finally:
                sp.close()
    finally:
        gpu_engine.set_schedule(DEFAULT)


@pytest.mark.parametrize("nvec", [1, 3, 4])
def test_sigma_batch_equals_single_vectors(fixtures, oracle_engine, gpu_engine, nvec):
    """mq_sigma_batch: each output row equals mq_sigma of its input row (the
    same tiles and K chunks per vector), in FP64 and in the mixed form."""
    from qcmaquis_amd.desc import MQ_PREC_FP32, MQ_PREC_FP64
    ham, psi, Lg, Rg, W = _benzene(oracle_engine, gpu_engine, fixtures)
    sp = gpu_engine.sigma_plan(ham.symm, psi, Lg, Rg, W)
    try:
        X = gpu_engine.torch.as_tensor(np.random.default_rng(nvec).standard_normal((nvec, sp.n))).to(gpu_engine.device)
        for prec in (MQ_PREC_FP64, MQ_PREC_FP32):
            print(f"batch nvec {nvec} prec {prec}", flush=True)
            Y = sp.apply_batch(X, prec=prec).cpu().numpy()
            for j in range(nvec):
                y = sp.apply(X[j].contiguous(), prec=prec).cpu().numpy()
                assert np.array_equal(Y[j], y), (prec, j, np.abs(Y[j] - y).max())
    finally:
        sp.close()


def test_plans_without_the_reference_count(fixtures, oracle_engine, gpu_engine):
    """MQ_OPT_REF_FLOPS = 0 skips the reference FLOP count (ref_flops -1) and
    changes nothing else: same sigma, same executed FLOPs."""
    ham, psi, Lg, Rg, W = _benzene(oracle_engine, gpu_engine, fixtures)
    x = gpu_engine.to_dev(np.random.default_rng(5).standard_normal(psi.to_flat().size))
    out = []
    try:
        for on in (True, False):
            gpu_engine.set_ref_flops(on)
            sp = gpu_engine.sigma_plan(ham.symm, psi, Lg, Rg, W)
            try:
                out.append((sp.ref_flops, sp.exec_flops, sp.apply(x).cpu().numpy()))
            finally:
                sp.close()
    finally:
        gpu_engine.set_ref_flops(True)
    assert out[0][0] > 0 and out[1][0] == -1 and out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2])


def test_release_frees_caches_and_calls_rebuild_them(fixtures, oracle_engine, gpu_engine):
    """mq_plan_release drops the batched lists and arenas, the Davidson vectors and
    the captured graph; the next batched sigma, graph replay and block Davidson
    rebuild them and give the same results."""
    from qcmaquis_amd.desc import MQ_PREC_FP32, MQ_PREC_FP64
    ham, psi, Lg, Rg, W = _benzene(oracle_engine, gpu_engine, fixtures)
    sp = gpu_engine.sigma_plan(ham.symm, psi, Lg, Rg, W)
    try:
        X = gpu_engine.torch.as_tensor(np.random.default_rng(6).standard_normal((3, sp.n))).to(gpu_engine.device)
        before = []
        for rep in range(2):
            out = [sp.apply_batch(X, prec=MQ_PREC_FP64).cpu().numpy(),
                   sp.apply_batch(X, prec=MQ_PREC_FP32).cpu().numpy(),
                   sp.apply_graph(X[0].contiguous(), prec=MQ_PREC_FP64).cpu().numpy()]
            th, _, its = sp.block_davidson(X, 100, 1e-10, prec=MQ_PREC_FP64)
            out.append(np.array(th))
            sp.release()
            before.append(out)
        for a, b in zip(*before):
            assert np.array_equal(a, b)
    finally:
        sp.close()
