"""Schedule autotuner for the sigma-vector plans.

A plan's schedule has four choices that the task lists depend on but the
results do not (they agree to round-off):
  - the fused-column rule: which MPO columns are summed on the left boundary
    first (mode 0 none, 1 by an entries/positions ratio and a minimum entry
    count, 2 all), trading W-application traffic for GEMM flops;
  - the split-K chunk of the output GEMM groups;
  - the GEMM row tile (0: 128 rows unless a group fits 64; or 64 for every group,
    twice the tiles with half the accumulators each);
  - the GEMM column tile (0: per group the width that wastes the fewest padded
    columns; or one width, 32, 64 or 128, for every group).
The tuner builds the plan of one site problem under each candidate schedule,
times a few sigma vectors in the requested precision with HIP events on the
library stream, and keeps the fastest.  Results are cached per site shape
(the MPO bond dimensions and the boundary and vector sizes), so a sweep tunes
each distinct problem once.
"""
from dataclasses import dataclass, asdict

import numpy as np

from .desc import MQ_PREC_FP64


@dataclass(frozen=True)
class Schedule:
    fuse: int = 1
    ratio: int = 4
    min_entries: int = 64
    kchunk: int = 2048
    tile_m: int = 0
    sched: int = 0  # MQ_OPT_SCHEDULE: 0 lbtm everywhere, 1 the reference's lbtm/rbtm choice
    tile_n: int = 0  # MQ_OPT_TILE_N: 0 the padding-minimising width per group, else 16..128

    def label(self):
        tm = (f" tile_m={self.tile_m}" if self.tile_m else "") + (f" tile_n={self.tile_n}" if self.tile_n else "")
        rule = " sched=reference" if self.sched else ""
        if self.fuse != 1:
            return f"fuse={self.fuse} kchunk={self.kchunk}{tm}{rule}"
        return f"fuse=1 ratio={self.ratio} min={self.min_entries} kchunk={self.kchunk}{tm}{rule}"


DEFAULT = Schedule()


def candidates():
    """Fused-column rules around the default (ratio x minimum entries), the two
    extremes, split-K chunks of 1024 / 2048 / 4096, the 64-row tile and fixed
    column tiles of 32 / 64 / 128 (and 64 x 64) for the default rule."""
    out = [DEFAULT, Schedule(fuse=0), Schedule(fuse=2)]
    for ratio in (2, 3, 6, 8):
        for mn in (1, 64):
            out.append(Schedule(1, ratio, mn))
    out.append(Schedule(1, 4, 1))
    out.append(Schedule(1, 4, 16))
    for kc in (1024, 4096):
        out.append(Schedule(kchunk=kc))
    out.append(Schedule(tile_m=64))
    for tn in (32, 64, 128):
        out.append(Schedule(tile_n=tn))
    out.append(Schedule(tile_m=64, tile_n=64))
    return out


def site_key(symm, psi, L, R, W):
    return (symm, W.row_dim, W.col_dim, psi.to_flat().size, L.layout.total, R.layout.total)


def time_schedule(engine, symm, psi, L, R, W, sched, x, prec, reps=5, warmup=2):
    """Median stage-summed sigma time (ms) of one schedule, and the plan's
    executed FLOPs and workspace bytes."""
    prev = getattr(engine, "schedule", DEFAULT)
    engine.set_schedule(sched)
    try:
        sp = engine.sigma_plan(symm, psi, L, R, W)
    finally:
        engine.set_schedule(prev)
    try:
        y = engine.torch.empty_like(x)
        for _ in range(warmup):
            sp.apply(x, y, prec=prec)
        ts = []
        for _ in range(reps):
            ms, _ = sp.timed(x, y, prec=prec)
            ts.append(sum(ms[1:]))
        return float(np.median(ts)), sp.exec_flops, sp.workspace_bytes
    finally:
        sp.close()


def tune(engine, symm, psi, L, R, W, prec=MQ_PREC_FP64, cands=None, reps=5, log=None, margin=0.05):
    """Fastest schedule for this site problem in precision `prec`, with the
    timing table [(schedule, ms, exec_flops, workspace_bytes)] sorted by time.
    The default schedule is kept unless another is faster by more than
    `margin` (a fraction of the default's time): candidates within timing
    noise of it would otherwise change the pick from run to run."""
    cache = engine.__dict__.setdefault("tuned", {})
    key = (site_key(symm, psi, L, R, W), prec)
    if key in cache and cands is None:
        return cache[key]
    x = engine.to_dev(np.random.default_rng(7).standard_normal(psi.to_flat().size))
    table = []
    for s in (cands or candidates()):
        ms, ex, ws = time_schedule(engine, symm, psi, L, R, W, s, x, prec, reps)
        table.append((s, ms, ex, ws))
        if log:
            log(f"  schedule {s.label()}: {ms:.3f} ms, executed {ex / 1e9:.1f} GF, workspace {ws / 1e9:.2f} GB")
    table.sort(key=lambda r: r[1])
    best = table[0][0]
    dflt = [r for r in table if r[0] == DEFAULT]
    if dflt and table[0][1] > (1.0 - margin) * dflt[0][1]:
        best = DEFAULT
    res = (best, table)
    cache[key] = res
    return res


def table_json(table):
    return [dict(asdict(s), ms=ms, exec_gflop=ex / 1e9, workspace_gb=ws / 1e9) for s, ms, ex, ws in table]
