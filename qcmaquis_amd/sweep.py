"""Two-site DMRG sweep driver (the caller of the hot path).

Restates the microiteration loop of
SweepBasedAlgorithms/GenericSweepSimulation.h:73-152 with the two-site
trait of SweepOptimizationTypeTrait.h:87-144, the local solve of
SweepBasedEnergyMinimization.h:75-100, the split/merge of
SweepMPSUpdater.h:180-254 and the boundary propagation of
BoundaryPropagator.h:45-142.  The engine argument provides the hot path
(sigma vector / Jacobi-Davidson / boundary steps): `GpuEngine` in this
package, or the CPU oracle in tests.
"""
from __future__ import annotations

import time

from .blocks import Index
from .desc import MQ_SYMM_SU2U1, MQ_SYMM_TWOU1
from .integrals import to_aligned
from .mpo import TaggedMPOMaker, make_twosite_mpo
from .mps import (MPSTensor, left_normalize_and_return, make_twosite_mps, multiply_from_left,
                  multiply_from_right, normalize_right, random_mps, right_normalize_and_return, split_twosite)
from .operators import OpTable, SU2Ops, TwoU1Ops, su2_phys_index, twou1_phys_index
from .terms import su2_terms
from .twou1 import twou1_terms


class Hamiltonian:
    """Quantum-chemistry Hamiltonian as an MPO (harness input)."""

    def __init__(self, L, nelec, spin, integrals, symm="su2u1", herm=True, progress=None):
        """spin is 2S.  symm "su2u1" (models/chem/su2u1) or "2u1" (models/chem/2u1,
        spin-orbital basis with N_alpha - N_beta = 2S).  herm=False builds the SU2
        MPO without hermitian-partner bonds (every boundary slice is stored).
        progress(msg), when given, reports the stages of a long construction."""
        self._progress = progress
        self.L, self.nelec, self.spin = L, nelec, spin
        table = OpTable()
        if symm == "su2u1":
            self.symm = MQ_SYMM_SU2U1
            self.su2 = True
            ops = SU2Ops(table)
            self.terms = su2_terms(to_aligned(integrals), L, nelec, ops)
            mk = TaggedMPOMaker(L, table, ops.ident, ops.fill, ops.ident_full, herm=herm)
            self.phys = su2_phys_index()
            self.right_end = (nelec, spin)
        elif symm == "2u1":
            if (nelec + spin) % 2:
                raise ValueError("N and 2S must have the same parity")
            self.symm = MQ_SYMM_TWOU1
            self.su2 = False
            ops = TwoU1Ops(table)
            self.terms = twou1_terms(integrals, L, ops, progress=progress)
            mk = TaggedMPOMaker(L, table, ops.ident, ops.fill, su2=False, herm=False)
            self.phys = twou1_phys_index()
            self.right_end = ((nelec + spin) // 2, (nelec - spin) // 2)
        else:
            raise ValueError(f"unknown symmetry {symm!r}")
        step = max(1, len(self.terms) // 8)
        for i, t in enumerate(self.terms):
            if progress and i and i % step == 0:
                progress(f"MPO: {i}/{len(self.terms)} terms placed")
            mk.add_term(t)
        self.mpo, self.ecore = mk.create_mpo()
        if progress:
            progress(f"MPO built: {len(self.terms)} terms, bond dimensions up to "
                     f"{max(w.col_dim for w in self.mpo)}")
        self._ts = {}

    def ts_mpo(self, p):
        w = self._ts.get(p)
        if w is None:
            w = self._make_ts(p)
            self._ts[p] = w
        return w

    def _make_ts(self, p):
        return make_twosite_mpo(self.mpo[p], self.mpo[p + 1], self.phys, self.phys, su2=self.su2,
                                progress=getattr(self, "_progress", None))

    def prepare_twosite(self, workers=None, log=None):
        """Build the flattened two-site MPO of every site pair (the arrays the GPU
        plans read) in `workers` forked processes, default min(16, CPUs).  ts_mpo(p)
        then returns a FlatTwoSiteMPO, which builds the full tensor in this process
        only if something other than a plan asks for it.  Call it before the GPU is
        initialised: the workers are forked from this process."""
        import multiprocessing as mp
        import os
        global _PREP_HAM
        todo = [p for p in range(self.L - 1) if p not in self._ts]
        n = min(len(todo), workers or min(16, os.cpu_count() or 1))
        if n <= 1:
            return
        t0 = time.perf_counter()
        _PREP_HAM = self
        try:
            with mp.get_context("fork").Pool(n) as pool:
                for p, arrays in pool.imap_unordered(_flat_twosite, todo):
                    self._ts[p] = FlatTwoSiteMPO(self, p, arrays)
        finally:
            _PREP_HAM = None
        if log:
            log(f"two-site MPOs of {len(todo)} site pairs built on {n} processes in {time.perf_counter() - t0:.1f} s")


_PREP_HAM = None  # the Hamiltonian whose two-site MPOs forked workers build


def _flat_twosite(p):
    from .desc import MPOArrays
    return p, MPOArrays(_PREP_HAM._make_ts(p))


class FlatTwoSiteMPO:
    """A two-site MPO tensor held as its flattened arrays (desc.MPOArrays: what the
    plans read); any other attribute builds the full tensor (make_twosite_mpo) on
    first use and is read from it."""

    def __init__(self, ham, p, arrays):
        self.row_dim, self.col_dim = arrays.row_dim, arrays.col_dim
        self._mq_arrays = arrays
        self._ham, self._p, self._full = ham, p, None

    def __getattr__(self, k):  # attributes not set in __init__
        if k.startswith("__") or "_full" not in self.__dict__:
            raise AttributeError(k)
        if self._full is None:
            self._full = self._ham._make_ts(self._p)
            self._full._mq_arrays = self._mq_arrays
        return getattr(self._full, k)


def n_microiterations(L):
    return 1 if L == 2 else 2 * (L - 2)


def micro_site(L, i):
    return i if i < L - 1 else 2 * L - 4 - i


class TwoSiteDMRG:
    def __init__(self, ham: Hamiltonian, engine, max_bond_dimension, init_bond_dimension=5, seed=42,
                 cutoff=1e-16, jd_tol=1e-8, jd_maxiter=10, mps=None, log=None, eigensolver="jd", split="engine"):
        """eigensolver: "jd" (Jacobi-Davidson, IETL_JCD) or "davidson" (diagonal-preconditioned
        Davidson, IETL_DAVIDSON, SU2 only), as parms["eigensolver"] selects in
        optimize/ts_optimize.hpp:139-147; both take ietl_jcd_maxiter / ietl_jcd_tol.
        split: "engine" truncates with the engine's split_twosite when it has one
        (the GPU engine: block SVD on the device), "host" with mps.split_twosite."""
        if eigensolver not in ("jd", "davidson"):
            raise ValueError(f"unknown eigensolver {eigensolver!r}")
        if split not in ("engine", "host"):
            raise ValueError(f"unknown split {split!r}")
        self.split = split
        self.solve = engine.jd if eigensolver == "jd" else engine.davidson
        self.ham = ham
        self.engine = engine
        self.m = max_bond_dimension
        self.cutoff = cutoff
        self.tol = jd_tol
        self.maxiter = jd_maxiter
        self.log = log
        L = ham.L
        if mps is None:
            mps = random_mps(L, ham.phys, ham.right_end, init_bond_dimension, seed=seed, su2=ham.su2)
        self.mps = mps
        normalize_right(self.mps)
        self.left = [None] * (L + 1)
        self.right = [None] * (L + 1)
        self.left[0] = engine.seed_boundary((0, 0))
        self.right[L] = engine.seed_boundary(ham.right_end)
        for i in range(L - 1, 0, -1):
            self.right[i] = engine.right_step(ham.symm, self.mps[i], self.right[i + 1], ham.mpo[i])
        self.energies = []
        self.records = []

    def site_problem(self, site):
        """Two-site psi0, boundaries and MPO of the microiteration at `site`."""
        psi = make_twosite_mps(self.mps[site], self.mps[site + 1], su2=self.ham.su2)
        return psi, self.left[site], self.right[site + 2], self.ham.ts_mpo(site)

    def microiteration(self, sweep, i):
        ham, L = self.ham, self.ham.L
        site = micro_site(L, i)
        forward = i < L - 1
        change = (i == L - 2) or (i == n_microiterations(L) - 1)
        t0 = time.perf_counter()
        psi, Lb, Rb, W = self.site_problem(site)
        ts = time.perf_counter()
        theta, u, its = self.solve(ham.symm, psi, Lb, Rb, W, self.maxiter, self.tol)
        t1 = time.perf_counter()
        E = theta + ham.ecore
        l2r = forward and not change
        dev_split = getattr(self.engine, "split_twosite", None) if self.split == "engine" else None
        if dev_split is not None:
            A, B, info = dev_split(u, ham.phys, ham.phys, self.mps[site].left, self.mps[site + 1].right,
                                   self.m, self.cutoff, l2r=l2r, su2=ham.su2)
        else:
            ut = MPSTensor.from_flat(psi.phys, psi.left, psi.right, u)
            A, B, info = split_twosite(ut, ham.phys, ham.phys, self.mps[site].left, self.mps[site + 1].right,
                                       self.m, self.cutoff, l2r=l2r, su2=ham.su2)
        tsplit = time.perf_counter()
        self.mps[site], self.mps[site + 1] = A, B
        if l2r:
            self.left[site + 1] = self.engine.left_step(ham.symm, self.mps[site], self.left[site], ham.mpo[site])
            if site + 2 < L:
                R = left_normalize_and_return(self.mps[site + 1])
                multiply_from_left(self.mps[site + 2], R)
            else:
                left_normalize_and_return(self.mps[site + 1])
        else:
            self.right[site + 1] = self.engine.right_step(ham.symm, self.mps[site + 1], self.right[site + 2],
                                                          ham.mpo[site + 1])
            if site > 0:
                Lf = right_normalize_and_return(self.mps[site])
                multiply_from_right(self.mps[site - 1], Lf)
            else:
                right_normalize_and_return(self.mps[site])
        t2 = time.perf_counter()
        # t_jd: site problem and local solve; t_rest: truncation and boundary step, of
        # which t_site (two-site tensor and MPO), t_split (the SVD truncation, on the
        # device with the GPU engine) and t_step (boundary step and normalisation) are parts
        rec = dict(sweep=sweep, micro=i, site=site, energy=E, jd_iters=its, bond=info["bond_dimension"],
                   truncated_weight=info["truncated_weight"], t_jd=t1 - t0, t_rest=t2 - t1,
                   t_site=ts - t0, t_split=tsplit - t1, t_step=t2 - tsplit, dim=len(u))
        rec.update(getattr(self.engine, "last_timing", None) or {})
        if dev_split is not None:
            rec.update(getattr(self.engine, "last_split", None) or {})
        self.records.append(rec)
        self.energies.append(E)
        if self.log:
            self.log(rec)
        return E

    def sweep(self, s):
        for i in range(n_microiterations(self.ham.L)):
            self.microiteration(s, i)
        return self.energies[-1]

    def run(self, nsweeps):
        for s in range(nsweeps):
            self.sweep(s)
        return self.energies[-1]


class StateAveragedDMRG(TwoSiteDMRG):
    """The nroots lowest states in one shared basis (state-averaged two-site DMRG,
    equal weights; no counterpart in the reference, which optimises excited
    states one at a time under orthogonality constraints).

    Every local problem is solved for its nroots lowest eigenpairs together
    (engine.block_solve: the batched-sigma block Davidson on the GPU, sequential
    constrained Jacobi-Davidson on the oracle), and the truncation keeps the
    leading singular vectors of the roots' two-site tensors side by side
    (split_twosite_sa).  The roots' own parts travel with the sweep as
    `centers` at `center_site`, one of the two sites of the next problem (the
    backward pass ends at site 1 and the next forward pass starts at site 0)."""

    def __init__(self, ham, engine, max_bond_dimension, nroots, seed=42, **kw):
        super().__init__(ham, engine, max_bond_dimension, seed=seed, **kw)
        import numpy as np
        rng = np.random.default_rng(seed + 1)
        c0 = self.mps[0]
        self.nroots = nroots
        self.centers = [c0.copy()]
        for _ in range(1, nroots):
            c = c0.copy()
            c.make_left_paired()
            for k in c.data.blocks:
                c.data.blocks[k] = rng.standard_normal(c.data.blocks[k].shape)
            self.centers.append(c)
        self.center_site = 0
        self.root_energies = []

    def microiteration(self, sweep, i):
        from .mps import split_twosite_sa
        ham, L = self.ham, self.ham.L
        site = micro_site(L, i)
        forward = i < L - 1
        change = (i == L - 2) or (i == n_microiterations(L) - 1)
        t0 = time.perf_counter()
        if self.center_site == site:
            psis = [make_twosite_mps(c, self.mps[site + 1], su2=ham.su2) for c in self.centers]
        elif self.center_site == site + 1:
            psis = [make_twosite_mps(self.mps[site], c, su2=ham.su2) for c in self.centers]
        else:
            raise RuntimeError(f"state-averaged sweep: roots at site {self.center_site}, problem at {site}")
        Lb, Rb, W = self.left[site], self.right[site + 2], ham.ts_mpo(site)
        thetas, us, its = self.engine.block_solve(ham.symm, psis, Lb, Rb, W, self.maxiter, self.tol)
        t1 = time.perf_counter()
        psi = psis[0]
        uts = [MPSTensor.from_flat(psi.phys, psi.left, psi.right, u) for u in us]
        l2r = forward and not change
        left_i, right_i = self.mps[site].left, self.mps[site + 1].right
        if l2r:
            A, Bs, info = split_twosite_sa(uts, ham.phys, ham.phys, left_i, right_i, self.m, self.cutoff, l2r=True,
                                           su2=ham.su2)
            self.mps[site], self.mps[site + 1] = A, Bs[0]
            self.centers, self.center_site = Bs, site + 1
            self.left[site + 1] = self.engine.left_step(ham.symm, A, self.left[site], ham.mpo[site])
        else:
            As, B, info = split_twosite_sa(uts, ham.phys, ham.phys, left_i, right_i, self.m, self.cutoff, l2r=False,
                                           su2=ham.su2)
            self.mps[site], self.mps[site + 1] = As[0], B
            self.centers, self.center_site = As, site
            self.right[site + 1] = self.engine.right_step(ham.symm, B, self.right[site + 2], ham.mpo[site + 1])
        energies = [t + ham.ecore for t in thetas]
        rec = dict(sweep=sweep, micro=i, site=site, energies=energies, iters=its, bond=info["bond_dimension"],
                   truncated_weight=info["truncated_weight"], t_solve=t1 - t0, t_rest=time.perf_counter() - t1)
        self.records.append(rec)
        self.root_energies.append(energies)
        self.energies.append(energies[0])
        if self.log:
            self.log(rec)
        return energies

    def run(self, nsweeps):
        for s in range(nsweeps):
            self.sweep(s)
        return self.root_energies[-1]


def expval(engine, ham: Hamiltonian, mps, from_right=False):
    """<mps|H|mps> / <mps|mps> by propagating a boundary through the whole MPS
    with the engine's boundary steps (mp_tensors/mps_mpo_ops.h:42-103: the trace
    of the fully propagated boundary plus E_core times the overlap)."""
    L = ham.L
    if not from_right:
        b = engine.seed_boundary((0, 0))
        for i in range(L):
            b = engine.left_step(ham.symm, mps[i], b, ham.mpo[i])
    else:
        b = engine.seed_boundary(ham.right_end)
        for i in range(L - 1, -1, -1):
            b = engine.right_step(ham.symm, mps[i], b, ham.mpo[i])
    if b.layout.aux != 1 or b.layout.nblocks != 1:
        raise RuntimeError("the propagated boundary is not a scalar")
    h = float(b.data[0])
    # overlap: norm of the last tensor once all others are left-normalised
    m = [t.copy() for t in mps]
    for i in range(L - 1):
        Rf = left_normalize_and_return(m[i])
        multiply_from_left(m[i + 1], Rf)
    nrm2 = m[-1].norm() ** 2
    return h / nrm2 + ham.ecore
