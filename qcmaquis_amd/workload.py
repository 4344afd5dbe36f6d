"""Centre-bond site problems of the L-orbital half-filled singlet workload.

The sigma-vector cost is set by the symmetry-sector structure of the MPS
bonds.  A recorded structure (per bond a list of (N, 2S, size), written by
tests/golden/make_sectors.py from a real DMRG ramp) is turned into a random
right-normalised MPS with exactly those sectors; the boundaries up to the
centre bond are then built with the engine's own boundary steps.
"""
from __future__ import annotations

import json

from .blocks import Index
from .integrals import synthetic
from .mps import (bounded_sectors, left_normalize_and_return, make_twosite_mps, multiply_from_left, normalize_right,
                  random_mps_on)
from .sweep import Hamiltonian


def load_sectors(path):
    rec = json.load(open(path))
    bonds = [Index([((n, s2), size) for n, s2, size in b]) for b in rec["bonds"]]
    return rec, bonds


def stretched_sectors(path, L, m, phys, right_end, su2=True):
    """Bond sectors of an L-orbital state modelled on a sector structure recorded
    from a DMRG run on a shorter chain (tests/golden/make_sectors.py).

    Bond i of the long chain takes the sectors of the recorded bond at the same
    relative position, moved in particle number by the even shift that carries
    the recorded mean filling there (scaled by N / N_recorded) to this chain's,
    with each size scaled by m / m_recorded and capped at the sector's number of
    states; sectors this chain does not allow are dropped.  The result keeps the
    shape of a real ground state's bond (a few tens of sectors around the mean
    filling, small spins) where the degeneracy-proportional structure of
    bounded_sectors spreads m over every allowed sector."""
    import numpy as np
    from .mps import allowed_sectors
    rec, bonds = load_sectors(path)
    L0, N0 = rec["L"], rec["nelec"]
    m0 = max(b.sum_of_sizes() for b in bonds)
    nbar0 = np.array([sum(c[0] * n for c, n in b) / max(1, b.sum_of_sizes()) for b in bonds])
    allowed = allowed_sectors(L, phys, right_end, 1 << 30, su2)
    out = []
    for i in range(L + 1):
        x = i * L0 / L
        j = int(round(x))
        target = float(np.interp(x, np.arange(L0 + 1), nbar0)) * right_end[0] / N0
        shift = 2 * int(round((target - nbar0[j]) / 2))
        d = {}
        for c, n in bonds[j]:
            q = (c[0] + shift, c[1])
            if allowed[i].has(q):
                d[q] = min(allowed[i].size_of_block(q), max(1, int(round(n * m / m0))))
        if not d:
            raise ValueError(f"no recorded sector of bond {j} fits bond {i} of the L={L} chain")
        out.append(Index.from_dict(d))
    return out


class SiteProblem:
    """Two-site problem: Hamiltonian, psi0, boundaries, two-site MPO, the
    site index and the MPS it was cut from (sites < site left-normalised,
    sites > site + 1 right-normalised)."""

    def __init__(self, ham, psi, L, R, W, site, mps):
        self.ham, self.psi, self.L, self.R, self.W, self.site, self.mps = ham, psi, L, R, W, site, mps


def centre_problem(engine, L, m=None, sectors=None, site=None, seed=42, ham=None, symm="su2u1", decay=None,
                   log=None):
    """Build the SiteProblem of the two-site problem at `site` (default
    L/2 - 1).  `sectors` is a list of Index (one per bond, L + 1 entries);
    when None the degeneracy-proportional structure capped at m is used.
    log(msg), when given, reports progress."""
    if ham is None:
        ham = Hamiltonian(L, L, 0, synthetic(L, seed=1000 + L, decay=decay), symm=symm, progress=log)
    if sectors is None:
        sectors = bounded_sectors(L, ham.phys, ham.right_end, m, su2=ham.su2)
    if len(sectors) != L + 1:
        raise ValueError(f"sector structure has {len(sectors)} bonds, expected {L + 1}")
    mps = random_mps_on(sectors, ham.phys, seed=seed)
    normalize_right(mps)
    c = L // 2 - 1 if site is None else site
    R = engine.seed_boundary(ham.right_end)
    for i in range(L - 1, c + 1, -1):
        R = engine.right_step(ham.symm, mps[i], R, ham.mpo[i])
        if log and i % 4 == 0:
            log(f"right boundary at site {i}")
    Lb = engine.seed_boundary((0, 0))
    for i in range(0, c):
        if log and i % 4 == 0:
            log(f"left boundary at site {i}")
        Rf = left_normalize_and_return(mps[i])
        multiply_from_left(mps[i + 1], Rf)
        Lb = engine.left_step(ham.symm, mps[i], Lb, ham.mpo[i])
    W = ham.ts_mpo(c)
    psi = make_twosite_mps(mps[c], mps[c + 1], su2=ham.su2)
    return SiteProblem(ham, psi, Lb, R, W, c, mps)
