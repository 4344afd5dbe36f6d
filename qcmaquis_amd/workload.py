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
