"""Measurements on an MPS through the engine's boundary steps.

rdm1: the spin-summed one-particle reduced density matrix
    gamma_pq = <psi| sum_sigma a+_{p sigma} a_{q sigma} |psi> / <psi|psi>.
Each element is the expectation value of a one-term MPO, evaluated with the
engine's boundary steps (expval, mp_tensors/mps_mpo_ops.h:42-103), the way the
reference's RDM measurements contract one MPO per index tuple through its
boundary kernels (measurements/tagged_nrank_rdm.h).  The one-term MPO of the
integral h_pq = h_qp = v is v (E_pq + E_qp), so v = 1/2 gives
(gamma_pq + gamma_qp) / 2 = gamma_pq for the real symmetric gamma of a real
wavefunction, and v = 1 on the diagonal gives gamma_pp.

rdm2: the spin-summed two-particle reduced density matrix
    Gamma_pqrs = <E_pq E_rs> - delta_qr <E_ps>,
so that E = sum_pq h_pq gamma_pq + 1/2 sum_pqrs (pq|rs) Gamma_pqrs + E_core,
averaged over the 8-fold permutation orbit of (pq|rs).  An orbit is measured
from the MPO of one two-electron integral (pq|rs) = 1, which the Hamiltonian
builder expands over the orbit, so the expectation value is 1/2 sum over the
orbit of Gamma.  For a real wavefunction Gamma itself is only 4-fold
symmetric (Gamma_pqrs = Gamma_rspq = Gamma_qpsr); the orbit average is the
part every real 8-fold-symmetric integral set contracts with, so it gives the
energy and the traces exactly, but not the exchange-type difference between
Gamma_pqrs and Gamma_qprs.
"""
from __future__ import annotations

import itertools

import numpy as np

from .sweep import Hamiltonian, expval


def _orbit(p, q, r, s):
    """The 8-fold permutation orbit of (pq|rs) for real orbitals."""
    return {(p, q, r, s), (q, p, r, s), (p, q, s, r), (q, p, s, r),
            (r, s, p, q), (s, r, p, q), (r, s, q, p), (s, r, q, p)}


def _one_term(ham, key, value):
    symm = "su2u1" if ham.su2 else "2u1"
    return Hamiltonian(ham.L, ham.nelec, ham.spin, [(key, value)], symm=symm)


def rdm1(engine, ham: Hamiltonian, mps):
    """(L, L) spin-summed one-particle RDM of `mps` (any normalisation)."""
    L = ham.L
    gamma = np.zeros((L, L))
    for p in range(L):
        for q in range(p + 1):
            v = 1.0 if p == q else 0.5
            e = expval(engine, _one_term(ham, (p + 1, q + 1, 0, 0), v), mps)
            gamma[p, q] = gamma[q, p] = e
    return gamma


def rdm2(engine, ham: Hamiltonian, mps):
    """(L, L, L, L) spin-summed two-particle RDM averaged over each permutation
    orbit of (pq|rs): one measurement per orbit."""
    L = ham.L
    Gamma = np.zeros((L, L, L, L))
    done = set()
    for p, q, r, s in itertools.product(range(L), repeat=4):
        if (p, q, r, s) in done:
            continue
        orb = _orbit(p, q, r, s)
        done |= orb
        e = expval(engine, _one_term(ham, (p + 1, q + 1, r + 1, s + 1), 1.0), mps)
        val = 2.0 * e / len(orb)
        for t in orb:
            Gamma[t] = val
    return Gamma
