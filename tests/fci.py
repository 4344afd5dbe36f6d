"""Brute-force FCI in a fixed (N_alpha, N_beta) sector -- test infrastructure.

H = sum_pq h_pq E_pq + 1/2 sum_pqrs (pq|rs) (E_pq E_rs - delta_qr E_ps) + E_core
with spin-summed excitation operators E_pq built from alpha/beta string
tables (determinants ordered alpha-string-major, alpha creators first).
"""
import itertools

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla


def _strings(L, n):
    out = []
    for occ in itertools.combinations(range(L), n):
        s = 0
        for o in occ:
            s |= 1 << o
        out.append(s)
    return out


def _exc_tables(L, strings):
    index = {s: i for i, s in enumerate(strings)}
    mats = {}
    for p in range(L):
        for q in range(L):
            rows, cols, vals = [], [], []
            for j, s in enumerate(strings):
                if not (s >> q) & 1:
                    continue
                t = s & ~(1 << q)
                if (t >> p) & 1:
                    continue
                # a_p^dag a_q sign: count occupied orbitals below q in s, below p in t
                sgn = (-1) ** (bin(s & ((1 << q) - 1)).count("1") + bin(t & ((1 << p) - 1)).count("1"))
                u = t | (1 << p)
                rows.append(index[u])
                cols.append(j)
                vals.append(sgn)
            n = len(strings)
            mats[(p, q)] = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    return mats


def fci_energies(h, g, ecore, L, na, nb, k=4):
    sa, sb = _strings(L, na), _strings(L, nb)
    Ea, Eb = _exc_tables(L, sa), _exc_tables(L, sb)
    Ia, Ib = sp.identity(len(sa), format="csr"), sp.identity(len(sb), format="csr")
    E = {}
    for p in range(L):
        for q in range(L):
            E[(p, q)] = (sp.kron(Ea[(p, q)], Ib) + sp.kron(Ia, Eb[(p, q)])).tocsr()
    n = len(sa) * len(sb)
    H = sp.csr_matrix((n, n))
    for p in range(L):
        for q in range(L):
            coef = h[p, q] - 0.5 * sum(g[p, r, r, q] for r in range(L))
            if coef != 0:
                H = H + coef * E[(p, q)]
    for p in range(L):
        for q in range(L):
            Mpq = sp.csr_matrix((n, n))
            for r in range(L):
                for s in range(L):
                    if g[p, q, r, s] != 0:
                        Mpq = Mpq + g[p, q, r, s] * E[(r, s)]
            H = H + 0.5 * (E[(p, q)] @ Mpq)
    H = 0.5 * (H + H.T)
    k = min(k, n - 1)
    if n <= 600:
        w = np.linalg.eigvalsh(H.toarray())[:k]
    else:
        w = np.sort(sla.eigsh(H, k=k, which="SA", tol=1e-12)[0])
    return w + ecore


def fci_ground_rdms(h, g, ecore, L, na, nb):
    """Ground state of the (na, nb) sector: energy, spin-summed one-particle RDM
    gamma_pq = <E_pq> and two-particle RDM Gamma_pqrs = <E_pq E_rs> - delta_qr <E_ps>."""
    sa, sb = _strings(L, na), _strings(L, nb)
    Ea, Eb = _exc_tables(L, sa), _exc_tables(L, sb)
    Ia, Ib = sp.identity(len(sa), format="csr"), sp.identity(len(sb), format="csr")
    E = {}
    for p in range(L):
        for q in range(L):
            E[(p, q)] = (sp.kron(Ea[(p, q)], Ib) + sp.kron(Ia, Eb[(p, q)])).tocsr()
    n = len(sa) * len(sb)
    H = sp.csr_matrix((n, n))
    for p in range(L):
        for q in range(L):
            coef = h[p, q] - 0.5 * sum(g[p, r, r, q] for r in range(L))
            if coef != 0:
                H = H + coef * E[(p, q)]
    for p in range(L):
        for q in range(L):
            Mpq = sp.csr_matrix((n, n))
            for r in range(L):
                for s in range(L):
                    if g[p, q, r, s] != 0:
                        Mpq = Mpq + g[p, q, r, s] * E[(r, s)]
            H = H + 0.5 * (E[(p, q)] @ Mpq)
    H = 0.5 * (H + H.T)
    w, v = np.linalg.eigh(H.toarray())
    psi = v[:, 0]
    Epsi = {k: m @ psi for k, m in E.items()}
    gamma = np.array([[psi @ Epsi[(p, q)] for q in range(L)] for p in range(L)])
    Gamma = np.zeros((L, L, L, L))
    for p in range(L):
        for q in range(L):
            for r in range(L):
                for s in range(L):
                    Gamma[p, q, r, s] = Epsi[(q, p)] @ Epsi[(r, s)] - (gamma[p, s] if q == r else 0.0)
    return w[0] + ecore, gamma, Gamma
